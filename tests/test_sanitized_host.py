"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU):
tests/sanitize/Makefile builds the oracle, input.cpp and aggregate_sweeps.cpp
with both sanitizers; tests/sanitize/drive.py runs them with the ASan runtime
preloaded.  Any out-of-bounds access, use after free or undefined behaviour
aborts the driver before it prints its last line.  The driver also checks the
library's host VMB / HEM sweeps against the oracle's aggregates."""
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _runtime():
    out = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True, check=True)
    return out.stdout.strip()


def test_host_code_under_address_and_ub_sanitizers():
    subprocess.run(["make", "-C", str(ROOT / "tests" / "sanitize")], check=True, capture_output=True)
    env = dict(os.environ, LD_PRELOAD=_runtime(), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "sanitize" / "drive.py")], env=env,
                       capture_output=True, text=True, timeout=600)
    tail = (r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, tail
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, tail
    assert "SANITIZED OK" in r.stdout, tail
