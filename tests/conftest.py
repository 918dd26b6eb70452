import ctypes as C
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def lib():
    from hazmath_amd import capi
    L = capi.load_library()
    if L.hz_device_count() < 1:
        pytest.fail("GPU test started without a visible HIP device")
    opts = capi.hz_options()
    L.hz_get_options(C.byref(opts))
    opts.return_on_fatal = 1
    opts.jacobi_weight = 1.0
    L.hz_set_options(C.byref(opts))
    return L


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle.load()


@pytest.fixture(scope="session")
def libc():
    return C.CDLL(None)


@pytest.fixture(scope="session")
def solvers_fixture():
    d = np.load(ROOT / "tests" / "golden" / "solvers_A_b.npz")
    return d["ia"], d["ja"], d["val"], d["b"]
