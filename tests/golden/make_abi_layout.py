"""Writes abi_layout_reference.json: sizeof and every field offset of the
reference's public structs on this platform, from the reference's own headers
(include/macro.h, sparse.h, vec.h, param.h) compiled with gcc.  Run in a
container that holds the reference tree:

    python tests/golden/make_abi_layout.py /root/reference

The field lists come from the ctypes mirrors in hazmath_amd/capi.py, which in
turn mirror include/hazmath_mi355x.h."""
import json
import subprocess
import sys
import tempfile
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from hazmath_amd import capi  # noqa: E402

STRUCTS = {"dCSRmat": capi.dCSRmat, "dvector": capi.dvector,
           "linear_itsolver_param": capi.linear_itsolver_param, "AMG_param": capi.AMG_param,
           "input_param": capi.input_param}


def layout_source(headers):
    lines = ["#include <stdio.h>", "#include <stddef.h>"] + [f'#include "{h}"' for h in headers]
    lines.append("int main(void){")
    for s, cls in STRUCTS.items():
        lines.append(f'printf("{s} sizeof %zu\\n", sizeof({s}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{s} {f} %zu\\n", offsetof({s}, {f}));')
    lines.append("return 0;}")
    return "\n".join(lines)


def compile_and_run(src, include_dir):
    with tempfile.TemporaryDirectory() as d:
        c = Path(d) / "layout.c"
        c.write_text(src)
        exe = Path(d) / "layout"
        subprocess.run(["gcc", "-w", "-I", str(include_dir), str(c), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    res = {}
    for line in out.splitlines():
        s, f, v = line.split()
        res.setdefault(s, {})[f] = int(v)
    return res


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    src = layout_source(["macro.h", "sparse.h", "vec.h", "param.h"])
    res = compile_and_run(src, ref / "include")
    out = Path(__file__).parent / "abi_layout_reference.json"
    out.write_text(json.dumps(res, indent=1, sort_keys=True))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
