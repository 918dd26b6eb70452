"""Run in a fresh process by test_gpu_parity.py: the very first library call
of the process is a HEC setup right after srand(1).  Prints one line:
aggregates-equal and libc-state-equal flags against the oracle."""
import ctypes as C
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
import oracle  # noqa: E402
from hazmath_amd import capi, problems  # noqa: E402


def main():
    libc = C.CDLL(None)
    lib = capi.load_library()
    orc = oracle.load()
    ia, ja, va = problems.poisson2d_p1(48)
    A = capi.HostCSR(ia, ja, va)
    out = []
    for setup, free, params in ((lib.hz_setup_hierarchy, lib.hz_hierarchy_free,
                                 lambda: capi.solver_params(lib, aggregation=capi.HEC, coarse_dof=50)),
                                (orc.orc_setup_hierarchy, orc.orc_hierarchy_free,
                                 lambda: oracle.params(aggregation=capi.HEC, coarse_dof=50))):
        _, amg = params()
        h = capi.hz_hierarchy()
        libc.srand(1)
        assert setup(A.ptr(), C.byref(amg), C.byref(h)) == 0
        lv = capi.hierarchy_to_numpy(h)
        out.append(([l["agg"] for l in lv[:-1]], libc.rand()))
        free(C.byref(h))
    (ag, rg), (ao, ro) = out
    same = len(ag) == len(ao) and all(np.array_equal(x, y) for x, y in zip(ag, ao))
    print(f"aggregates_equal={int(same)} rand_equal={int(rg == ro)}", flush=True)


if __name__ == "__main__":
    main()
