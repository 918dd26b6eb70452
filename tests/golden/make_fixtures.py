"""Writes the committed fixtures for the UA-AMG-PCG parity tests.

Run in the build container (it reads the reference tree):
    python tests/golden/make_fixtures.py

Inputs: examples/solvers/A.dat and b.dat of the reference (COO "m n nnz" header
then "i j a_ij" lines, 0-based; b: "n" then values).  The COO triplets are turned
into CSR row-stably, like dcoo_2_dcsr_p (src/utilities/format.c:785-827), and
stored as solvers_A_b.npz (data only: ia, ja, val, b).

probe_reference.json holds the level sizes and PCG iteration counts that the
compiled reference produced on that fixture (examples/solvers harness, PCG,
STOP_REL_RES, x0 = 0, UA V(1,1), coarse_dof=100, coarse_solver=32, theta=0,
max_aggregation=20, 20 levels), as recorded when the reference was built and
run in this container; it is written here verbatim.
"""
import json
from pathlib import Path

import numpy as np

REF = Path("/root/reference/examples/solvers")
OUT = Path(__file__).resolve().parent


def read_coo(path):
    with open(path) as f:
        m, n, nnz = (int(t) for t in f.readline().split())
        data = np.loadtxt(f, dtype=np.float64, max_rows=nnz)
    rows = data[:, 0].astype(np.int64)
    cols = data[:, 1].astype(np.int32)
    vals = data[:, 2]
    order = np.argsort(rows, kind="stable")   # row-stable, like dcoo_2_dcsr_p
    ia = np.zeros(m + 1, np.int64)
    np.add.at(ia, rows + 1, 1)
    ia = np.cumsum(ia).astype(np.int32)
    return ia, cols[order], vals[order], n


def read_vec(path):
    with open(path) as f:
        n = int(f.readline())
        v = np.loadtxt(f, dtype=np.float64, max_rows=n)
    return v


PROBE = {
    "source": "compiled reference, examples/solvers/solvers.c on A.dat/b.dat",
    "settings": {"itsolver": "CG", "precond": "AMG", "stop_type": 1, "x0": 0.0,
                 "AMG_type": "UA", "cycle": "V", "presmooth": 1, "postsmooth": 1,
                 "AMG_maxit": 1, "coarse_dof": 100, "coarse_solver": 32,
                 "strong_coupled": 0.0, "max_aggregation": 20, "AMG_levels": 20,
                 "linear_maxit_at_1e-8": 100},
    "initial_residual_norm": 1.682070e+02,
    "b_norm2": 168.20703796526038,
    "levels": {
        "VMB": [16384, 2752, 330, 46],
        "HEC": [16384, 8193, 3110, 1071, 349, 101, 30],
        "HEM": [16384, 8192, 4096, 2048, 1024, 512, 256, 128, 64],
    },
    "operator_complexity": {"VMB": 1.261, "HEC": 1.978, "HEM": 1.988},
    "iterations_tol_1e-8": {"VMB+JACOBI": 46, "HEC+JACOBI": 41, "HEM+JACOBI": ">100",
                            "HEM+L1DIAG": 28, "HEC+L1DIAG": 48, "VMB+GS": 38,
                            "VMB+JACOBI+coarseCG": 46},
    "iterations_tol_1e-10": {"VMB+JACOBI": 60, "VMB+L1DIAG": 67, "HEC+JACOBI": 52,
                             "HEC+L1DIAG": 59, "HEM+JACOBI": 215, "HEM+L1DIAG": 37},
    "hec_rng_state": "process-initial glibc rand() state (equivalent to srand(1))",
}


def main():
    ia, ja, val, ncol = read_coo(REF / "A.dat")
    b = read_vec(REF / "b.dat")
    np.savez_compressed(OUT / "solvers_A_b.npz", ia=ia, ja=ja, val=val, b=b)
    (OUT / "probe_reference.json").write_text(json.dumps(PROBE, indent=1) + "\n")
    print("n =", ia.size - 1, "nnz =", ia[-1], "|b| =", float(np.linalg.norm(b)))


if __name__ == "__main__":
    main()
