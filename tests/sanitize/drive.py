"""Runs the sanitizer build (TEST INFRASTRUCTURE; started by
tests/test_sanitized_host.py with the ASan runtime preloaded): the oracle's
solver paths, failure paths, the library's host VMB / HEM sweeps against the
oracle's aggregates, input.dat parsing plus the raw-pointer wrapper, and the
batched autotuner's cache file.
Prints "SANITIZED OK" at the end; any sanitizer report aborts before that."""
import ctypes as C
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

import oracle  # noqa: E402
from hazmath_amd import capi, problems  # noqa: E402

LIB = oracle.load(ROOT / "tests" / "sanitize" / "_build" / "libsan_host.so")
P = C.POINTER
LIB.san_vmb.argtypes = [C.c_int, P(C.c_int), P(C.c_int), C.c_int, P(C.c_int), P(C.c_int)]
LIB.san_hem.argtypes = [C.c_int, P(C.c_int), P(C.c_int), P(C.c_double), P(C.c_int), P(C.c_int)]
LIB.param_input.argtypes = [C.c_char_p, P(capi.input_param)]
LIB.param_input_init.argtypes = [P(capi.input_param)]
LIB.param_linear_solver_set.argtypes = [P(capi.linear_itsolver_param), P(capi.input_param)]
LIB.param_amg_set.argtypes = [P(capi.AMG_param), P(capi.input_param)]
LIB.python_wrapper_krylov_amg.argtypes = [P(capi.INT), P(capi.INT), P(capi.INT), P(capi.INT), P(capi.REAL),
                                          P(capi.REAL), P(capi.REAL), P(capi.REAL), P(capi.INT), P(capi.INT)]
LIB.san_tune_reset.argtypes = [C.c_char_p]
LIB.san_tune_lookup.argtypes = [C.c_char_p]
LIB.san_tune_store.argtypes = [C.c_char_p, C.c_int]
LIBC = C.CDLL(None)
LIBC.free.argtypes = [C.c_void_p]


def solve(A, b, x0=None, **kw):
    it, amg = oracle.params(**{k: v for k, v in kw.items() if k not in ("amg_type", "steps")})
    if "amg_type" in kw:
        amg.AMG_type = kw["amg_type"]
    if "steps" in kw:
        amg.presmooth_iter = amg.postsmooth_iter = kw["steps"]
    x = capi.HostVec(np.zeros(b.size) if x0 is None else x0.copy())
    LIBC.srand(1)
    return LIB.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                              C.byref(capi.orc_trace()))


def main():
    d = np.load(ROOT / "tests" / "golden" / "solvers_A_b.npz")
    A = capi.HostCSR(d["ia"], d["ja"], d["val"])
    b = d["b"]
    # solver paths
    for agg in (capi.HEC, capi.HEM, capi.VMB):
        for sm in (1, 10):
            assert solve(A, b, tol=1e-8, maxit=300, aggregation=agg, smoother=sm) > 0
    for kw in (dict(cycle=2), dict(coarse_solver=0), dict(coarse_scaling=1), dict(stop_type=2),
               dict(stop_type=3), dict(amg_maxit=2), dict(amg_type=2, aggregation=capi.VMB, smoother=91, steps=2),
               dict(amg_type=2, aggregation=capi.HEC)):
        assert solve(A, b, tol=1e-8, maxit=300, **kw) > 0, kw
    for sm in (2, 3, 5, 6, 7, 8):
        assert solve(A, b, tol=1e-8, maxit=300, aggregation=capi.HEC, smoother=sm) > 0
    it, amg = oracle.params(tol=1e-8, maxit=50, aggregation=capi.HEM, smoother=10)
    amg.tol, amg.maxit = 1e-8, 50
    x = capi.HostVec(np.zeros(b.size))
    assert LIB.orc_linear_solver_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(amg)) > 0
    _, amg = oracle.params(aggregation=capi.HEC)
    h = LIB.orc_precond_create(A.ptr(), C.byref(amg))
    z = np.zeros(b.size)
    LIB.orc_precond_apply(h, capi.dptr(b), capi.dptr(z))
    LIB.orc_precond_free(h)
    # failure paths
    from test_gpu_semantics import _code_case, _isolated_duplicate_diagonal
    for case, code in (("maxit", -48), ("solstag", -43), ("tolsmall", -44), ("stag", -42)):
        ia, ja, va, bb, x0, kw, _ = _code_case(case)
        assert solve(capi.HostCSR(ia, ja, va), bb, x0, **kw) == code, case
    ia, ja, va = _isolated_duplicate_diagonal(16, 100)
    for agg, code in ((capi.HEC, -1016), (capi.HEM, -1017)):
        assert solve(capi.HostCSR(ia, ja, va), np.ones(ia.size - 1), tol=1e-8, maxit=50, aggregation=agg) == code
    # host sweeps against the oracle's aggregates
    mats = [(d["ia"], d["ja"], d["val"]), problems.poisson3d_kuhn(12), problems.aniso2d_p1(40, 0.7, 1e-2)]
    for ia, ja, va in mats:
        M = capi.HostCSR(ia, ja, va)
        n = M.n
        for theta in (0.0, 0.04, 0.25):
            N = capi.dCSRmat()
            LIB.orc_strength(M.ptr(), theta, C.byref(N))
            nnz = N.IA[n]
            for typ in (capi.VMB, capi.HEM):
                ref = np.zeros(n, np.int32)
                nref = capi.INT()
                sref = LIB.orc_aggregate(M.ptr(), typ, theta, 20, capi.dptr(ref, capi.INT), C.byref(nref))
                got = np.zeros(n, np.int32)
                ng = C.c_int()
                if typ == capi.VMB:
                    st = LIB.san_vmb(n, N.IA, N.JA, 20, capi.dptr(got, C.c_int), C.byref(ng))
                else:
                    st = LIB.san_hem(n, N.IA, N.JA, N.val, capi.dptr(got, C.c_int), C.byref(ng))
                assert st == sref and ng.value == nref.value, (typ, theta, n, nnz, st, sref)
                if sref == 0:
                    assert np.array_equal(got, ref), (typ, theta, n, nnz)
            for ptr in (N.IA, N.JA, N.val):
                LIBC.free(C.cast(ptr, C.c_void_p))
    # input.dat and the raw-pointer wrapper
    with tempfile.TemporaryDirectory() as td:
        Path(td, "input.dat").write_text(
            "%% comment\nprint_level = 0\nlinear_itsolver_type = 1\nlinear_stop_type = 1\nAMG_type = UA\n"
            "AMG_cycle_type = W\nAMG_levels = 20\nAMG_smoother = L1DIAG\nAMG_coarse_dof = 100\n"
            "AMG_coarse_solver = 32\nAMG_aggregation_type = 5\nAMG_strong_coupled = 0.0\n"
            "AMG_smooth_filter = OFF\nnot_a_keyword = 3\n")
        p = capi.input_param()
        LIB.param_input(str(Path(td, "input.dat")).encode(), C.byref(p))
        it, amg = capi.linear_itsolver_param(), capi.AMG_param()
        LIB.param_linear_solver_set(C.byref(it), C.byref(p))
        LIB.param_amg_set(C.byref(amg), C.byref(p))
        assert p.AMG_cycle_type == 2 and amg.smoother == 10
        old = os.getcwd()
        os.chdir(td)
        try:
            ia, ja, va = problems.poisson2d_p1(24)
            n = ia.size - 1
            u = np.zeros(n)
            rhs = problems.uniform_rhs(n, seed=2)
            LIB.python_wrapper_krylov_amg(C.byref(capi.INT(n)), C.byref(capi.INT(int(ia[-1]))),
                                          capi.dptr(ia, capi.INT), capi.dptr(ja, capi.INT), capi.dptr(va),
                                          capi.dptr(rhs), capi.dptr(u), C.byref(capi.REAL(1e-10)),
                                          C.byref(capi.INT(300)), C.byref(capi.INT(0)))
        finally:
            os.chdir(old)
        import scipy.sparse as sp
        assert np.linalg.norm(rhs - sp.csr_matrix((va, ja, ia)) @ u) < 1e-9 * np.linalg.norm(rhs)
    # the autotuner's cache file: malformed lines skipped, decisions appended
    # (directories created), read back by a fresh reset
    with tempfile.TemporaryDirectory() as td:
        f = Path(td, "seed.txt")
        f.write_text("gfx950-256cu/bsweep/14/60 208\n# comment 1\nbad\ngfx950-256cu/spmm64/14/60 404 x\n"
                     "gfx950-256cu/spmm64/9/56 -3\n" + "k" * 5000 + " 104\ngfx950-256cu/spmm64/14/60 104\n")
        assert LIB.san_tune_reset(str(f).encode()) == 3
        assert LIB.san_tune_lookup(b"gfx950-256cu/bsweep/14/60") == 208
        assert LIB.san_tune_lookup(b"gfx950-256cu/spmm64/14/60") == 104
        assert LIB.san_tune_lookup(b"gfx950-256cu/spmm64/9/56") == -1
        g = Path(td, "a", "b", "tune.txt")
        assert LIB.san_tune_reset(str(g).encode()) == 0
        LIB.san_tune_store(b"dev/bresidual/12/58", 408)
        LIB.san_tune_store(b"dev/bresidual/12/58", 204)
        assert LIB.san_tune_reset(str(g).encode()) == 2
        assert LIB.san_tune_lookup(b"dev/bresidual/12/58") == 204
        assert LIB.san_tune_reset(None) == 0 and LIB.san_tune_lookup(b"dev/bresidual/12/58") == -1
    print("SANITIZED OK", flush=True)


if __name__ == "__main__":
    main()
