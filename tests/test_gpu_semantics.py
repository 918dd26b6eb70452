"""GPU parity of the reference's failure paths and observable output, through
the C ABI, against the CPU oracle.

- PCG return codes (krylov.c:441-466, 517-629, 714-717): ERROR_SOLVER_MAXIT
  (-48), SOLSTAG (-43, check I), TOLSMALL (-44, check III) and STAG (-42,
  check II) on inputs that reach each of them by a wide margin, so that the
  GPU's different reduction rounding cannot change the outcome;
- the initial-exit rule (zero right-hand side, a guess that already solves);
- stdout at print levels 0-3 (itsolver.c:135-211, 846-931; message.c:28-136;
  amg_setup_ua.c:1426-1431), numbers compared to the precision they are
  printed with (timings excepted);
- the fatal aggregation paths exit(16) / exit(17) (amg_setup_ua.c:820-823,
  1029-1033) returned as -1016 / -1017 with the reference's stderr message;
- one-entry rows under HEC, HEM and VMB;
- HEM's one-launch GPU pipeline handing its chunks to the per-round kernel.
"""
import ctypes as C
import re

import numpy as np
import pytest

import oracle
from hazmath_amd import capi, problems
from test_gpu_parity import _hierarchies

pytestmark = pytest.mark.gpu


def _neumann2d(m):
    """5-point graph Laplacian of an m x m grid: every row sums to exactly 0,
    so A (c 1) = 0 exactly for any representable c (integer entries)."""
    import scipy.sparse as sp
    rows, cols, vals = [], [], []
    for j in range(m):
        for i in range(m):
            k = j * m + i
            nb = [q for q, ok in ((k - 1, i > 0), (k + 1, i < m - 1), (k - m, j > 0), (k + m, j < m - 1)) if ok]
            rows.append(k), cols.append(k), vals.append(float(len(nb)))
            for q in nb:
                rows.append(k), cols.append(q), vals.append(-1.0)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m * m, m * m))
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


def _run(which, lib, orc, libc, A, b, x0, **kw):
    x = capi.HostVec(np.array(x0, dtype=np.float64))
    bb = capi.HostVec(np.array(b, dtype=np.float64))
    libc.srand(1)
    if which == "gpu":
        it, amg = capi.solver_params(lib, **kw)
        r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg))
        info = capi.hz_solve_info()
        lib.hz_get_last_info(C.byref(info))
        cond = info.condest
    else:
        it, amg = oracle.params(**kw)
        tr = capi.orc_trace()
        r = orc.orc_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg), C.byref(tr))
        cond = tr.condest
    return int(r), x.arr.copy(), cond


def _code_case(case):
    """(A, b, x0, solver settings, expected return code)."""
    if case == "stag":
        # singular A with A (1e22 * 1) = 0 exactly: r0 = b, and every PCG
        # update (far below the 2^21 spacing of doubles near 1e22) is lost,
        # so check II restarts with the true residual unchanged until MAX_STAG.
        # b has mean zero, so every coarse system stays consistent and the
        # coarse CG returns corrections of the size of b
        ia, ja, va = _neumann2d(40)
        n = ia.size - 1
        b = problems.uniform_rhs(n, seed=5)
        b -= b.mean()
        return ia, ja, va, b, np.full(n, 1e22), dict(tol=1e-6, maxit=100, coarse_solver=0), -42
    ia, ja, va = problems.poisson2d_p1(64)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=7)
    if case == "maxit":
        return ia, ja, va, b, np.zeros(n), dict(tol=1e-10, maxit=5), -48
    if case == "solstag":
        # |b| ~ 1e-25 passes the initial exit (tol 1e-30): the first update has ||u||_inf < 1e-20 (check I)
        return ia, ja, va, b * 1e-25, np.zeros(n), dict(tol=1e-30, maxit=100), -43
    if case == "tolsmall":
        # the recursive residual falls below 1e-17 while the true one stays ~1e-14: check III, 20 times
        return ia, ja, va, b, np.zeros(n), dict(tol=1e-17, maxit=1000, aggregation=capi.HEM,
                                                smoother=capi.SMOOTHER_L1DIAG), -44
    raise ValueError(case)


@pytest.mark.parametrize("case", ["maxit", "solstag", "tolsmall", "stag"])
def test_pcg_return_codes_match_oracle(lib, orc, libc, case):
    ia, ja, va, b, x0, kw, expected = _code_case(case)
    A = capi.HostCSR(ia, ja, va)
    rg, xg, _ = _run("gpu", lib, orc, libc, A, b, x0, **kw)
    ro, xo, _ = _run("oracle", lib, orc, libc, A, b, x0, **kw)
    assert ro == expected, f"oracle returned {ro}"
    assert rg == expected, f"GPU returned {rg}"
    if case == "stag":
        # no update ever reaches u
        assert np.array_equal(xg, x0) and np.array_equal(xo, x0)


def test_initial_exit_rule(lib, orc, libc):
    """krylov.c:441-466: relres < tol or ||r0|| < 1e-3 tol returns 0 iterations
    with x untouched: a zero right-hand side, and a guess that already solves."""
    ia, ja, va = problems.poisson2d_p1(48)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=2)
    for which in ("gpu", "oracle"):
        r, x, _ = _run(which, lib, orc, libc, A, np.zeros(n), np.zeros(n), tol=1e-6, maxit=100)
        assert r == 0 and np.all(x == 0.0), which
    import scipy.sparse as sp
    _, xs, _ = _run("oracle", lib, orc, libc, A, b, np.zeros(n), tol=1e-14, maxit=300)
    r0 = np.linalg.norm(b - sp.csr_matrix((va, ja, ia), shape=(n, n)) @ xs)
    assert r0 < 1e-9   # below 1e-3 * tol for tol = 1e-6
    for which in ("gpu", "oracle"):
        r, x, _ = _run(which, lib, orc, libc, A, b, xs, tol=1e-6, maxit=100)
        assert r == 0 and np.array_equal(x, xs), which


@pytest.mark.parametrize("stop_type", [capi.STOP_REL_RES, capi.STOP_MOD_REL_RES])
def test_nonzero_initial_guess_parity(lib, orc, libc, stop_type):
    ia, ja, va = problems.poisson3d_kuhn(20)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=4)
    x0 = problems.uniform_rhs(n, seed=9) * 3.0
    kw = dict(tol=1e-10, maxit=300, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, stop_type=stop_type)
    rg, xg, _ = _run("gpu", lib, orc, libc, A, b, x0, **kw)
    ro, xo, _ = _run("oracle", lib, orc, libc, A, b, x0, **kw)
    assert ro > 0 and abs(rg - ro) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?")


def _compare_stdout(gpu_out, orc_out):
    """Same lines, same text around the numbers; integers equal, printed reals
    within the relative precision of their format (7 significant digits for
    %e, four decimals for the convergence factors and complexities), timing
    values ("... costs X seconds.") skipped."""
    gl = [ln for ln in gpu_out.splitlines() if not ln.startswith("[hz]")]
    ol = orc_out.splitlines()
    assert len(gl) == len(ol), "\n".join(["GPU:"] + gl + ["oracle:"] + ol)
    for g, o in zip(gl, ol):
        assert _NUM.sub("#", g) == _NUM.sub("#", o), (g, o)
        if " costs " in g:
            continue
        for a, c in zip(_NUM.findall(g), _NUM.findall(o)):
            if re.fullmatch(r"[-+]?\d+", a):
                assert a == c, (g, o)
            else:
                fa, fc = float(a), float(c)
                assert abs(fa - fc) <= 2e-5 * max(abs(fc), 1e-300) + 2e-4 * ("e" not in a.lower()), (g, o)


@pytest.mark.parametrize("prt", [0, 1, 2, 3])
def test_stdout_matches_oracle(lib, orc, libc, capfd, prt):
    """print_level 0-3 of itparam and amgparam: the setup table and time, the
    iteration table, ITS_FINAL, the condition number line and the timing
    lines of solver_dcsr_linear_itsolver / linear_solver_dcsr_krylov_amg."""
    ia, ja, va = problems.poisson2d_p1(32)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=7)
    kw = dict(tol=1e-8, maxit=100, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, print_level=prt)
    capfd.readouterr()
    rg, _, cg = _run("gpu", lib, orc, libc, A, b, np.zeros(n), **kw)
    gpu_out = capfd.readouterr().out
    ro, _, co = _run("oracle", lib, orc, libc, A, b, np.zeros(n), **kw)
    orc_out = capfd.readouterr().out
    assert rg == ro
    assert abs(cg - co) <= 1e-8 * co
    assert "Estimated condition number" in gpu_out
    _compare_stdout(gpu_out, orc_out)


def test_condition_estimate_matches_oracle(lib, orc, libc):
    """The Lanczos condition estimate (krylov.c:663-705) on a 3-D problem."""
    ia, ja, va = problems.poisson3d_kuhn(16)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=3)
    kw = dict(tol=1e-10, maxit=300, aggregation=capi.HEC)
    rg, _, cg = _run("gpu", lib, orc, libc, A, b, np.zeros(n), **kw)
    ro, _, co = _run("oracle", lib, orc, libc, A, b, np.zeros(n), **kw)
    assert abs(rg - ro) <= 1
    assert co > 1.0 and abs(cg - co) <= 1e-6 * co


def _isolated_duplicate_diagonal(m, k):
    """2-D P1 matrix whose row k keeps no coupling and stores its diagonal as
    two entries: the strength row has two entries and no neighbour."""
    import scipy.sparse as sp
    ia, ja, va = problems.poisson2d_p1(m)
    n = ia.size - 1
    A = sp.csr_matrix((va, ja, ia), shape=(n, n)).tolil()
    for j in list(A.rows[k]):
        A[k, j] = 0.0
        A[j, k] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    rows = []
    for i in range(n):
        cols = list(A.indices[A.indptr[i]:A.indptr[i + 1]])
        vals = list(A.data[A.indptr[i]:A.indptr[i + 1]])
        if i == k:
            cols, vals = [k, k], [2.0, 2.0]
        rows.append((cols, vals))
    ia2 = np.zeros(n + 1, np.int32)
    ia2[1:] = np.cumsum([len(c) for c, _ in rows])
    ja2 = np.concatenate([np.array(c, np.int32) for c, _ in rows])
    va2 = np.concatenate([np.array(v, np.float64) for _, v in rows])
    return ia2, ja2, va2


@pytest.mark.parametrize("agg,code,message", [
    (capi.HEC, -1016, "%% ERROR Stop: negative index encountered in aggregation_hec()"),
    (capi.HEM, -1017, "%%ERROR in aggregation_hem: isolated point (=100) on the second matching pass"),
])
def test_fatal_aggregation_paths(lib, orc, libc, capfd, agg, code, message):
    """A strength row of two entries without a neighbour: HEC finds no
    strongest neighbour (exit(16)), HEM's second pass finds no aggregate for
    it (exit(17)); with return_on_fatal both return the exit code as -1000 -
    status and print the reference's message to stderr."""
    ia, ja, va = _isolated_duplicate_diagonal(16, 100)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=1)
    capfd.readouterr()
    rg, _, _ = _run("gpu", lib, orc, libc, A, b, np.zeros(n), tol=1e-8, maxit=50, aggregation=agg)
    gerr = capfd.readouterr().err
    ro, _, _ = _run("oracle", lib, orc, libc, A, b, np.zeros(n), tol=1e-8, maxit=50, aggregation=agg)
    oerr = capfd.readouterr().err
    assert ro == code and rg == code
    assert message in oerr and message in gerr


def _with_unit_rows(m, every):
    """2-D P1 matrix in which every `every`-th row is a lone diagonal 1.0 (its
    couplings removed symmetrically): one-entry strength rows."""
    import scipy.sparse as sp
    ia, ja, va = problems.poisson2d_p1(m)
    n = ia.size - 1
    A = sp.csr_matrix((va, ja, ia), shape=(n, n)).tolil()
    for k in range(0, n, every):
        for j in list(A.rows[k]):
            A[k, j] = 0.0
            A[j, k] = 0.0
        A[k, k] = 1.0
    A = A.tocsr()
    A.eliminate_zeros()
    A.sort_indices()
    return A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)


@pytest.mark.parametrize("agg", [capi.HEC, capi.HEM, capi.VMB])
def test_one_entry_rows(lib, orc, libc, agg):
    """Rows whose strength row is the diagonal alone are UNPT (HEC
    amg_setup_ua.c:792-793, HEM 950-956, VMB): hierarchies bit-exact and PCG
    within one iteration of the oracle."""
    ia, ja, va = _with_unit_rows(40, 37)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    (g, thg, rg), (o, tho, ro) = _hierarchies(lib, orc, libc, A, agg, 0.04)
    assert len(g) == len(o)
    for gl, ol in zip(g, o):
        if "agg" in gl:
            assert np.array_equal(gl["agg"], ol["agg"])
        for key in ("ia", "ja", "val"):
            assert np.array_equal(gl[key], ol[key]), key
    assert thg == tho and rg == ro
    assert np.any(g[0]["agg"] == -1)
    b = problems.uniform_rhs(n, seed=8)
    kw = dict(tol=1e-10, maxit=300, aggregation=agg, smoother=capi.SMOOTHER_L1DIAG)
    ig, xg, _ = _run("gpu", lib, orc, libc, A, b, np.zeros(n), **kw)
    io, xo, _ = _run("oracle", lib, orc, libc, A, b, np.zeros(n), **kw)
    assert io > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("cap,host_max", [(1, -1), (-1, -1), (-1, 2000)])
def test_hem_pipeline_give_up_path_bitexact(lib, orc, libc, cap, host_max):
    """hem_pipeline_max_sweeps = +-1: every chunk of the one-launch pipeline
    that is not decided within its first 16 sweeps raises the give-up flag
    (the kernel's own bounded wait) and the workgroups stop.  With cap 1 the
    per-round kernel finishes pass 1; with cap -1 the level goes to the host
    sweep over preference lists -- on every level (hem_host_max_rows -1) or
    on the levels the size rule gives the pipeline (hem_host_max_rows 2000).
    The aggregates stay the reference's."""
    ia, ja, va = problems.poisson3d_kuhn(24)
    A = capi.HostCSR(ia, ja, va)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = (opts.hem_pipeline_max_sweeps, opts.hem_host_max_rows)
    opts.hem_pipeline_max_sweeps = cap
    opts.hem_host_max_rows = host_max
    lib.hz_set_options(C.byref(opts))
    try:
        (g, thg, _), (o, tho, _) = _hierarchies(lib, orc, libc, A, capi.HEM, 0.04)
        info = capi.hz_solve_info()
        lib.hz_get_last_info(C.byref(info))
    finally:
        lib.hz_get_options(C.byref(opts))
        opts.hem_pipeline_max_sweeps, opts.hem_host_max_rows = old
        lib.hz_set_options(C.byref(opts))
    print(f"[semantics] cap {cap}, hem_host_max_rows {host_max}: {info.hem_pipeline_giveups} levels gave up",
          flush=True)
    assert info.hem_pipeline_giveups >= 1
    assert len(g) == len(o)
    for gl, ol in zip(g, o):
        if "agg" in gl:
            assert np.array_equal(gl["agg"], ol["agg"])
        for key in ("ia", "ja", "val"):
            assert np.array_equal(gl[key], ol[key]), key
    assert thg == tho


def test_dense_coarsest_size_limit(lib, libc, capfd):
    """A coarsest level beyond the dense solve's 32,768 rows (here max_levels
    2 on a 3-D problem: level 1 has about 55,000 rows) is refused with
    ERROR_ALLOC_MEM and a message, instead of a device allocation failure."""
    ia, ja, va = problems.poisson3d_kuhn(48)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = capi.HostVec(problems.uniform_rhs(n, seed=3))
    x = capi.HostVec(np.zeros(n))
    it, amg = capi.solver_params(lib, tol=1e-8, maxit=50, aggregation=capi.HEM, max_levels=2)
    libc.srand(1)
    r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), b.ptr(), x.ptr(), C.byref(it), C.byref(amg))
    err = capfd.readouterr().err
    print(f"[semantics] max_levels 2, dense coarsest: {r}; {err.strip()}", flush=True)
    assert r == capi.ERROR_ALLOC_MEM
    assert "the dense coarsest solve takes at most 32768" in err
    assert not np.any(x.arr)


def _convdiff(m, bx, by):
    """Upwind convection-diffusion on an m x m grid: a nonsymmetric M-matrix."""
    import scipy.sparse as sp
    n = m * m
    rows, cols, vals = [], [], []
    for j in range(m):
        for i in range(m):
            k = j * m + i
            for q, v, ok in ((k - m, -1.0 - by, j > 0), (k - 1, -1.0 - bx, i > 0), (k, 4.0 + bx + by, True),
                             (k + 1, -1.0, i < m - 1), (k + m, -1.0, j < m - 1)):
                if ok:
                    rows.append(k), cols.append(q), vals.append(v)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


def _coarse_gpu(lib, A, b):
    Ah = capi.HostCSR(A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64))
    x = np.zeros(A.shape[0])
    xcg = np.zeros(A.shape[0])
    st_cg = capi.INT()
    st = lib.hz_coarse_itsolver(Ah.ptr(), capi.dptr(b), capi.dptr(x), 1e-8, C.byref(st_cg), capi.dptr(xcg))
    return st, st_cg.value, x, xcg, Ah


def test_coarse_itsolver_cg_on_spd_matches_oracle(lib, orc):
    """coarse_itsolver on an SPD level (mgcycle.c:34-53): CG alone, as the
    oracle's dcsr_pcg (iterations within one, solution within 1e-8)."""
    import scipy.sparse as sp
    ia, ja, va = problems.poisson2d_p1(24)
    A = sp.csr_matrix((va, ja, ia))
    b = problems.uniform_rhs(A.shape[0], seed=2)
    st, st_cg, x, xcg, Ah = _coarse_gpu(lib, A, b)
    xo = np.zeros(A.shape[0])
    so = orc.orc_coarse_itsolver(Ah.ptr(), capi.dptr(b), capi.dptr(xo), 1e-8)
    print(f"[semantics] coarse CG on SPD: GPU {st} (CG {st_cg}), oracle {so}", flush=True)
    assert st == st_cg > 0 and abs(st - so) <= 1
    assert np.array_equal(x, xcg)
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_coarse_itsolver_gmres_fallback_vs_oracle(lib, orc):
    """On a nonsymmetric level the coarse CG fails (as the oracle's does) and
    GMRES(30) with dcsr_pvgmres's variable restart continues from CG's last
    iterate: the GPU result reaches 1e-8 relative to that iterate's residual,
    and the oracle's GMRES from the same iterate takes the same number of
    iterations (within two) to a solution within 1e-6."""
    A = _convdiff(10, 6.0, 4.0)
    n = A.shape[0]
    b = problems.uniform_rhs(n, seed=3)
    st, st_cg, x, xcg, Ah = _coarse_gpu(lib, A, b)
    xo_cg = np.zeros(n)
    so_cg = orc.orc_cg(Ah.ptr(), capi.dptr(b), capi.dptr(xo_cg), 1e-8, 20 * n)
    xo = xcg.copy()
    so = orc.orc_pvgmres(Ah.ptr(), capi.dptr(b), capi.dptr(xo), 1e-8, 20 * n, 30)
    r_cg = np.linalg.norm(b - A @ xcg)
    print(f"[semantics] coarse solve, nonsymmetric: GPU CG {st_cg} then GMRES {st}; oracle CG {so_cg}, "
          f"oracle GMRES from the GPU's CG iterate {so}; CG residual {r_cg:.3e}", flush=True)
    assert st_cg < 0 and so_cg < 0
    assert st > 0 and abs(st - so) <= 2
    assert np.linalg.norm(b - A @ x) < 1e-8 * r_cg
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_release_after_call_frees_the_pool(lib, libc):
    """hz_options.release_after_call returns the library's device pool to HIP
    when each call returns, and hz_release_device_memory does it on demand
    (hz_device_pool_bytes: the pool's chunks); results are the same bits as
    with the pool kept."""
    ia, ja, va = problems.poisson3d_kuhn(40)
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(ia.size - 1, seed=4)

    def solve():
        x = capi.HostVec(np.zeros(ia.size - 1))
        it, amg = capi.solver_params(lib, tol=1e-10, maxit=200, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG)
        libc.srand(1)
        r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg))
        assert r > 0
        return x.arr.copy()

    x_kept = solve()
    kept = lib.hz_device_pool_bytes()
    lib.hz_release_device_memory()
    released = lib.hz_device_pool_bytes()
    x_again = solve()
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = opts.release_after_call
    opts.release_after_call = 1
    lib.hz_set_options(C.byref(opts))
    try:
        x_rel = solve()
        after = lib.hz_device_pool_bytes()
    finally:
        opts.release_after_call = old
        lib.hz_set_options(C.byref(opts))
    print(f"[semantics] pool bytes: kept {kept >> 20} MiB, after release {released}, after a releasing call {after}",
          flush=True)
    assert kept >= 256 << 20 and released == 0 and after == 0
    assert np.array_equal(x_kept, x_again) and np.array_equal(x_kept, x_rel)
