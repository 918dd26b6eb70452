"""GPU parity of the MI355X library against the CPU oracle, through the C ABI.

Bars: integer/structure outputs bit-exact; per-row kernels (SpMV, residual,
Jacobi/L1 sweeps, coarse products) bit-exact as well (same summation order,
no FMA contraction on either side); PCG iteration counts within +-1 and
solutions within 1e-10 relative of the oracle's.
"""
import ctypes as C

import numpy as np
import pytest

from hazmath_amd import capi, problems

pytestmark = pytest.mark.gpu


def random_csr(n, seed, maxlen=20, long_rows=(), empty_rows=(), shuffle=True, ncol=None):
    rng = np.random.default_rng(seed)
    ncol = ncol or n
    rows = []
    for i in range(n):
        if i in empty_rows:
            rows.append(np.zeros(0, np.int32))
            continue
        L = 150 if i in long_rows else int(rng.integers(1, maxlen))
        cols = rng.choice(ncol, size=min(L, ncol), replace=False).astype(np.int32)
        if i < ncol and i not in cols:
            cols[0] = i
        if not shuffle:
            cols.sort()
        rows.append(cols)
    ia = np.zeros(n + 1, np.int32)
    ia[1:] = np.cumsum([r.size for r in rows])
    ja = np.concatenate(rows).astype(np.int32)
    va = rng.standard_normal(ja.size)
    va[rng.random(ja.size) < 0.05] = 0.0
    return ia, ja, va


def _spd_like(ia, ja, va):
    """Make the diagonal dominant so sweeps are well defined."""
    va = va.copy()
    for i in range(ia.size - 1):
        for k in range(ia[i], ia[i + 1]):
            if ja[k] == i:
                va[k] = 10.0 + abs(va[k])
    return va


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_spmv_bitexact(lib, orc, seed):
    ia, ja, va = random_csr(3000, seed, long_rows={5, 700, 2999}, empty_rows={10, 11})
    A = capi.HostCSR(ia, ja, va)
    x = np.random.default_rng(seed + 10).standard_normal(3000)
    yg = np.zeros(3000)
    yo = np.zeros(3000)
    assert lib.hz_dcsr_mxv(A.ptr(), capi.dptr(x), capi.dptr(yg)) == 0
    orc.orc_dcsr_mxv(A.ptr(), capi.dptr(x), capi.dptr(yo))
    assert np.array_equal(yg, yo)


def test_residual_bitexact(lib, orc):
    ia, ja, va = random_csr(2500, 5, long_rows={3})
    A = capi.HostCSR(ia, ja, va)
    rng = np.random.default_rng(3)
    x, b = rng.standard_normal(2500), rng.standard_normal(2500)
    wg, wo = np.zeros(2500), np.zeros(2500)
    assert lib.hz_dcsr_residual(A.ptr(), capi.dptr(b), capi.dptr(x), capi.dptr(wg)) == 0
    orc.orc_dcsr_residual(A.ptr(), capi.dptr(b), capi.dptr(x), capi.dptr(wo))
    assert np.array_equal(wg, wo)


@pytest.mark.parametrize("smoother,w", [(1, 1.0), (1, 2.0 / 3.0), (10, 1.0)])
def test_smoother_bitexact(lib, orc, smoother, w):
    ia, ja, va = random_csr(2000, 9)
    va = _spd_like(ia, ja, va)
    A = capi.HostCSR(ia, ja, va)
    rng = np.random.default_rng(4)
    x0, b = rng.standard_normal(2000), rng.standard_normal(2000)
    xg, xo = x0.copy(), x0.copy()
    assert lib.hz_smoother_sweep(A.ptr(), capi.dptr(b), capi.dptr(xg), smoother, w) == 0
    orc.orc_smoother_sweep(A.ptr(), capi.dptr(b), capi.dptr(xo), smoother, w)
    assert np.array_equal(xg, xo)


@pytest.mark.parametrize("seed", [0, 1])
def test_rap_agg_bitexact(lib, orc, seed):
    n = 4000
    ia, ja, va = random_csr(n, seed + 20, long_rows={1, 2})
    rng = np.random.default_rng(seed)
    nagg = 900
    agg = rng.integers(-1, nagg, size=n).astype(np.int32)
    agg[:nagg] = np.arange(nagg)               # every aggregate non-empty
    # one very large aggregate exercises the global-memory path
    agg[rng.choice(n, 120, replace=False)] = 7
    A = capi.HostCSR(ia, ja, va)
    Cg, Co = capi.dCSRmat(), capi.dCSRmat()
    assert lib.hz_rap_agg(A.ptr(), capi.dptr(agg, capi.INT), nagg, C.byref(Cg)) == 0
    orc.orc_rap_agg(A.ptr(), capi.dptr(agg, capi.INT), nagg, C.byref(Co))
    g, o = capi.csr_to_numpy(Cg), capi.csr_to_numpy(Co)
    for a, b_ in zip(g, o):
        assert np.array_equal(a, b_)


def _hierarchies(lib, orc, libc, A, agg_type, theta, coarse_solver=32, coarse_dof=100, max_aggregation=None):
    out = []
    for which in ("gpu", "oracle"):
        it, amg = capi.solver_params(lib, aggregation=agg_type, theta=theta,
                                     coarse_solver=coarse_solver, coarse_dof=coarse_dof)
        if max_aggregation is not None:
            amg.max_aggregation = max_aggregation
        h = capi.hz_hierarchy()
        libc.srand(1)
        if which == "gpu":
            st = lib.hz_setup_hierarchy(A.ptr(), C.byref(amg), C.byref(h))
        else:
            st = orc.orc_setup_hierarchy(A.ptr(), C.byref(amg), C.byref(h))
        assert st == 0
        levels = capi.hierarchy_to_numpy(h)
        after = libc.rand()   # libc state after the setup
        (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
        out.append((levels, amg.strong_coupled, after))
    return out


class _HemPath:
    """HEM levels through the default choice by level size ("host": these
    problems are small enough for the full host sweep), or every level through
    one form (hz_options.hem_host_max_rows): the GPU pipeline (-1), the full
    host sweep over the whole strength graph (-2), or the host sweep over
    GPU-sorted preference lists with pass 2 on the GPU (-3).  All must give
    the reference's aggregates."""

    def __init__(self, lib, path):
        self.lib, self.path = lib, path

    def __enter__(self):
        opts = capi.hz_options()
        self.lib.hz_get_options(C.byref(opts))
        self.old = opts.hem_host_max_rows
        opts.hem_host_max_rows = {"gpu": -1, "full": -2, "staged": -3}.get(self.path, 0)
        self.lib.hz_set_options(C.byref(opts))

    def __exit__(self, *exc):
        opts = capi.hz_options()
        self.lib.hz_get_options(C.byref(opts))
        opts.hem_host_max_rows = self.old
        self.lib.hz_set_options(C.byref(opts))


@pytest.mark.parametrize("agg_type,hem_path", [(capi.HEC, "host"), (capi.HEM, "host"), (capi.HEM, "gpu"),
                                               (capi.HEM, "full"), (capi.HEM, "staged"), (capi.VMB, "host")])
@pytest.mark.parametrize("theta", [0.0, 0.04, 0.25])
def test_hierarchy_bitexact_solvers_fixture(lib, orc, libc, solvers_fixture, agg_type, hem_path, theta):
    ia, ja, va, _ = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    with _HemPath(lib, hem_path):
        (lg, thg, rg), (lo, tho, ro) = _hierarchies(lib, orc, libc, A, agg_type, theta)
    assert len(lg) == len(lo)
    for g, o in zip(lg, lo):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(g[key], o[key]), key
        assert g["theta"] == o["theta"]
        if "agg" in o:
            assert g["nagg"] == o["nagg"]
            assert np.array_equal(g["agg"], o["agg"])
    assert thg == tho          # caller's strong_coupled mutated identically
    assert rg == ro            # libc rand() state advanced identically (HEC)


@pytest.mark.parametrize("m", [16, 24])
@pytest.mark.parametrize("agg_type,hem_path", [(capi.HEC, "host"), (capi.HEM, "host"), (capi.HEM, "gpu"),
                                               (capi.HEM, "full"), (capi.HEM, "staged"), (capi.VMB, "host")])
def test_hierarchy_bitexact_kuhn3d(lib, orc, libc, m, agg_type, hem_path):
    ia, ja, va = problems.poisson3d_kuhn(m)
    A = capi.HostCSR(ia, ja, va)
    with _HemPath(lib, hem_path):
        (lg, _, rg), (lo, _, ro) = _hierarchies(lib, orc, libc, A, agg_type, 0.04, coarse_dof=50)
    assert len(lg) == len(lo)
    for g, o in zip(lg, lo):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(g[key], o[key]), key
        if "agg" in o:
            assert np.array_equal(g["agg"], o["agg"])
    assert rg == ro


@pytest.mark.parametrize("problem,maxagg", [("poisson", 20), ("aniso", 20), ("poisson", 4)])
def test_vmb_second_pass_on_the_gpu_bitexact(lib, orc, libc, problem, maxagg):
    """VMB levels of 65,536 rows or more run the second pass (join the first
    neighbouring pass-1 aggregate) on the GPU when no aggregate would pass
    max_aggregation with every first choice, and the host pass otherwise
    (max_aggregation 4: pass 1 fills its aggregates); both give the
    reference's aggregates and coarse matrices (47^3 = 103,823 rows)."""
    if problem == "poisson":
        ia, ja, va = problems.poisson3d_kuhn(48)
    else:
        ia, ja, va = problems.aniso3d_kuhn(48, 0.7, 1e-3)
    A = capi.HostCSR(ia, ja, va)
    (lg, thg, _), (lo, tho, _) = _hierarchies(lib, orc, libc, A, capi.VMB, 0.04, coarse_dof=100,
                                              max_aggregation=maxagg)
    print(f"VMB {problem} max_aggregation {maxagg}: levels {[g['ia'].size - 1 for g in lg]}", flush=True)
    assert len(lg) == len(lo)
    for g, o in zip(lg, lo):
        if "agg" in o:
            assert g["nagg"] == o["nagg"]
            assert np.array_equal(g["agg"], o["agg"])
        for key in ("ia", "ja", "val"):
            assert np.array_equal(g[key], o[key]), key
    assert thg == tho


@pytest.mark.parametrize("m", [64, 106])
def test_hierarchy_bitexact_kuhn3d_hem_by_size(lib, orc, libc, m):
    """HEM with the default form per level size: at m = 106 (1,157,625 rows)
    level 0 runs the GPU pipeline, the levels down to 32,768 rows the host
    sweep over GPU-sorted preference lists, the rest the full host sweep."""
    ia, ja, va = problems.poisson3d_kuhn(m)
    A = capi.HostCSR(ia, ja, va)
    with _HemPath(lib, "host"):
        (lg, thg, _), (lo, tho, _) = _hierarchies(lib, orc, libc, A, capi.HEM, 0.04, coarse_dof=100)
    assert len(lg) == len(lo)
    for g, o in zip(lg, lo):
        if "agg" in o:
            assert g["nagg"] == o["nagg"]
            assert np.array_equal(g["agg"], o["agg"])
        for key in ("ia", "ja", "val"):
            assert np.array_equal(g[key], o[key]), key
    assert thg == tho


def _solve_both(lib, orc, libc, A, b, **kw):
    res = {}
    for which in ("gpu", "oracle"):
        x = capi.HostVec(np.zeros(b.size))
        bb = capi.HostVec(b)
        libc.srand(1)
        if which == "gpu":
            it, amg = capi.solver_params(lib, **kw)
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg))
        else:
            import oracle
            it, amg = oracle.params(**kw)
            tr = capi.orc_trace()
            r = orc.orc_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg), C.byref(tr))
        res[which] = (r, x.arr.copy())
    return res


@pytest.mark.parametrize("agg_type", [capi.HEC, capi.VMB, capi.HEM])
@pytest.mark.parametrize("smoother", [capi.SMOOTHER_JACOBI, capi.SMOOTHER_L1DIAG])
@pytest.mark.parametrize("coarse_solver", [32, 0])
def test_pcg_parity_solvers_fixture(lib, orc, libc, solvers_fixture, agg_type, smoother, coarse_solver):
    ia, ja, va, b = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    res = _solve_both(lib, orc, libc, A, b, tol=1e-10, maxit=400, aggregation=agg_type,
                      smoother=smoother, coarse_solver=coarse_solver)
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert ig >= 0 and io >= 0
    assert abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("m", [16, 32])
def test_pcg_parity_kuhn3d(lib, orc, libc, m):
    ia, ja, va = problems.poisson3d_kuhn(m)
    b = problems.uniform_rhs(ia.size - 1, seed=3)
    A = capi.HostCSR(ia, ja, va)
    res = _solve_both(lib, orc, libc, A, b, tol=1e-10, maxit=400, aggregation=capi.HEC,
                      coarse_dof=100)
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("variant", [
    dict(cycle=capi.W_CYCLE),
    dict(coarse_scaling=1),
    dict(stop_type=capi.STOP_REL_PRECRES),
    dict(stop_type=capi.STOP_MOD_REL_RES),
    dict(amg_maxit=2),
    dict(aggregation=capi.VMB, cycle=capi.W_CYCLE, coarse_solver=0),
])
def test_pcg_parity_cycle_variants(lib, orc, libc, solvers_fixture, variant):
    ia, ja, va, b = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    kw = dict(tol=1e-10, maxit=400, aggregation=capi.HEC, smoother=capi.SMOOTHER_L1DIAG)
    kw.update(variant)
    res = _solve_both(lib, orc, libc, A, b, **kw)
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert ig >= 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("sweeps", [(2, 2), (0, 1), (1, 0)])
def test_pcg_parity_sweep_counts(lib, orc, libc, sweeps):
    import oracle
    ia, ja, va = problems.poisson2d_p1(48)
    b = problems.uniform_rhs(ia.size - 1, seed=11)
    A = capi.HostCSR(ia, ja, va)
    res = {}
    for which in ("gpu", "oracle"):
        x = capi.HostVec(np.zeros(b.size))
        bb = capi.HostVec(b)
        libc.srand(1)
        if which == "gpu":
            it, amg = capi.solver_params(lib, tol=1e-10, maxit=400, aggregation=capi.HEM,
                                         smoother=capi.SMOOTHER_L1DIAG)
        else:
            it, amg = oracle.params(tol=1e-10, maxit=400, aggregation=capi.HEM, smoother=10)
        amg.presmooth_iter, amg.postsmooth_iter = sweeps
        if which == "gpu":
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg))
        else:
            tr = capi.orc_trace()
            r = orc.orc_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg), C.byref(tr))
        res[which] = (r, x.arr.copy())
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert abs(ig - io) <= 1
    if io > 0:
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


def test_jacobi_weight_extension(lib, orc, libc):
    """Damped Jacobi (w = 2/3) through hz_options against the oracle with the
    same weight."""
    import oracle
    ia, ja, va = problems.poisson2d_p1(64)
    b = problems.uniform_rhs(ia.size - 1, seed=2)
    A = capi.HostCSR(ia, ja, va)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    opts.jacobi_weight = 2.0 / 3.0
    lib.hz_set_options(C.byref(opts))
    orc.orc_set_jacobi_weight(2.0 / 3.0)
    try:
        res = _solve_both(lib, orc, libc, A, b, tol=1e-10, maxit=400, aggregation=capi.HEM)
    finally:
        opts.jacobi_weight = 1.0
        lib.hz_set_options(C.byref(opts))
        orc.orc_set_jacobi_weight(1.0)
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert ig > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


def test_solution_quality_kuhn3d_64(lib):
    """Size-independent check at a size the oracle does not run in the test
    budget: the true FP64 residual of the returned solution meets the tolerance."""
    import scipy.sparse as sp
    ia, ja, va = problems.poisson3d_kuhn(64)
    b = problems.uniform_rhs(ia.size - 1, seed=5)
    A = capi.HostCSR(ia, ja, va)
    x = capi.HostVec(np.zeros(b.size))
    bb = capi.HostVec(b)
    it, amg = capi.solver_params(lib, tol=1e-10, maxit=400, aggregation=capi.HEM,
                                 smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200)
    r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg))
    assert r > 0
    Asp = sp.csr_matrix((va, ja, ia))
    assert np.linalg.norm(b - Asp @ x.arr) / np.linalg.norm(b) < 1e-10


def test_hec_turn_order_host_rand_matches_gpu_stream(lib, orc, libc):
    """hec_rng_source = 1 (rand() called on the host, the reference loop) and the
    default GPU generation of the glibc stream give the same hierarchy and leave
    libc's state in the same place."""
    ia, ja, va = problems.poisson3d_kuhn(20)
    A = capi.HostCSR(ia, ja, va)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    saved = opts.hec_rng_source
    per_source = {}
    try:
        for src in (0, 1):
            opts.hec_rng_source = src
            lib.hz_set_options(C.byref(opts))
            (lg, _, rg), (lo, _, ro) = _hierarchies(lib, orc, libc, A, capi.HEC, 0.0, coarse_dof=40)
            assert rg == ro
            assert len(lg) == len(lo)
            for g, o in zip(lg, lo):
                if "agg" in o:
                    assert np.array_equal(g["agg"], o["agg"])
            per_source[src] = (lg, rg)
    finally:
        opts.hec_rng_source = saved
        lib.hz_set_options(C.byref(opts))
    assert per_source[0][1] == per_source[1][1]


def test_dense_inverse_uses_matrix_cores(lib):
    """The FP64 MFMA lane layout assumed by the GEMM passes its self-check."""
    assert lib.hz_dense_mfma_active() == 1


@pytest.mark.parametrize("n", [1, 17, 64, 65, 200, 257, 640, 1000])
def test_dense_inverse_blocked(lib, n):
    """Block Gauss-Jordan inverse (FP64 matrix cores) of an SPD matrix against
    LAPACK: A X = I and X equal to numpy's inverse to 1e-12 relative."""
    rng = np.random.default_rng(n)
    G = rng.standard_normal((n, n))
    A = G @ G.T / n + np.eye(n)
    X = np.ascontiguousarray(A.copy())
    assert lib.hz_dense_inverse(n, capi.dptr(X)) == 0
    Xr = np.linalg.inv(A)
    assert np.linalg.norm(X - Xr) / np.linalg.norm(Xr) < 1e-12
    assert np.linalg.norm(A @ X - np.eye(n)) < 1e-11 * n


def test_dense_inverse_zero_pivot_reports_error(lib):
    A = np.ones((70, 70))   # first pivot block singular after one step
    assert lib.hz_dense_inverse(70, capi.dptr(A)) == -47


def test_pcg_parity_large_coarsest(lib, orc, libc):
    """A coarsest level above 256 rows takes the blocked inverse; iteration count
    and solution match the oracle's dense LDL^T coarse solve."""
    ia, ja, va = problems.poisson2d_p1(64)
    b = problems.uniform_rhs(ia.size - 1, seed=4)
    A = capi.HostCSR(ia, ja, va)
    res = _solve_both(lib, orc, libc, A, b, tol=1e-10, maxit=400, aggregation=capi.HEM,
                      coarse_dof=1000)
    info = capi.hz_solve_info()
    lib.hz_get_last_info(C.byref(info))
    assert info.level_rows[info.num_levels - 1] > 256
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert ig > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("agg_type,smoother,coarse_solver,tol", [
    (capi.HEM, capi.SMOOTHER_L1DIAG, 32, 1e-10), (capi.HEC, capi.SMOOTHER_L1DIAG, 0, 1e-6),
    (capi.VMB, capi.SMOOTHER_JACOBI, 0, 1e-6)])
def test_linear_solver_amg_parity(lib, orc, libc, agg_type, smoother, coarse_solver, tol):
    """AMG as the solver (linear_solver_amg / amg_solve) from a nonzero guess:
    iteration count within +-1 and solution within 1e-10 of the oracle's."""
    import oracle
    ia, ja, va = problems.poisson2d_p1(64)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=6)
    x0 = np.random.default_rng(1).standard_normal(n) * 0.01
    A = capi.HostCSR(ia, ja, va)
    out = {}
    for which in ("gpu", "oracle"):
        kw = dict(aggregation=agg_type, smoother=smoother, coarse_dof=60, coarse_solver=coarse_solver)
        _, amg = (capi.solver_params(lib, **kw) if which == "gpu" else oracle.params(**kw))
        amg.tol, amg.maxit = tol, 400
        x = capi.HostVec(x0.copy())
        bb = capi.HostVec(b)
        libc.srand(1)
        f = lib.linear_solver_amg if which == "gpu" else orc.orc_linear_solver_amg
        out[which] = (f(A.ptr(), bb.ptr(), x.ptr(), C.byref(amg)), x.arr.copy())
    (ig, xg), (io, xo) = out["gpu"], out["oracle"]
    assert 0 < io <= 400 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


def test_linear_solver_amg_edge_semantics(lib, libc):
    ia, ja, va = problems.poisson2d_p1(16)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    _, amg = capi.solver_params(lib, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=20)
    amg.tol, amg.maxit = 1e-14, 3
    x = capi.HostVec(np.zeros(n))
    assert lib.linear_solver_amg(A.ptr(), capi.HostVec(np.ones(n)).ptr(), x.ptr(), C.byref(amg)) == 4
    _, amg = capi.solver_params(lib, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=20)
    amg.tol, amg.maxit = 1e-8, 10
    x = capi.HostVec(np.full(n, 0.5))
    assert lib.linear_solver_amg(A.ptr(), capi.HostVec(np.zeros(n)).ptr(), x.ptr(), C.byref(amg)) == 1
    assert np.all(x.arr == 0.5)


def test_precond_object_parity_and_isolation(lib, orc, libc):
    """create_precond_amg / apply_precond against the oracle's precond_amg; the
    hierarchy survives other library calls (its own device arena) and the
    device-array variant gives the same z."""
    import oracle
    ia, ja, va = problems.poisson2d_p1(80)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    _, amg = capi.solver_params(lib, aggregation=capi.HEC, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=80)
    libc.srand(1)
    pc = lib.create_precond_amg(A.ptr(), C.byref(amg))
    assert pc
    _, amgo = oracle.params(aggregation=capi.HEC, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=80)
    libc.srand(1)
    h = orc.orc_precond_create(A.ptr(), C.byref(amgo))
    assert amg.strong_coupled == amgo.strong_coupled
    rng = np.random.default_rng(5)
    try:
        for k in range(3):
            r = rng.standard_normal(n)
            zg, zo = np.zeros(n), np.zeros(n)
            lib.apply_precond(capi.dptr(r), capi.dptr(zg), pc)
            orc.orc_precond_apply(h, capi.dptr(r), capi.dptr(zo))
            assert np.linalg.norm(zg - zo) / np.linalg.norm(zo) < 1e-12
            # an unrelated call reuses the library's working memory
            y = np.zeros(n)
            assert lib.hz_dcsr_mxv(A.ptr(), capi.dptr(r), capi.dptr(y)) == 0
        hip = _hip_runtime()
        dr, dz = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(dr), C.c_size_t(8 * n)) == 0
        assert hip.hipMalloc(C.byref(dz), C.c_size_t(8 * n)) == 0
        try:
            assert hip.hipMemcpy(dr, capi.dptr(r), C.c_size_t(8 * n), 1) == 0   # host to device
            lib.hz_precond_amg_apply_device(dr, dz, pc.contents.data)
            zd = np.zeros(n)
            assert hip.hipMemcpy(capi.dptr(zd), dz, C.c_size_t(8 * n), 2) == 0  # device to host
            assert np.array_equal(zd, zg)
        finally:
            hip.hipFree(dr)
            hip.hipFree(dz)
    finally:
        lib.hz_precond_free(pc)
        orc.orc_precond_free(h)


def _hip_runtime():
    """The HIP runtime the library itself is linked against (device buffers for
    the device-pointer entry points)."""
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _set_precision(lib, p):
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = opts.vcycle_precision
    opts.vcycle_precision = p
    lib.hz_set_options(C.byref(opts))
    return old


@pytest.mark.parametrize("precision,ztol", [(1, 2e-5), (2, 5e-2)])
def test_mixed_precision_cycle_vs_fp64_cycle(lib, libc, precision, ztol):
    """One preconditioner application with the FP32 cycle (FP32 or BF16
    values) against the FP64 cycle on the same hierarchy: relative difference
    within the storage precision's reach (FP32: 2e-5, BF16 values: 5e-2)."""
    ia, ja, va, _ = problems.poisson3d_kuhn_coef(24)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    r = np.random.default_rng(2).standard_normal(n)
    z = {}
    for p in (0, precision):
        old = _set_precision(lib, p)
        try:
            _, amg = capi.solver_params(lib, aggregation=capi.HEC, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=100)
            libc.srand(1)
            pc = lib.create_precond_amg(A.ptr(), C.byref(amg))
            assert pc
            zz = np.zeros(n)
            lib.apply_precond(capi.dptr(r), capi.dptr(zz), pc)
            lib.hz_precond_free(pc)
            z[p] = zz
        finally:
            _set_precision(lib, old)
    rel = np.linalg.norm(z[precision] - z[0]) / np.linalg.norm(z[0])
    assert rel < ztol, rel


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("problem", ["poisson", "coef"])
def test_mixed_precision_pcg_reaches_fp64_tolerance(lib, orc, libc, precision, problem):
    """PCG with the FP32 / BF16-valued cycle reaches 1e-10 in the FP64 true
    residual, within a few iterations of the FP64 cycle, and its solution is
    within 1e-8 of the FP64 oracle's."""
    import scipy.sparse as sp
    if problem == "poisson":
        ia, ja, va = problems.poisson3d_kuhn(32)
    else:
        ia, ja, va, _ = problems.poisson3d_kuhn_coef(32)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=4)
    A = capi.HostCSR(ia, ja, va)
    kw = dict(tol=1e-10, maxit=500, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200,
              theta=0.04)
    res = {}
    for p in (0, precision):
        old = _set_precision(lib, p)
        try:
            res[p] = _solve_both(lib, orc, libc, A, b, **kw)["gpu"] if p == 0 else None
            if p:
                x = capi.HostVec(np.zeros(n))
                it, amg = capi.solver_params(lib, **kw)
                libc.srand(1)
                res[p] = (lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(),
                                                            C.byref(it), C.byref(amg)), x.arr.copy())
        finally:
            _set_precision(lib, old)
    import oracle
    it, amg = oracle.params(**kw)
    xo = capi.HostVec(np.zeros(n))
    libc.srand(1)
    io = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), xo.ptr(), C.byref(it), C.byref(amg),
                            C.byref(capi.orc_trace()))
    (i64, _), (im, xm) = res[0], res[precision]
    assert io > 0 and abs(i64 - io) <= 1
    assert 0 < im <= i64 + 4, (im, i64)
    Asp = sp.csr_matrix((va, ja, ia), shape=(n, n))
    assert np.linalg.norm(b - Asp @ xm) / np.linalg.norm(b) < 1e-10
    assert np.linalg.norm(xm - xo.arr) / np.linalg.norm(xo.arr) < 1e-8


def test_python_wrapper_reads_input_dat(lib, libc, tmp_path, monkeypatch):
    """python_wrapper_krylov_amg (wrapper.c:36-79): parameters from
    ./input.dat, tol/maxit/print level from the arguments, solve in place on
    raw arrays; equals the drop-in called with the same parameters."""
    (tmp_path / "input.dat").write_text(
        "linear_itsolver_type = 1\nlinear_stop_type = 1\nAMG_type = UA\nAMG_cycle_type = V\n"
        "AMG_levels = 20\nAMG_smoother = L1DIAG\nAMG_coarse_dof = 100\nAMG_coarse_solver = 32\n"
        "AMG_aggregation_type = 5\nAMG_strong_coupled = 0.0\nAMG_maxit = 1\nAMG_tol = 1e-6\n")
    monkeypatch.chdir(tmp_path)
    ia, ja, va = problems.poisson2d_p1(48)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=9)
    u = np.zeros(n)
    args = [C.byref(INTv) for INTv in (capi.INT(n), capi.INT(int(ia[-1])))]
    libc.srand(1)
    lib.python_wrapper_krylov_amg(args[0], args[1], capi.dptr(ia, capi.INT), capi.dptr(ja, capi.INT),
                                  capi.dptr(va), capi.dptr(b), capi.dptr(u), C.byref(capi.REAL(1e-10)),
                                  C.byref(capi.INT(300)), C.byref(capi.INT(0)))
    x = capi.HostVec(np.zeros(n))
    it, amg = capi.solver_params(lib, tol=1e-10, maxit=300, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG,
                                 coarse_dof=100, theta=0.0)
    libc.srand(1)
    r = lib.linear_solver_dcsr_krylov_amg(capi.HostCSR(ia, ja, va).ptr(), capi.HostVec(b).ptr(), x.ptr(),
                                          C.byref(it), C.byref(amg))
    assert r > 0
    assert np.array_equal(u, x.arr)


def _sa_param(params_fn, agg_type, filt, omega=0.67, **kw):
    it, amg = params_fn(aggregation=agg_type, **kw)
    amg.AMG_type = capi.SA_AMG
    amg.smooth_filter = filt
    amg.tentative_smooth = omega
    return it, amg


@pytest.mark.parametrize("agg_type", [capi.HEC, capi.HEM, capi.VMB])
@pytest.mark.parametrize("filt", [1, 0])
def test_sa_hierarchy_bitexact(lib, orc, libc, solvers_fixture, agg_type, filt):
    """Smoothed aggregation setup (amg_setup_smoothP_smoothR): aggregates,
    smoothed P (structure, column order and values) and every coarse matrix
    bit-identical to the oracle.  (Pairwise HEM aggregates make smoothed
    coarse levels nearly dense, so HEM runs on a smaller 2-D problem.)"""
    import oracle
    if agg_type == capi.HEM:
        ia, ja, va = problems.poisson2d_p1(40)
    else:
        ia, ja, va, _ = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    out = []
    for which in ("gpu", "oracle"):
        fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
        _, amg = _sa_param(fn, agg_type, filt, theta=0.04, coarse_dof=100)
        h = capi.hz_hierarchy()
        libc.srand(1)
        st = (lib.hz_setup_hierarchy if which == "gpu" else orc.orc_setup_hierarchy)(A.ptr(), C.byref(amg),
                                                                                      C.byref(h))
        assert st == 0
        out.append(capi.hierarchy_to_numpy(h))
        (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
    g, o = out
    assert len(g) == len(o) >= 3
    for lg, lo in zip(g, o):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(lg[key], lo[key]), key
        if "agg" in lo:
            assert np.array_equal(lg["agg"], lo["agg"])
            for a, b_ in zip(lg["P"], lo["P"]):
                assert np.array_equal(a, b_)


@pytest.mark.parametrize("agg_type,smoother", [(capi.HEC, capi.SMOOTHER_L1DIAG), (capi.HEM, capi.SMOOTHER_L1DIAG),
                                               (capi.VMB, capi.SMOOTHER_JACOBI)])
def test_sa_pcg_parity(lib, orc, libc, solvers_fixture, agg_type, smoother):
    import oracle
    if agg_type == capi.HEM:   # smoothed pairwise aggregates: nearly dense coarse levels
        ia, ja, va = problems.poisson2d_p1(40)
        b = problems.uniform_rhs(ia.size - 1, seed=3)
    else:
        ia, ja, va, b = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    res = {}
    for which in ("gpu", "oracle"):
        fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
        it, amg = _sa_param(fn, agg_type, 1, tol=1e-10, maxit=300, smoother=smoother, coarse_dof=100)
        x = capi.HostVec(np.zeros(b.size))
        libc.srand(1)
        if which == "gpu":
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                                                  C.byref(amg))
        else:
            r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                   C.byref(capi.orc_trace()))
        res[which] = (r, x.arr.copy())
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert io > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


def test_sa_kuhn3d_hierarchy_and_solve(lib, orc, libc):
    """SA on the 3-D Kuhn problem (m = 16): hierarchy bit-exact (large rows
    exercise the Galerkin tables) and PCG parity."""
    import oracle
    ia, ja, va = problems.poisson3d_kuhn(16)
    b = problems.uniform_rhs(ia.size - 1, seed=2)
    A = capi.HostCSR(ia, ja, va)
    out, res = [], {}
    for which in ("gpu", "oracle"):
        fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
        it, amg = _sa_param(fn, capi.HEC, 1, tol=1e-10, maxit=300, smoother=capi.SMOOTHER_L1DIAG,
                            coarse_dof=60, theta=0.04)
        h = capi.hz_hierarchy()
        libc.srand(1)
        st = (lib.hz_setup_hierarchy if which == "gpu" else orc.orc_setup_hierarchy)(A.ptr(), C.byref(amg),
                                                                                      C.byref(h))
        assert st == 0
        out.append(capi.hierarchy_to_numpy(h))
        (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
        it, amg = _sa_param(fn, capi.HEC, 1, tol=1e-10, maxit=300, smoother=capi.SMOOTHER_L1DIAG,
                            coarse_dof=60, theta=0.04)
        x = capi.HostVec(np.zeros(b.size))
        libc.srand(1)
        if which == "gpu":
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                                                  C.byref(amg))
        else:
            r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                   C.byref(capi.orc_trace()))
        res[which] = (r, x.arr.copy())
    for lg, lo in zip(*out):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(lg[key], lo[key]), key
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert io > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("agg,wide", [(capi.VMB, False), (capi.HEC, True)])
def test_sa_galerkin_narrow_tables_first_bitexact(lib, orc, libc, agg, wide):
    """The default form of the finest level's general Galerkin product first
    runs with 128-column tables and, when a row has more columns, runs again
    with the full 256-column ones: 3-D Kuhn 16^3 with VMB aggregates (coarse
    rows within 128 columns) and with HEC pairs (wider rows: the second run).
    Every coarse matrix bit-identical to the oracle's."""
    ia, ja, va = problems.poisson3d_kuhn(16)
    A = capi.HostCSR(ia, ja, va)
    out = []
    for which in ("gpu", "oracle"):
        import oracle
        fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
        _, amg = _sa_param(fn, agg, 1, theta=0.04, coarse_dof=60)
        h = capi.hz_hierarchy()
        libc.srand(1)
        st = (lib.hz_setup_hierarchy if which == "gpu" else orc.orc_setup_hierarchy)(A.ptr(), C.byref(amg),
                                                                                      C.byref(h))
        assert st == 0
        out.append(capi.hierarchy_to_numpy(h))
        (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
    lg, lo = out
    assert len(lg) == len(lo) and len(lg) > 1
    widest = int(np.diff(lg[1]["ia"]).max())
    print(f"aggregation {agg}: widest row of the level-1 operator {widest}", flush=True)
    assert (widest > 128) == wide
    for g, o in zip(lg, lo):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(g[key], o[key]), key


@pytest.mark.parametrize("problem", ["kuhn16_hec", "p2d40_hem"])
@pytest.mark.parametrize("form", [-1, -2, 100])
def test_sa_galerkin_product_forms_bitexact(lib, orc, libc, problem, form):
    """dcsr_rap (sparse.c:2401-2538) through each form of the GPU product
    (hz_options.rap_wave_stage_pairs): every row one thread per row (-1), the
    wavefront kernel limited to 16 columns so that rows overflow its column
    table and fall back (-2), and stages of 100 (R entry, A entry) pairs, so
    that rows span many stages and rows with a longer A row fall back (100).
    Every coarse matrix bit-identical to the oracle's."""
    import oracle
    if problem == "kuhn16_hec":
        (ia, ja, va), agg, cdof = problems.poisson3d_kuhn(16), capi.HEC, 60
    else:
        (ia, ja, va), agg, cdof = problems.poisson2d_p1(40), capi.HEM, 100
    A = capi.HostCSR(ia, ja, va)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = opts.rap_wave_stage_pairs
    out = []
    try:
        opts.rap_wave_stage_pairs = form
        lib.hz_set_options(C.byref(opts))
        for which in ("gpu", "oracle"):
            fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
            _, amg = _sa_param(fn, agg, 1, theta=0.04, coarse_dof=cdof)
            h = capi.hz_hierarchy()
            libc.srand(1)
            st = (lib.hz_setup_hierarchy if which == "gpu" else orc.orc_setup_hierarchy)(A.ptr(), C.byref(amg),
                                                                                          C.byref(h))
            assert st == 0
            out.append(capi.hierarchy_to_numpy(h))
            (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
    finally:
        opts.rap_wave_stage_pairs = old
        lib.hz_set_options(C.byref(opts))
    g, o = out
    assert len(g) == len(o) >= 3
    for lg, lo in zip(g, o):
        for key in ("ia", "ja", "val"):
            assert np.array_equal(lg[key], lo[key]), key


def test_first_call_in_process_keeps_caller_rand_stream():
    """The first library call of a process brings up the HIP runtime, which
    consumes libc random() values; the caller's stream must be untouched so
    that HEC (iarray_shuffle with rand(), array.c:302-318) sees srand(1)."""
    import subprocess
    import sys
    from pathlib import Path
    script = Path(__file__).parent / "first_call_hec.py"
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "aggregates_equal=1 rand_equal=1" in r.stdout, r.stdout


def _mlp_data(B, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, 12)).astype(np.float32)
    # a smooth target of the features (the network must be able to fit it)
    t = np.stack([np.tanh(x[:, 0] - 0.5 * x[:, 3]) + 0.1 * x[:, 5] ** 2,
                  np.sin(x[:, 1]) * 0.5 + 0.2 * x[:, 7]], axis=1).astype(np.float32)
    return x, t


@pytest.mark.parametrize("B", [1, 63, 64, 200, 1000])
def test_mlp_forward_and_gradient_vs_oracle(lib, B):
    """FP32 forward/backward on the matrix cores against the FP64 oracle:
    outputs within 2e-5 relative, gradient within 1e-4 relative (FP32
    accumulation over the batch)."""
    from oracle import mlp
    theta = np.zeros(capi.MLP_PARAMS, np.float32)
    lib.hz_mlp_init(capi.fptr(theta), 11)
    theta[mlp.OFF_B1:mlp.OFF_W2] = 0.05   # nonzero biases
    x, t = _mlp_data(B, B)
    y = np.zeros((B, 2), np.float32)
    assert lib.hz_mlp_forward(capi.fptr(theta), capi.fptr(x), B, capi.fptr(y)) == 0
    yo, _ = mlp.forward(theta.astype(np.float64), x.astype(np.float64))
    assert np.linalg.norm(y - yo) / np.linalg.norm(yo) < 2e-5
    g = np.zeros(capi.MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    assert lib.hz_mlp_loss_grad(capi.fptr(theta), capi.fptr(x), capi.fptr(t), B, capi.fptr(g), capi.fptr(loss)) == 0
    lo, go = mlp.loss_grad(theta.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss[0] - lo) < 1e-5 * max(1.0, lo)
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 1e-4
    # padding entries get exactly zero gradient
    assert np.all(g[mlp.OFF_W1:mlp.OFF_B1].reshape(16, 64)[12:] == 0)
    assert np.all(g[mlp.OFF_W3:mlp.OFF_B3].reshape(64, 16)[:, 2:] == 0)


def test_mlp_training_fits_smooth_target(lib):
    """Adam on the GPU kernels reduces the loss of a smooth regression target
    by more than 10x, and the trained network generalises to fresh samples."""
    from oracle import mlp
    theta = np.zeros(capi.MLP_PARAMS, np.float32)
    lib.hz_mlp_init(capi.fptr(theta), 5)
    x, t = _mlp_data(4096, 1)
    hist = np.zeros(41, np.float32)
    assert lib.hz_mlp_train(capi.fptr(theta), capi.fptr(x), capi.fptr(t), 4096, 40, 256, 3e-3, 9,
                            capi.fptr(hist)) == 0
    assert hist[-1] < hist[0] / 10, hist
    xv, tv = _mlp_data(1024, 2)
    assert mlp.loss(theta.astype(np.float64), xv.astype(np.float64), tv.astype(np.float64)) < 3 * hist[-1] + 1e-3


def _band_spd(n, half, seed):
    """Symmetric banded matrix with 2*half+1 entries a row (fewer near the
    ends), strictly diagonally dominant: long rows for the 16-lane kernels."""
    rng = np.random.default_rng(seed)
    off = {k: -rng.uniform(0.1, 1.0, n - k) for k in range(1, half + 1)}
    rows, cols, vals = [], [], []
    for i in range(n):
        s = 0.0
        for j in range(max(0, i - half), min(n, i + half + 1)):
            if j == i:
                continue
            v = off[abs(i - j)][min(i, j)]
            rows.append(i), cols.append(j), vals.append(v)
            s += abs(v)
        rows.append(i), cols.append(i), vals.append(s + 1.0)
    order = np.lexsort((np.array(cols), np.array(rows)))
    ja = np.array(cols, np.int32)[order]
    va = np.array(vals)[order]
    ia = np.zeros(n + 1, np.int32)
    np.add.at(ia, np.array(rows)[order] + 1, 1)
    return np.cumsum(ia).astype(np.int32), ja, va


@pytest.mark.parametrize("problem", ["kuhn_coef12", "band41"])
def test_chebyshev_smoother_bitexact(lib, orc, problem):
    """Chebyshev steps against the oracle, bit-exact: on the SELL kernel
    (15-entry rows) and on the 16-lanes-per-row CSR kernel that small levels
    with long rows take (41-entry band)."""
    if problem == "kuhn_coef12":
        ia, ja, va, _ = problems.poisson3d_kuhn_coef(12)
    else:
        ia, ja, va = _band_spd(3000, 20, 5)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    rng = np.random.default_rng(3)
    b, x0 = rng.standard_normal(n), rng.standard_normal(n)
    for steps, scale in ((1, 1.0), (3, 0.8)):
        xg, xo = x0.copy(), x0.copy()
        assert lib.hz_cheby_smooth(A.ptr(), capi.dptr(b), capi.dptr(xg), steps, scale) == 0
        orc.orc_cheby_smooth(A.ptr(), capi.dptr(b), capi.dptr(xo), steps, scale)
        assert np.array_equal(xg, xo)


def test_sa_prolongation_per_aggregate_damping_bitexact(lib, orc, libc):
    ia, ja, va = problems.aniso2d_p1(40, 0.7, 0.01)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    agg = (np.arange(n) // 4).astype(np.int32)
    agg[::17] = -1
    nagg = int(agg.max()) + 1
    om = (0.2 + np.random.default_rng(2).random(nagg)).astype(np.float32)
    for filt, omp in ((1, om), (0, om), (1, None)):
        Pg, Po = capi.dCSRmat(), capi.dCSRmat()
        ptr = capi.fptr(omp) if omp is not None else None
        assert lib.hz_sa_prolongation(A.ptr(), capi.dptr(agg, capi.INT), nagg, 0.1, 0.67, ptr, filt,
                                      C.byref(Pg)) == 0
        orc.orc_sa_prolongation(A.ptr(), capi.dptr(agg, capi.INT), nagg, 0.1, 0.67, ptr, filt, C.byref(Po))
        for a, b_ in zip(capi.csr_to_numpy(Pg), capi.csr_to_numpy(Po)):
            assert np.array_equal(a, b_)


def test_aggregate_features_vs_oracle(lib):
    from oracle import learned
    ia, ja, va, _ = problems.poisson3d_kuhn_coef(10)
    n = ia.size - 1
    agg = (np.arange(n) // 5).astype(np.int32)
    agg[3::11] = -1
    nagg = int(agg.max()) + 1
    F = np.zeros((nagg, 12), np.float32)
    assert lib.hz_aggregate_features(capi.HostCSR(ia, ja, va).ptr(), capi.dptr(agg, capi.INT), nagg, 2,
                                     capi.fptr(F)) == 0
    Fo = learned.aggregate_features(ia, ja, va, agg, nagg, 2)
    assert np.allclose(F, Fo, rtol=1e-5, atol=1e-6)


def test_sa_chebyshev_pcg_parity(lib, orc, libc, solvers_fixture):
    import oracle
    ia, ja, va, b = solvers_fixture
    A = capi.HostCSR(ia, ja, va)
    res = {}
    for which in ("gpu", "oracle"):
        fn = (lambda **k: capi.solver_params(lib, **k)) if which == "gpu" else oracle.params
        it, amg = _sa_param(fn, capi.HEC, 1, tol=1e-10, maxit=300, smoother=capi.SMOOTHER_CHEBY, coarse_dof=100)
        amg.presmooth_iter = amg.postsmooth_iter = 2
        x = capi.HostVec(np.zeros(b.size))
        libc.srand(1)
        if which == "gpu":
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                                                  C.byref(amg))
        else:
            r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                   C.byref(capi.orc_trace()))
        res[which] = (r, x.arr.copy())
    (ig, xg), (io, xo) = res["gpu"], res["oracle"]
    assert io > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


def test_two_grid_rate_behaves(lib, libc):
    """Two-grid contraction of SA + Chebyshev(2,2) on a 2-D Poisson problem
    against the spectral radius of the dense error propagator I - B A built
    from the oracle's cycle: 0.700 (damping 0) and 0.596 (damping 0.67)."""
    ia, ja, va = problems.poisson2d_p1(32)
    A = capi.HostCSR(ia, ja, va)
    _, amg = capi.solver_params(lib, aggregation=capi.HEC, coarse_dof=10)
    amg.presmooth_iter = amg.postsmooth_iter = 2
    rates = {}
    for om in (0.0, 0.67):
        r = capi.REAL()
        libc.srand(1)
        assert lib.hz_two_grid_rate(A.ptr(), C.byref(amg), om, 1.0, 40, C.byref(r)) == 0
        rates[om] = r.value
    assert abs(rates[0.0] - 0.700) < 0.03 and abs(rates[0.67] - 0.596) < 0.03, rates


def _aniso_plain(lib, libc, ia, ja, va, b, agg, sm, typ=capi.UA_AMG, steps=1):
    A = capi.HostCSR(ia, ja, va)
    x = capi.HostVec(np.zeros(b.size))
    it, amg = capi.solver_params(lib, tol=1e-10, maxit=1000, aggregation=agg, smoother=sm, coarse_dof=200,
                                 theta=0.04)
    amg.AMG_type = typ
    amg.presmooth_iter = amg.postsmooth_iter = steps
    libc.srand(1)
    r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg))
    return r, x.arr


@pytest.mark.parametrize("m,th", [(32, 1.1), (48, 0.4)])
def test_learned_parameters_gain(lib, libc, m, th):
    """The trained network (hazmath_amd/models) on held-out 3-D anisotropic
    problems (eps = 1e-3): the learned SA + Chebyshev(2) method reaches the
    FP64 tolerance without fallback, in fewer PCG iterations than the same
    method with the default damping 0.67 / bound scale 1, and in fewer than
    the plain drop-in (UA, HEM, L1)."""
    import scipy.sparse as sp
    from hazmath_amd import learned
    theta = learned.load_model()
    ia, ja, va = problems.aniso3d_kuhn(m, th, 1e-3)
    b = problems.uniform_rhs(ia.size - 1, seed=1)
    libc.srand(1)
    il, xl, info = learned.solve_learned(lib, ia, ja, va, b, theta, budget=400)
    assert info.fallback_used == 0 and il > 0
    Asp = sp.csr_matrix((va, ja, ia))
    assert np.linalg.norm(b - Asp @ xl) / np.linalg.norm(b) < 1e-10
    idef, _ = _aniso_plain(lib, libc, ia, ja, va, b, capi.VMB, capi.SMOOTHER_CHEBY, capi.SA_AMG, 2)
    iplain, _ = _aniso_plain(lib, libc, ia, ja, va, b, capi.HEM, capi.SMOOTHER_L1DIAG)
    assert il < idef and il < iplain, (il, idef, iplain)


def test_learned_fallback_on_budget(lib, libc):
    """Over budget, the plain method produces x (fallback flag set) and the
    result meets the tolerance."""
    import scipy.sparse as sp
    from hazmath_amd import learned
    theta = learned.load_model()
    ia, ja, va = problems.aniso2d_p1(64, 0.7, 1e-2)
    b = problems.uniform_rhs(ia.size - 1, seed=2)
    libc.srand(1)
    r, x, info = learned.solve_learned(lib, ia, ja, va, b, theta, budget=3)
    assert info.fallback_used == 1 and r > 3
    Asp = sp.csr_matrix((va, ja, ia))
    assert np.linalg.norm(b - Asp @ x) / np.linalg.norm(b) < 1e-10


# ---- Gauss-Seidel family: multicolour sweeps (smoother.c:109-307) ----------
# The GPU sweeps run Gauss-Seidel in a colour-major row order.  Bar 1: for
# the order the library reports, its sweeps are bit-identical to the oracle's
# restatement of the reference's loops run in that order (and the colouring
# is valid: no two rows of a colour coupled in A or A^T).  Bar 2: against the
# reference's natural order, PCG iteration counts within 25% (at least +-2)
# and the same tolerance reached; AMG as a stationary solver within 35%: its
# count grows like 1 / (1 - rho) in the cycle's contraction rho, PCG's only
# like the square root of the condition number, so the same change of
# smoothing order moves it further.

GS_FAMILY = [capi.SMOOTHER_GS, capi.SMOOTHER_SGS, capi.SMOOTHER_SOR, capi.SMOOTHER_SSOR, capi.SMOOTHER_GSOR,
             capi.SMOOTHER_SGSOR]


def _gs_gpu(lib, A, b, x0, sm, post, sweeps, relax):
    n = b.size
    x = x0.copy()
    order = np.zeros(n, np.int32)
    cptr = np.zeros(n + 1, np.int32)
    nc = capi.INT()
    assert lib.hz_gs_smooth(A.ptr(), capi.dptr(b), capi.dptr(x), sm, post, sweeps, relax, capi.dptr(order, capi.INT),
                            capi.dptr(cptr, capi.INT), C.byref(nc)) == 0
    return x, order, cptr[:nc.value + 1].copy()


def _check_colouring(ia, ja, order, cptr):
    n = ia.size - 1
    assert np.array_equal(np.sort(order), np.arange(n))
    col = np.empty(n, np.int64)
    for c in range(cptr.size - 1):
        blk = order[cptr[c]:cptr[c + 1]]
        assert np.all(np.diff(blk) > 0)
        col[blk] = c
    rows = np.repeat(np.arange(n), np.diff(ia))
    off = rows != ja
    assert not np.any(col[rows[off]] == col[ja[off]])


@pytest.mark.parametrize("case", ["random", "kuhn", "random_large", "kuhn_large"])
def test_gs_family_sweeps_bitexact_in_colour_order(lib, orc, case):
    # up to 16384 rows a sweep is one single-workgroup launch, above it one
    # launch per colour: both paths are covered
    if case.startswith("random"):   # structurally nonsymmetric, shuffled rows
        ia, ja, va = random_csr(2500 if case == "random" else 40000, 21, maxlen=12)
        va = _spd_like(ia, ja, va)
    else:
        ia, ja, va = problems.poisson3d_kuhn(14 if case == "kuhn" else 30)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    rng = np.random.default_rng(5)
    x0, b = rng.standard_normal(n), rng.standard_normal(n)
    for sm in GS_FAMILY:
        for post in (0, 1):
            for sweeps, relax in ((1, 1.0), (2, 1.3)):
                xg, order, cptr = _gs_gpu(lib, A, b, x0, sm, post, sweeps, relax)
                _check_colouring(ia, ja, order, cptr)
                xo = x0.copy()
                assert orc.orc_gs_smooth_ordered(A.ptr(), capi.dptr(b), capi.dptr(xo), sm, post, sweeps, relax,
                                                 capi.dptr(order, capi.INT), capi.dptr(cptr, capi.INT),
                                                 cptr.size - 1) == 0
                assert np.array_equal(xg, xo), (sm, post, sweeps, relax)


def _gs_tolerance(io, frac=0.25):
    return max(2, int(np.ceil(frac * io)))


@pytest.mark.parametrize("smoother", GS_FAMILY)
@pytest.mark.parametrize("problem", ["fixture", "kuhn"])
def test_gs_family_pcg_iterations_vs_oracle(lib, orc, libc, solvers_fixture, smoother, problem):
    import oracle
    if problem == "fixture":
        ia, ja, va, b = solvers_fixture
        agg = capi.HEM
    else:
        ia, ja, va = problems.poisson3d_kuhn(32)
        b = problems.uniform_rhs(ia.size - 1, seed=3)
        agg = capi.HEC
    A = capi.HostCSR(ia, ja, va)
    out = {}
    for which in ("gpu", "oracle"):
        kw = dict(tol=1e-10, maxit=400, aggregation=agg, smoother=smoother, coarse_dof=100)
        it, amg = capi.solver_params(lib, **kw) if which == "gpu" else oracle.params(**kw)
        amg.relaxation = 1.2
        x = capi.HostVec(np.zeros(b.size))
        libc.srand(1)
        if which == "gpu":
            r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg))
        else:
            r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                   C.byref(capi.orc_trace()))
        out[which] = (r, x.arr.copy())
    (ig, xg), (io, xo) = out["gpu"], out["oracle"]
    print(f"[gs] {problem} smoother {smoother}: gpu {ig} oracle {io}", flush=True)
    assert 0 < io < 400 and 0 < ig < 400
    assert abs(ig - io) <= _gs_tolerance(io)
    Asp = __import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix((va, ja, ia))
    assert np.linalg.norm(b - Asp @ xg) / np.linalg.norm(b) < 1e-9
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-7


@pytest.mark.parametrize("smoother", [capi.SMOOTHER_GS, capi.SMOOTHER_SGS, capi.SMOOTHER_SSOR])
def test_gs_family_linear_solver_amg_vs_oracle(lib, orc, libc, smoother):
    import oracle
    ia, ja, va = problems.poisson2d_p1(64)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=6)
    A = capi.HostCSR(ia, ja, va)
    out = {}
    for which in ("gpu", "oracle"):
        kw = dict(aggregation=capi.HEM, smoother=smoother, coarse_dof=60, coarse_solver=32)
        _, amg = (capi.solver_params(lib, **kw) if which == "gpu" else oracle.params(**kw))
        amg.tol, amg.maxit, amg.relaxation = 1e-8, 400, 1.1
        x = capi.HostVec(np.zeros(n))
        libc.srand(1)
        f = lib.linear_solver_amg if which == "gpu" else orc.orc_linear_solver_amg
        out[which] = (f(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(amg)), x.arr.copy())
    (ig, xg), (io, xo) = out["gpu"], out["oracle"]
    print(f"[gs] amg_solve smoother {smoother}: gpu {ig} oracle {io}", flush=True)
    assert 0 < io <= 400 and 0 < ig <= 400
    assert abs(ig - io) <= _gs_tolerance(io, 0.35)
    Asp = __import__("scipy.sparse", fromlist=["csr_matrix"]).csr_matrix((va, ja, ia))
    assert np.linalg.norm(b - Asp @ xg) / np.linalg.norm(b) < 1e-8


# ---- batched systems x 8 right-hand sides (hz_batch_krylov_amg) -------------
# Bar: every (system, right-hand side) pair behaves as the single-system
# solver with the same FP32 cycle (linear_solver_dcsr_krylov_amg with
# vcycle_precision = 1): iteration counts within +-1 and solutions within
# 1e-9 relative; against the FP64 oracle on a seeded subset: iteration counts
# within +-2 (the cycle's rounding differs) and the FP64 true residual below
# the tolerance (check III recomputes it before a pair is declared converged).

def _batch_case(nrhs=8, tiny=True):
    specs = problems.batch_anisotropic_2d(7, seed=9, side_range=(33, 97))
    systems, B = [], []
    for d in specs:
        ia, ja, va = problems.aniso2d_p1(d["side"] - 1, d["theta"], d["eps"])
        systems.append((ia, ja, va))
        B.append(problems.batch_rhs(ia.size - 1, d["index"], nrhs=nrhs))
    if tiny:   # 121 rows <= coarse_dof: solved by its dense inverse at level 0
        ia, ja, va = problems.aniso2d_p1(12, 0.3, 0.01)
        systems.insert(2, (ia, ja, va))
        B.insert(2, problems.batch_rhs(ia.size - 1, 99, nrhs=nrhs))
    return systems, B


def _batch_params(lib, tol=1e-10, **kw):
    it, amg = capi.solver_params(lib, tol=tol, maxit=500, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG,
                                 coarse_dof=200, theta=0.04, **kw)
    return it, amg


def test_batch_matches_single_system_fp32_cycle(lib, libc):
    systems, B = _batch_case(nrhs=5)
    B[3][1] = 0.0      # a zero right-hand side finishes at once
    X = [np.zeros_like(b) for b in B]
    it, amg = _batch_params(lib)
    r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
    assert r == 0
    assert iters[3, 1] == 0 and not X[3][1].any()
    # dense-inverse-only system: exact up to the cycle's FP32 output, so the
    # second iteration reaches 1e-10
    assert iters[2].max() <= 2
    old = _set_precision(lib, 1)
    try:
        for k, (ia, ja, va) in enumerate(systems):
            A = capi.HostCSR(ia, ja, va)
            for q in range(B[k].shape[0]):
                if not B[k][q].any():
                    continue
                its, amgs = _batch_params(lib)
                x = capi.HostVec(np.zeros(ia.size - 1))
                libc.srand(1)
                rs = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(B[k][q]).ptr(), x.ptr(), C.byref(its),
                                                       C.byref(amgs))
                print(f"[batch] system {k} rhs {q}: batched {iters[k, q]} single {rs}", flush=True)
                assert abs(int(iters[k, q]) - rs) <= 1
                assert np.linalg.norm(X[k][q] - x.arr) <= 1e-9 * np.linalg.norm(x.arr)
    finally:
        _set_precision(lib, old)


def test_batch_vs_oracle_subset(lib, orc, libc):
    import oracle
    import scipy.sparse as sp
    systems, B = _batch_case()
    X = [np.zeros_like(b) for b in B]
    it, amg = _batch_params(lib)
    r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
    assert r == 0
    rng = np.random.default_rng(3)
    for k, q in zip(rng.integers(0, len(systems), 5), rng.integers(0, 8, 5)):
        ia, ja, va = systems[k]
        A = capi.HostCSR(ia, ja, va)
        ito, amgo = oracle.params(tol=1e-10, maxit=500, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG,
                                  coarse_dof=200, theta=0.04)
        x = capi.HostVec(np.zeros(ia.size - 1))
        libc.srand(1)
        io = orc.orc_krylov_amg(A.ptr(), capi.HostVec(B[k][q]).ptr(), x.ptr(), C.byref(ito), C.byref(amgo),
                                C.byref(capi.orc_trace()))
        print(f"[batch] system {k} rhs {q}: batched {iters[k, q]} oracle {io}", flush=True)
        assert io > 0 and abs(int(iters[k, q]) - io) <= 2
        M = sp.csr_matrix((va, ja, ia))
        assert np.linalg.norm(B[k][q] - M @ X[k][q]) / np.linalg.norm(B[k][q]) < 1e-10
        assert np.linalg.norm(X[k][q] - x.arr) <= 1e-8 * np.linalg.norm(x.arr)


def test_batch_device_entry_matches_host_entry_and_repeats(lib):
    """The device form (stacked CSR, interleaved RHS) gives the host form's
    bits, and a second call (graph rebuilt) repeats them."""
    systems, B = _batch_case(tiny=False)
    X = [np.zeros_like(b) for b in B]
    it, amg = _batch_params(lib)
    r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
    assert r == 0
    off, ias, jas, vas, bs = [0], [], [], [], []
    nz = 0
    for (ia, ja, va), b in zip(systems, B):
        ias.append(ia[:-1] + nz)
        jas.append(ja + off[-1])
        vas.append(va)
        bs.append(b.T)
        nz += int(ia[-1])
        off.append(off[-1] + ia.size - 1)
    host = [np.ascontiguousarray(np.concatenate(ias + [np.array([nz])]).astype(np.int32)),
            np.ascontiguousarray(np.concatenate(jas).astype(np.int32)),
            np.ascontiguousarray(np.concatenate(vas)),
            np.ascontiguousarray(np.concatenate(bs))]
    hip = _hip_runtime()
    dev = []
    try:
        for h in host + [np.zeros_like(host[3])]:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), C.c_size_t(h.nbytes)) == 0
            dev.append(p)
            assert hip.hipMemcpy(p, h.ctypes.data_as(C.c_void_p), C.c_size_t(h.nbytes), 1) == 0
        outs = []
        for rep in range(2):
            zero = np.zeros_like(host[3])
            assert hip.hipMemcpy(dev[4], zero.ctypes.data_as(C.c_void_p), C.c_size_t(zero.nbytes), 1) == 0
            its = np.zeros(len(systems) * 8, np.int32)
            it2, amg2 = _batch_params(lib)
            r = lib.hz_batch_krylov_amg_device(len(systems), (capi.INT * len(off))(*off), nz, dev[0], dev[1],
                                               dev[2], dev[3], dev[4], C.byref(it2), C.byref(amg2),
                                               capi.dptr(its, capi.INT))
            assert r == 0
            x = np.zeros_like(host[3])
            assert hip.hipMemcpy(x.ctypes.data_as(C.c_void_p), dev[4], C.c_size_t(x.nbytes), 2) == 0
            outs.append((x, its))
    finally:
        for p in dev:
            hip.hipFree(p)
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    x = outs[0][0]
    for k in range(len(systems)):
        assert np.array_equal(x[off[k]:off[k + 1]].T, X[k])
    assert np.array_equal(outs[0][1].reshape(len(systems), 8), iters)


_SHAPES = (104, 108, 204, 208, 404, 408)


def test_batch_every_kernel_shape_gives_identical_bits(lib, monkeypatch):
    """Each shape the autotuner chooses from (1, 2 or 4 lanes per row, loads
    issued 4 or 8 entries at a time), forced on the FP32 cycle kernels and the
    FP64 SpMM alike, adds every row's entries in stored order: the same
    iterations and solution bits."""
    systems, B = _batch_case(tiny=False)
    ref = None
    for code in _SHAPES:
        monkeypatch.setenv("HZ_BATCH_SHAPE", str(code))
        monkeypatch.setenv("HZ_BATCH_SPMM_SHAPE", str(code))
        X = [np.zeros_like(b) for b in B]
        it, amg = _batch_params(lib)
        r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
        st = capi.hz_batch_stats()
        lib.hz_batch_get_stats(C.byref(st))
        print(f"[batch] shape {code}: max iterations {iters.max()}, shapes {st.sweep_shape0} / {st.spmm_shape0}",
              flush=True)
        assert r == 0 and st.sweep_shape0 == code and st.spmm_shape0 == code
        if ref is None:
            ref = (iters.copy(), [x.copy() for x in X])
        else:
            assert np.array_equal(iters, ref[0])
            assert all(np.array_equal(a, b) for a, b in zip(X, ref[1]))


def test_batch_tuner_cache_file_round_trip(lib, tmp_path):
    """The tuner times every shape once per (device, kernel, matrix class),
    appends its choices to the cache file, and a fresh process state read from
    that file (hz_batch_tune_cache) takes them without timing anything and
    reproduces the solution bits."""
    path = tmp_path / "cache" / "tune.txt"
    systems, B = _batch_case(tiny=False)
    runs = []
    try:
        assert lib.hz_batch_tune_cache(str(path).encode()) == 0
        for rep in range(2):
            X = [np.zeros_like(b) for b in B]
            it, amg = _batch_params(lib)
            r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
            st = capi.hz_batch_stats()
            lib.hz_batch_get_stats(C.byref(st))
            print(f"[batch] cache pass {rep}: timed {st.tuned} ({st.tune_ms:.1f} ms), cached {st.tune_cached}, "
                  f"level-0 shapes {st.sweep_shape0} / {st.spmm_shape0}", flush=True)
            assert r == 0
            runs.append((st.tuned, st.tune_cached, iters.copy(), X))
            if rep == 0:
                lines = path.read_text().splitlines()
                assert len(lines) == st.tuned > 0
                for ln in lines:
                    key, code = ln.split()
                    assert key.count("/") == 3 and int(code) in _SHAPES
                assert lib.hz_batch_tune_cache(str(path).encode()) == len(lines)
    finally:
        lib.hz_batch_tune_cache(None)
    assert runs[1][0] == 0 and runs[1][1] == runs[0][0] + runs[0][1]
    assert np.array_equal(runs[0][2], runs[1][2])
    assert all(np.array_equal(a, b) for a, b in zip(runs[0][3], runs[1][3]))


@pytest.mark.parametrize("smoother,permute", [("l1", False), ("jacobi", False), ("l1", True)])
def test_fp64_cycle_16bit_column_offsets_bitexact(lib, libc, monkeypatch, smoother, permute):
    """On levels of 65,536 rows or more the FP64 Jacobi / L1 cycle and the
    PCG's product read the columns as 16-bit offsets from the row, two to a
    32-bit word, where every offset fits: the 103,823-row Poisson solve is
    bit-identical (iterations and every solution entry) with them switched off
    (HZ_COL16=0).  A random symmetric permutation has offsets beyond the
    16-bit range, so its levels keep the int32 columns, with the same result."""
    import scipy.sparse as sp
    ia, ja, va = problems.poisson3d_kuhn(48)
    n = ia.size - 1
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    if permute:
        perm = np.random.default_rng(11).permutation(n)
        M = M[perm][:, perm].tocsr()
        M.sort_indices()
        ia, ja, va = M.indptr.astype(ia.dtype), M.indices.astype(ja.dtype), M.data.astype(va.dtype)
        rows = np.repeat(np.arange(n), np.diff(ia))
        assert np.abs(ja.astype(np.int64) - rows).max() > 32767
    b = problems.uniform_rhs(n, seed=4)
    A = capi.HostCSR(ia, ja, va)
    sm = capi.SMOOTHER_L1DIAG if smoother == "l1" else capi.SMOOTHER_JACOBI
    kw = dict(tol=1e-10, maxit=500, aggregation=capi.HEM, smoother=sm, coarse_dof=1000, theta=0.04)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("HZ_COL16", flag)
        x = capi.HostVec(np.zeros(n))
        it, amg = capi.solver_params(lib, **kw)
        libc.srand(1)
        r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg))
        rel = np.linalg.norm(b - M @ x.arr) / np.linalg.norm(b)
        print(f"[col16] HZ_COL16={flag} {smoother} permuted {permute}: {r} iterations, residual {rel:.2e}",
              flush=True)
        assert r > 0 and rel < 1e-10
        out[flag] = (r, x.arr.copy())
    assert out["0"][0] == out["1"][0]
    assert np.array_equal(out["0"][1], out["1"][1])
