"""The workloads' own configurations at their full sizes, on the GPU, against
the CPU oracle (the serial restatement of the reference path).

1. 2-D P1 Poisson, 129 x 129 nodes (n = 16,129, 5-point), seeded RHS,
   pairwise (HEM) UA-AMG, V(1,1) Jacobi w = 2/3, PCG to 1e-8 from x0 = 0.
2. 3-D Kuhn P1 Poisson, 128^3 cells (n = 2,048,383, 30,340,171 nonzeros):
   every level's aggregates and CSR bit-exact, PCG to 1e-10 within one
   iteration, solution within 1e-10.
3. The same mesh with kappa piecewise constant on 8^3 blocks (log-uniform in
   [1e-3, 1e3]): FP32 and BF16-valued cycles under the FP64 PCG reach 1e-10
   in the FP64 true residual; their distance to the FP64 oracle's solution is
   printed and bounded.
4. 256 anisotropic 2-D P1 systems x 8 right-hand sides in one batched call:
   every one of the 2,048 pairs meets 1e-10 in the FP64 true residual, and a
   subset that holds the largest and the most anisotropic systems matches
   the oracle.
5. The learned method on 3-D Kuhn P1 anisotropic (eps = 1e-3) at 128^3,
   with the FP64 and the FP32 cycle: no fallback, 1e-10 in the FP64 true
   residual (its oracle composition runs at smaller sizes in
   test_gpu_learned.py).
"""
import ctypes as C
import time

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from hazmath_amd import capi, problems

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[configs] {msg}", flush=True)


def _oracle_solve(orc, libc, A, b, **kw):
    it, amg = oracle.params(**kw)
    x = capi.HostVec(np.zeros(b.size))
    libc.srand(1)
    t0 = time.perf_counter()
    r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                           C.byref(capi.orc_trace()))
    return int(r), x.arr.copy(), time.perf_counter() - t0


def _gpu_solve(lib, libc, A, b, **kw):
    it, amg = capi.solver_params(lib, **kw)
    x = capi.HostVec(np.zeros(b.size))
    libc.srand(1)
    r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg))
    return int(r), x.arr.copy()


def _set_options(lib, **kv):
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = {k: getattr(opts, k) for k in kv}
    for k, v in kv.items():
        setattr(opts, k, v)
    lib.hz_set_options(C.byref(opts))
    return old


def test_config1_poisson2d_129_jacobi_two_thirds(lib, orc, libc):
    ia, ja, va = problems.poisson2d_p1(128)
    n = ia.size - 1
    assert n == 16129 and int(ia[-1]) == 80137
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=1)
    kw = dict(tol=1e-8, maxit=500, aggregation=capi.HEM, smoother=capi.SMOOTHER_JACOBI, coarse_dof=100,
              coarse_solver=32, max_levels=20, theta=0.0)
    old = _set_options(lib, jacobi_weight=2.0 / 3.0)
    orc.orc_set_jacobi_weight(2.0 / 3.0)
    try:
        ig, xg = _gpu_solve(lib, libc, A, b, **kw)
        io, xo, _ = _oracle_solve(orc, libc, A, b, **kw)
    finally:
        _set_options(lib, **old)
        orc.orc_set_jacobi_weight(1.0)
    _say(f"config 1: GPU {ig} iterations, oracle {io}")
    assert io > 0 and abs(ig - io) <= 1
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    assert np.linalg.norm(b - M @ xg) / np.linalg.norm(b) < 1e-8


_KUHN_KW = dict(aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200, coarse_solver=32,
                max_levels=20, theta=0.04)


def test_config2_kuhn128_hierarchy_bitexact_and_pcg(lib, orc, libc):
    ia, ja, va = problems.poisson3d_kuhn(128)
    n = ia.size - 1
    assert n == 2048383 and int(ia[-1]) == 30340171
    A = capi.HostCSR(ia, ja, va)
    levels = {}
    thetas = {}
    for which in ("gpu", "oracle"):
        _say(f"config 2: {which} setup")
        it, amg = (capi.solver_params(lib, **_KUHN_KW) if which == "gpu" else oracle.params(**_KUHN_KW))
        h = capi.hz_hierarchy()
        libc.srand(1)
        st = (lib.hz_setup_hierarchy if which == "gpu" else orc.orc_setup_hierarchy)(A.ptr(), C.byref(amg),
                                                                                   C.byref(h))
        assert st == 0
        levels[which] = capi.hierarchy_to_numpy(h)
        thetas[which] = amg.strong_coupled
        (lib.hz_hierarchy_free if which == "gpu" else orc.orc_hierarchy_free)(C.byref(h))
    g, o = levels["gpu"], levels["oracle"]
    _say(f"config 2: {len(g)} GPU levels, {len(o)} oracle levels, rows {[l_['ia'].size - 1 for l_ in g]}")
    assert len(g) == len(o) == 15
    for lv, (gl, ol) in enumerate(zip(g, o)):
        if "agg" in gl:
            assert np.array_equal(gl["agg"], ol["agg"]), f"aggregates of level {lv}"
        for key in ("ia", "ja", "val"):
            assert np.array_equal(gl[key], ol[key]), f"{key} of level {lv}"
    assert thetas["gpu"] == thetas["oracle"]
    del levels, g, o
    b = problems.uniform_rhs(n, seed=1)
    # coarse_dof 200 (param_amg_init's default) and 1000 (the benchmark's: the
    # same hierarchy cut three levels earlier, a dense coarsest inverse of 992 rows)
    for coarse_dof in (200, 1000):
        kw = dict(tol=1e-10, maxit=1000, **dict(_KUHN_KW, coarse_dof=coarse_dof))
        _say(f"config 2: GPU solve, coarse_dof {coarse_dof}")
        ig, xg = _gpu_solve(lib, libc, A, b, **kw)
        _say(f"config 2: oracle solve, coarse_dof {coarse_dof}")
        io, xo, t = _oracle_solve(orc, libc, A, b, **kw)
        _say(f"config 2: coarse_dof {coarse_dof}: GPU {ig} iterations, oracle {io} ({t:.1f} s)")
        assert io > 0 and abs(ig - io) <= 1
        assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) < 1e-10


@pytest.mark.parametrize("precision,bound", [(1, 1e-8), (2, 1e-8)])
def test_config3_kuhn128_coefficient_mixed_precision(lib, orc, libc, precision, bound):
    ia, ja, va, _ = problems.poisson3d_kuhn_coef(128)
    n = ia.size - 1
    A = capi.HostCSR(ia, ja, va)
    b = problems.uniform_rhs(n, seed=1)
    kw = dict(tol=1e-10, maxit=1000, **_KUHN_KW)
    _say(f"config 3: FP64 GPU solve (precision {precision} run follows)")
    i64, x64 = _gpu_solve(lib, libc, A, b, **kw)
    old = _set_options(lib, vcycle_precision=precision)
    try:
        _say(f"config 3: precision {precision} GPU solve")
        im, xm = _gpu_solve(lib, libc, A, b, **kw)
    finally:
        _set_options(lib, **old)
    _say("config 3: oracle solve")
    io, xo, t = _oracle_solve(orc, libc, A, b, **kw)
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    res = np.linalg.norm(b - M @ xm) / np.linalg.norm(b)
    err = np.linalg.norm(xm - xo) / np.linalg.norm(xo)
    _say(f"config 3: precision {precision}: {im} iterations (FP64 GPU {i64}, oracle {io}, {t:.1f} s); "
         f"FP64 true residual {res:.3e}; distance to the FP64 oracle's solution {err:.3e}")
    assert io > 0 and abs(i64 - io) <= 1
    assert 0 < im <= io + 4
    assert res < 1e-10
    assert err < bound


def test_config4_full_batch_every_pair_and_oracle_subset(lib, orc, libc):
    specs = problems.batch_anisotropic_2d(256)
    _say("config 4: generating 256 systems")
    systems, B = [], []
    for d in specs:
        ia, ja, va = problems.aniso2d_p1(d["side"] - 1, d["theta"], d["eps"])
        systems.append((ia, ja, va))
        B.append(problems.batch_rhs(ia.size - 1, d["index"]))
    X = [np.zeros_like(b_) for b_ in B]
    it, amg = capi.solver_params(lib, tol=1e-10, maxit=1000, aggregation=capi.HEM,
                                 smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200, coarse_solver=32,
                                 max_levels=20, theta=0.04)
    _say("config 4: batched solve")
    r, iters = capi.batch_solve(lib, systems, B, X, it, amg)
    assert r == 0
    assert iters.shape == (256, 8) and iters.min() > 0
    worst = 0.0
    for k, (ia, ja, va) in enumerate(systems):
        M = sp.csr_matrix((va, ja, ia))
        R = B[k] - (M @ X[k].T).T
        rel = np.linalg.norm(R, axis=1) / np.linalg.norm(B[k], axis=1)
        worst = max(worst, float(rel.max()))
    _say(f"config 4: worst FP64 true relative residual over 2048 pairs {worst:.3e}, "
         f"iterations max {iters.max()} mean {iters.mean():.1f}")
    assert worst < 1e-10
    sides = np.array([d["side"] for d in specs])
    eps = np.array([d["eps"] for d in specs])
    picks = [int(np.argmax(sides)), int(np.argmin(eps)), 17, 200]
    assert sides[picks[0]] >= 450 and eps[picks[1]] < 2e-4
    for k in picks:
        ia, ja, va = systems[k]
        q = k % 8
        A = capi.HostCSR(ia, ja, va)
        _say(f"config 4: oracle, system {k} (side {sides[k]}, eps {eps[k]:.2e}) rhs {q}")
        io, xo, t = _oracle_solve(orc, libc, A, B[k][q].copy(), tol=1e-10, maxit=1000, aggregation=capi.HEM,
                                  smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200, coarse_solver=32, max_levels=20,
                                  theta=0.04)
        err = np.linalg.norm(X[k][q] - xo) / np.linalg.norm(xo)
        _say(f"config 4: system {k}: batched {iters[k, q]} oracle {io} ({t:.1f} s), distance {err:.2e}")
        assert io > 0 and abs(int(iters[k, q]) - io) <= max(2, int(0.05 * io))
        assert err < 1e-8


@pytest.mark.parametrize("precision", [0, 1])
def test_config5_learned_kuhn128_anisotropic(lib, libc, precision):
    """5. The learned method's evaluation problem: 3-D Kuhn P1 anisotropic,
    eps = 1e-3, seeded rotation, 128^3 cells: the learned SA + Chebyshev(4)
    method with the committed network (FP64 cycle, and the FP32 cycle under
    the FP64 PCG) converges without falling back and meets 1e-10 in the FP64
    true residual; the plain drop-in runs beside it for the record."""
    from hazmath_amd import learned
    theta_dir = float(np.random.default_rng(5).uniform(0.0, np.pi))
    ia, ja, va = problems.aniso3d_kuhn(128, theta_dir, 1e-3)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=1)
    weights = learned.load_model()
    old = _set_options(lib, vcycle_precision=precision)
    try:
        _say(f"config 5: learned solve, precision {precision}")
        libc.srand(1)
        r, x, info = learned.solve_learned(lib, ia, ja, va, b, weights, budget=400, cheb_steps=4)
    finally:
        _set_options(lib, **old)
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    res = np.linalg.norm(b - M @ x) / np.linalg.norm(b)
    scales = [round(info.cheb_scale[l], 4) for l in range(info.num_levels - 1)]
    _say(f"config 5: precision {precision}: {r} iterations, fallback {info.fallback_used}, "
         f"{info.total_ms:.1f} ms, FP64 true residual {res:.3e}, scales {scales}")
    assert info.fallback_used == 0 and 0 < r <= 400
    assert res < 1e-10
