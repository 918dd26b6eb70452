"""The learned-parameter method (hz_krylov_amg_learned) against an oracle
composition, the estimate behind its Chebyshev floor, and its mixed-precision
form.

Composition: the oracle runs the whole method -- VMB aggregation, smoothed
prolongation with per-aggregate damping (the reference's
amg_setup_smoothP_smoothR, amg_setup_ua.c:1461-1687, with omega per
aggregate), the general Galerkin product, the Chebyshev smoother and PCG --
while the damping of every aggregate comes from the library's feature kernel
and network forward pass (each pinned against the oracle on its own:
test_aggregate_features_vs_oracle, test_mlp_forward_and_gradient_vs_oracle),
and the Chebyshev bound scale of every level is the one the library reports.
The scales are checked against their own rule: the network's mean output,
floored at min(1, 1.1 lambda_est / Gershgorin), lambda_est the Lanczos
estimate, which is checked against closed forms and dense eigenvalues.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from hazmath_amd import capi, learned, problems

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[learned] {msg}", flush=True)


def test_lanczos_estimate_closed_form_2d_poisson(lib):
    """D^-1 A of the 5-point Laplacian on an m x m grid of cells (Dirichlet):
    lambda_max = 1 + cos(pi / m); Gershgorin bound 2.  The Ritz value is a
    lower bound; on the dense top of a large grid's spectrum 30 steps leave
    it up to about 1 % low (measured 0.5 % at m = 64), which the floor's
    factor 1.1 covers."""
    for m in (16, 64, 128, 256):
        ia, ja, va = problems.poisson2d_p1(m)
        A = capi.HostCSR(ia, ja, va)
        lam, ger = capi.REAL(), capi.REAL()
        assert lib.hz_lambda_max_dinv(A.ptr(), 30, C.byref(lam), C.byref(ger)) == 0
        exact = 1.0 + np.cos(np.pi / m)
        _say(f"2-D m={m}: lanczos {lam.value:.10f} exact {exact:.10f} gershgorin {ger.value}")
        assert ger.value == 2.0
        assert lam.value <= exact * (1 + 1e-12)
        assert lam.value * 1.1 >= exact * 1.05
        if m <= 16:
            assert lam.value >= exact * (1 - 1e-6)   # the Krylov space nearly spans the spectrum


def _dense_lambda_dinv(ia, ja, va):
    n = ia.size - 1
    A = sp.csr_matrix((va, ja, ia), shape=(n, n)).toarray()
    d = np.diag(A).copy()
    s = 1.0 / np.sqrt(d)
    return float(np.linalg.eigvalsh(A * s[:, None] * s[None, :]).max())


@pytest.mark.parametrize("eps", [1.0, 1e-3])
def test_lanczos_estimate_vs_dense_eigenvalues(lib, eps):
    ia, ja, va = problems.aniso3d_kuhn(12, 0.9, eps)
    A = capi.HostCSR(ia, ja, va)
    lam, ger = capi.REAL(), capi.REAL()
    assert lib.hz_lambda_max_dinv(A.ptr(), 30, C.byref(lam), C.byref(ger)) == 0
    exact = _dense_lambda_dinv(ia, ja, va)
    _say(f"3-D aniso eps={eps}: lanczos {lam.value:.10f} dense {exact:.10f} gershgorin {ger.value:.6f}")
    assert lam.value <= exact * (1 + 1e-10)
    assert lam.value >= exact * (1 - 2e-3)
    assert ger.value >= exact * (1 - 1e-12)


def _constant_network(omega, scale):
    """All weights zero, final bias (omega, scale): outputs exactly (omega, scale)."""
    th = np.zeros(capi.MLP_PARAMS, np.float32)
    off_b3 = 16 * 64 + 64 + 64 * 64 + 64 + 64 * 16
    th[off_b3] = omega
    th[off_b3 + 1] = scale
    return th


def test_learned_scale_floor_binds(lib, libc):
    """A network that asks for bound scale 0.5 everywhere: each level's scale
    is raised to min(1, 1.1 lambda_est / Gershgorin) (the library reports both
    per level), the solve needs no fallback and meets the tolerance, and the
    level-0 estimate is within 0.2 % of the dense eigenvalue.  Where the floor
    reaches 1 the estimate may stop early: it is then a Ritz value of fewer
    steps, at most the 30-step one and at least 1/1.1 of the bound."""
    ia, ja, va = problems.aniso3d_kuhn(12, 0.7, 1e-3)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=1)
    th = _constant_network(0.67, 0.5)
    libc.srand(1)
    r, x, info = learned.solve_learned(lib, ia, ja, va, b, th, budget=400)
    assert info.fallback_used == 0 and r > 0
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    assert np.linalg.norm(b - M @ x) / np.linalg.norm(b) < 1e-10
    bound = False
    for l in range(info.num_levels - 1):
        ratio, scale = info.lambda_ratio[l], info.cheb_scale[l]
        expect = min(1.5, max(0.5, min(1.0, 1.1 * ratio)))
        _say(f"level {l}: lambda ratio {ratio:.6f}, scale {scale:.6f} (expected {expect:.6f})")
        assert 0.0 < ratio <= 1.0 + 1e-12
        assert scale == pytest.approx(expect, rel=1e-12)
        bound = bound or scale > 0.5
    assert bound
    lam, ger = capi.REAL(), capi.REAL()
    assert lib.hz_lambda_max_dinv(capi.HostCSR(ia, ja, va).ptr(), 30, C.byref(lam), C.byref(ger)) == 0
    exact = _dense_lambda_dinv(ia, ja, va) / ger.value
    full = lam.value / ger.value
    _say(f"level 0: 30-step ratio {full:.6f}, in the solve {info.lambda_ratio[0]:.6f}, dense {exact:.6f}")
    assert abs(full - exact) <= 2e-3 * exact
    if info.cheb_scale[0] == 1.0:
        assert 1.0 / 1.1 * (1 - 1e-12) <= info.lambda_ratio[0] <= full * (1 + 1e-12)
    else:
        assert info.lambda_ratio[0] == pytest.approx(full, rel=1e-12)


def _set_precision(lib, p, cheb_coarse=None):
    """Set hz_options.vcycle_precision (and cheb_coarse_steps when given);
    returns the previous (precision, cheb_coarse_steps)."""
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    old = (opts.vcycle_precision, opts.cheb_coarse_steps)
    opts.vcycle_precision = p
    if cheb_coarse is not None:
        opts.cheb_coarse_steps = cheb_coarse
    lib.hz_set_options(C.byref(opts))
    return old


def _restore(lib, old):
    _set_precision(lib, old[0], old[1])


def _oracle_composition(lib, orc, libc, ia, ja, va, b, theta, scales, cheb_steps=2, cheb_coarse=0):
    """orc_krylov_amg with the learned method's settings (SA on VMB, Chebyshev
    pre/post steps, the drop-in's other parameters) and the per-aggregate
    damping of the library's feature kernel + network on each oracle level."""
    omegas = {}

    def omega_fn(level, Ap, aggp, nagg, out):
        feat = np.zeros((nagg, capi.MLP_INPUTS), np.float32)
        assert lib.hz_aggregate_features(Ap, aggp, nagg, level, capi.fptr(feat)) == 0
        y = np.zeros((nagg, capi.MLP_OUTPUTS), np.float32)
        assert lib.hz_mlp_forward(capi.fptr(theta), capi.fptr(feat), nagg, capi.fptr(y)) == 0
        om = np.minimum(np.maximum(y[:, 0], np.float32(0.0)), np.float32(1.6)).astype(np.float32)
        omegas[level] = (om, float(np.clip(y[:, 1].astype(np.float64).mean(), 0.2, 1.5)))
        C.memmove(out, om.ctypes.data, 4 * nagg)

    cb = oracle.ORC_OMEGA_FN(omega_fn)
    sc = (capi.REAL * capi.MAX_AMG_LVL)(*scales)
    orc.orc_set_learned(cb, sc)
    orc.orc_set_cheb_coarse_steps(cheb_coarse)
    try:
        it, amg = oracle.params(tol=1e-10, maxit=1000, aggregation=capi.VMB, smoother=91, coarse_dof=200,
                                theta=0.04)
        amg.AMG_type = capi.SA_AMG
        amg.presmooth_iter = amg.postsmooth_iter = cheb_steps
        x = capi.HostVec(np.zeros(b.size))
        libc.srand(1)
        r = orc.orc_krylov_amg(capi.HostCSR(ia, ja, va).ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                               C.byref(amg), C.byref(capi.orc_trace()))
    finally:
        orc.orc_set_learned(C.cast(None, oracle.ORC_OMEGA_FN), None)
        orc.orc_set_cheb_coarse_steps(0)
    return int(r), x.arr.copy(), omegas


@pytest.mark.parametrize("m,th,steps,coarse", [(16, 0.3, 2, 0), (20, 2.2, 2, 0), (20, 2.2, 4, 2)])
@pytest.mark.parametrize("precision,bar", [(0, 1e-10), (1, 1e-8)])
def test_learned_entry_vs_oracle_composition(lib, orc, libc, m, th, steps, coarse, precision, bar):
    """hz_krylov_amg_learned (SA + Chebyshev(steps), committed network) on a
    3-D anisotropic problem (eps = 1e-3) against the oracle composition: PCG
    iterations within one, solution within 1e-10 with the FP64 cycle and 1e-8
    with the FP32 cycle; every level's scale follows the floor rule.  With
    hz_options.cheb_coarse_steps the levels below the finest take that many
    steps (the oracle's orc_set_cheb_coarse_steps)."""
    theta = learned.load_model()
    ia, ja, va = problems.aniso3d_kuhn(m, th, 1e-3)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=1)
    old = _set_precision(lib, precision, coarse)
    try:
        libc.srand(1)
        rg, xg, info = learned.solve_learned(lib, ia, ja, va, b, theta, budget=400, cheb_steps=steps)
    finally:
        _restore(lib, old)
    assert info.fallback_used == 0 and rg > 0
    nl = info.num_levels
    scales = [info.cheb_scale[l] if l < nl - 1 else 1.0 for l in range(capi.MAX_AMG_LVL)]
    ro, xo, omegas = _oracle_composition(lib, orc, libc, ia, ja, va, b, theta, scales, cheb_steps=steps,
                                         cheb_coarse=coarse)
    rel = np.linalg.norm(xg - xo) / np.linalg.norm(xo)
    _say(f"m={m} precision {precision}: GPU {rg} iterations, oracle composition {ro}, distance {rel:.2e}")
    assert ro > 0 and abs(rg - ro) <= 1
    assert rel < bar
    assert sorted(omegas) == list(range(nl - 1))
    for l in range(nl - 1):
        net = omegas[l][1]
        expect = min(1.5, max(net, min(1.0, 1.1 * info.lambda_ratio[l])))
        assert info.cheb_scale[l] == pytest.approx(expect, rel=1e-9), (l, net, info.lambda_ratio[l])


def test_learned_fp32_cycle_reaches_fp64_tolerance(lib, libc):
    """The learned mixed-precision method: the FP32 cycle over the SA levels
    and the Chebyshev smoother, under the FP64 PCG, reaches 1e-10 in the FP64
    true residual without fallback, within a few iterations of the FP64
    cycle; BF16 coarse values likewise."""
    theta = learned.load_model()
    ia, ja, va = problems.aniso3d_kuhn(32, 1.1, 1e-3)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=1)
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    res = {}
    for p in (0, 1, 2):
        old = _set_precision(lib, p)
        try:
            libc.srand(1)
            r, x, info = learned.solve_learned(lib, ia, ja, va, b, theta, budget=400)
        finally:
            _restore(lib, old)
        rel = np.linalg.norm(b - M @ x) / np.linalg.norm(b)
        _say(f"precision {p}: {r} iterations, fallback {info.fallback_used}, residual {rel:.2e}, "
             f"{info.total_ms:.1f} ms")
        assert info.fallback_used == 0 and r > 0
        assert rel < 1e-10
        res[p] = (r, x)
    assert res[1][0] <= res[0][0] + 3 and res[2][0] <= res[0][0] + 6
    x64 = res[0][1]
    assert np.linalg.norm(res[1][1] - x64) / np.linalg.norm(x64) < 1e-8


@pytest.mark.parametrize("precision,permute", [(1, False), (2, False), (1, True)])
def test_fp32_cycle_16bit_column_offsets_bitexact(lib, libc, monkeypatch, precision, permute):
    """The FP32 cycle's SELL Chebyshev steps and residuals read 16-bit column
    offsets from the row where every stored entry's fits (padding -32768): the
    learned mixed-precision solve (FP32 values, and BF16 below level 0) is
    bit-identical with them switched off (HZ_COL16=0).  A random symmetric
    permutation of a 59,319-row problem has offsets beyond the 16-bit range,
    so its levels keep the int32 columns, with the same result."""
    theta = learned.load_model()
    ia, ja, va = problems.aniso3d_kuhn(40 if permute else 32, 1.1, 1e-3)
    n = ia.size - 1
    M = sp.csr_matrix((va, ja, ia), shape=(n, n))
    if permute:
        perm = np.random.default_rng(7).permutation(n)
        M = M[perm][:, perm].tocsr()
        M.sort_indices()
        ia, ja, va = M.indptr.astype(ia.dtype), M.indices.astype(ja.dtype), M.data.astype(va.dtype)
        rows = np.repeat(np.arange(n), np.diff(ia))
        assert np.abs(ja.astype(np.int64) - rows).max() > 32767
    b = problems.uniform_rhs(n, seed=1)
    out = {}
    for flag in ("0", "1"):
        monkeypatch.setenv("HZ_COL16", flag)
        old = _set_precision(lib, precision, cheb_coarse=2)
        try:
            libc.srand(1)
            r, x, info = learned.solve_learned(lib, ia, ja, va, b, theta, budget=400, cheb_steps=4)
        finally:
            _restore(lib, old)
        rel = np.linalg.norm(b - M @ x) / np.linalg.norm(b)
        _say(f"HZ_COL16={flag} precision {precision} permuted {permute}: {r} iterations, residual {rel:.2e}")
        assert info.fallback_used == 0 and rel < 1e-10
        out[flag] = (r, x.copy())
    assert out["0"][0] == out["1"][0]
    assert np.array_equal(out["0"][1], out["1"][1])
