"""Training and use of the per-aggregate parameter network.

Pipeline (all numerical work through the C ABI, on the GPU):
  1. seeded anisotropic P1 problems (2-D 257^2 and 3-D 64^3 grids, rotation
     theta ~ U[0, pi), anisotropy eps log-uniform in [1e-4, 1]);
  2. per-aggregate labels: every level-0 aggregate a gets the damping
     omega*_a that minimises the energy of its smoothed basis function
     (I - omega D_F^-1 A_F) 1_a in the A-norm (closed form per aggregate);
     a sweep then measures the cycle's error contraction for c * omega*_a
     over a grid of multipliers c and Chebyshev bound scales s
     (hz_cycle_rate_agg), and the best (c, s) labels every aggregate of the
     problem with (c * omega*_a, s) -- damping labels that vary per
     aggregate, chosen by measured convergence;
  3. features of the problem's level-0 aggregates (hz_aggregate_features);
  4. training of the MLP with the library's FP32 forward/backward kernels and
     Adam (hz_mlp_train);
  5. use: hz_krylov_amg_learned, which falls back to the plain method when the
     learned one misses its iteration budget or the FP64 residual check.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import capi, problems

MODEL_PATH = Path(__file__).resolve().parent / "models" / "sa_damping_cheb_mlp.npy"
OMEGAS = (0.0, 0.3, 0.5, 0.67, 0.85, 1.0, 1.2)
SCALES = (0.5, 0.7, 0.85, 1.0)


@dataclass
class Problem:
    ia: np.ndarray
    ja: np.ndarray
    va: np.ndarray
    dim: int
    m: int
    theta: float
    eps: float


def sample_problems(count: int, seed: int, sizes_2d=(33, 49, 65), sizes_3d=(12, 16), frac_3d=0.4):
    """Seeded anisotropic problems: theta ~ U[0, pi), eps = 10^U[-4, 0]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        th = float(rng.uniform(0.0, np.pi))
        eps = float(10.0 ** rng.uniform(-4.0, 0.0))
        if rng.random() < frac_3d:
            m = int(rng.choice(sizes_3d))
            ia, ja, va = problems.aniso3d_kuhn(m, th, eps)
            out.append(Problem(ia, ja, va, 3, m, th, eps))
        else:
            m = int(rng.choice(sizes_2d))
            ia, ja, va = problems.aniso2d_p1(m, th, eps)
            out.append(Problem(ia, ja, va, 2, m, th, eps))
    return out


def amg_params(lib, sweeps=2):
    """The learned method's setup: smoothed aggregation on VMB aggregates with
    the Chebyshev smoother (what hz_krylov_amg_learned runs)."""
    _, amg = capi.solver_params(lib, aggregation=capi.VMB, smoother=capi.SMOOTHER_CHEBY, coarse_dof=200,
                                theta=0.04)
    amg.AMG_type = capi.SA_AMG
    amg.presmooth_iter = amg.postsmooth_iter = sweeps
    return amg


C_GRID = (0.5, 0.75, 1.0, 1.25)
SCALE_GRID = (0.6, 0.8, 1.0)


def level0_aggregates(lib, libc, p: Problem):
    """(aggregate map, count) of level 0 under the learned method's setup."""
    A = capi.HostCSR(p.ia, p.ja, p.va)
    amg = amg_params(lib)
    amg.max_levels = 2
    h = capi.hz_hierarchy()
    libc.srand(1)
    assert lib.hz_setup_hierarchy(A.ptr(), C.byref(amg), C.byref(h)) == 0
    lv = capi.hierarchy_to_numpy(h)
    lib.hz_hierarchy_free(C.byref(h))
    return lv[0]["agg"].astype(np.int32), int(lv[0]["nagg"])


def energy_damping(lib, p: Problem, agg, nagg, theta=0.04, filt=1):
    """omega*_a = argmin_w ||(I - w D_F^-1 A_F) 1_a||_A for every aggregate a
    (D_F, A_F: the smoothed-aggregation filter of amg_setup_smoothP_smoothR):
    with T the tentative prolongation and U = D_F^-1 A_F T, omega*_a =
    (T^T A U)_aa / (U^T A U)_aa.  U is read off the library's own smoothed
    prolongation at omega = 1 (P = T - U)."""
    import scipy.sparse as sp
    n = p.ia.size - 1
    A = capi.HostCSR(p.ia, p.ja, p.va)
    P1 = capi.dCSRmat()
    assert lib.hz_sa_prolongation(A.ptr(), capi.dptr(agg, capi.INT), nagg, theta, 1.0, None, filt,
                                  C.byref(P1)) == 0
    pia, pja, pva = capi.csr_to_numpy(P1)
    lib.dcsr_free(C.byref(P1))
    P = sp.csr_matrix((pva, pja, pia), shape=(n, nagg))
    rows = np.nonzero(agg >= 0)[0]
    T = sp.csr_matrix((np.ones(rows.size), (rows, agg[rows])), shape=(n, nagg))
    U = (T - P).tocsr()
    Asp = sp.csr_matrix((p.va, p.ja, p.ia), shape=(n, n))
    AU = (Asp @ U).tocsr()
    num = np.asarray(T.multiply(AU).sum(axis=0)).ravel()
    den = np.asarray(U.multiply(AU).sum(axis=0)).ravel()
    w = np.where(den > 0, num / np.where(den > 0, den, 1.0), 0.67)
    return np.clip(w, 0.0, 1.6).astype(np.float32)


def label_per_aggregate(lib, libc, p: Problem, iters=24):
    """Per-aggregate labels of one problem: (omega[nagg], scale, rate grid,
    aggregate map, count)."""
    A = capi.HostCSR(p.ia, p.ja, p.va)
    agg, nagg = level0_aggregates(lib, libc, p)
    wstar = energy_damping(lib, p, agg, nagg)
    grid = np.ones((len(C_GRID), len(SCALE_GRID)))
    for a, c in enumerate(C_GRID):
        om = np.ascontiguousarray(np.clip(c * wstar, 0.0, 1.6).astype(np.float32))
        for s, sc in enumerate(SCALE_GRID):
            amg = amg_params(lib)
            r = capi.REAL()
            libc.srand(1)
            st = lib.hz_cycle_rate_agg(A.ptr(), C.byref(amg), capi.fptr(om), nagg, sc, iters, C.byref(r))
            grid[a, s] = r.value if st == 0 and np.isfinite(r.value) else 1.0
    a, s = np.unravel_index(np.argmin(grid), grid.shape)
    omega = np.clip(C_GRID[a] * wstar, 0.0, 1.6).astype(np.float32)
    return omega, SCALE_GRID[s], grid, agg, nagg


def label(lib, libc, p: Problem, iters=30):
    """(best omega, best scale, rate grid) by two-grid contraction."""
    A = capi.HostCSR(p.ia, p.ja, p.va)
    grid = np.zeros((len(OMEGAS), len(SCALES)))
    for a, om in enumerate(OMEGAS):
        for s, sc in enumerate(SCALES):
            amg = amg_params(lib)
            r = capi.REAL()
            libc.srand(1)
            st = lib.hz_two_grid_rate(A.ptr(), C.byref(amg), om, sc, iters, C.byref(r))
            grid[a, s] = r.value if st == 0 else 1.0
    a, s = np.unravel_index(np.argmin(grid), grid.shape)
    return OMEGAS[a], SCALES[s], grid


def level0_features(lib, libc, p: Problem):
    A = capi.HostCSR(p.ia, p.ja, p.va)
    amg = amg_params(lib)
    amg.max_levels = 2
    h = capi.hz_hierarchy()
    libc.srand(1)
    assert lib.hz_setup_hierarchy(A.ptr(), C.byref(amg), C.byref(h)) == 0
    lv = capi.hierarchy_to_numpy(h)
    lib.hz_hierarchy_free(C.byref(h))
    agg, nagg = lv[0]["agg"], lv[0]["nagg"]
    F = np.zeros((nagg, capi.MLP_INPUTS), np.float32)
    assert lib.hz_aggregate_features(A.ptr(), capi.dptr(agg, capi.INT), nagg, 0, capi.fptr(F)) == 0
    return F


def build_dataset_per_aggregate(lib, probs, per_problem=400, seed=11, log=print):
    """Features and per-aggregate labels of every problem (at most
    `per_problem` aggregates sampled from each), plus one metadata row per
    problem (dim, m, theta, eps, multiplier, scale, best rate, rate at
    multiplier 1 / scale 1, mean and spread of omega*)."""
    X, T, meta = [], [], []
    libc = C.CDLL(None)
    rng = np.random.default_rng(seed)
    for k, p in enumerate(probs):
        omega, sc, grid, agg, nagg = label_per_aggregate(lib, libc, p)
        A = capi.HostCSR(p.ia, p.ja, p.va)
        F = np.zeros((nagg, capi.MLP_INPUTS), np.float32)
        assert lib.hz_aggregate_features(A.ptr(), capi.dptr(agg, capi.INT), nagg, 0, capi.fptr(F)) == 0
        pick = rng.choice(nagg, size=min(per_problem, nagg), replace=False)
        X.append(F[pick])
        T.append(np.stack([omega[pick], np.full(pick.size, sc, np.float32)], axis=1))
        a, s = np.unravel_index(np.argmin(grid), grid.shape)
        meta.append((p.dim, p.m, p.theta, p.eps, C_GRID[a], sc, float(grid.min()), float(grid[2, 2]),
                     float(omega.mean()), float(omega.std())))
        if log:
            log(f"[learn] problem {k}: dim {p.dim} m {p.m} theta {p.theta:.3f} eps {p.eps:.2e} -> "
                f"multiplier {C_GRID[a]} scale {sc} rate {grid.min():.3f} (1.0/1.0: {grid[2, 2]:.3f}), "
                f"omega mean {omega.mean():.3f} std {omega.std():.3f}")
    return np.concatenate(X), np.concatenate(T), meta


def build_dataset(lib, probs, log=print):
    X, T, meta = [], [], []
    libc = C.CDLL(None)
    for k, p in enumerate(probs):
        om, sc, grid = label(lib, libc, p)
        F = level0_features(lib, libc, p)
        X.append(F)
        T.append(np.tile(np.array([om, sc], np.float32), (F.shape[0], 1)))
        meta.append((p.dim, p.m, p.theta, p.eps, om, sc, float(grid.min()), float(grid[3, 3])))
        if log:
            log(f"[learn] problem {k}: dim {p.dim} m {p.m} theta {p.theta:.3f} eps {p.eps:.2e} -> "
                f"omega {om} scale {sc} rate {grid.min():.3f} (0.67/1.0: {grid[3, 3]:.3f})")
    return np.concatenate(X), np.concatenate(T), meta


def train(lib, X, T, epochs=60, batch=512, lr=2e-3, seed=7):
    theta = np.zeros(capi.MLP_PARAMS, np.float32)
    lib.hz_mlp_init(capi.fptr(theta), seed)
    hist = np.zeros(epochs + 1, np.float32)
    X = np.ascontiguousarray(X, np.float32)
    T = np.ascontiguousarray(T, np.float32)
    assert lib.hz_mlp_train(capi.fptr(theta), capi.fptr(X), capi.fptr(T), X.shape[0], epochs, batch, lr, seed,
                            capi.fptr(hist)) == 0
    return theta, hist


def load_model(path: Path = MODEL_PATH) -> np.ndarray:
    theta = np.load(path, allow_pickle=False).astype(np.float32)
    assert theta.shape == (capi.MLP_PARAMS,)
    return np.ascontiguousarray(theta)


def solve_learned(lib, ia, ja, va, b, theta, tol=1e-10, maxit=1000, budget=400, x0=None, plain_kw=None,
                  cheb_steps=2):
    """hz_krylov_amg_learned: SA + Chebyshev(cheb_steps) with learned
    parameters; the plain method (UA, HEM, L1, V(1,1) unless plain_kw says
    otherwise) is the fallback.  Returns (iterations, x, info)."""
    A = capi.HostCSR(ia, ja, va)
    x = capi.HostVec(np.zeros(b.size) if x0 is None else x0.copy())
    kw = dict(tol=tol, maxit=maxit, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200,
              theta=0.04)
    kw.update(plain_kw or {})
    it, amg = capi.solver_params(lib, **kw)
    r = lib.hz_krylov_amg_learned(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                  capi.fptr(theta), budget, cheb_steps)
    info = capi.hz_solve_info()
    lib.hz_get_last_info(C.byref(info))
    return r, x.arr.copy(), info


def main():
    import argparse
    import time
    ap = argparse.ArgumentParser(description="label, train and save the parameter network")
    ap.add_argument("--problems", type=int, default=120)
    ap.add_argument("--epochs", type=int, default=80)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--sizes-2d", type=int, nargs="+", default=[256], help="cells per side of the 2-D problems")
    ap.add_argument("--sizes-3d", type=int, nargs="+", default=[64], help="cells per side of the 3-D problems")
    ap.add_argument("--frac-3d", type=float, default=0.25)
    ap.add_argument("--per-problem", type=int, default=400, help="aggregates sampled per problem")
    ap.add_argument("--time-limit", type=float, default=0.0,
                    help="stop labelling new problems after this many seconds (0: no limit)")
    ap.add_argument("--out", type=Path, default=MODEL_PATH)
    ap.add_argument("--meta", type=Path, default=None, help="where to save the per-problem labelling record")
    a = ap.parse_args()
    lib = capi.load_library()
    t0 = time.time()
    rng = np.random.default_rng(a.seed)
    X, T, meta = [], [], []
    for k in range(a.problems):
        if a.time_limit and time.time() - t0 > a.time_limit:
            print(f"[learn] time limit: {k} problems labelled", flush=True)
            break
        p = sample_problems(1, int(rng.integers(1 << 31)), sizes_2d=tuple(a.sizes_2d), sizes_3d=tuple(a.sizes_3d),
                            frac_3d=a.frac_3d)[0]
        x, t, m = build_dataset_per_aggregate(lib, [p], per_problem=a.per_problem, seed=k,
                                              log=lambda s, k=k: print(s.replace("problem 0", f"problem {k}"),
                                                                       flush=True))
        X.append(x)
        T.append(t)
        meta.extend(m)
    X, T = np.concatenate(X), np.concatenate(T)
    print(f"[learn] {X.shape[0]} labelled aggregates from {len(meta)} problems in {time.time() - t0:.1f}s",
          flush=True)
    theta, hist = train(lib, X, T, epochs=a.epochs)
    print(f"[learn] training loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    a.out.parent.mkdir(parents=True, exist_ok=True)
    np.save(a.out, theta)
    if a.meta:
        # one row per labelled problem: dim, m, theta, eps, multiplier, scale, best rate,
        # rate at multiplier 1 / scale 1, mean and spread of the damping labels
        np.save(a.meta, np.array(meta, dtype=np.float64))
    print(f"[learn] saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
