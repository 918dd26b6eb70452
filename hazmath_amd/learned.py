"""Training and use of the per-aggregate parameter network.

Pipeline (all numerical work through the C ABI, on the GPU):
  1. seeded anisotropic P1 problems (2-D and 3-D, rotation theta ~ U[0, pi),
     anisotropy eps log-uniform in [1e-4, 1]);
  2. labels: for each problem the (prolongator damping, Chebyshev bound scale)
     pair of a grid with the smallest two-grid contraction
     (hz_two_grid_rate: smoothed aggregation + Chebyshev, exact coarse solve);
  3. features of the problem's level-0 aggregates (hz_aggregate_features), each
     labelled with its problem's pair;
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
    ap.add_argument("--out", type=Path, default=MODEL_PATH)
    a = ap.parse_args()
    lib = capi.load_library()
    t0 = time.time()
    probs = sample_problems(a.problems, a.seed)
    X, T, meta = build_dataset(lib, probs, log=lambda s: print(s, flush=True))
    print(f"[learn] {X.shape[0]} labelled aggregates from {len(probs)} problems in {time.time() - t0:.1f}s",
          flush=True)
    theta, hist = train(lib, X, T, epochs=a.epochs)
    print(f"[learn] training loss {hist[0]:.4f} -> {hist[-1]:.4f}", flush=True)
    a.out.parent.mkdir(parents=True, exist_ok=True)
    np.save(a.out, theta)
    np.save(a.out.with_name(a.out.stem + "_labels.npy"), np.array(meta, dtype=np.float64))
    print(f"[learn] saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
