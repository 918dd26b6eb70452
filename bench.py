#!/usr/bin/env python3
"""Benchmark: UA-AMG-preconditioned CG time to 1e-10 on the 3-D Kuhn P1 Poisson
problem with 128^3 cells (n = 2,048,383, nnz = 30,340,171, FP64), one MI355X.

A step is one call of the device-resident drop-in (hz_krylov_amg_device): AMG
setup on the GPU plus PCG to relative residual 1e-10 from x0 = 0, inputs already
resident in HBM.  Prints ONE JSON line (rank 0).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

from hazmath_amd import capi, problems  # noqa: E402

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E, 8 TB/s
COPY_CAL_DOUBLES = 8 << 20
CAL_COPIES = 3
SPMV_KERNEL = "k_spmv<true, true, 1>"


def host_cpu_model():
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return os.uname().machine


def _counter_rows(outdir):
    """(dispatch id, kernel name, counter name, value, grid size) rows of every
    *counter_collection.csv that rocprofv3 wrote under outdir."""
    import csv
    rows = []
    for f in sorted(Path(outdir).rglob("*counter_collection.csv")):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                try:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], r["Counter_Name"],
                                 float(r["Counter_Value"]), int(r.get("Grid_Size", 0) or 0)))
                except (KeyError, ValueError):
                    continue
    return rows


# fine-level kernels whose HBM traffic the counter passes report: the graded
# SpMV, and the V-cycle's fine-level residual and last sweep (largest grid)
TRAFFIC_KERNELS = {"spmv": r"\bk_spmv<true", "residual": r"\bk_residual<", "sweep_dot": r"\bk_sweep<10, false, true"}


def measure_traffic(a):
    """HBM bytes per fine-level PCG SpMV launch, measured now: two rocprofv3
    passes (FETCH_SIZE, then WRITE_SIZE) over a child run of this script in
    --traffic-probe mode (one solve, then CAL_COPIES device copies of 64 MiB as
    its last dispatches).  Each counter is scaled by what it reads for the copy
    (known bytes read and written).  Runs before this process touches the GPU.
    Returns (dict, None) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None, "rocprofv3 not found"
    per = {}
    cal = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            out = Path(td) / counter
            cmd = [prof, "--pmc", counter, "-f", "csv", "-d", str(out), "-o", "c", "--",
                   sys.executable, str(ROOT / "bench.py"), "--traffic-probe", "--m", str(a.m),
                   "--agg", a.agg, "--smoother", a.smoother, "--theta", str(a.theta),
                   "--coarse-dof", str(a.coarse_dof), "--tol", str(a.tol), "--problem", a.problem,
                   "--precision", a.precision, "--jacobi-weight", str(a.jacobi_weight)]
            env = dict(os.environ, TMPDIR="/tmp")
            try:
                r = subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL,
                                   stderr=subprocess.PIPE, timeout=240, start_new_session=True)
            except subprocess.TimeoutExpired:
                return None, f"rocprofv3 --pmc {counter} timed out"
            rows = _counter_rows(out)
            if not rows:
                tail = r.stderr.decode(errors="replace")[-300:] if r.stderr else ""
                return None, f"rocprofv3 --pmc {counter} wrote no counters (rc {r.returncode}): {tail}"
            vals = {}
            names = {}
            grids = {}
            for d, k, c, v, g in rows:
                if c == counter:
                    vals[d] = vals.get(d, 0.0) + v
                    names[d] = k
                    grids[d] = g
            ids = sorted(vals)
            cal_ids = ids[-CAL_COPIES:]
            if len(cal_ids) < CAL_COPIES or not all("opy" in names[d] for d in cal_ids):
                seen = sorted({names[d][:60] for d in ids})[:8]
                return None, (f"{counter}: calibration copies not found "
                              f"({len(ids)} dispatches; last {[names[d][:40] for d in cal_ids]}; seen {seen})")
            cal[counter] = float(np.mean([vals[d] for d in cal_ids]))
            for key, pat in TRAFFIC_KERNELS.items():
                # demangled names: "void hz::(anonymous namespace)::k_spmv<true>(int, ...)"
                ds = [d for d in ids if re.search(pat, names[d])]
                if not ds:
                    return None, f"{counter}: no {key} dispatches"
                gmax = max(grids[d] for d in ds)
                sel = [vals[d] for d in ds if grids[d] == gmax]
                per.setdefault(key, {})[counter] = (float(np.mean(sel)), len(sel))
    cal_bytes = COPY_CAL_DOUBLES * 8
    f_fetch = cal_bytes / (cal["FETCH_SIZE"] * 1024.0)
    f_write = cal_bytes / (cal["WRITE_SIZE"] * 1024.0)
    out = {"copy_fetch_kib": cal["FETCH_SIZE"], "copy_write_kib": cal["WRITE_SIZE"], "fetch_factor": f_fetch,
           "write_factor": f_write}
    for key, d in per.items():
        out[key] = {"traffic_bytes": d["FETCH_SIZE"][0] * 1024.0 * f_fetch + d["WRITE_SIZE"][0] * 1024.0 * f_write,
                    "fetch_kib": d["FETCH_SIZE"][0], "write_kib": d["WRITE_SIZE"][0],
                    "dispatches": d["FETCH_SIZE"][1]}
    out["traffic_bytes"] = out["spmv"]["traffic_bytes"]
    return out, None

AGG = {"hem": capi.HEM, "hec": capi.HEC, "vmb": capi.VMB}
SMOOTHER = {"jacobi": capi.SMOOTHER_JACOBI, "l1": capi.SMOOTHER_L1DIAG, "gs": capi.SMOOTHER_GS,
            "sgs": capi.SMOOTHER_SGS, "sor": capi.SMOOTHER_SOR, "ssor": capi.SMOOTHER_SSOR,
            "gsor": capi.SMOOTHER_GSOR, "sgsor": capi.SMOOTHER_SGSOR}
PRECISION = {"fp64": 0, "fp32": 1, "bf16": 2}
DTYPE = {"fp64": "f64", "fp32": "f32 cycle, f64 PCG", "bf16": "bf16 values/f32 cycle, f64 PCG"}


def make_problem(a, m):
    if a.problem == "coef":
        ia, ja, va, _ = problems.poisson3d_kuhn_coef(m)
    else:
        ia, ja, va = problems.poisson3d_kuhn(m)
    return ia, ja, va


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--m", type=int, default=128, help="cells per direction")
    p.add_argument("--agg", choices=AGG, default="hem")
    p.add_argument("--smoother", choices=SMOOTHER, default="l1")
    p.add_argument("--jacobi-weight", type=float, default=1.0)
    p.add_argument("--relaxation", type=float, default=1.0, help="SOR-type relaxation weight")
    p.add_argument("--coarse-dof", type=int, default=1000,
                   help="rows at or below which the coarsest level is solved by its dense inverse")
    p.add_argument("--theta", type=float, default=0.04)
    p.add_argument("--tol", type=float, default=1e-10)
    p.add_argument("--kernel-reps", type=int, default=50)
    p.add_argument("--problem", choices=("poisson", "coef"), default="poisson",
                   help="poisson: -Laplace; coef: kappa piecewise constant on 8^3 blocks, "
                        "log-uniform in [1e-3, 1e3]")
    p.add_argument("--precision", choices=PRECISION, default="fp64",
                   help="arithmetic of the AMG cycle (PCG and coarsest solve stay FP64)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--batched", action="store_true",
                   help="the batched workload instead: 256 anisotropic 2-D P1 systems x 8 RHS")
    p.add_argument("--batch-count", type=int, default=256)
    p.add_argument("--batch-cpu-pairs", type=int, default=24,
                   help="(system, RHS) pairs solved by the serial oracle for the CPU baseline")
    p.add_argument("--batch-cpu-procs", type=int, default=8)
    p.add_argument("--learned", action="store_true",
                   help="the learned-parameter evaluation: 3-D Kuhn 128^3 anisotropic (eps=1e-3, seeded theta), "
                        "learned vs plain AMG-PCG")
    p.add_argument("--learned-kinds", default="plain,sa,learned,learned_fp32",
                   help="methods of --learned, comma separated (all four print the full line)")
    p.add_argument("--learned-problem", choices=("aniso", "poisson"), default="aniso",
                   help="--learned on the anisotropic problem (eps=1e-3) or on the Poisson problem")
    p.add_argument("--cheb-steps", type=int, default=4,
                   help="Chebyshev steps per pre- and post-smoothing of the SA carrier's finest level (constant "
                        "and learned); measured on the 128^3 anisotropic problem, learned FP32 (every level the "
                        "same): 2 -> 151.7 ms, 3 -> 141.2, 4 -> 139.2, 5 -> 140.1, 6 -> 143.7, 8 -> 155.1")
    p.add_argument("--cheb-coarse-steps", type=int, default=2,
                   help="Chebyshev steps per smoothing on the SA carrier's coarser levels (hz_options."
                        "cheb_coarse_steps; 0: as --cheb-steps); 128^3 anisotropic, learned FP32 with "
                        "--cheb-steps 4: 4 -> 96 iterations, PCG 78.3 ms; 3 -> 96, 73.6; 2 -> 96, 68.8; 1 -> 110, 73.2")
    p.add_argument("--cpu-sample-m", type=int, default=128,
                   help="cells per direction of the CPU-baseline sample")
    p.add_argument("--cpu-runs", type=int, default=5, help="CPU-baseline solves (median reported)")
    p.add_argument("--no-learned", action="store_true",
                   help="skip the learned-vs-plain comparison on the anisotropic problem in the default line")
    p.add_argument("--no-traffic", action="store_true",
                   help="skip the two rocprofv3 counter passes that measure the SpMV's HBM traffic")
    p.add_argument("--traffic-probe", action="store_true",
                   help="(internal) one solve, then the counter calibration copies, nothing printed")
    return p.parse_args()


def set_params(lib, a):
    it, amg = capi.solver_params(lib, tol=a.tol, maxit=1000, aggregation=AGG[a.agg],
                                 smoother=SMOOTHER[a.smoother], coarse_dof=a.coarse_dof,
                                 coarse_solver=capi.SOLVER_UMFPACK, max_levels=20, theta=a.theta)
    amg.relaxation = a.relaxation
    return it, amg


def cpu_baseline(a, ia, ja, va, b):
    """Serial C restatement of the reference path (oracle/liboracle.so), 1 core."""
    import oracle
    orc = oracle.load()
    orc.orc_set_jacobi_weight(a.jacobi_weight)
    libc = C.CDLL(None)
    it, amg = oracle.params(tol=a.tol, maxit=1000, aggregation=AGG[a.agg],
                            smoother=SMOOTHER[a.smoother], coarse_dof=a.coarse_dof,
                            coarse_solver=32, max_levels=20, theta=a.theta)
    amg.relaxation = a.relaxation
    A = capi.HostCSR(ia, ja, va)
    x = capi.HostVec(np.zeros(b.size))
    bb = capi.HostVec(b)
    tr = capi.orc_trace()
    libc.srand(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    sys.stdout.flush()
    os.dup2(devnull, 1)
    try:
        t0 = time.perf_counter()
        r = orc.orc_krylov_amg(A.ptr(), bb.ptr(), x.ptr(), C.byref(it), C.byref(amg), C.byref(tr))
        t = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
    return t, r, tr, x.arr


def _batch_pair_worker(job):
    """Serial oracle solve of one (system, RHS) pair; returns (seconds, iterations)."""
    side, theta, eps, index, q, tol = job
    import oracle
    orc = oracle.load()
    libc = C.CDLL(None)
    ia, ja, va = problems.aniso2d_p1(side - 1, theta, eps)
    b = problems.batch_rhs(ia.size - 1, index)[q]
    it, amg = oracle.params(tol=tol, maxit=1000, aggregation=capi.HEM, smoother=capi.SMOOTHER_L1DIAG,
                            coarse_dof=200, coarse_solver=32, max_levels=20, theta=0.04)
    A = capi.HostCSR(ia, ja, va)
    x = capi.HostVec(np.zeros(b.size))
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)
    sys.stdout.flush()
    os.dup2(devnull, 1)
    try:
        libc.srand(1)
        t0 = time.perf_counter()
        r = orc.orc_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                               C.byref(capi.orc_trace()))
        t = time.perf_counter() - t0
    finally:
        os.dup2(saved, 1)
        os.close(devnull)
        os.close(saved)
    return t, int(r), int(ia.size - 1)


def main_batched(a, lib):
    """Batched workload: 256 independent anisotropic 2-D P1 systems (sides
    log-U[33,513], theta ~ U[0,pi), eps = 10^U[-4,0]), 8 seeded RHS each, one
    call of hz_batch_krylov_amg_device (stacked CSR and RHS resident in HBM)."""
    import torch
    specs = problems.batch_anisotropic_2d(a.batch_count)
    t_gen = time.perf_counter()
    ias, jas, vas, off, bs = [], [], [], [0], []
    nnz_off = 0
    for d in specs:
        ia, ja, va = problems.aniso2d_p1(d["side"] - 1, d["theta"], d["eps"])
        n_s = ia.size - 1
        ias.append(ia[:-1].astype(np.int64) + nnz_off)
        jas.append(ja.astype(np.int64) + off[-1])
        vas.append(va)
        bs.append(problems.batch_rhs(n_s, d["index"]).T)
        nnz_off += int(ia[-1])
        off.append(off[-1] + n_s)
    ia = np.concatenate(ias + [np.array([nnz_off])]).astype(np.int32)
    ja = np.concatenate(jas).astype(np.int32)
    va = np.concatenate(vas)
    b = np.ascontiguousarray(np.concatenate(bs))          # [n][8]
    n, nnz = int(off[-1]), int(nnz_off)
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] batched: {len(specs)} systems, n={n}, nnz={nnz}, generated in {t_gen:.1f}s",
          file=sys.stderr, flush=True)
    dev = torch.device("cuda:0")
    dia, dja, dva = (torch.from_numpy(v).to(dev) for v in (ia, ja, va))
    db = torch.from_numpy(b).to(dev)
    dx = torch.zeros_like(db)
    offs = (capi.INT * len(off))(*off)
    nsys = len(specs)
    iters = np.zeros(nsys * capi.BATCH_RHS, np.int32)
    libc = C.CDLL(None)

    def step():
        dx.zero_()
        torch.cuda.synchronize()
        it, amg = capi.solver_params(lib, tol=a.tol, maxit=1000, aggregation=capi.HEM,
                                     smoother=capi.SMOOTHER_L1DIAG, coarse_dof=200,
                                     coarse_solver=capi.SOLVER_UMFPACK, max_levels=20, theta=0.04)
        libc.srand(1)
        r = lib.hz_batch_krylov_amg_device(nsys, offs, nnz, dia.data_ptr(), dja.data_ptr(), dva.data_ptr(),
                                           db.data_ptr(), dx.data_ptr(), C.byref(it), C.byref(amg),
                                           capi.dptr(iters, capi.INT))
        if r < 0:
            raise RuntimeError(f"batched solver returned {r}")

    tune_first = None
    for w in range(max(1, a.warmup)):
        step()
        if tune_first is None:
            st = capi.hz_batch_stats()
            lib.hz_batch_get_stats(C.byref(st))
            tune_first = (st.tuned, st.tune_ms)
        print(f"[bench] warmup {w}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(a.steps):
        step()
        st = capi.hz_batch_stats()
        lib.hz_batch_get_stats(C.byref(st))
        stats.append((st.setup_ms, st.solve_ms, st.graph_launches))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / a.steps
    pairs = nsys * capi.BATCH_RHS
    print(f"[bench] {a.steps} steps, {ms_per_step:.1f} ms/step", file=sys.stderr, flush=True)
    # true FP64 residuals of every pair (CPU)
    import scipy.sparse as sp
    x = dx.cpu().numpy()
    A = sp.csr_matrix((va, ja, ia), shape=(n, n))
    R = b - A @ x
    worst = 0.0
    for s_ in range(nsys):
        rs = np.linalg.norm(R[off[s_]:off[s_ + 1]], axis=0) / np.linalg.norm(b[off[s_]:off[s_ + 1]], axis=0)
        worst = max(worst, float(rs.max()))
    # level-0 kernels timed live (HIP events on the library's stream)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    opts.batch_time_kernels = 1
    lib.hz_set_options(C.byref(opts))
    step()
    opts.batch_time_kernels = 0
    lib.hz_set_options(C.byref(opts))
    ks = capi.hz_batch_stats()
    lib.hz_batch_get_stats(C.byref(ks))
    # algorithmic bytes of y = A x with 8 FP64 RHS: 12 nnz + 4 (n+1) + 64 n (x) + 64 n (y)
    bytes_spmm = 12 * nnz + 4 * (n + 1) + 128 * n
    spmm_gbs = bytes_spmm / (ks.spmm_ms * 1e-3) / 1e9
    # FP32 sweep with 8 RHS: 8 nnz (FP32 value + col) + 32 n each for x, b, xout + n (live)
    bytes_sweep = 8 * nnz + 96 * n + n
    sweep_gbs = bytes_sweep / (ks.sweep_ms * 1e-3) / 1e9
    value = float(np.mean([s_[0] + s_[1] for s_ in stats])) * 1e-3
    cpu = None
    if not a.no_cpu_baseline:
        import multiprocessing as mp
        rng = np.random.default_rng(11)
        pick = rng.choice(pairs, size=min(a.batch_cpu_pairs, pairs), replace=False)
        jobs = [(specs[p // 8]["side"], specs[p // 8]["theta"], specs[p // 8]["eps"], specs[p // 8]["index"],
                 int(p % 8), a.tol) for p in pick]
        procs = max(1, min(a.batch_cpu_procs, os.cpu_count() or 1))
        print(f"[bench] cpu baseline: {len(jobs)} pairs on {procs} processes ...", file=sys.stderr, flush=True)
        tc = time.perf_counter()
        # spawned workers: a forked child would inherit this process's GPU state
        with mp.get_context("spawn").Pool(procs) as pool:
            res = pool.map(_batch_pair_worker, jobs, chunksize=1)
        tc = time.perf_counter() - tc
        rows_done = sum(r_[2] for r_ in res)
        # throughput in pairs/s, scaled by the sample's share of the batch's rows
        mean_rows = n / nsys
        cpu_pairs_per_s = len(jobs) / tc * (rows_done / len(jobs)) / mean_rows
        gi = {int(p): r_[1] for p, r_ in zip(pick, res)}
        cpu = {"value": pairs / cpu_pairs_per_s, "unit": "s", "cores": procs, "kind": "port",
               "sample": (f"{len(jobs)} seeded (system, RHS) pairs of the batch solved one by one by the serial "
                          f"C restatement (FP64 cycle) on {procs} processes in {tc:.1f}s; value = time for all "
                          f"{pairs} pairs at that rate, scaled by rows"),
               "pairs_per_s": cpu_pairs_per_s,
               "iterations_vs_gpu": [[int(p // 8), int(p % 8), gi[int(p)], int(iters[int(p)])] for p in pick[:8]],
               "host_nproc": os.cpu_count()}
    line = {
        "metric": "batched AMG-PCG time-to-1e-10 for 256 systems x 8 RHS (s)",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cycle, f64 PCG",
        "data": "synthetic",
        "config": {
            "workload": (f"{nsys} 2-D P1 anisotropic diffusion systems (sides log-U[33,513], theta~U[0,pi), "
                         f"eps=10^U[-4,0], seed 4), 8 seeded RHS each, n={n}, nnz={nnz}; UA-AMG HEM theta=0.04 "
                         f"V(1,1) L1, dense coarsest <=200, FP32 cycle / FP64 PCG to {a.tol:g}, one PCG "
                         f"iteration per hipGraph replay"),
        },
        "pairs": pairs,
        "pairs_per_s": pairs / value,
        "setup_ms": float(np.mean([s_[0] for s_ in stats])),
        "solve_ms": float(np.mean([s_[1] for s_ in stats])),
        "graph_launches": stats[-1][2],
        "iterations_max": int(iters.max()),
        "iterations_mean": float(iters.mean()),
        "worst_true_rel_residual": worst,
        "roofline": {
            "bound": "hbm",
            "kernel": "FP64 SpMM y=Ax, 8 interleaved RHS, level 0 (PCG)",
            "achieved": spmm_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmm_gbs / HBM_PEAK_GBS,
            "traffic": None,
            "algorithmic_bytes_per_launch": bytes_spmm,
            "avg_launch_ms": ks.spmm_ms,
        },
        "kernels": {"spmm_ms": ks.spmm_ms, "sweep_f32_ms": ks.sweep_ms, "sweep_f32_gbs": sweep_gbs,
                    "shape_sweep_level0": ks.sweep_shape0, "shape_spmm_level0": ks.spmm_shape0,
                    "shape_code": "100 x lanes per row + entries per load round", "levels": ks.num_levels,
                    "op_complexity": ks.op_complexity},
        "autotuner": {"first_call_timed_decisions": tune_first[0], "first_call_tune_ms": tune_first[1],
                      "timed_call_cached_decisions": ks.tune_cached, "timed_call_timed_decisions": ks.tuned,
                      "cache_file": os.environ.get("HZ_TUNE_CACHE", "default (~/.cache/hazmath_mi355x/tune.txt)")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


LEARNED_EPS = 1e-3


def learned_comparison(a, lib, steps, warmup, kinds=("plain", "sa", "learned", "learned_fp32")):
    """Learned per-aggregate SA parameters against the plain drop-in on the
    3-D Kuhn P1 anisotropic problem (K = R_z(theta) diag(1, eps, eps) R_z^T,
    eps = 1e-3, theta = U[0, pi) seed 5), m^3 cells, seeded RHS, tol 1e-10.
    Times are device times of the whole call (setup + PCG; for the learned
    method the attempt plus, if it fell back, the plain solve), inputs
    uploaded before the timed region."""
    from hazmath_amd import learned
    theta_dir = float(np.random.default_rng(5).uniform(0.0, np.pi))
    poisson = getattr(a, "learned_problem", "aniso") == "poisson"
    cheb_steps = getattr(a, "cheb_steps", 4)
    cheb_coarse = getattr(a, "cheb_coarse_steps", 2)
    if poisson:
        ia, ja, va = problems.poisson3d_kuhn(a.m)
    else:
        ia, ja, va = problems.aniso3d_kuhn(a.m, theta_dir, LEARNED_EPS)
    n = ia.size - 1
    b = problems.uniform_rhs(n, seed=1)
    weights = learned.load_model()
    A = capi.HostCSR(ia, ja, va)
    libc = C.CDLL(None)
    import scipy.sparse as sp
    Asp = sp.csr_matrix((va, ja, ia), shape=(n, n))
    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)

    def quiet(fn):
        sys.stdout.flush()
        os.dup2(devnull, 1)
        try:
            return fn()
        finally:
            os.dup2(saved, 1)

    def run(kind):
        x = capi.HostVec(np.zeros(n))
        it, amg = capi.solver_params(lib, tol=a.tol, maxit=1000, aggregation=capi.HEM,
                                     smoother=capi.SMOOTHER_L1DIAG, coarse_dof=a.coarse_dof,
                                     coarse_solver=capi.SOLVER_UMFPACK, max_levels=20, theta=0.04)
        opts = capi.hz_options()
        lib.hz_get_options(C.byref(opts))
        old = opts.vcycle_precision
        old_coarse = opts.cheb_coarse_steps
        opts.vcycle_precision = 1 if kind == "learned_fp32" else 0
        # the SA carrier's levels below the finest: Chebyshev(cheb_coarse) (plain: no Chebyshev)
        opts.cheb_coarse_steps = cheb_coarse
        lib.hz_set_options(C.byref(opts))
        libc.srand(1)
        try:
            if kind == "plain":
                r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                                                      C.byref(amg))
            elif kind == "sa":
                amg.AMG_type = capi.SA_AMG
                amg.aggregation_type = capi.VMB
                amg.smoother = capi.SMOOTHER_CHEBY
                amg.presmooth_iter = amg.postsmooth_iter = cheb_steps
                r = lib.linear_solver_dcsr_krylov_amg(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it),
                                                      C.byref(amg))
            else:
                r = lib.hz_krylov_amg_learned(A.ptr(), capi.HostVec(b).ptr(), x.ptr(), C.byref(it), C.byref(amg),
                                              capi.fptr(weights), 400, cheb_steps)
        finally:
            opts.vcycle_precision = old
            opts.cheb_coarse_steps = old_coarse
            lib.hz_set_options(C.byref(opts))
        info = capi.hz_solve_info()
        lib.hz_get_last_info(C.byref(info))
        rel = float(np.linalg.norm(b - Asp @ x.arr) / np.linalg.norm(b))
        out = dict(iterations=int(r), setup_ms=info.setup_ms, solve_ms=info.solve_ms, total_ms=info.total_ms,
                   true_rel_residual=rel)
        if kind.startswith("learned"):
            out["fallback"] = int(info.fallback_used)
            out["learned_attempt"] = dict(iterations=int(info.learned_iterations), total_ms=info.learned_total_ms)
            out["cheb_scale"] = [round(info.cheb_scale[l], 4) for l in range(max(0, info.num_levels - 1))]
            out["lambda_ratio"] = [round(info.lambda_ratio[l], 4) for l in range(max(0, info.num_levels - 1))]
            if info.fallback_used:
                out["total_ms"] = info.learned_total_ms + info.total_ms
        return out

    res = {}
    for kind in kinds:
        for _ in range(warmup):
            quiet(lambda: run(kind))
        runs = [quiet(lambda: run(kind)) for _ in range(steps)]
        best = dict(runs[-1])
        for key in ("total_ms", "setup_ms", "solve_ms"):
            best[key] = float(np.mean([r_[key] for r_ in runs]))
        res[kind] = best
        print(f"[bench] {kind}: {best}", file=sys.stderr, flush=True)
    os.close(devnull)
    os.close(saved)
    problem = (f"3-D Kuhn P1 Poisson, {a.m}^3 cells" if poisson else
               f"3-D Kuhn P1 -div(K grad u), K = R_z(theta) diag(1, eps, eps) R_z^T, eps={LEARNED_EPS}, "
               f"theta={theta_dir:.4f}, {a.m}^3 cells")
    workload = (f"{problem}, n={n}, seeded RHS, PCG to {a.tol:g}; "
                f"coarse_dof={a.coarse_dof}; plain = UA HEM L1 V(1,1) FP64 drop-in; sa = SA on VMB with "
                f"Chebyshev({cheb_steps}) on the finest level and Chebyshev({cheb_coarse or cheb_steps}) below, constant "
                "parameters; learned = the same with per-aggregate damping and bound scale from "
                "the committed network, plain method as fallback; learned_fp32 = learned with the FP32 cycle "
                "under the FP64 PCG")
    return res, workload


def main_learned(a, lib):
    kinds = tuple(a.learned_kinds.split(","))
    res, workload = learned_comparison(a, lib, a.steps, a.warmup, kinds=kinds)
    if len(kinds) < 4:
        print(json.dumps({"partial_learned_comparison": res}))
        return
    line = {
        "metric": ("learned vs plain AMG-PCG time-to-1e-10 (s), 3-D Kuhn P1 "
                   + ("Poisson" if a.learned_problem == "poisson" else "anisotropic eps=1e-3")),
        "value": res["learned_fp32"]["total_ms"] * 1e-3,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res["learned_fp32"]["total_ms"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 cycle, f64 PCG (learned network f32)",
        "data": "synthetic",
        "config": {"workload": workload},
        "plain": res["plain"], "sa_constant": res["sa"], "learned": res["learned"],
        "learned_fp32": res["learned_fp32"],
        "speedup_learned_vs_plain": res["plain"]["total_ms"] / res["learned"]["total_ms"],
        "speedup_learned_fp32_vs_plain": res["plain"]["total_ms"] / res["learned_fp32"]["total_ms"],
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {a.gpus} is not supported "
              "(the single-GPU throughput is the reported value)", file=sys.stderr)
        sys.exit(2)
    if a.batched:
        import torch
        lib = capi.load_library()
        if lib.hz_device_count() < 1 or not torch.cuda.is_available():
            print("bench.py: no HIP device", file=sys.stderr)
            sys.exit(1)
        return main_batched(a, lib)
    if a.learned:
        lib = capi.load_library()
        if lib.hz_device_count() < 1:
            print("bench.py: no HIP device", file=sys.stderr)
            sys.exit(1)
        return main_learned(a, lib)
    traffic, traffic_note = None, "not measured (--no-traffic)"
    if not a.traffic_probe and not a.no_traffic:
        # before this process initialises the GPU: the counter passes run in children
        t_tr = time.perf_counter()
        traffic, traffic_note = measure_traffic(a)
        print(f"[bench] traffic passes {time.perf_counter() - t_tr:.1f}s: "
              f"{traffic['traffic_bytes'] / 1e6 if traffic else traffic_note}", file=sys.stderr, flush=True)
    import torch
    lib = capi.load_library()
    if lib.hz_device_count() < 1 or not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    opts = capi.hz_options()
    lib.hz_get_options(C.byref(opts))
    opts.jacobi_weight = a.jacobi_weight
    opts.return_on_fatal = 1
    opts.vcycle_precision = PRECISION[a.precision]
    lib.hz_set_options(C.byref(opts))

    t_gen = time.perf_counter()
    ia, ja, va = make_problem(a, a.m)
    n, nnz = ia.size - 1, int(ia[-1])
    b = problems.uniform_rhs(n, seed=1)
    t_gen = time.perf_counter() - t_gen
    print(f"[bench] generated n={n} nnz={nnz} in {t_gen:.1f}s", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    dia = torch.from_numpy(ia).to(dev)
    dja = torch.from_numpy(ja).to(dev)
    dva = torch.from_numpy(va).to(dev)
    db = torch.from_numpy(b).to(dev)
    dx = torch.zeros(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    libc = C.CDLL(None)

    def step():
        dx.zero_()
        torch.cuda.synchronize()
        it, amg = set_params(lib, a)
        libc.srand(1)
        r = lib.hz_krylov_amg_device(n, nnz, dia.data_ptr(), dja.data_ptr(), dva.data_ptr(),
                                     db.data_ptr(), dx.data_ptr(), C.byref(it), C.byref(amg))
        if r < 0:
            raise RuntimeError(f"solver returned {r}")
        return r

    devnull = os.open(os.devnull, os.O_WRONLY)
    saved = os.dup(1)

    def quiet(fn):
        sys.stdout.flush()
        os.dup2(devnull, 1)
        try:
            return fn()
        finally:
            os.dup2(saved, 1)

    if a.traffic_probe:
        # one solve, then the calibration copies as the run's last dispatches
        quiet(step)
        cal_src = torch.ones(COPY_CAL_DOUBLES, dtype=torch.float64, device=dev)
        cal_dst = torch.empty_like(cal_src)
        torch.cuda.synchronize()
        for _ in range(CAL_COPIES):
            cal_dst.copy_(cal_src)
        torch.cuda.synchronize()
        return

    for w in range(a.warmup):
        r = quiet(step)
        print(f"[bench] warmup {w}: iterations {r}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for s in range(a.steps):
        r = quiet(step)
        info = capi.hz_solve_info()
        lib.hz_get_last_info(C.byref(info))
        infos.append((r, info.setup_ms, info.solve_ms, info.total_ms, info.num_levels,
                      info.final_relres, list(info.level_rows[:info.num_levels]),
                      list(info.level_nnz[:info.num_levels]), info.solve_alg_bytes))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / a.steps
    print(f"[bench] {a.steps} steps, {ms_per_step:.2f} ms/step", file=sys.stderr, flush=True)

    # true relative residual of the last solution (FP64, on the CPU)
    x = dx.cpu().numpy()
    import scipy.sparse as sp
    Asp = sp.csr_matrix((va, ja, ia), shape=(n, n))
    true_rel = float(np.linalg.norm(b - Asp @ x) / np.linalg.norm(b))

    # The fine-level SpMV inside the solve: one more (untimed) solve with HIP
    # events around each PCG SpMV launch on the library's stream.
    opts.time_pcg_spmv = 1
    lib.hz_set_options(C.byref(opts))
    quiet(step)
    opts.time_pcg_spmv = 0
    lib.hz_set_options(C.byref(opts))
    pinfo = capi.hz_solve_info()
    lib.hz_get_last_info(C.byref(pinfo))
    pcg_spmv_ms, pcg_spmv_count = pinfo.pcg_spmv_ms, pinfo.pcg_spmv_count
    res_ms, sweep_ms = pinfo.cycle_residual_ms, pinfo.cycle_sweep_ms
    # fine-level kernels back to back (the matrix partly stays in the 256 MB
    # last-level cache between repetitions, so these read faster than in a solve)
    kt = capi.hz_kernel_times()
    if a.kernel_reps > 0:   # 0: skip (profiling runs that should see only the solve's launches)
        rc = lib.hz_time_fine_kernels(n, nnz, dia.data_ptr(), dja.data_ptr(), dva.data_ptr(),
                                      a.kernel_reps, C.byref(kt))
        if rc != 0:
            raise RuntimeError("hz_time_fine_kernels failed")
    # algorithmic bytes: 12 nnz (val + col) + 4 (n+1) (row pointers) + 8 n (x) + 8 n (y).
    # The solve's fine-level kernels of an FP64 Jacobi / L1 cycle read the columns
    # as 16-bit offsets from the row, two to a 32-bit word (10 bytes an entry), on
    # levels of 65,536 rows or more whose offsets all fit (the library's rule;
    # HZ_COL16=0 keeps the int32 columns); the back-to-back kernels read int32.
    rows = np.repeat(np.arange(n, dtype=np.int64), np.diff(ia))
    paired = (a.precision == "fp64" and a.smoother in ("l1", "jacobi") and n >= 65536
              and os.environ.get("HZ_COL16", "1") != "0"
              and int(np.abs(ja.astype(np.int64) - rows).max()) <= 32767)
    eb = 10 if paired else 12
    bytes_spmv = eb * nnz + 4 * (n + 1) + 16 * n
    bytes_jacobi = 12 * nnz + 4 * (n + 1) + 24 * n
    # residual w = b - A x: matrix, x and b read, w written; last sweep: matrix, b (= r) and x
    # read, x' written ((x', r) needs no other array)
    bytes_res = eb * nnz + 4 * (n + 1) + 24 * n
    bytes_sweep_dot = eb * nnz + 4 * (n + 1) + 24 * n
    spmv_gbs = bytes_spmv / (pcg_spmv_ms * 1e-3) / 1e9
    jac_gbs = bytes_jacobi / (kt.jacobi_ms * 1e-3) / 1e9 if kt.jacobi_ms > 0 else None

    iters = [i[0] for i in infos]
    value = float(np.mean([i[3] for i in infos])) * 1e-3   # device time per solve (s)
    lvp = None
    if not a.no_learned and a.problem == "poisson" and a.precision == "fp64":
        # the metric's second half: learned vs plain on the anisotropic 128^3 problem
        res, workload = learned_comparison(a, lib, steps=3, warmup=1, kinds=("plain", "learned", "learned_fp32"))
        lvp = {"workload": workload, "plain": res["plain"], "learned": res["learned"],
               "learned_fp32": res["learned_fp32"],
               "speedup_learned_vs_plain": res["plain"]["total_ms"] / res["learned"]["total_ms"],
               "speedup_learned_fp32_vs_plain": res["plain"]["total_ms"] / res["learned_fp32"]["total_ms"]}
    cpu = None
    if not a.no_cpu_baseline:
        if a.cpu_sample_m == a.m:
            ci, cj, cv, cb = ia, ja, va, b
        else:
            ci, cj, cv = make_problem(a, a.cpu_sample_m)
            cb = problems.uniform_rhs(ci.size - 1, seed=1)
        runs = []
        for k in range(max(1, a.cpu_runs)):
            print(f"[bench] cpu baseline run {k} ...", file=sys.stderr, flush=True)
            t_cpu, r_cpu, tr, x_cpu = cpu_baseline(a, ci, cj, cv, cb)
            runs.append((t_cpu, tr.setup_s, tr.solve_s))
            print(f"[bench] cpu baseline {t_cpu:.2f}s, iterations {r_cpu}", file=sys.stderr, flush=True)
        t_med = float(np.median([r_[0] for r_ in runs]))
        sample = (f"median of {len(runs)} full solves (setup + PCG to {a.tol:g}) of the same workload with "
                  f"{a.cpu_sample_m}^3 cells, serial C restatement of the reference path, "
                  f"{r_cpu} iterations each")
        scale = 1.0 if a.cpu_sample_m == a.m else (a.m / a.cpu_sample_m) ** 3
        cpu = {"value": t_med * scale, "unit": "s", "cores": 1, "kind": "port", "sample": sample,
               "iterations": int(r_cpu), "runs_s": [r_[0] for r_ in runs],
               "setup_s": float(np.median([r_[1] for r_ in runs])),
               "solve_s": float(np.median([r_[2] for r_ in runs])),
               "host_cpu": host_cpu_model(), "host_nproc": os.cpu_count()}
        if a.cpu_sample_m == a.m:
            # error of the GPU solution against the FP64 oracle's
            cpu["x_rel_diff_vs_gpu"] = float(np.linalg.norm(x_cpu - x) / np.linalg.norm(x_cpu))

    line = {
        "metric": "AMG-PCG time-to-1e-10 rel. residual (s) + SpMV GB/s, 128³ P1; learned vs plain",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[a.precision],
        "data": "synthetic",
        "config": {
            "workload": (f"3-D Kuhn P1 {'Poisson' if a.problem == 'poisson' else 'diffusion, kappa on 8^3 blocks log-U[1e-3,1e3]'}, "
                         f"{a.m}^3 cells, n={n}, nnz={nnz} (15-point rows, "
                         f"zeros stored), seeded U[-1,1) RHS, x0=0, UA-AMG ({a.agg.upper()} "
                         f"aggregation, theta={a.theta}, V(1,1) {a.smoother} w={a.jacobi_weight}, "
                         f"coarse_dof={a.coarse_dof}, dense coarsest solve) PCG to {a.tol:g}, "
                         + ("plain FP64 drop-in" if a.precision == "fp64" else
                            f"{a.precision.upper()} cycle under FP64 PCG")),
            "levels": infos[-1][4],
            "level_rows": infos[-1][6],
            "level_nnz": infos[-1][7],
        },
        "iterations": iters[-1],
        "setup_ms": float(np.mean([i[1] for i in infos])),
        "solve_ms": float(np.mean([i[2] for i in infos])),
        "true_rel_residual": true_rel,
        "whole_solve": {
            "what": "algorithmic bytes of every PCG kernel launch (counted by the library) over the PCG's "
                    "device time (gaps between launches included)",
            "bytes": infos[-1][8],
            "solve_ms": infos[-1][2],
            "gbs": infos[-1][8] / (infos[-1][2] * 1e-3) / 1e9 if infos[-1][2] > 0 else None,
            "frac": infos[-1][8] / (infos[-1][2] * 1e-3) / 1e9 / HBM_PEAK_GBS if infos[-1][2] > 0 else None,
        },
        "spmv_gbs": spmv_gbs,
        "roofline": {
            "bound": "hbm",
            "kernel": ("k_spmv<true, true, 1, unsigned int>: SELL-64 SpMV t=Ap with (t,p) partials, columns as "
                       "paired 16-bit offsets (10 bytes an entry)" if paired else
                       "k_spmv<true, true, 1>: SELL-64 SpMV t=Ap with (t,p) partials") + ", fine level, in the PCG",
            "achieved": spmv_gbs,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": spmv_gbs / HBM_PEAK_GBS,
            "traffic": traffic["traffic_bytes"] if traffic else None,
            "traffic_detail": traffic if traffic else traffic_note,
            "algorithmic_bytes_per_launch": bytes_spmv,
            "avg_launch_ms": pcg_spmv_ms,
            "launches_timed": pcg_spmv_count,
        },
        "fine_kernels_in_solve": {
            "what": "fine-level kernels timed with HIP events inside one solve; algorithmic bytes per launch; "
                    "traffic from the FETCH_SIZE / WRITE_SIZE passes",
            "spmv": {"us": pcg_spmv_ms * 1e3, "bytes": bytes_spmv,
                     "frac": bytes_spmv / (pcg_spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if pcg_spmv_ms else None,
                     "traffic": traffic["spmv"]["traffic_bytes"] if traffic else None},
            "residual": {"us": res_ms * 1e3, "bytes": bytes_res,
                         "frac": bytes_res / (res_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if res_ms else None,
                         "traffic": traffic["residual"]["traffic_bytes"] if traffic else None},
            "last_sweep": {"us": sweep_ms * 1e3, "bytes": bytes_sweep_dot,
                           "frac": bytes_sweep_dot / (sweep_ms * 1e-3) / 1e9 / HBM_PEAK_GBS if sweep_ms else None,
                           "traffic": traffic["sweep_dot"]["traffic_bytes"] if traffic else None},
        },
        "kernels": {
            "back_to_back_spmv_ms": kt.spmv_ms,
            "back_to_back_spmv_gbs": (12 * nnz + 4 * (n + 1) + 16 * n) / (kt.spmv_ms * 1e-3) / 1e9
            if kt.spmv_ms > 0 else None,
            "spmv_dot_ms": kt.spmv_dot_ms, "residual_ms": kt.residual_ms,
            "jacobi_ms": kt.jacobi_ms, "jacobi_gbs": jac_gbs, "sell_entries": kt.sell_entries,
        },
        "cpu_baseline": cpu,
        "learned_vs_plain": lvp,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
