"""ctypes mirror of include/hazmath_mi355x.h.

The structures below have the same field order and C types as HAZmath's public
types (include/sparse.h:25-45, include/vec.h:20-42, include/param.h:138-266), so
the same Python objects can be handed to the MI355X library and to the CPU
oracle.  Loading the library is explicit (`load_library`) and fails loudly when
the built shared object is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

INT = C.c_int
SHORT = C.c_short
REAL = C.c_double
MAX_AMG_LVL = 20

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libhazmath_mi355x.so"

# Error codes (include/macro.h:38-75)
SUCCESS = 0
ERROR_INPUT_PAR = -13
ERROR_ALLOC_MEM = -20
ERROR_AMG_SMOOTH_TYPE = -31
ERROR_AMG_AGG_TYPE = -32
ERROR_AMG_COARSEING = -33
ERROR_SOLVER_TYPE = -40
ERROR_SOLVER_STAG = -42
ERROR_SOLVER_SOLSTAG = -43
ERROR_SOLVER_TOLSMALL = -44
ERROR_SOLVER_MAXIT = -48
ERROR_DEVICE = -90

# Enumerations (include/macro.h:126-221)
SOLVER_CG = 1
SOLVER_UMFPACK = 32
STOP_REL_RES, STOP_REL_PRECRES, STOP_MOD_REL_RES = 1, 2, 3
PREC_AMG = 2
UA_AMG = 1
SA_AMG = 2
VMB, HEC, HEM = 1, 4, 5
V_CYCLE, W_CYCLE = 1, 2
SMOOTHER_JACOBI, SMOOTHER_GS, SMOOTHER_L1DIAG = 1, 2, 10
SMOOTHER_CHEBY = 91
SMOOTHER_GS, SMOOTHER_SGS, SMOOTHER_SOR, SMOOTHER_SSOR, SMOOTHER_GSOR, SMOOTHER_SGSOR = 2, 3, 5, 6, 7, 8


class dCSRmat(C.Structure):
    _fields_ = [("row", INT), ("col", INT), ("nnz", INT),
                ("IA", C.POINTER(INT)), ("JA", C.POINTER(INT)), ("val", C.POINTER(REAL))]


class dvector(C.Structure):
    _fields_ = [("row", INT), ("val", C.POINTER(REAL))]


class linear_itsolver_param(C.Structure):
    _fields_ = [("linear_itsolver_type", SHORT), ("linear_precond_type", SHORT),
                ("linear_stop_type", SHORT), ("linear_maxit", INT), ("linear_tol", REAL),
                ("linear_restart", INT), ("linear_print_level", SHORT),
                ("HX_smooth_iter", SHORT), ("AL_scaling_param", REAL)]


class AMG_param(C.Structure):
    _fields_ = [("AMG_type", SHORT), ("print_level", SHORT), ("maxit", INT), ("tol", REAL),
                ("max_levels", SHORT), ("coarse_dof", INT), ("cycle_type", SHORT),
                ("smoother", SHORT), ("presmooth_iter", SHORT), ("postsmooth_iter", SHORT),
                ("relaxation", REAL), ("polynomial_degree", SHORT), ("coarse_solver", SHORT),
                ("coarse_scaling", SHORT), ("amli_degree", SHORT), ("amli_coef", C.POINTER(REAL)),
                ("nl_amli_krylov_type", SHORT), ("aggregation_type", SHORT),
                ("strong_coupled", REAL), ("max_aggregation", INT), ("smooth_filter", SHORT),
                ("tentative_smooth", REAL), ("Schwarz_levels", INT), ("Schwarz_mmsize", INT),
                ("Schwarz_maxlvl", INT), ("Schwarz_type", INT), ("Schwarz_blksolver", INT),
                ("HAZDIR", C.c_char_p), ("Schwarz_on_blk", C.POINTER(INT)),
                ("Schwarz_patch_type", C.POINTER(INT)), ("damping_param", REAL),
                ("BSR_alpha", REAL), ("BSR_omega", REAL), ("fpwr", REAL),
                ("smoother_function", C.c_void_p)]


class hz_options(C.Structure):
    _fields_ = [("jacobi_weight", REAL), ("hec_rng_source", INT), ("verbose_timing", INT),
                ("return_on_fatal", INT), ("vcycle_precision", INT), ("batch_time_kernels", INT),
                ("time_pcg_spmv", INT), ("hem_host_max_rows", INT),
                ("rap_wave_stage_pairs", INT), ("hem_pipeline_max_sweeps", INT), ("release_after_call", INT),
                ("cheb_coarse_steps", INT)]


class hz_batch_stats(C.Structure):
    _fields_ = [("setup_ms", REAL), ("solve_ms", REAL), ("op_complexity", REAL), ("num_levels", INT),
                ("graph_launches", INT), ("shape_level0", INT), ("spmm_ms", REAL), ("sweep_ms", REAL),
                ("rows", C.c_longlong), ("nnz", C.c_longlong), ("sell_entries", C.c_longlong),
                ("sweep_shape0", INT), ("spmm_shape0", INT), ("tuned", INT), ("tune_cached", INT),
                ("tune_ms", REAL)]


class hz_solve_info(C.Structure):
    _fields_ = [("iterations", INT), ("num_levels", INT),
                ("level_rows", INT * MAX_AMG_LVL), ("level_nnz", INT * MAX_AMG_LVL),
                ("setup_ms", REAL), ("solve_ms", REAL), ("total_ms", REAL),
                ("final_relres", REAL), ("condest", REAL), ("relres_count", INT),
                ("relres_hist", REAL * 4096), ("fallback_used", INT), ("pcg_spmv_ms", REAL),
                ("pcg_spmv_count", INT), ("learned_iterations", INT), ("learned_total_ms", REAL),
                ("hem_pipeline_giveups", INT), ("cheb_scale", REAL * MAX_AMG_LVL),
                ("lambda_ratio", REAL * MAX_AMG_LVL), ("solve_alg_bytes", REAL), ("cycle_residual_ms", REAL),
                ("cycle_sweep_ms", REAL)]


class hz_hierarchy(C.Structure):
    _fields_ = [("num_levels", INT), ("A", dCSRmat * MAX_AMG_LVL),
                ("aggregates", C.POINTER(INT) * MAX_AMG_LVL),
                ("num_aggregates", INT * MAX_AMG_LVL), ("theta", REAL * MAX_AMG_LVL),
                ("P", dCSRmat * MAX_AMG_LVL)]


class input_param(C.Structure):
    _fields_ = [("print_level", SHORT), ("inifile", C.c_char * 256), ("gridfile", C.c_char * 256),
                ("output_dir", C.c_char * 256), ("read_mesh_from_file", INT), ("spatial_dim", INT),
                ("refinement_type", INT), ("refinement_levels", INT), ("boundary_codes", INT),
                ("nquad", INT), ("FE_type", INT), ("Mass_lump", INT), ("time_start", REAL),
                ("time_step_type", INT), ("time_steps", INT), ("time_step_size", REAL),
                ("rhs_time_dep", INT), ("nonlinear_itsolver_type", INT),
                ("nonlinear_itsolver_maxit", INT), ("nonlinear_itsolver_tol", REAL),
                ("nonlinear_itsolver_toltype", INT), ("fas_presmoothers", INT),
                ("fas_postsmoothers", INT), ("fas_smooth_tol", INT), ("linear_itsolver_type", INT),
                ("linear_itsolver_maxit", INT), ("linear_itsolver_tol", REAL),
                ("linear_stop_type", SHORT), ("linear_restart", INT), ("linear_precond_type", INT),
                ("Schwarz_mmsize", INT), ("Schwarz_maxlvl", INT), ("Schwarz_type", INT),
                ("Schwarz_blksolver", INT), ("AMG_type", SHORT), ("AMG_levels", SHORT),
                ("AMG_cycle_type", SHORT), ("AMG_smoother", SHORT), ("AMG_relaxation", REAL),
                ("AMG_presmooth_iter", SHORT), ("AMG_postsmooth_iter", SHORT),
                ("AMG_polynomial_degree", SHORT), ("AMG_coarse_dof", INT), ("AMG_tol", REAL),
                ("AMG_maxit", INT), ("AMG_coarse_solver", SHORT), ("AMG_coarse_scaling", SHORT),
                ("AMG_amli_degree", SHORT), ("AMG_nl_amli_krylov_type", SHORT),
                ("AMG_Schwarz_levels", INT), ("AMG_fpwr", REAL), ("AMG_aggregation_type", SHORT),
                ("AMG_strong_coupled", REAL), ("AMG_max_aggregation", INT),
                ("AMG_smooth_filter", SHORT), ("AMG_tentative_smooth", REAL),
                ("HX_smooth_iter", SHORT), ("BSR_alpha", REAL), ("BSR_omega", REAL)]


PRECOND_FN = C.CFUNCTYPE(None, C.POINTER(REAL), C.POINTER(REAL), C.c_void_p)


class precond(C.Structure):
    _fields_ = [("data", C.c_void_p), ("fct", PRECOND_FN), ("setup_time", REAL)]


class hz_kernel_times(C.Structure):
    _fields_ = [("reps", INT), ("spmv_ms", REAL), ("residual_ms", REAL), ("jacobi_ms", REAL),
                ("spmv_dot_ms", REAL), ("sell_entries", C.c_longlong)]


class orc_trace(C.Structure):
    _fields_ = [("num_levels", INT), ("level_rows", INT * MAX_AMG_LVL),
                ("level_nnz", INT * MAX_AMG_LVL), ("theta", REAL * MAX_AMG_LVL),
                ("iterations", INT), ("condest", REAL), ("final_relres", REAL),
                ("relres_count", INT), ("relres_hist", REAL * 4096),
                ("setup_s", REAL), ("solve_s", REAL)]


# ---------------------------------------------------------------------------
# numpy <-> struct helpers (the numpy arrays must outlive the structs)
# ---------------------------------------------------------------------------
class HostCSR:
    """Keeps the numpy arrays of a CSR matrix alive next to its dCSRmat view."""

    def __init__(self, ia, ja, val, ncol=None):
        self.ia = np.ascontiguousarray(ia, dtype=np.int32)
        self.ja = np.ascontiguousarray(ja, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        n = self.ia.size - 1
        self.mat = dCSRmat(n, n if ncol is None else ncol, int(self.ia[-1]),
                           self.ia.ctypes.data_as(C.POINTER(INT)),
                           self.ja.ctypes.data_as(C.POINTER(INT)),
                           self.val.ctypes.data_as(C.POINTER(REAL)))

    @property
    def n(self):
        return self.ia.size - 1

    @property
    def nnz(self):
        return int(self.ia[-1])

    def ptr(self):
        return C.byref(self.mat)


class HostVec:
    def __init__(self, arr):
        self.arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.vec = dvector(self.arr.size, self.arr.ctypes.data_as(C.POINTER(REAL)))

    def ptr(self):
        return C.byref(self.vec)


def csr_to_numpy(m: dCSRmat):
    n, nnz = m.row, m.nnz
    ia = np.ctypeslib.as_array(m.IA, shape=(n + 1,)).copy() if n > 0 else np.zeros(1, np.int32)
    ja = np.ctypeslib.as_array(m.JA, shape=(nnz,)).copy() if nnz > 0 else np.zeros(0, np.int32)
    val = np.ctypeslib.as_array(m.val, shape=(nnz,)).copy() if nnz > 0 else np.zeros(0)
    return ia, ja, val


def hierarchy_to_numpy(h: hz_hierarchy):
    """Returns a list of per-level dicts with A (ia, ja, val), aggregates, theta."""
    out = []
    for level in range(h.num_levels):
        ia, ja, val = csr_to_numpy(h.A[level])
        d = {"ia": ia, "ja": ja, "val": val, "theta": h.theta[level]}
        if level < h.num_levels - 1 and bool(h.aggregates[level]):
            d["agg"] = np.ctypeslib.as_array(h.aggregates[level], shape=(ia.size - 1,)).copy()
            d["nagg"] = h.num_aggregates[level]
            if bool(h.P[level].IA):
                d["P"] = csr_to_numpy(h.P[level])
        out.append(d)
    return out


def default_amg_param(lib=None) -> AMG_param:
    p = AMG_param()
    (lib or load_library()).param_amg_init(C.byref(p))
    return p


def default_itsolver_param(lib=None) -> linear_itsolver_param:
    p = linear_itsolver_param()
    (lib or load_library()).param_linear_solver_init(C.byref(p))
    return p


def solver_params(lib, *, tol=1e-8, maxit=500, aggregation=HEC, smoother=SMOOTHER_JACOBI,
                  coarse_dof=100, coarse_solver=SOLVER_UMFPACK, max_levels=20, theta=0.0,
                  coarse_scaling=0, cycle=V_CYCLE, print_level=0, max_aggregation=20,
                  stop_type=STOP_REL_RES, amg_maxit=1):
    """The UA-AMG-PCG settings of the reference's examples/solvers harness."""
    it = linear_itsolver_param()
    amg = AMG_param()
    lib.param_linear_solver_init(C.byref(it))
    lib.param_amg_init(C.byref(amg))
    it.linear_itsolver_type = SOLVER_CG
    it.linear_precond_type = PREC_AMG
    it.linear_stop_type = stop_type
    it.linear_maxit = maxit
    it.linear_tol = tol
    it.linear_print_level = print_level
    amg.AMG_type = UA_AMG
    amg.cycle_type = cycle
    amg.max_levels = max_levels
    amg.coarse_dof = coarse_dof
    amg.coarse_solver = coarse_solver
    amg.coarse_scaling = coarse_scaling
    amg.smoother = smoother
    amg.aggregation_type = aggregation
    amg.strong_coupled = theta
    amg.max_aggregation = max_aggregation
    amg.maxit = amg_maxit
    amg.print_level = print_level
    return it, amg


BATCH_RHS = 8


def batch_solve(lib, systems, B, X, it, amg):
    """hz_batch_krylov_amg on host data: systems = [(ia, ja, va)], B[s] and X[s]
    arrays of shape (nrhs, n_s) (X: initial guesses, overwritten).  Returns
    (status, iters[nsys, nrhs])."""
    nsys = len(systems)
    nrhs = B[0].shape[0]
    mats = [HostCSR(*sysm) for sysm in systems]
    Aptr = (C.POINTER(dCSRmat) * nsys)(*[C.pointer(m.mat) for m in mats])
    Bc = [np.ascontiguousarray(b, dtype=np.float64) for b in B]
    for x in X:
        assert x.flags.c_contiguous and x.dtype == np.float64
    Bptr = (C.POINTER(REAL) * nsys)(*[dptr(b) for b in Bc])
    Xptr = (C.POINTER(REAL) * nsys)(*[dptr(x) for x in X])
    iters = np.zeros(nsys * nrhs, np.int32)
    r = lib.hz_batch_krylov_amg(nsys, Aptr, nrhs, Bptr, Xptr, C.byref(it), C.byref(amg), dptr(iters, INT))
    return r, iters.reshape(nsys, nrhs)


_LIB = None


def _declare(lib):
    P = C.POINTER
    lib.linear_solver_dcsr_krylov_amg.argtypes = [P(dCSRmat), P(dvector), P(dvector),
                                                  P(linear_itsolver_param), P(AMG_param)]
    lib.linear_solver_dcsr_krylov_amg.restype = INT
    lib.param_amg_init.argtypes = [P(AMG_param)]
    lib.param_amg_init.restype = None
    lib.param_linear_solver_init.argtypes = [P(linear_itsolver_param)]
    lib.param_linear_solver_init.restype = None
    lib.hz_get_options.argtypes = [P(hz_options)]
    lib.hz_set_options.argtypes = [P(hz_options)]
    lib.hz_get_last_info.argtypes = [P(hz_solve_info)]
    lib.hz_krylov_amg_device.argtypes = [INT, INT, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, P(linear_itsolver_param),
                                         P(AMG_param)]
    lib.hz_krylov_amg_device.restype = INT
    lib.hz_setup_hierarchy.argtypes = [P(dCSRmat), P(AMG_param), P(hz_hierarchy)]
    lib.hz_setup_hierarchy.restype = INT
    lib.hz_hierarchy_free.argtypes = [P(hz_hierarchy)]
    lib.hz_dcsr_mxv.argtypes = [P(dCSRmat), P(REAL), P(REAL)]
    lib.hz_dcsr_mxv.restype = INT
    lib.hz_dcsr_residual.argtypes = [P(dCSRmat), P(REAL), P(REAL), P(REAL)]
    lib.hz_dcsr_residual.restype = INT
    lib.hz_smoother_sweep.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, REAL]
    lib.hz_smoother_sweep.restype = INT
    lib.hz_rap_agg.argtypes = [P(dCSRmat), P(INT), INT, P(dCSRmat)]
    lib.hz_rap_agg.restype = INT
    lib.hz_dense_inverse.argtypes = [INT, P(REAL)]
    lib.hz_dense_inverse.restype = INT
    lib.hz_dense_mfma_active.restype = INT
    lib.param_input_init.argtypes = [P(input_param)]
    lib.param_input.argtypes = [C.c_char_p, P(input_param)]
    lib.param_linear_solver_set.argtypes = [P(linear_itsolver_param), P(input_param)]
    lib.param_amg_set.argtypes = [P(AMG_param), P(input_param)]
    lib.python_wrapper_krylov_amg.argtypes = [P(INT), P(INT), P(INT), P(INT), P(REAL), P(REAL), P(REAL),
                                              P(REAL), P(INT), P(INT)]
    F = C.c_float
    lib.hz_mlp_init.argtypes = [P(F), C.c_ulonglong]
    lib.hz_mlp_forward.argtypes = [P(F), P(F), INT, P(F)]
    lib.hz_mlp_forward.restype = INT
    lib.hz_cycle_rate_agg.argtypes = [P(dCSRmat), P(AMG_param), P(F), INT, REAL, INT, P(REAL)]
    lib.hz_cycle_rate_agg.restype = INT
    lib.hz_lambda_max_dinv.argtypes = [P(dCSRmat), INT, P(REAL), P(REAL)]
    lib.hz_lambda_max_dinv.restype = INT
    lib.hz_mlp_forward_noalloc.argtypes = [C.c_void_p, C.c_void_p, INT, C.c_void_p, C.c_void_p]
    lib.hz_mlp_forward_noalloc.restype = INT
    lib.hz_mlp_loss_grad.argtypes = [P(F), P(F), P(F), INT, P(F), P(F)]
    lib.hz_mlp_loss_grad.restype = INT
    lib.hz_mlp_train.argtypes = [P(F), P(F), P(F), INT, INT, INT, F, C.c_ulonglong, P(F)]
    lib.hz_mlp_train.restype = INT
    lib.hz_krylov_amg_learned.argtypes = [P(dCSRmat), P(dvector), P(dvector), P(linear_itsolver_param),
                                          P(AMG_param), P(F), INT, INT]
    lib.hz_krylov_amg_learned.restype = INT
    lib.hz_two_grid_rate.argtypes = [P(dCSRmat), P(AMG_param), REAL, REAL, INT, P(REAL)]
    lib.hz_two_grid_rate.restype = INT
    lib.hz_aggregate_features.argtypes = [P(dCSRmat), P(INT), INT, INT, P(F)]
    lib.hz_aggregate_features.restype = INT
    lib.hz_cheby_smooth.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, REAL]
    lib.hz_cheby_smooth.restype = INT
    lib.hz_batch_krylov_amg.argtypes = [INT, P(P(dCSRmat)), INT, P(P(REAL)), P(P(REAL)),
                                        P(linear_itsolver_param), P(AMG_param), P(INT)]
    lib.hz_batch_krylov_amg.restype = INT
    lib.hz_batch_get_stats.argtypes = [P(hz_batch_stats)]
    lib.hz_batch_get_stats.restype = None
    lib.hz_release_device_memory.argtypes = []
    lib.hz_release_device_memory.restype = None
    lib.hz_device_pool_bytes.argtypes = []
    lib.hz_device_pool_bytes.restype = C.c_longlong
    lib.hz_coarse_itsolver.argtypes = [P(dCSRmat), P(REAL), P(REAL), REAL, P(INT), P(REAL)]
    lib.hz_coarse_itsolver.restype = INT
    lib.hz_batch_tune_cache.argtypes = [C.c_char_p]
    lib.hz_batch_tune_cache.restype = INT
    lib.hz_batch_krylov_amg_device.argtypes = [INT, P(INT), INT, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_void_p, P(linear_itsolver_param), P(AMG_param),
                                               P(INT)]
    lib.hz_batch_krylov_amg_device.restype = INT
    lib.hz_gs_smooth.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, INT, INT, REAL, P(INT), P(INT), P(INT)]
    lib.hz_gs_smooth.restype = INT
    lib.hz_sa_prolongation.argtypes = [P(dCSRmat), P(INT), INT, REAL, REAL, P(F), INT, P(dCSRmat)]
    lib.hz_sa_prolongation.restype = INT
    lib.linear_solver_amg.argtypes = [P(dCSRmat), P(dvector), P(dvector), P(AMG_param)]
    lib.linear_solver_amg.restype = INT
    lib.create_precond_amg.argtypes = [P(dCSRmat), P(AMG_param)]
    lib.create_precond_amg.restype = P(precond)
    lib.apply_precond.argtypes = [P(REAL), P(REAL), P(precond)]
    lib.hz_precond_amg_apply.argtypes = [P(REAL), P(REAL), C.c_void_p]
    lib.hz_precond_amg_apply_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.hz_precond_free.argtypes = [P(precond)]
    lib.hz_time_fine_kernels.argtypes = [INT, INT, C.c_void_p, C.c_void_p, C.c_void_p, INT,
                                         P(hz_kernel_times)]
    lib.hz_time_fine_kernels.restype = INT
    lib.hz_version.restype = C.c_char_p
    lib.hz_device_count.restype = INT
    lib.dcsr_free.argtypes = [P(dCSRmat)]


def load_library(path: os.PathLike | None = None):
    """Load libhazmath_mi355x.so (built in-tree by __graft_entry__.build())."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(str(p), mode=C.RTLD_LOCAL)
    _declare(lib)
    if path is None:
        _LIB = lib
    return lib


MLP_INPUTS, MLP_OUTPUTS, MLP_PARAMS = 12, 2, 6288
MLP_NOALLOC_MAX_ROWS = 16384   # HZ_MLP_NOALLOC_MAX_ROWS


def fptr(arr: np.ndarray):
    """float32 pointer of a contiguous float32 array."""
    assert arr.dtype == np.float32 and arr.flags.c_contiguous
    return arr.ctypes.data_as(C.POINTER(C.c_float))


def dptr(arr: np.ndarray, ctype=REAL):
    return arr.ctypes.data_as(C.POINTER(ctype))


def libc():
    return C.CDLL(None)
