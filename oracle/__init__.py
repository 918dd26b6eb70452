"""CPU oracle loader -- TEST INFRASTRUCTURE ONLY.

Loads oracle/liboracle.so, the serial C restatement of HAZmath's UA-AMG-PCG path
(hz_oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg import this module; the product package never does.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from hazmath_amd.capi import (AMG_param, INT, REAL, dCSRmat, dvector, hz_hierarchy,
                              linear_itsolver_param, orc_trace)

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"
_LIB = None


def load(path=None):
    """The oracle library (oracle/liboracle.so), or another build of the same
    source at `path` (the sanitizer build of tests/sanitize), declared alike."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} missing: run `make -C oracle`")
    lib = C.CDLL(str(p), mode=C.RTLD_LOCAL)
    P = C.POINTER
    lib.orc_krylov_amg.argtypes = [P(dCSRmat), P(dvector), P(dvector), P(linear_itsolver_param),
                                   P(AMG_param), P(orc_trace)]
    lib.orc_krylov_amg.restype = INT
    lib.orc_param_amg_init.argtypes = [P(AMG_param)]
    lib.orc_param_linear_solver_init.argtypes = [P(linear_itsolver_param)]
    lib.orc_setup_hierarchy.argtypes = [P(dCSRmat), P(AMG_param), P(hz_hierarchy)]
    lib.orc_setup_hierarchy.restype = INT
    lib.orc_hierarchy_free.argtypes = [P(hz_hierarchy)]
    lib.orc_dcsr_mxv.argtypes = [P(dCSRmat), P(REAL), P(REAL)]
    lib.orc_dcsr_residual.argtypes = [P(dCSRmat), P(REAL), P(REAL), P(REAL)]
    lib.orc_smoother_sweep.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, REAL]
    lib.orc_rap_agg.argtypes = [P(dCSRmat), P(INT), INT, P(dCSRmat)]
    lib.orc_rap_agg.restype = INT
    lib.orc_strength.argtypes = [P(dCSRmat), REAL, P(dCSRmat)]
    lib.orc_aggregate.argtypes = [P(dCSRmat), INT, REAL, INT, P(INT), P(INT)]
    lib.orc_aggregate.restype = INT
    lib.orc_tridiag_condest.argtypes = [INT, P(REAL), P(REAL)]
    lib.orc_pvgmres.argtypes = [P(dCSRmat), P(REAL), P(REAL), REAL, INT, INT]
    lib.orc_pvgmres.restype = INT
    lib.orc_coarse_itsolver.argtypes = [P(dCSRmat), P(REAL), P(REAL), REAL]
    lib.orc_coarse_itsolver.restype = INT
    lib.orc_cg.argtypes = [P(dCSRmat), P(REAL), P(REAL), REAL, INT]
    lib.orc_cg.restype = INT
    lib.orc_tridiag_condest.restype = REAL
    lib.orc_set_jacobi_weight.argtypes = [REAL]
    lib.orc_sa_prolongation.argtypes = [P(dCSRmat), P(INT), INT, REAL, REAL, C.POINTER(C.c_float), INT,
                                        P(dCSRmat)]
    lib.orc_cheby_smooth.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, REAL]
    lib.orc_linear_solver_amg.argtypes = [P(dCSRmat), P(dvector), P(dvector), P(AMG_param)]
    lib.orc_linear_solver_amg.restype = INT
    lib.orc_precond_create.argtypes = [P(dCSRmat), P(AMG_param)]
    lib.orc_precond_create.restype = C.c_void_p
    lib.orc_precond_apply.argtypes = [C.c_void_p, P(REAL), P(REAL)]
    lib.orc_precond_free.argtypes = [C.c_void_p]
    lib.orc_gs_smooth_ordered.argtypes = [P(dCSRmat), P(REAL), P(REAL), INT, INT, INT, REAL, P(INT), P(INT),
                                          INT]
    lib.orc_gs_smooth_ordered.restype = C.c_int
    lib.orc_set_learned.argtypes = [ORC_OMEGA_FN, P(REAL)]
    lib.orc_set_learned.restype = None
    lib.orc_set_cheb_coarse_steps.argtypes = [INT]
    lib.orc_set_cheb_coarse_steps.restype = None
    if path is None:
        _LIB = lib
    return lib


# fn(level, A_l, agg_l, nagg_l, omega[nagg_l]) of orc_set_learned
ORC_OMEGA_FN = C.CFUNCTYPE(None, INT, C.POINTER(dCSRmat), C.POINTER(INT), INT, C.POINTER(C.c_float))


def params(*, tol=1e-8, maxit=500, aggregation=4, smoother=1, coarse_dof=100, coarse_solver=32,
           max_levels=20, theta=0.0, coarse_scaling=0, cycle=1, print_level=0,
           max_aggregation=20, stop_type=1, amg_maxit=1):
    """Oracle-side twin of hazmath_amd.capi.solver_params (same field values)."""
    lib = load()
    it = linear_itsolver_param()
    amg = AMG_param()
    lib.orc_param_linear_solver_init(C.byref(it))
    lib.orc_param_amg_init(C.byref(amg))
    it.linear_itsolver_type = 1
    it.linear_precond_type = 2
    it.linear_stop_type = stop_type
    it.linear_maxit = maxit
    it.linear_tol = tol
    it.linear_print_level = print_level
    amg.AMG_type = 1
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
