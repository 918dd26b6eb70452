/*
 * hz_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * Serial CPU restatement of HAZmath's unsmoothed-aggregation AMG-preconditioned
 * CG path.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg
 * may load it; the product library never links or calls it.
 *
 * Every entry point mirrors one reference function (file:line in hz_oracle.c).
 * Function names carry the prefix orc_ so the oracle can share a process with
 * the product library (which exports the unprefixed HAZmath names).
 */
#ifndef HZ_ORACLE_H
#define HZ_ORACLE_H

#include "hazmath_mi355x.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_EXIT_HEC   (-1016)  /* reference calls exit(16) (amg_setup_ua.c:820-823) */
#define ORC_EXIT_HEM   (-1017)  /* reference calls exit(17) (amg_setup_ua.c:1031-1033) */
#define ORC_EXIT_PARAM (-1013)  /* reference calls check_error -> exit (message.c:191-282) */

/* Trace of one solve (filled when non-NULL). */
typedef struct {
    INT  num_levels;
    INT  level_rows[HZ_MAX_AMG_LVL];
    INT  level_nnz[HZ_MAX_AMG_LVL];
    REAL theta[HZ_MAX_AMG_LVL];
    INT  iterations;
    REAL condest;
    REAL final_relres;
    INT  relres_count;
    REAL relres_hist[4096];
    REAL setup_s;   /* CLOCK_MONOTONIC wall seconds */
    REAL solve_s;
} orc_trace;

INT  orc_krylov_amg(dCSRmat *A, dvector *b, dvector *x,
                    linear_itsolver_param *itparam, AMG_param *amgparam,
                    orc_trace *trace);
void orc_param_amg_init(AMG_param *p);
void orc_param_linear_solver_init(linear_itsolver_param *p);

/* Setup only, exporting the hierarchy (host arrays; free with orc_hierarchy_free). */
INT  orc_setup_hierarchy(dCSRmat *A, AMG_param *amgparam, hz_hierarchy *out);
void orc_hierarchy_free(hz_hierarchy *h);

/* Kernel-level restatements. */
void orc_dcsr_mxv(const dCSRmat *A, const REAL *x, REAL *y);
void orc_dcsr_residual(const dCSRmat *A, const REAL *b, const REAL *x, REAL *w);
void orc_smoother_sweep(const dCSRmat *A, const REAL *b, REAL *x, INT smoother, REAL w);
INT  orc_rap_agg(const dCSRmat *A, const INT *agg, INT nagg, dCSRmat *Ac);
void orc_strength(const dCSRmat *A, REAL theta, dCSRmat *N);
INT  orc_aggregate(const dCSRmat *A, INT type, REAL theta, INT max_aggregation,
                   INT *agg, INT *nagg);
REAL orc_tridiag_condest(INT n, const REAL *alpha, const REAL *beta);
/* dcsr_pvgmres without preconditioner, stop type 1 (krylov.c:4612-4931), and
 * coarse_itsolver: CG, then GMRES(30) when CG fails (mgcycle.c:34-53). */
INT  orc_pvgmres(const dCSRmat *A, const REAL *b, REAL *x, REAL tol, INT MaxIt, INT restart);
INT  orc_coarse_itsolver(const dCSRmat *A, const REAL *b, REAL *x, REAL ctol);
INT  orc_cg(const dCSRmat *A, const REAL *b, REAL *x, REAL tol, INT MaxIt);
/* Jacobi damping used by the V-cycle (reference: fixed 1.0, smoother.c:47). */
void orc_set_jacobi_weight(REAL w);
/* smoothed prolongation of one level (uniform or per-aggregate damping) */
void orc_sa_prolongation(const dCSRmat *A, const INT *agg, INT nagg, REAL theta, REAL omega,
                         const float *omega_agg, INT filter, dCSRmat *P);
/* Chebyshev smoother, `steps` steps in place (build-owned method) */
/* dcsr_presmoothing / dcsr_postsmoothing (mgcycle.c:76-223) for GS 2, SGS 3,
 * SOR 5, SSOR 6, GSOR 7, SGSOR 8 over a row ordering in blocks (NULL: natural
 * order, the reference's loops).  Returns -1 for other smoothers. */
int orc_gs_smooth_ordered(const dCSRmat *A, const REAL *b, REAL *x, INT smoother, INT post, INT nsweeps,
                          REAL relax, const INT *ord, const INT *cptr, INT nblk);
int orc_is_gs_family(INT smoother);
void orc_cheby_smooth(const dCSRmat *A, const REAL *b, REAL *x, INT steps, REAL scale);
/* Chebyshev steps per smoothing below the finest level in the cycle (0: the
 * cycle's presmooth / postsmooth counts); no reference counterpart. */
void orc_set_cheb_coarse_steps(INT steps);
/* linear_solver_amg + amg_solve (itsolver.c:524-640, amg_solve.c:109-161) */
INT  orc_linear_solver_amg(dCSRmat *A, dvector *b, dvector *x, AMG_param *param);
/* create_precond_amg / precond_amg (helper.c:206-283, precond.c:164-183) */
void *orc_precond_create(dCSRmat *A, AMG_param *amgparam);
void orc_precond_apply(void *h, const REAL *r, REAL *z);
void orc_precond_free(void *h);

/* Learned-parameter composition (no reference counterpart): fn(level, A_l,
 * agg_l, nagg_l, omega[nagg_l]) fills the per-aggregate damping of every SA
 * level; scales[HZ_MAX_AMG_LVL] (may be NULL) the Chebyshev bound scale per
 * level.  orc_set_learned(NULL, NULL) restores the uniform parameters. */
typedef void (*orc_omega_fn)(INT level, const dCSRmat *A, const INT *agg, INT nagg, float *omega);
void orc_set_learned(orc_omega_fn fn, const REAL *scales);

#ifdef __cplusplus
}
#endif
#endif
