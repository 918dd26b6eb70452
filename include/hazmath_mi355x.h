/*
 * hazmath_mi355x.h -- public C ABI of the MI355X UA-AMG-PCG library.
 *
 * The types below are laid out field-for-field like HAZmath's public types so
 * that a caller compiled against HAZmath can link against this library
 * unchanged (same field names, types and order; INT=int, SHORT=short,
 * REAL=double).  Every entry point cites the HAZmath function it replaces.
 *
 * The solver entry points take HOST arrays (HAZmath's contract): A, b, x live
 * in host memory; the library uploads them, runs setup and PCG on the GPU, and
 * writes the solution back into x->val.  A device-resident variant
 * (hz_krylov_amg_device) takes device pointers for callers whose data already
 * lives in HBM.
 */
#ifndef HAZMATH_MI355X_H
#define HAZMATH_MI355X_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar types: include/macro.h:102-111 (SHORT short, INT int, REAL double). */
#ifndef SHORT
#define SHORT short
#endif
#ifndef INT
#define INT int
#endif
#ifndef REAL
#define REAL double
#endif

/* Constants with the reference's values: include/macro.h. */
#define HZ_SUCCESS                0
#define HZ_ERROR_OPEN_FILE      -10
#define HZ_ERROR_WRONG_FILE     -11
#define HZ_ERROR_INPUT_PAR      -13
#define HZ_ERROR_MAT_SIZE       -15
#define HZ_ERROR_ALLOC_MEM      -20
#define HZ_ERROR_DATA_STRUCTURE -21
#define HZ_ERROR_AMG_SMOOTH_TYPE -31
#define HZ_ERROR_AMG_AGG_TYPE   -32
#define HZ_ERROR_AMG_COARSEING  -33
#define HZ_ERROR_SOLVER_TYPE    -40
#define HZ_ERROR_SOLVER_STAG    -42
#define HZ_ERROR_SOLVER_SOLSTAG -43
#define HZ_ERROR_SOLVER_TOLSMALL -44
#define HZ_ERROR_SOLVER_MISC    -47
#define HZ_ERROR_SOLVER_MAXIT   -48
#define HZ_ERROR_DEVICE         -90   /* no counterpart: HIP runtime failure */

/* include/macro.h:126-221 */
#define HZ_SOLVER_DEFAULT   0
#define HZ_SOLVER_CG        1
#define HZ_SOLVER_UMFPACK  32
#define HZ_SOLVER_AMG      21
#define HZ_STOP_REL_RES     1
#define HZ_STOP_REL_PRECRES 2
#define HZ_STOP_MOD_REL_RES 3
#define HZ_PREC_AMG         2
#define HZ_UA_AMG           1
#define HZ_SA_AMG           2
#define HZ_VMB              1
#define HZ_HEC              4
#define HZ_HEM              5
#define HZ_V_CYCLE          1
#define HZ_W_CYCLE          2
#define HZ_SMOOTHER_JACOBI  1
#define HZ_SMOOTHER_GS      2   /* GS family: multicolour sweeps on the GPU */
#define HZ_SMOOTHER_SGS     3
#define HZ_SMOOTHER_SOR     5
#define HZ_SMOOTHER_SSOR    6
#define HZ_SMOOTHER_GSOR    7
#define HZ_SMOOTHER_SGSOR   8
#define HZ_SMOOTHER_L1DIAG 10
#define HZ_SMOOTHER_CHEBY  91   /* no counterpart: Chebyshev on D^-1 A (build-owned) */
#define HZ_UNPT            -1
#define HZ_MAX_AMG_LVL     20

/* dCSRmat: include/sparse.h:25-45 (0-based CSR). */
typedef struct dCSRmat {
    INT row;
    INT col;
    INT nnz;
    INT *IA;
    INT *JA;
    REAL *val;
} dCSRmat;

/* dvector / ivector: include/vec.h:20-42. */
typedef struct dvector {
    INT row;
    REAL *val;
} dvector;

typedef struct ivector {
    INT row;
    INT *val;
} ivector;

/* linear_itsolver_param: include/param.h:138-154. */
typedef struct {
    SHORT linear_itsolver_type;
    SHORT linear_precond_type;
    SHORT linear_stop_type;
    INT   linear_maxit;
    REAL  linear_tol;
    INT   linear_restart;
    SHORT linear_print_level;
    SHORT HX_smooth_iter;
    REAL  AL_scaling_param;
} linear_itsolver_param;

/* AMG_param: include/param.h:163-266. */
typedef struct {
    SHORT AMG_type;
    SHORT print_level;
    INT   maxit;
    REAL  tol;
    SHORT max_levels;
    INT   coarse_dof;
    SHORT cycle_type;
    SHORT smoother;
    SHORT presmooth_iter;
    SHORT postsmooth_iter;
    REAL  relaxation;
    SHORT polynomial_degree;
    SHORT coarse_solver;
    SHORT coarse_scaling;
    SHORT amli_degree;
    REAL *amli_coef;
    SHORT nl_amli_krylov_type;
    SHORT aggregation_type;
    REAL  strong_coupled;
    INT   max_aggregation;
    SHORT smooth_filter;
    REAL  tentative_smooth;
    INT   Schwarz_levels;
    INT   Schwarz_mmsize;
    INT   Schwarz_maxlvl;
    INT   Schwarz_type;
    INT   Schwarz_blksolver;
    char *HAZDIR;
    INT  *Schwarz_on_blk;
    INT  *Schwarz_patch_type;
    REAL  damping_param;
    REAL  BSR_alpha;
    REAL  BSR_omega;
    REAL  fpwr;
    void *smoother_function;
} AMG_param;

/* ---------------------------------------------------------------------------
 * Drop-in entry points (same prototypes as HAZmath).
 * ------------------------------------------------------------------------- */

/* Replaces linear_solver_dcsr_krylov_amg, src/solver/itsolver.c:846-931.
 * Returns the PCG iteration count (>=0) or a negative HAZmath error code.
 * x->val is overwritten with the solution; amgparam->strong_coupled is
 * modified per level exactly like src/solver/amg_setup_ua.c:1325-1328; HEC
 * aggregation draws from libc rand() like src/utilities/array.c:302-318. */
INT linear_solver_dcsr_krylov_amg(dCSRmat *A, dvector *b, dvector *x,
                                  linear_itsolver_param *itparam,
                                  AMG_param *amgparam);

/* Replaces param_amg_init, src/utilities/parameters.c:158-205. */
void param_amg_init(AMG_param *amgparam);

/* Replaces param_linear_solver_init, src/utilities/parameters.c:238-251. */
void param_linear_solver_init(linear_itsolver_param *itsparam);

/* Replace dcsr_create / dcsr_free, src/utilities/sparse.c:27-57, 573-592. */
dCSRmat dcsr_create(const INT m, const INT n, const INT nnz);
void dcsr_free(dCSRmat *A);

/* Replace dvec_create / dvec_free / dvec_set, src/utilities/vec.c:68-79, 157-163, 280-295. */
dvector dvec_create(const INT m);
void dvec_free(dvector *u);
void dvec_set(INT n, dvector *x, const REAL val);

/* Replace dcoo_read_dcsr_p / dvector_read_p, src/utilities/io.c:2296-2346, 2427-2445
 * ("m n nnz" header then "i j a_ij" lines, 0-based; row-stable COO->CSR like
 * src/utilities/format.c:785-827). Returned pointers are freed with
 * hz_free_dcsr_p / hz_free_dvector_p (or by free()-ing arrays then struct). */
dCSRmat *dcoo_read_dcsr_p(FILE *fp);
dvector *dvector_read_p(FILE *fp);

/* Replaces linear_solver_amg, src/solver/itsolver.c:524-640: UA-AMG setup, then
 * AMG cycles as the solver (amg_solve, src/solver/amg_solve.c:109-161) from the
 * initial guess in x until ||b-Ax||/||b|| < param->tol.  Returns the iteration
 * count (param->maxit + 1 when it did not converge, like the reference), or a
 * negative error code.  Where the reference falls back to GMRES after a failed
 * setup, this returns the setup status. */
INT linear_solver_amg(dCSRmat *A, dvector *b, dvector *x, AMG_param *param);

/* precond: include/solver.h:25-36. */
typedef struct {
    void *data;
    void (*fct)(REAL *, REAL *, void *);
    REAL setup_time;
} precond;

/* Replaces create_precond_amg, src/haznics/helper.c:206-283: UA-AMG setup once
 * on the GPU (the hierarchy stays in device memory), returns a preconditioner
 * whose fct applies precond_amg (src/solver/precond.c:164-183: amgparam->maxit
 * cycles from zero) to HOST arrays.  NULL when setup fails.  Like the
 * reference, amgparam->strong_coupled is updated per level. */
precond *create_precond_amg(dCSRmat *A, AMG_param *amgparam);

/* Replaces apply_precond, src/haznics/helper.c:1254-1260: pc->fct(r, z, pc->data). */
void apply_precond(REAL *r, REAL *z, precond *pc);

/* input_param: include/param.h:23-132 (field for field). */
typedef struct {
    SHORT print_level;
    char  inifile[256];
    char  gridfile[256];
    char  output_dir[256];
    INT   read_mesh_from_file;
    INT   spatial_dim;
    INT   refinement_type;
    INT   refinement_levels;
    INT   boundary_codes;
    INT   nquad;
    INT   FE_type;
    INT   Mass_lump;
    REAL  time_start;
    INT   time_step_type;
    INT   time_steps;
    REAL  time_step_size;
    INT   rhs_time_dep;
    INT   nonlinear_itsolver_type;
    INT   nonlinear_itsolver_maxit;
    REAL  nonlinear_itsolver_tol;
    INT   nonlinear_itsolver_toltype;
    INT   fas_presmoothers;
    INT   fas_postsmoothers;
    INT   fas_smooth_tol;
    INT   linear_itsolver_type;
    INT   linear_itsolver_maxit;
    REAL  linear_itsolver_tol;
    SHORT linear_stop_type;
    INT   linear_restart;
    INT   linear_precond_type;
    INT   Schwarz_mmsize;
    INT   Schwarz_maxlvl;
    INT   Schwarz_type;
    INT   Schwarz_blksolver;
    SHORT AMG_type;
    SHORT AMG_levels;
    SHORT AMG_cycle_type;
    SHORT AMG_smoother;
    REAL  AMG_relaxation;
    SHORT AMG_presmooth_iter;
    SHORT AMG_postsmooth_iter;
    SHORT AMG_polynomial_degree;
    INT   AMG_coarse_dof;
    REAL  AMG_tol;
    INT   AMG_maxit;
    SHORT AMG_coarse_solver;
    SHORT AMG_coarse_scaling;
    SHORT AMG_amli_degree;
    SHORT AMG_nl_amli_krylov_type;
    INT   AMG_Schwarz_levels;
    REAL  AMG_fpwr;
    SHORT AMG_aggregation_type;
    REAL  AMG_strong_coupled;
    INT   AMG_max_aggregation;
    SHORT AMG_smooth_filter;
    REAL  AMG_tentative_smooth;
    SHORT HX_smooth_iter;
    REAL  BSR_alpha;
    REAL  BSR_omega;
} input_param;

/* Replaces param_input_init, src/utilities/parameters.c:39-141. */
void param_input_init(input_param *inparam);
/* Replaces param_input, src/utilities/input.c:42-830: defaults, then the
 * "keyword = value" entries of filenm (NULL: defaults only).  A file that
 * cannot be opened is fatal like check_error(ERROR_OPEN_FILE). */
void param_input(const char *filenm, input_param *inparam);
/* Replace param_linear_solver_set / param_amg_set, src/utilities/parameters.c:263-348. */
void param_linear_solver_set(linear_itsolver_param *itsparam, input_param *inparam);
void param_amg_set(AMG_param *amgparam, input_param *inparam);
/* Replaces python_wrapper_krylov_amg, src/utilities/wrapper.c:36-79: parameters
 * from ./input.dat with tol, maxit and print level from the arguments; solves
 * in place on the caller's raw CSR arrays (u holds the guess and the result). */
void python_wrapper_krylov_amg(INT *n, INT *nnz, INT *ia, INT *ja, REAL *a, REAL *b, REAL *u,
                               REAL *tol, INT *maxit, INT *print_lvl);

/* ---------------------------------------------------------------------------
 * Build-owned extensions (no HAZmath counterpart).
 * ------------------------------------------------------------------------- */

/* Side options that AMG_param cannot express without changing its layout.
 * Process-global; set before calling the solver. */
typedef struct {
    REAL jacobi_weight;    /* damping w of SMOOTHER_JACOBI; the reference fixes w=1.0
                              (src/solver/smoother.c:47).  1.0 = reference behaviour. */
    INT  hec_rng_source;   /* HEC turn order (iarray_shuffle with libc rand()):
                              0: the caller's glibc random() state is read, its stream is
                                 generated and the shuffle run on the GPU, and the libc state
                                 is advanced exactly as rand() would have been (checked; falls
                                 back to 1 with any other libc);
                              1: rand() called on the host, the reference's own loop. */
    INT  verbose_timing;   /* >0: print setup/solve wall times from hipEvents */
    INT  return_on_fatal;  /* 0: exit() like the reference on its fatal paths
                              (exit(16)/exit(17)/check_error); 1: return the negative
                              code (-1016/-1017/-1013) instead. */
    INT  vcycle_precision; /* arithmetic of the AMG preconditioner's cycle; the PCG
                              (residuals, dot products, SpMV) and the coarsest solve
                              stay FP64 in every mode:
                              0: FP64 (the reference's);
                              1: FP32 vectors and matrix values, FP32 sums;
                              2: as 1 with BF16 matrix values on levels below the finest.
                              Modes 1 and 2 need coarse_scaling = OFF. */
    INT  batch_time_kernels; /* >0: the batched solver also times its level-0 SpMM and
                              sweep after the solve (hz_batch_get_stats) */
    INT  time_pcg_spmv;      /* >0: record HIP events around each PCG SpMV of the solve
                              (hz_solve_info.pcg_spmv_ms) */
    INT  hem_host_max_rows;  /* HEM levels up to this many rows run pass 1 as a sequential sweep
                              on the host (up to 32768 rows over the whole strength graph, above
                              over GPU-sorted preference lists), larger ones the GPU pipeline;
                              all give the same aggregates.  0: 1100000 (or the HZ_HEM_HOST_MAX
                              environment variable); every level through one form: -1 the GPU
                              pipeline, -2 the full host sweep, -3 the preference lists */
    INT  rap_wave_stage_pairs; /* SA's general Galerkin product (dcsr_rap) runs one wavefront per
                              coarse row, staging the row's (R entry, A entry) pairs in LDS this
                              many at a time; rows of more than 256 columns, or with one A row
                              longer than a stage, run one thread per row.  All forms give the
                              same matrix.  0: 256 (the stage's LDS capacity; larger values are
                              capped there), -1: every row one thread per row, -2: the default with a
                              16-column limit */
    INT  hem_pipeline_max_sweeps; /* HEM's one-launch GPU pipeline gives up on a chunk of vertices
                              after this many sweeps (checked every 16).  > 0: the remaining
                              pass-1 decisions run as one launch per round; < 0: after -value
                              sweeps the level goes to the host sweep over preference lists.
                              All forms give the same aggregates.  0: levels the pipeline takes
                              by size go to the host sweep after 128 sweeps (long dependency
                              chains, e.g. strongly anisotropic problems); a pipeline forced on
                              every level (hem_host_max_rows -1) finishes by rounds after 8192 */
    INT  release_after_call;  /* 1: every solver call returns the library's device memory pool to
                              HIP before it returns (the reference frees its work arrays); 0
                              (default): the pool is kept, so a warm call performs no hipMalloc.
                              Preconditioner objects keep their own memory until hz_precond_free */
    INT  cheb_coarse_steps;   /* Chebyshev smoother (amg.smoother 91): steps per smoothing on the levels
                              below the finest; 0 (default): presmooth_iter / postsmooth_iter on every
                              level.  No counterpart in the reference; measured on the learned method
                              (128^3 anisotropic, Chebyshev(4) on the finest level): 2 keeps the
                              iteration count and takes 12 % off the PCG */
} hz_options;

void hz_get_options(hz_options *opt);
void hz_set_options(const hz_options *opt);
/* Returns the library's device memory pool to HIP now (no counterpart in the
 * reference, which allocates per call); the next call allocates again.
 * hz_device_pool_bytes: device bytes the pool holds now. */
void hz_release_device_memory(void);
long long hz_device_pool_bytes(void);

/* Per-solve statistics of the last call on this thread. */
typedef struct {
    INT    iterations;        /* return value of the PCG */
    INT    num_levels;
    INT    level_rows[HZ_MAX_AMG_LVL];
    INT    level_nnz[HZ_MAX_AMG_LVL];
    REAL   setup_ms;          /* GPU wall time of the AMG setup */
    REAL   solve_ms;          /* GPU wall time of the PCG */
    REAL   total_ms;          /* setup + solve (device-resident inputs) */
    REAL   final_relres;
    REAL   condest;
    INT    relres_count;      /* entries valid in relres_hist (<= 4096) */
    REAL   relres_hist[4096]; /* relres after each PCG iteration */
    INT    fallback_used;     /* hz_krylov_amg_learned: 1 when the plain method produced x */
    REAL   pcg_spmv_ms;       /* with hz_options.time_pcg_spmv: mean duration of the PCG's fine-level
                                 SpMV launches in this solve (HIP events on the library's stream) */
    INT    pcg_spmv_count;
    INT    learned_iterations;  /* hz_krylov_amg_learned: the learned attempt's result (iterations or
                                   error code) and its device time, also when it fell back */
    REAL   learned_total_ms;
    INT    hem_pipeline_giveups;   /* HEM levels whose one-launch GPU pipeline stopped at its sweep
                                      cap and handed the rest to the per-round kernel or the host
                                      sweep (hem_pipeline_max_sweeps) */
    REAL   cheb_scale[HZ_MAX_AMG_LVL];   /* learned SA: Chebyshev bound scale in force per level
                                            (after the floor), 0 on levels without one */
    REAL   lambda_ratio[HZ_MAX_AMG_LVL]; /* learned SA: estimated lambda_max(D^-1 A) over the
                                            Gershgorin bound per level, 0 where not estimated
                                            (levels whose network scale is 1 or more: the
                                            floor, at most 1, cannot bind there); 30 Lanczos
                                            steps, run while the host aggregates the level */
    REAL   solve_alg_bytes;  /* algorithmic bytes of every kernel the PCG launched (each distinct
                                input read once, each output written once; the gathered coarse
                                corrections of the transfers not counted): over solve_ms, the
                                whole-solve bandwidth */
    REAL   cycle_residual_ms;  /* with hz_options.time_pcg_spmv: mean duration of the cycle's
                                  fine-level residual and of its last fine-level sweep (the one that */
    REAL   cycle_sweep_ms;     /* also forms (z, r)) in this solve; 0 where not launched */
} hz_solve_info;

void hz_get_last_info(hz_solve_info *info);

/* Device-resident drop-in: same semantics as linear_solver_dcsr_krylov_amg
 * but IA/JA/val, b and x are device pointers (hipMalloc'd).  Timed region of
 * the benchmark.  Returns iterations or a negative error code. */
INT hz_krylov_amg_device(INT n, INT nnz, const INT *dIA, const INT *dJA,
                         const REAL *dval, const REAL *db, REAL *dx,
                         linear_itsolver_param *itparam, AMG_param *amgparam);

/* Hierarchy export for parity checks: runs only the AMG setup of
 * linear_solver_dcsr_krylov_amg on the GPU and copies level data to host
 * buffers allocated with malloc (release with hz_hierarchy_free).  aggregates[l]
 * has level_rows[l] entries (vertex -> aggregate id, -1 = UNPT). */
typedef struct {
    INT num_levels;
    dCSRmat A[HZ_MAX_AMG_LVL];
    INT *aggregates[HZ_MAX_AMG_LVL];
    INT num_aggregates[HZ_MAX_AMG_LVL];
    REAL theta[HZ_MAX_AMG_LVL];   /* strong_coupled in force when level l was aggregated */
    dCSRmat P[HZ_MAX_AMG_LVL];    /* smoothed aggregation: prolongation of level l (empty for UA) */
} hz_hierarchy;

INT hz_setup_hierarchy(dCSRmat *A, AMG_param *amgparam, hz_hierarchy *out);
void hz_hierarchy_free(hz_hierarchy *h);

/* Single-kernel parity hooks on host arrays (upload, run one kernel, download).
 * y = A x  (dcsr_mxv, src/utilities/sparse.c:1456-1590). */
INT hz_dcsr_mxv(dCSRmat *A, const REAL *x, REAL *y);
/* y = y - A x ... residual form of dcsr_aAxpy(-1.0), sparse.c:1693-1730. */
INT hz_dcsr_residual(dCSRmat *A, const REAL *b, const REAL *x, REAL *w);
/* One Jacobi (smoother=1, weight w) or L1DIAG (smoother=10) sweep in place,
 * smoother.c:35-92 / 324-381. */
INT hz_smoother_sweep(dCSRmat *A, const REAL *b, REAL *x, INT smoother, REAL w);
/* Coarse product dcsr_rap_agg(R,A,P) with R=P^T built from aggregates
 * (sparse.c:2557-2683, sparse.c:770-830). Result arrays are malloc'd. */
INT hz_rap_agg(dCSRmat *A, const INT *aggregates, INT num_aggregates, dCSRmat *Ac);
/* In-place inverse of a dense row-major n x n host matrix with the kernels that
 * build the coarsest-level inverse of coarse_solver 32 (stands in for the
 * U^T D U factorization, hazmath_direct.c:633-854).  Returns
 * HZ_ERROR_SOLVER_MISC on a zero pivot. */
INT hz_dense_inverse(INT n, REAL *M);
/* 1 when the one-time lane-layout self-check of the FP64 MFMA GEMM passed (the
 * dense inverse runs its rank-64 updates on the matrix cores), 0 when it
 * failed: then every dense inverse returns HZ_ERROR_DEVICE. */
INT hz_dense_mfma_active(void);

/* Average device duration of the fine-level solve kernels on the library's
 * stream (hipEvents around `reps` back-to-back launches), for the roofline.
 * IA/JA/val are device pointers (same matrix as hz_krylov_amg_device). */
typedef struct {
    INT  reps;
    REAL spmv_ms;        /* y = A x              (dcsr_mxv) */
    REAL residual_ms;    /* w = b - A x          (array_cp + dcsr_aAxpy) */
    REAL jacobi_ms;      /* one Jacobi sweep     (smoother_dcsr_jacobi) */
    REAL spmv_dot_ms;    /* y = A x with (y, x)  (PCG step) */
    long long sell_entries;  /* padded entries streamed by the SELL-64 kernels */
} hz_kernel_times;

INT hz_time_fine_kernels(INT n, INT nnz, const INT *dIA, const INT *dJA, const REAL *dval,
                         INT reps, hz_kernel_times *out);

/* The fct installed by create_precond_amg (host arrays r, z of length A->row). */
void hz_precond_amg_apply(REAL *r, REAL *z, void *data);
/* Same action on device arrays (pc->data of a create_precond_amg result);
 * returns after the result is in dz. */
void hz_precond_amg_apply_device(const REAL *dr, REAL *dz, void *data);
/* Releases a preconditioner from create_precond_amg (device memory included);
 * the reference leaves this to its Python bindings. */
void hz_precond_free(precond *pc);

/* ---------------------------------------------------------------------------
 * Per-aggregate parameter network (no HAZmath counterpart): MLP 12 -> 64 -> 64
 * -> 2 with GELU in FP32, forward/backward on the FP32 matrix cores.  theta
 * holds HZ_MLP_PARAMS floats: W1[16][64] (input rows 12..15 unused), b1[64],
 * W2[64][64], b2[64], W3[64][16] (output columns 2..15 unused), b3[16].
 * Loss: mean squared error over the two outputs.  Host arrays throughout.
 * ------------------------------------------------------------------------- */
#define HZ_MLP_INPUTS   12
#define HZ_MLP_OUTPUTS  2
#define HZ_MLP_PARAMS   6288
/* Glorot-uniform weights from splitmix64(seed), zero biases and padding. */
void hz_mlp_init(float *theta, unsigned long long seed);
/* y[B][2] = MLP(x[B][12]). */
INT hz_mlp_forward(const float *theta, const float *x, INT B, float *y);
/* grad[HZ_MLP_PARAMS] = dLoss/dtheta and *loss for targets t[B][2]. */
INT hz_mlp_loss_grad(const float *theta, const float *x, const float *t, INT B, float *grad, float *loss);
/* Adam (lr, beta 0.9/0.999) over `epochs` shuffled passes with mini-batches of
 * `batch` rows (shuffles from splitmix64(seed)); theta is updated in place and
 * loss_hist[e] (epochs+1 entries) holds the full-data loss before epoch e and
 * after the last one. */
INT hz_mlp_train(float *theta, const float *x, const float *t, INT B, INT epochs, INT batch, float lr,
                 unsigned long long seed, float *loss_hist);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hz_mlp_forward with DEVICE pointers theta, x [B][12] and
 * y [B][2], queued on `stream` (a hipStream_t; NULL = the null stream), returning
 * without waiting.  It allocates nothing: its intermediate rows live in a
 * static device array of HZ_MLP_NOALLOC_MAX_ROWS rows (10 MiB, held on every
 * GPU the library runs on until the process ends).  B above that maximum
 * returns HZ_ERROR_INPUT_PAR.  All calls on one GPU share that array, so calls
 * that overlap on one GPU (different streams, or threads) race; calls queued
 * on one stream do not.  Same results as hz_mlp_forward, bit for bit. */
#define HZ_MLP_NOALLOC_MAX_ROWS 16384
INT hz_mlp_forward_noalloc(const float *theta, const float *x, INT B, float *y, void *stream);

/* ---------------------------------------------------------------------------
 * Learned-parameter AMG (no HAZmath counterpart).
 * ------------------------------------------------------------------------- */
/* PCG with smoothed aggregation whose per-aggregate prolongator damping and
 * per-level Chebyshev bound scale come from the parameter network (weights
 * theta, HZ_MLP_PARAMS floats; the scale is raised to at least
 * min(1, 1.1 lambda_est / Gershgorin bound), lambda_est from 60 power
 * iterations on D^-1 A, so the interval covers lambda_max), VMB aggregation and `cheb_steps` pre- and
 * post-smoothing steps of the HZ_SMOOTHER_CHEBY smoother (amgparam's strength
 * threshold, max_aggregation, coarse_dof and cycle are used).  If it
 * needs more than `budget` iterations, fails, or its FP64 true residual
 * ||b - A x|| exceeds itparam->linear_tol times the stop type's normaliser
 * (max(1e-20, ||b - A x0||) for stop type 1, ||x|| for stop type 3; stop
 * type 2 returns ERROR_INPUT_PAR), x is reset to the caller's guess and
 * the plain method of itparam/amgparam (linear_solver_dcsr_krylov_amg) solves
 * instead, its result returned as that solver returns it;
 * hz_solve_info.fallback_used records which.  Returns the iteration
 * count of the method that produced x (or a negative error code). */
INT hz_krylov_amg_learned(dCSRmat *A, dvector *b, dvector *x, linear_itsolver_param *itparam,
                          AMG_param *amgparam, const float *theta, INT budget, INT cheb_steps);
/* Two-grid error contraction for labelling: smoothed aggregation (uniform
 * damping omega, filter from amgparam) on two levels with an exact coarse
 * solve and the Chebyshev smoother (bound scale lscale); the cycle iterated
 * `iters` times on A x = 0 from a seeded random x_0; *rate =
 * (||x_K||_A/||x_{K/2}||_A)^(1/(K-K/2)) estimates the asymptotic factor. */
INT hz_two_grid_rate(dCSRmat *A, AMG_param *amgparam, REAL omega, REAL lscale, INT iters, REAL *rate);
/* Labelling with per-aggregate damping: the same contraction estimate for the
 * cycle of amgparam's whole SA hierarchy (Chebyshev smoother with bound scale
 * lscale on every level, V-cycle), with level 0's prolongator damped per
 * aggregate by omega_agg[nagg] (nagg must equal level 0's aggregate count,
 * else HZ_ERROR_INPUT_PAR) and coarser levels by amgparam->tentative_smooth. */
INT hz_cycle_rate_agg(dCSRmat *A, AMG_param *amgparam, const float *omega_agg, INT nagg, REAL lscale, INT iters,
                      REAL *rate);
/* Parity hooks: feat[nagg][12] of the aggregates of level `level`; `steps`
 * Chebyshev steps on x in place; the smoothed prolongation with strength
 * threshold theta, damping omega or per-aggregate omega_agg (may be NULL). */
INT hz_aggregate_features(dCSRmat *A, const INT *aggregates, INT num_aggregates, INT level, float *feat);
INT hz_cheby_smooth(dCSRmat *A, const REAL *b, REAL *x, INT steps, REAL scale);
/* Parity hook of the coarsest level's iterative solve (coarse_solver other
 * than 32; replaces coarse_itsolver, mgcycle.c:34-53): CG from x to the
 * relative tolerance ctol, and when it fails GMRES(30) with the variable
 * restart of dcsr_pvgmres (krylov.c:4612-4931) from CG's last iterate, both
 * with 20 n iterations at most, in one workgroup on the device.  Returns the
 * last method's result (iterations, or a negative code); cg_status and x_cg
 * (either may be NULL) receive CG's result and last iterate. */
INT hz_coarse_itsolver(dCSRmat *A, const REAL *b, REAL *x, REAL ctol, INT *cg_status, REAL *x_cg);
/* The estimate behind the learned method's Chebyshev floor: *lambda = the
 * largest Ritz value of `steps` Lanczos steps on D^-1/2 A D^-1/2 (D the first
 * diagonal entries; a lower bound of lambda_max(D^-1 A) for SPD A) and
 * *gershgorin = max_i sum_j |a_ij| / |a_ii| (an upper bound). */
INT hz_lambda_max_dinv(dCSRmat *A, INT steps, REAL *lambda, REAL *gershgorin);
INT hz_sa_prolongation(dCSRmat *A, const INT *aggregates, INT num_aggregates, REAL theta, REAL omega,
                       const float *omega_agg, INT filter, dCSRmat *P);
/* No counterpart in the reference: many independent systems, up to
 * HZ_BATCH_RHS right-hand sides each, through one AMG-PCG -- the semantics of
 * linear_solver_dcsr_krylov_amg (itsolver.c:846-931) per (system, right-hand
 * side) pair, with the V-cycle in FP32 and CG in FP64.  Settings: CG, stop
 * type 1, UA with HEM aggregation, V-cycle with JACOBI or L1DIAG, dense
 * coarsest solve (coarse_solver 32), maxit 1, no coarse scaling; anything
 * else returns ERROR_INPUT_PAR.  Host form: b[s] and x[s] hold nrhs vectors of
 * A[s]->row entries one after the other (x: initial guesses in, solutions
 * out); iters[s * nrhs + q] receives the iteration count of pair (s, q), or
 * ERROR_SOLVER_MAXIT / ERROR_SOLVER_TOLSMALL.  The caller's
 * amgparam->strong_coupled is not modified (each system keeps its own
 * threshold sequence internally).  Device form: the systems stacked into one
 * block-diagonal CSR (row_offsets[nsys+1], global column indices) and b, x as
 * n x HZ_BATCH_RHS row-major arrays (entry (i, q) at i * HZ_BATCH_RHS + q);
 * iters has nsys * HZ_BATCH_RHS entries. */
#define HZ_BATCH_RHS 8
/* Measurements of the last batched call (no counterpart in the reference). */
typedef struct hz_batch_stats {
    double setup_ms, solve_ms, op_complexity;
    INT num_levels, graph_launches, shape_level0;   /* shape_level0 = 10 x (lanes per row of the FP32
                                                       sweep) + (lanes per row of the FP64 SpMM) */
    double spmm_ms, sweep_ms;      /* level-0 FP64 SpMM and FP32 sweep, when timed */
    long long rows, nnz, sell_entries;
    INT sweep_shape0, spmm_shape0; /* level-0 kernel shapes: 100 x lanes per row + entries per
                                      load round (104, 108, 204, 208, 404 or 408) */
    INT tuned, tune_cached;        /* autotuner decisions timed in this call / read from its cache */
    double tune_ms;                /* host wall time of the timed decisions */
} hz_batch_stats;
void hz_batch_get_stats(hz_batch_stats *out);
/* The batched kernels' autotuner (no counterpart in the reference) times the
 * six shapes of each kernel on each level's matrix class (device, kernel,
 * rounded log2 rows, mean row length) once and keeps the fastest in a cache
 * that persists in a text file: this call's path, else the HZ_TUNE_CACHE
 * environment variable ("" or "off": memory only), else
 * ${XDG_CACHE_HOME:-$HOME/.cache}/hazmath_mi355x/tune.txt.  The call drops
 * the decisions in memory, reads those of the file at path (NULL: memory
 * only from now on) and returns how many it read.  HZ_BATCH_SHAPE and
 * HZ_BATCH_SPMM_SHAPE force a shape code.  Every shape gives bit-identical
 * results. */
INT hz_batch_tune_cache(const char *path);
INT hz_batch_krylov_amg(INT nsys, dCSRmat *const *A, INT nrhs, REAL *const *b, REAL *const *x,
                        linear_itsolver_param *itparam, AMG_param *amgparam, INT *iters);
INT hz_batch_krylov_amg_device(INT nsys, const INT *row_offsets, INT nnz, const INT *dIA, const INT *dJA,
                               const REAL *dval, const REAL *db, REAL *dx, linear_itsolver_param *itparam,
                               AMG_param *amgparam, INT *iters);
/* Parity hook: dcsr_presmoothing (post = 0) or dcsr_postsmoothing (post = 1)
 * of mgcycle.c:76-223 with a Gauss-Seidel-type smoother (2, 3, 5, 6, 7, 8) as
 * the library's multicolour sweeps, x updated in place; order[n] and cptr[n+1]
 * (may be NULL) receive the colour-major row order, colour c at positions
 * cptr[c] .. cptr[c+1]-1, and *ncolors the colour count. */
INT hz_gs_smooth(dCSRmat *A, const REAL *b, REAL *x, INT smoother, INT post, INT nsweeps, REAL relax, INT *order,
                 INT *cptr, INT *ncolors);

/* Runtime information. */
const char *hz_version(void);
INT hz_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* HAZMATH_MI355X_H */
