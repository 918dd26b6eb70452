/*
 * hz_oracle.c -- TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Serial CPU restatement of HAZmath's UA-AMG-preconditioned CG path, written
 * from the reference's semantics.  Each function names the reference function
 * and file:line it follows.  Built by oracle/Makefile with
 * -ffp-contract=off so that every per-row expression rounds exactly like the
 * reference built with gcc -O2 on x86-64 (no FMA contraction).
 *
 * Pinning (see tests/test_oracle_golden.py):
 *  - level sizes and PCG iteration counts recorded from the compiled reference
 *    on examples/solvers/A.dat + b.dat (VMB/HEC/HEM x JACOBI/L1DIAG);
 *  - closed forms and brute force on small inputs for the kernels;
 *  - the coarsest UᵀDU solve (hazmath_direct.c:633-854) is restated as a dense
 *    LDLᵀ solve: equal to rounding, not bit-identical (parity unpinned at the
 *    last-bit level, pinned by residual checks);
 *  - the GMRES fallback of the coarse CG (mgcycle.c:45-47, dcsr_pvgmres
 *    krylov.c:4612-4931) is pinned by the minimal-residual property of its
 *    first restart cycle (least squares over the Krylov space, numpy) and by
 *    convergence on nonsymmetric systems where the CG fails.
 */
#include "hz_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define ORC_SMALLREAL 1e-20
#define ORC_BIGREAL   1e+20
#define ORC_MAX_STAG      20
#define ORC_MAX_RESTART   20
#define ORC_STAG_RATIO  1e-4
#define ORC_MIN_CDOF       1

static double orc_now(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}

static void csr_free(dCSRmat *A) {
    free(A->IA); free(A->JA); free(A->val);
    A->IA = NULL; A->JA = NULL; A->val = NULL;
    A->row = A->col = A->nnz = 0;
}

static void csr_copy(const dCSRmat *A, dCSRmat *B) {
    B->row = A->row; B->col = A->col; B->nnz = A->nnz;
    B->IA = (INT *)xcalloc((size_t)A->row + 1, sizeof(INT));
    B->JA = (INT *)xcalloc((size_t)A->nnz, sizeof(INT));
    B->val = (REAL *)xcalloc((size_t)A->nnz, sizeof(REAL));
    memcpy(B->IA, A->IA, sizeof(INT) * ((size_t)A->row + 1));
    memcpy(B->JA, A->JA, sizeof(INT) * (size_t)A->nnz);
    memcpy(B->val, A->val, sizeof(REAL) * (size_t)A->nnz);
}

/* Counting-sort transpose: row i of AT lists source rows in increasing order.
 * Follows dcsr_trans / dcsr_transz (sparse.c:770-830, 864-965). */
static void csr_transpose(const dCSRmat *A, dCSRmat *AT) {
    const INT n = A->row, m = A->col, nnz = A->IA[A->row];
    AT->row = m; AT->col = n; AT->nnz = nnz;
    AT->IA = (INT *)xcalloc((size_t)m + 1, sizeof(INT));
    AT->JA = (INT *)xcalloc((size_t)nnz, sizeof(INT));
    AT->val = (REAL *)xcalloc((size_t)nnz, sizeof(REAL));
    for (INT k = 0; k < nnz; ++k) AT->IA[A->JA[k] + 1]++;
    for (INT i = 0; i < m; ++i) AT->IA[i + 1] += AT->IA[i];
    INT *next = (INT *)xcalloc((size_t)m + 1, sizeof(INT));
    memcpy(next, AT->IA, sizeof(INT) * ((size_t)m + 1));
    for (INT i = 0; i < n; ++i)
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
            INT pos = next[A->JA[k]]++;
            AT->JA[pos] = i;
            AT->val[pos] = A->val[k];
        }
    free(next);
}

/* ------------------------------------------------------------------ */
/* Parameters: parameters.c:158-205, 238-251                           */
/* ------------------------------------------------------------------ */
void orc_param_amg_init(AMG_param *p) {
    memset(p, 0, sizeof(*p));
    p->AMG_type = HZ_UA_AMG;
    p->print_level = 0;
    p->maxit = 1;
    p->tol = 1e-6;
    p->max_levels = 10;
    p->coarse_dof = 200;
    p->cycle_type = HZ_W_CYCLE;
    p->smoother = HZ_SMOOTHER_GS;
    p->presmooth_iter = 1;
    p->postsmooth_iter = 1;
    p->coarse_solver = HZ_SOLVER_DEFAULT;
    p->relaxation = 1.0;
    p->polynomial_degree = 2;
    p->coarse_scaling = 0;
    p->amli_degree = 2;
    p->amli_coef = NULL;
    p->nl_amli_krylov_type = 4;
    p->fpwr = 1.0;
    p->aggregation_type = HZ_HEC;
    p->strong_coupled = 0.04;
    p->max_aggregation = 20;
    p->tentative_smooth = 0.67;
    p->smooth_filter = 1;
    p->Schwarz_levels = 0;
    p->Schwarz_mmsize = 200;
    p->Schwarz_maxlvl = -1;
    p->Schwarz_type = 1;
    p->Schwarz_blksolver = HZ_SOLVER_UMFPACK;
    p->HAZDIR = NULL;
    p->Schwarz_on_blk = NULL;
    p->Schwarz_patch_type = NULL;
    p->damping_param = 1.0;
    p->BSR_alpha = -1000.;
    p->BSR_omega = -1000.;
}

void orc_param_linear_solver_init(linear_itsolver_param *p) {
    memset(p, 0, sizeof(*p));
    p->linear_print_level = 0;
    p->linear_itsolver_type = 4;
    p->linear_precond_type = 0;
    p->linear_stop_type = HZ_STOP_REL_RES;
    p->linear_maxit = 500;
    p->linear_restart = 100;
    p->linear_tol = 1e-6;
    p->HX_smooth_iter = 1;
}

/* ------------------------------------------------------------------ */
/* Kernels: sparse.c / array.c / smoother.c                            */
/* ------------------------------------------------------------------ */

/* y = A x, row sums in storage order starting from 0.0 (sparse.c:1456-1590). */
void orc_dcsr_mxv(const dCSRmat *A, const REAL *x, REAL *y) {
    for (INT i = 0; i < A->row; ++i) {
        REAL t = 0.0;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) t += A->val[k] * x[A->JA[k]];
        y[i] = t;
    }
}

/* w = b; w -= A x  (array_cp + dcsr_aAxpy(-1.0), mgcycle.c:294-296, sparse.c:1711-1718). */
void orc_dcsr_residual(const dCSRmat *A, const REAL *b, const REAL *x, REAL *w) {
    for (INT i = 0; i < A->row; ++i) {
        REAL t = 0.0;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) t += A->val[k] * x[A->JA[k]];
        w[i] = b[i] - t;
    }
}

/* y += alpha A x with the alpha==1 / -1 special cases (sparse.c:1693-1730). */
static void orc_dcsr_aAxpy(REAL alpha, const dCSRmat *A, const REAL *x, REAL *y) {
    for (INT i = 0; i < A->row; ++i) {
        REAL t = 0.0;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) t += A->val[k] * x[A->JA[k]];
        if (alpha == 1.0) y[i] += t;
        else if (alpha == -1.0) y[i] -= t;
        else y[i] += t * alpha;
    }
}

/* y^T A x (sparse.c:1797-1816). */
static REAL orc_dcsr_vmv(const dCSRmat *A, const REAL *x, const REAL *y) {
    REAL v = 0.0;
    for (INT i = 0; i < A->row; ++i) {
        REAL t = 0.0;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) t += A->val[k] * x[A->JA[k]];
        v += y[i] * t;
    }
    return v;
}

static REAL orc_dot(INT n, const REAL *x, const REAL *y) {   /* array.c:448-461 */
    REAL v = 0.0;
    for (INT i = 0; i < n; ++i) v += x[i] * y[i];
    return v;
}
static REAL orc_norm2(INT n, const REAL *x) {                /* array.c:498-507 */
    REAL v = 0.0;
    for (INT i = 0; i < n; ++i) v += x[i] * x[i];
    return sqrt(v);
}
static REAL orc_norminf(INT n, const REAL *x) {              /* array.c:521-530 */
    REAL v = 0.0;
    for (INT i = 0; i < n; ++i) { REAL a = fabs(x[i]); v = (v > a) ? v : a; }
    return v;
}
static void orc_axpy(INT n, REAL a, const REAL *x, REAL *y) {  /* array.c:357-373 */
    if (a == 1.0) { for (INT i = 0; i < n; ++i) y[i] += x[i]; }
    else if (a == -1.0) { for (INT i = 0; i < n; ++i) y[i] -= x[i]; }
    else { for (INT i = 0; i < n; ++i) y[i] += a * x[i]; }
}
static void orc_axpby(INT n, REAL a, const REAL *x, REAL b, REAL *y) {  /* array.c:423-433 */
    for (INT i = 0; i < n; ++i) y[i] = a * x[i] + b * y[i];
}

/* One Jacobi sweep with weight w (smoother.c:35-92, w fixed to 1.0 there) or one
 * L1-diagonal sweep (smoother.c:324-381).  The sweep direction does not change
 * the result (all t_i are formed from the old iterate first). */
void orc_smoother_sweep(const dCSRmat *A, const REAL *b, REAL *x, INT smoother, REAL w) {
    const INT n = A->row;
    REAL *t = (REAL *)xcalloc((size_t)n, sizeof(REAL));
    REAL *d = (REAL *)xcalloc((size_t)n, sizeof(REAL));
    if (smoother == HZ_SMOOTHER_L1DIAG) {
        for (INT i = 0; i < n; ++i) {
            t[i] = b[i]; d[i] = 0.0;
            for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
                t[i] -= A->val[k] * x[A->JA[k]];
                d[i] += fabs(A->val[k]);
            }
        }
        for (INT i = 0; i < n; ++i)
            if (fabs(d[i]) > ORC_SMALLREAL) x[i] += t[i] / d[i];
    } else {
        for (INT i = 0; i < n; ++i) {
            t[i] = b[i];
            for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
                INT j = A->JA[k];
                if (i != j) t[i] -= A->val[k] * x[j];
                else d[i] = A->val[k];
            }
        }
        for (INT i = 0; i < n; ++i)
            if (fabs(d[i]) > ORC_SMALLREAL) x[i] = (1 - w) * x[i] + w * t[i] / d[i];
    }
    free(t); free(d);
}

/* Gauss-Seidel-type sweeps, smoother_dcsr_gs (smoother.c:109-190),
 * smoother_dcsr_sgs (smoother.c:205-250) and smoother_dcsr_sor
 * (smoother.c:266-307), generalised to a row ordering: `ord` lists the rows
 * in blocks ord[cptr[c] .. cptr[c+1]), c < nblk.  A forward sweep visits the
 * blocks in increasing order, a backward sweep in decreasing order, rows
 * inside a block in list order.  ord == NULL is the natural order with one
 * row per block, which is exactly the reference's loop i = 0..n-1 (forward)
 * or n-1..0 (backward).  When the rows of every block are pairwise uncoupled
 * (a colouring), the order inside a block does not change the result.
 *   mode 0  (GS)  d = 1/a_ii when |a_ii| > SMALLREAL, x_i = t * d; d keeps its
 *                 last value across rows and sweeps of one call, as in
 *                 smoother.c:140-146;
 *   mode 1  (SGS) x_i = t / a_ii when |a_ii| > SMALLREAL;
 *   mode 2  (SOR) x_i = w (t / a_ii) + (1 - w) x_i when |a_ii| > SMALLREAL.
 * skip_first: the backward half of SGS starts one block early (n-2 in the
 * natural order, smoother.c:236). */
typedef struct {
    const INT *ord, *cptr;
    INT nblk;
} orc_order;

static void orc_gs_row(const dCSRmat *A, const REAL *b, REAL *x, INT i, int mode, REAL w, REAL *dstate) {
    REAL t = b[i], d = (mode == 0) ? *dstate : 0.0;
    for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
        const INT j = A->JA[k];
        if (i != j) t -= A->val[k] * x[j];
        else if (mode == 0) { if (fabs(A->val[k]) > ORC_SMALLREAL) d = 1.e+0 / A->val[k]; }
        else d = A->val[k];
    }
    if (mode == 0) { x[i] = t * d; *dstate = d; }
    else if (fabs(d) > ORC_SMALLREAL) x[i] = (mode == 1) ? t / d : w * (t / d) + (1 - w) * x[i];
}

static void orc_gs_pass(const dCSRmat *A, const REAL *b, REAL *x, const orc_order *o, int forward,
                        int skip_first, int mode, REAL w, REAL *dstate) {
    const INT nb = o->ord ? o->nblk : A->row;
    for (INT q = skip_first; q < nb; ++q) {
        const INT c = forward ? q : nb - 1 - q;
        if (!o->ord) { orc_gs_row(A, b, x, c, mode, w, dstate); continue; }
        for (INT p = o->cptr[c]; p < o->cptr[c + 1]; ++p) orc_gs_row(A, b, x, o->ord[p], mode, w, dstate);
    }
}

/* smoother_dcsr_gs / smoother_dcsr_sor with L sweeps in one direction */
static void orc_gs_call(const dCSRmat *A, const REAL *b, REAL *x, const orc_order *o, int forward, int mode,
                        REAL w, INT L) {
    REAL dstate = 0.0;   /* smoother.c:124: d = 0.0 at entry */
    while (L-- > 0) orc_gs_pass(A, b, x, o, forward, 0, mode, w, &dstate);
}

/* smoother_dcsr_sgs: L times (forward sweep, backward sweep from n-2) */
static void orc_sgs_call(const dCSRmat *A, const REAL *b, REAL *x, const orc_order *o, INT L) {
    REAL dstate = 0.0;
    while (L-- > 0) {
        orc_gs_pass(A, b, x, o, 1, 0, 1, 1.0, &dstate);
        orc_gs_pass(A, b, x, o, 0, 1, 1, 1.0, &dstate);
    }
}

/* dcsr_presmoothing (mgcycle.c:76-137, called with istart 0, iend n-1,
 * istep 1) and dcsr_postsmoothing (mgcycle.c:162-223, called with istart 0,
 * iend n-1, istep -1) for smoothers GS 2, SGS 3, SOR 5, SSOR 6, GSOR 7 and
 * SGSOR 8.  Returns 0, or -1 for any other smoother. */
int orc_gs_smooth_ordered(const dCSRmat *A, const REAL *b, REAL *x, INT smoother, INT post, INT nsweeps,
                          REAL relax, const INT *ord, const INT *cptr, INT nblk) {
    orc_order o = {ord, cptr, nblk};
    const int F = 1, B = 0;
    switch (smoother) {
        case 2:   /* GS: forward before, backward after */
            orc_gs_call(A, b, x, &o, post ? B : F, 0, 1.0, nsweeps);
            break;
        case 3:
            orc_sgs_call(A, b, x, &o, nsweeps);
            break;
        case 5:
            orc_gs_call(A, b, x, &o, post ? B : F, 2, relax, nsweeps);
            break;
        case 6:   /* SSOR: forward then backward, before and after */
            orc_gs_call(A, b, x, &o, F, 2, relax, nsweeps);
            orc_gs_call(A, b, x, &o, B, 2, relax, nsweeps);
            break;
        case 7:   /* GSOR */
            if (!post) {
                orc_gs_call(A, b, x, &o, F, 0, 1.0, nsweeps);
                orc_gs_call(A, b, x, &o, B, 2, relax, nsweeps);
            } else {
                orc_gs_call(A, b, x, &o, F, 2, relax, nsweeps);
                orc_gs_call(A, b, x, &o, B, 0, 1.0, nsweeps);
            }
            break;
        case 8:   /* SGSOR */
            if (!post) {
                orc_gs_call(A, b, x, &o, F, 0, 1.0, nsweeps);
                orc_gs_call(A, b, x, &o, B, 0, 1.0, nsweeps);
                orc_gs_call(A, b, x, &o, F, 2, relax, nsweeps);
                orc_gs_call(A, b, x, &o, B, 2, relax, nsweeps);
            } else {
                orc_gs_call(A, b, x, &o, F, 2, relax, nsweeps);
                orc_gs_call(A, b, x, &o, B, 2, relax, nsweeps);
                orc_gs_call(A, b, x, &o, F, 0, 1.0, nsweeps);
                orc_gs_call(A, b, x, &o, B, 0, 1.0, nsweeps);
            }
            break;
        default:
            return -1;
    }
    return 0;
}

int orc_is_gs_family(INT smoother) {
    return smoother == 2 || smoother == 3 || smoother == 5 || smoother == 6 || smoother == 7 || smoother == 8;
}

/* ------------------------------------------------------------------ */
/* Strength graph: construct_strongly_coupled, amg_setup_ua.c:375-452  */
/* ------------------------------------------------------------------ */
void orc_strength(const dCSRmat *A, REAL theta, dCSRmat *N) {
    const INT n = A->row;
    REAL th2 = theta * theta;
    /* diagonal: first occurrence in the row, 0.0 when absent (dcsr_getdiag, sparse.c:2168-2189) */
    INT nd = A->row < A->col ? A->row : A->col;
    REAL *d = (REAL *)xcalloc((size_t)nd, sizeof(REAL));
    for (INT i = 0; i < nd; ++i)
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k)
            if (A->JA[k] == i) { d[i] = A->val[k]; break; }
    N->row = A->row; N->col = A->col;
    N->IA = (INT *)xcalloc((size_t)n + 1, sizeof(INT));
    INT cap = A->IA[n] - A->IA[0];
    N->JA = (INT *)xcalloc((size_t)cap, sizeof(INT));
    N->val = (REAL *)xcalloc((size_t)cap, sizeof(REAL));
    INT pos = 0;
    for (INT i = 0; i < n; ++i) {
        N->IA[i] = pos;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
            INT j = A->JA[k];
            REAL a = A->val[k];
            if (j == i || ((a * a) >= th2 * fabs(d[i] * d[j]) && a < 0e0)) {
                N->JA[pos] = j; N->val[pos] = a; ++pos;
            }
        }
    }
    N->IA[n] = pos;
    N->nnz = pos;
    free(d);
}

/* ------------------------------------------------------------------ */
/* Aggregation                                                         */
/* ------------------------------------------------------------------ */

/* VMB three-pass greedy: aggregation_vmb, amg_setup_ua.c:596-727. */
static INT orc_agg_vmb(const dCSRmat *N, INT maxagg, INT *v, INT *nagg) {
    const INT n = N->row;
    const INT *ia = N->IA, *ja = N->JA;
    INT left = n, na = 0;
    for (INT i = 0; i < n; ++i) v[i] = -2;
    /* pass 1: reverse index order; a vertex whose whole strong neighbourhood is
     * untouched becomes a root and takes up to maxagg-1 neighbours */
    for (INT i = n - 1; i >= 0; --i) {
        if (ia[i + 1] - ia[i] == 1) { v[i] = HZ_UNPT; --left; continue; }
        int free_nbhd = 1;
        for (INT k = ia[i]; k < ia[i + 1]; ++k)
            if (v[ja[k]] >= HZ_UNPT) { free_nbhd = 0; break; }
        if (!free_nbhd) continue;
        INT cnt = 1;
        v[i] = na; --left;
        for (INT k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] != i && cnt < maxagg) { v[ja[k]] = na; --left; ++cnt; }
        ++na;
    }
    *nagg = na;
    if (na < ORC_MIN_CDOF) return HZ_ERROR_AMG_COARSEING;
    /* pass 2: reverse order; untouched vertices join the first neighbouring
     * pass-1 aggregate that still has room */
    INT *c1 = (INT *)xcalloc((size_t)n, sizeof(INT));
    INT *size = (INT *)xcalloc((size_t)na, sizeof(INT));
    for (INT i = 0; i < n; ++i) { c1[i] = v[i]; if (v[i] >= 0) size[v[i]]++; }
    for (INT i = n - 1; i >= 0; --i) {
        if (v[i] >= HZ_UNPT) continue;
        for (INT k = ia[i]; k < ia[i + 1]; ++k) {
            INT a = c1[ja[k]];
            if (a > HZ_UNPT && size[a] < maxagg) { v[i] = a; --left; size[a]++; break; }
        }
    }
    /* pass 3: forward order; each remaining vertex seeds a new aggregate */
    while (left > 0) {
        for (INT i = 0; i < n; ++i) {
            if (v[i] >= HZ_UNPT) continue;
            INT cnt = 1;
            v[i] = na; --left;
            for (INT k = ia[i]; k < ia[i + 1]; ++k)
                if (ja[k] != i && v[ja[k]] < HZ_UNPT && cnt < maxagg) { v[ja[k]] = na; --left; ++cnt; }
            ++na;
        }
    }
    *nagg = na;
    free(c1); free(size);
    return HZ_SUCCESS;
}

/* HEC: random turn order from libc rand() (iarray_shuffle, array.c:302-318), each
 * vertex pairs with or joins its strongest neighbour (aggregation_hec,
 * amg_setup_ua.c:752-845). */
static INT orc_agg_hec(const dCSRmat *N, INT *v, INT *nagg) {
    const INT n = N->row;
    const INT *ia = N->IA, *ja = N->JA;
    const REAL *a = N->val;
    INT *perm = (INT *)xcalloc((size_t)n, sizeof(INT));
    for (INT i = 0; i < n; ++i) { perm[i] = i; v[i] = -2; }
    if (n > 1)
        for (INT i = 0; i < n - 1; ++i) {
            INT j = i + rand() / (RAND_MAX / (n - i) + 1);
            INT t = perm[j]; perm[j] = perm[i]; perm[i] = t;
        }
    INT na = 0;
    for (INT ii = 0; ii < n; ++ii) {
        INT i = perm[ii];
        if (ia[i + 1] - ia[i] == 1) { v[i] = HZ_UNPT; continue; }
        INT best = -1;
        REAL m = 0.0;
        for (INT k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] != i && fabs(a[k]) > m) { best = k; m = fabs(a[k]); }
        if (best < 0) {
            m = -1e0;
            for (INT k = ia[i]; k < ia[i + 1]; ++k)
                if (ja[k] != i && fabs(a[k]) > m) { best = k; m = fabs(a[k]); }
        }
        if (best < 0) {
            fprintf(stderr, "\n\n%%%% ERROR Stop: negative index encountered in %s()\n\n", "aggregation_hec");
            free(perm);
            return ORC_EXIT_HEC;
        }
        INT j = ja[best];
        if (v[j] < HZ_UNPT) { v[j] = na; v[i] = na; ++na; }
        else v[i] = v[j];
    }
    *nagg = na;
    free(perm);
    return HZ_SUCCESS;
}

/* heavy_edg (amg_setup_ua.c:864-891): scan from n0 towards n1, keep the first
 * element and replace it only on a strictly larger weight. */
static INT orc_heavy(const REAL *w, const INT *id, INT n0, INT n1) {
    if (n0 < 0 || n1 < 0) return -1;
    INT step = (n0 > n1) ? -1 : 1, end = n1 + step;
    INT best = id[n0];
    REAL bw = w[n0];
    for (INT j = n0; j != end; j += step)
        if (w[j] > bw && id[j] >= 0) { bw = w[j]; best = id[j]; }
    return best;
}

/* HEM: index-order greedy heavy-edge matching, leftovers join the smallest
 * neighbouring aggregate (aggregation_hem, amg_setup_ua.c:913-1051). */
static INT orc_agg_hem(const dCSRmat *N, INT *v, INT *nagg) {
    const INT n = N->row;
    const INT *ia = N->IA, *ja = N->JA;
    const REAL *a = N->val;
    INT maxdeg = 0;
    for (INT k = 0; k < n; ++k) {
        INT l = ia[k + 1] - ia[k];
        if (l > maxdeg) maxdeg = l;
        v[k] = (l > 1) ? -1 : -2;
    }
    REAL *w = (REAL *)xcalloc((size_t)maxdeg, sizeof(REAL));
    INT *id = (INT *)xcalloc((size_t)maxdeg, sizeof(INT));
    INT *size = (INT *)xcalloc((size_t)n, sizeof(INT));
    INT nc = 0;
    for (INT k = 0; k < n; ++k) {
        if (v[k] != -1) continue;
        INT iz = 0;
        for (INT jk = ia[k]; jk < ia[k + 1]; ++jk) {
            INT j = ja[jk];
            if (v[j] < 0 && j != k) { w[iz] = fabs(a[jk]); id[iz] = j; ++iz; }
        }
        if (iz) {
            INT pick = orc_heavy(w, id, iz - 1, 0);
            v[k] = nc; v[pick] = nc; size[nc] += 2; ++nc;
        }
    }
    for (INT k = 0; k < n; ++k) {
        if (v[k] != -1) continue;
        INT iz = 0;
        for (INT jk = ia[k]; jk < ia[k + 1]; ++jk) {
            INT c = v[ja[jk]];
            if (c >= 0) { w[iz] = (REAL)(-size[c]); id[iz] = c; ++iz; }
        }
        if (!iz) {
            fprintf(stderr, "%%%%ERROR in %s: isolated point (=%d) on the second matching pass", "aggregation_hem", k);
            free(w); free(id); free(size);
            return ORC_EXIT_HEM;
        }
        INT c = orc_heavy(w, id, 0, iz - 1);
        v[k] = c; size[c]++;
    }
    for (INT k = 0; k < n; ++k) if (v[k] < 0) v[k] = HZ_UNPT;
    *nagg = nc;
    free(w); free(id); free(size);
    return HZ_SUCCESS;
}

INT orc_aggregate(const dCSRmat *A, INT type, REAL theta, INT max_aggregation,
                  INT *agg, INT *nagg) {
    dCSRmat N;
    orc_strength(A, theta, &N);
    INT st;
    *nagg = 0;
    switch (type) {
        case HZ_VMB: st = orc_agg_vmb(&N, max_aggregation, agg, nagg); break;
        case HZ_HEC: st = orc_agg_hec(&N, agg, nagg); break;
        case HZ_HEM: st = orc_agg_hem(&N, agg, nagg); break;
        default: st = HZ_ERROR_AMG_AGG_TYPE; break;
    }
    csr_free(&N);
    return st;
}

/* ------------------------------------------------------------------ */
/* Galerkin product with unit R = P^T: dcsr_rap_agg, sparse.c:2557-2683 */
/* Row ic: diagonal slot first (starts at 0.0), then coarse columns in  */
/* the order first met while scanning members ascending x A-row order; */
/* a new slot starts at the contributing value, later ones are added.  */
/* ------------------------------------------------------------------ */
INT orc_rap_agg(const dCSRmat *A, const INT *agg, INT nagg, dCSRmat *Ac) {
    const INT n = A->row;
    /* R = P^T: members of each aggregate in ascending fine index (dcsr_trans) */
    INT *rp = (INT *)xcalloc((size_t)nagg + 1, sizeof(INT));
    INT *rm = (INT *)xcalloc((size_t)n, sizeof(INT));
    for (INT i = 0; i < n; ++i) if (agg[i] > HZ_UNPT) rp[agg[i] + 1]++;
    for (INT c = 0; c < nagg; ++c) rp[c + 1] += rp[c];
    INT *fill = (INT *)xcalloc((size_t)nagg + 1, sizeof(INT));
    memcpy(fill, rp, sizeof(INT) * ((size_t)nagg + 1));
    for (INT i = 0; i < n; ++i) if (agg[i] > HZ_UNPT) rm[fill[agg[i]]++] = i;
    free(fill);

    INT *slot = (INT *)xcalloc((size_t)nagg, sizeof(INT));
    for (INT c = 0; c < nagg; ++c) slot[c] = -1;
    /* symbolic pass */
    INT *ia = (INT *)xcalloc((size_t)nagg + 1, sizeof(INT));
    INT cnt = 0;
    for (INT ic = 0; ic < nagg; ++ic) {
        INT start = cnt;
        slot[ic] = cnt++;
        for (INT q = rp[ic]; q < rp[ic + 1]; ++q) {
            INT i1 = rm[q];
            for (INT k = A->IA[i1]; k < A->IA[i1 + 1]; ++k) {
                INT c = agg[A->JA[k]];
                if (c <= HZ_UNPT) continue;
                if (slot[c] < start) slot[c] = cnt++;
            }
        }
        ia[ic] = start;
    }
    ia[nagg] = cnt;
    Ac->row = nagg; Ac->col = nagg; Ac->nnz = cnt; Ac->IA = ia;
    Ac->JA = (INT *)xcalloc((size_t)cnt, sizeof(INT));
    Ac->val = (REAL *)xcalloc((size_t)cnt, sizeof(REAL));
    for (INT c = 0; c < nagg; ++c) slot[c] = -1;
    /* numeric pass */
    cnt = 0;
    for (INT ic = 0; ic < nagg; ++ic) {
        INT start = cnt;
        slot[ic] = cnt;
        Ac->JA[cnt] = ic; Ac->val[cnt] = 0.0; ++cnt;
        for (INT q = rp[ic]; q < rp[ic + 1]; ++q) {
            INT i1 = rm[q];
            for (INT k = A->IA[i1]; k < A->IA[i1 + 1]; ++k) {
                INT c = agg[A->JA[k]];
                if (c <= HZ_UNPT) continue;
                if (slot[c] < start) {
                    slot[c] = cnt;
                    Ac->JA[cnt] = c; Ac->val[cnt] = A->val[k]; ++cnt;
                } else {
                    Ac->val[slot[c]] += A->val[k];
                }
            }
        }
    }
    free(slot); free(rp); free(rm);
    return HZ_SUCCESS;
}

/* ------------------------------------------------------------------ */
/* Hierarchy                                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    dCSRmat A;
    dCSRmat P, R; /* smoothed aggregation: P = S P_tent and R = P^T (empty for UA) */
    INT *agg;     /* P: row i -> column agg[i], empty when agg[i] == UNPT */
    INT nagg;
    INT *rp, *rm; /* R = P^T rows (members ascending) */
    REAL *b, *x, *w;
    REAL *L;      /* coarsest: dense LDL^T factor (unit lower, diag in D) */
    REAL *D;
    REAL theta;
} orc_level;

typedef struct {
    INT nl;
    orc_level lv[HZ_MAX_AMG_LVL];
} orc_amg;

static void orc_amg_free(orc_amg *m) {
    for (INT l = 0; l < HZ_MAX_AMG_LVL; ++l) {
        orc_level *L = &m->lv[l];
        csr_free(&L->A);
        csr_free(&L->P);
        csr_free(&L->R);
        free(L->agg); free(L->rp); free(L->rm);
        free(L->b); free(L->x); free(L->w); free(L->L); free(L->D);
        memset(L, 0, sizeof(*L));
    }
    m->nl = 0;
}

/* Dense LDL^T (= U^T D U with U = L^T) of the coarsest matrix, standing in for
 * run_hazmath_factorize (hazmath_direct.c:633-732, identity permutation). */
static void orc_coarse_factor(orc_level *L) {
    const INT n = L->A.row;
    REAL *M = (REAL *)xcalloc((size_t)n * n, sizeof(REAL));
    for (INT i = 0; i < n; ++i)
        for (INT k = L->A.IA[i]; k < L->A.IA[i + 1]; ++k) M[(size_t)i * n + L->A.JA[k]] += L->A.val[k];
    REAL *D = (REAL *)xcalloc((size_t)n, sizeof(REAL));
    for (INT j = 0; j < n; ++j) {
        REAL dj = M[(size_t)j * n + j];
        for (INT k = 0; k < j; ++k) dj -= M[(size_t)j * n + k] * M[(size_t)j * n + k] * D[k];
        D[j] = dj;
        for (INT i = j + 1; i < n; ++i) {
            REAL s = M[(size_t)i * n + j];
            for (INT k = 0; k < j; ++k) s -= M[(size_t)i * n + k] * M[(size_t)j * n + k] * D[k];
            M[(size_t)i * n + j] = s / dj;
        }
    }
    L->L = M; L->D = D;
}

static void orc_coarse_solve(const orc_level *L, const REAL *b, REAL *x) {
    const INT n = L->A.row;
    const REAL *M = L->L;
    for (INT i = 0; i < n; ++i) {
        REAL s = b[i];
        for (INT k = 0; k < i; ++k) s -= M[(size_t)i * n + k] * x[k];
        x[i] = s;
    }
    for (INT i = 0; i < n; ++i) x[i] /= L->D[i];
    for (INT i = n - 1; i >= 0; --i) {
        REAL s = x[i];
        for (INT k = i + 1; k < n; ++k) s -= M[(size_t)k * n + i] * x[k];
        x[i] = s;
    }
}

/* ------------------------------------------------------------------ */
/* Smoothed aggregation (amg_setup_smoothP_smoothR, amg_setup_ua.c:1461-1687) */
/* ------------------------------------------------------------------ */

/* dcsr_mxm (sparse.c:1832-1918): row i of C lists the columns of B reached
 * from row i of A in order of first discovery (A row order, then B row
 * order); C[i][j] = 0 + sum of a_ik b_kj in the same order. */
static void orc_mxm(const dCSRmat *A, const dCSRmat *B, dCSRmat *C) {
    const INT n = A->row;
    INT *mark = (INT *)xcalloc((size_t)B->col + 1, sizeof(INT));
    for (INT j = 0; j < B->col; ++j) mark[j] = -1;
    C->row = n; C->col = B->col;
    C->IA = (INT *)xcalloc((size_t)n + 1, sizeof(INT));
    INT cap = 16, pos = 0;
    C->JA = (INT *)xcalloc((size_t)cap, sizeof(INT));
    C->val = (REAL *)xcalloc((size_t)cap, sizeof(REAL));
    for (INT i = 0; i < n; ++i) {
        INT row0 = pos;
        C->IA[i] = pos;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
            INT r = A->JA[k];
            for (INT l = B->IA[r]; l < B->IA[r + 1]; ++l) {
                INT j = B->JA[l];
                if (mark[j] < row0) {
                    if (pos == cap) {
                        cap *= 2;
                        C->JA = (INT *)realloc(C->JA, sizeof(INT) * (size_t)cap);
                        C->val = (REAL *)realloc(C->val, sizeof(REAL) * (size_t)cap);
                    }
                    mark[j] = pos;
                    C->JA[pos] = j;
                    C->val[pos] = 0.0;
                    ++pos;
                }
                C->val[mark[j]] += A->val[k] * B->val[l];
            }
        }
    }
    C->IA[n] = pos;
    C->nnz = pos;
    free(mark);
}

/* dcsr_rap (sparse.c:2401-2538): row ic of RAP starts with ic (value 0.0 plus
 * its products), then the coarse columns in order of first discovery over
 * (R row ic, A row, P row); every entry accumulates its products
 * (r_ij a_jk) p_kl in that traversal order, the first one assigned. */
static void orc_rap(const dCSRmat *R, const dCSRmat *A, const dCSRmat *P, dCSRmat *C) {
    const INT nc = R->row;
    INT *mark = (INT *)xcalloc((size_t)P->col + 1, sizeof(INT));
    for (INT j = 0; j < P->col; ++j) mark[j] = -1;
    C->row = nc; C->col = P->col;
    C->IA = (INT *)xcalloc((size_t)nc + 1, sizeof(INT));
    INT cap = 16, pos = 0;
    C->JA = (INT *)xcalloc((size_t)cap, sizeof(INT));
    C->val = (REAL *)xcalloc((size_t)cap, sizeof(REAL));
    for (INT ic = 0; ic < nc; ++ic) {
        INT row0 = pos;
        C->IA[ic] = pos;
        if (pos == cap) {
            cap *= 2;
            C->JA = (INT *)realloc(C->JA, sizeof(INT) * (size_t)cap);
            C->val = (REAL *)realloc(C->val, sizeof(REAL) * (size_t)cap);
        }
        mark[ic] = pos;
        C->JA[pos] = ic;
        C->val[pos] = 0.0;
        ++pos;
        for (INT q1 = R->IA[ic]; q1 < R->IA[ic + 1]; ++q1) {
            const REAL r = R->val[q1];
            const INT i1 = R->JA[q1];
            for (INT q2 = A->IA[i1]; q2 < A->IA[i1 + 1]; ++q2) {
                const REAL ra = r * A->val[q2];
                const INT i2 = A->JA[q2];
                for (INT q3 = P->IA[i2]; q3 < P->IA[i2 + 1]; ++q3) {
                    const REAL rap = ra * P->val[q3];
                    const INT i3 = P->JA[q3];
                    if (mark[i3] < row0) {
                        if (pos == cap) {
                            cap *= 2;
                            C->JA = (INT *)realloc(C->JA, sizeof(INT) * (size_t)cap);
                            C->val = (REAL *)realloc(C->val, sizeof(REAL) * (size_t)cap);
                        }
                        mark[i3] = pos;
                        C->JA[pos] = i3;
                        C->val[pos] = rap;
                        ++pos;
                    } else {
                        C->val[mark[i3]] += rap;
                    }
                }
            }
        }
    }
    C->IA[nc] = pos;
    C->nnz = pos;
    free(mark);
}

/* smooth_aggregation_p (amg_setup_ua.c:472-572): S = I - omega D^-1 M with M = A
 * (filter off) or the filtered strength graph N (filter on: the diagonal of N
 * absorbs the dropped off-diagonal row sum of A; N is modified in place), D
 * the first diagonal entry of M (1.0 where |d| < 1e-6); then P = S P_tent. */
static void orc_smooth_p(const dCSRmat *A, dCSRmat *N, const INT *agg, INT nagg, REAL omega, SHORT filter,
                         dCSRmat *P, const float *omega_agg) {
    const INT n = A->row;
    const dCSRmat *M = A;
    if (filter == 1) {
        for (INT i = 0; i < n; ++i) {
            REAL sa = 0.0, sn = 0.0;
            for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) if (A->JA[k] != i) sa += A->val[k];
            for (INT k = N->IA[i]; k < N->IA[i + 1]; ++k) if (N->JA[k] != i) sn += N->val[k];
            for (INT k = N->IA[i]; k < N->IA[i + 1]; ++k) if (N->JA[k] == i) N->val[k] += sa - sn;
        }
        M = N;
    }
    REAL *d = (REAL *)xcalloc((size_t)n, sizeof(REAL));
    for (INT i = 0; i < n; ++i)
        for (INT k = M->IA[i]; k < M->IA[i + 1]; ++k)
            if (M->JA[k] == i) { d[i] = M->val[k]; break; }
    for (INT i = 0; i < n; ++i) if (fabs(d[i]) < 1e-6) d[i] = 1.0;
    dCSRmat S;
    csr_copy(M, &S);
    for (INT i = 0; i < n; ++i) {
        /* per-aggregate damping (learned parameters): row i uses its aggregate's value */
        const REAL w = (omega_agg && agg[i] > HZ_UNPT) ? (REAL)omega_agg[agg[i]] : omega;
        for (INT k = S.IA[i]; k < S.IA[i + 1]; ++k)
            S.val[k] = (S.JA[k] == i) ? 1 - w * M->val[k] / d[i] : -w * M->val[k] / d[i];
    }
    free(d);
    /* tentative P (form_tentative_p, amg_setup_ua.c:251-295): (agg[i], 1.0) */
    dCSRmat T;
    T.row = n; T.col = nagg;
    T.IA = (INT *)xcalloc((size_t)n + 1, sizeof(INT));
    INT cnt = 0;
    for (INT i = 0; i < n; ++i) if (agg[i] > HZ_UNPT) ++cnt;
    T.JA = (INT *)xcalloc((size_t)cnt + 1, sizeof(INT));
    T.val = (REAL *)xcalloc((size_t)cnt + 1, sizeof(REAL));
    cnt = 0;
    for (INT i = 0; i < n; ++i) {
        T.IA[i] = cnt;
        if (agg[i] > HZ_UNPT) { T.JA[cnt] = agg[i]; T.val[cnt] = 1.0; ++cnt; }
    }
    T.IA[n] = cnt;
    T.nnz = cnt;
    orc_mxm(&S, &T, P);
    csr_free(&S);
    csr_free(&T);
}

/* Smoothed prolongation of one level for a given aggregation (strength
 * threshold theta), with a uniform or per-aggregate damping. */
void orc_sa_prolongation(const dCSRmat *A, const INT *agg, INT nagg, REAL theta, REAL omega,
                         const float *omega_agg, INT filter, dCSRmat *P) {
    dCSRmat N;
    orc_strength(A, theta, &N);
    orc_smooth_p(A, &N, agg, nagg, omega, (SHORT)filter, P, omega_agg);
    csr_free(&N);
}

/* Chebyshev smoother (no reference counterpart; Saad, Iterative Methods for
 * Sparse Linear Systems, 2nd ed., Alg. 12.1) on D^-1 A, D the first diagonal
 * entries, interval [lmax/30, lmax], lmax = scale * max_i sum_j |a_ij|/|a_ii|;
 * `steps` steps on x in place. */
void orc_cheby_smooth(const dCSRmat *A, const REAL *b, REAL *x, INT steps, REAL scale) {
    const INT n = A->row;
    REAL *dg = (REAL *)xcalloc((size_t)n + 1, sizeof(REAL));
    REAL rho_g = 0.0;
    for (INT i = 0; i < n; ++i) {
        REAL di = 0.0, s = 0.0;
        INT found = 0;
        for (INT k = A->IA[i]; k < A->IA[i + 1]; ++k) {
            s += fabs(A->val[k]);
            if (!found && A->JA[k] == i) { di = A->val[k]; found = 1; }
        }
        dg[i] = di;
        if (di != 0.0 && s / fabs(di) > rho_g) rho_g = s / fabs(di);
    }
    const REAL lmax = scale * rho_g, lmin = lmax / 30.0;
    const REAL theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    REAL rho = 1.0 / sigma;
    REAL *d = (REAL *)xcalloc((size_t)n + 1, sizeof(REAL));
    REAL *xn = (REAL *)xcalloc((size_t)n + 1, sizeof(REAL));
    for (INT k = 0; k < steps; ++k) {
        REAL c1 = 0.0, c2 = 1.0 / theta;
        if (k > 0) {
            const REAL rn = 1.0 / (2.0 * sigma - rho);
            c1 = rn * rho;
            c2 = 2.0 * rn / delta;
            rho = rn;
        }
        for (INT i = 0; i < n; ++i) {
            REAL t = b[i];
            for (INT q = A->IA[i]; q < A->IA[i + 1]; ++q) t = t - A->val[q] * x[A->JA[q]];
            const REAL dn = (c1 == 0.0 ? 0.0 : c1 * d[i]) + (dg[i] != 0.0 ? c2 * t / dg[i] : 0.0);
            d[i] = dn;
            xn[i] = x[i] + dn;
        }
        memcpy(x, xn, sizeof(REAL) * (size_t)n);
    }
    free(dg); free(d); free(xn);
}

/* Learned-parameter composition (no reference counterpart): per-aggregate
 * prolongator damping supplied level by level by a callback, and per-level
 * Chebyshev bound scales (instead of the uniform tentative_smooth and scale 1). */
static orc_omega_fn g_orc_omega_fn = NULL;
static REAL g_orc_cheb_scale[HZ_MAX_AMG_LVL];
static int g_orc_cheb_scale_set = 0;
void orc_set_learned(orc_omega_fn fn, const REAL *scales) {
    g_orc_omega_fn = fn;
    g_orc_cheb_scale_set = scales != NULL;
    for (INT l = 0; l < HZ_MAX_AMG_LVL; ++l) g_orc_cheb_scale[l] = scales ? scales[l] : 1.0;
}

/* amg_setup_unsmoothP_unsmoothR, amg_setup_ua.c:1228-1443, and for
 * AMG_type = SA amg_setup_smoothP_smoothR, amg_setup_ua.c:1461-1687. */
/* print_cputime, message.c:73-77 (wall seconds) */
static void orc_print_cputime(const char *message, double seconds) {
    printf("%s costs %.4f seconds.\n", message, seconds);
}

static INT orc_setup(orc_amg *m, const dCSRmat *A, AMG_param *param) {
    const double setup_start = orc_now();
    memset(m, 0, sizeof(*m));
    const SHORT min_cdof = (SHORT)(param->coarse_dof > 1 ? param->coarse_dof : 1);
    SHORT max_levels = param->max_levels;
    if (max_levels < 1) max_levels = 1;
    if (max_levels > HZ_MAX_AMG_LVL) max_levels = HZ_MAX_AMG_LVL;
    INT lvl = 0;
    INT status = HZ_SUCCESS;
    csr_copy(A, &m->lv[0].A);
    while (m->lv[lvl].A.row > min_cdof && lvl < max_levels - 1) {
        orc_level *L = &m->lv[lvl];
        const INT n = L->A.row;
        L->theta = param->strong_coupled;
        L->agg = (INT *)xcalloc((size_t)n, sizeof(INT));
        INT st;
        if (param->aggregation_type != HZ_VMB && param->aggregation_type != HZ_HEC &&
            param->aggregation_type != HZ_HEM) {
            printf("\n!!! ERROR HAZMATH DANGER: in function '%s' -- Unknown AMG aggregation type !!!\n\n",
                   "amg_setup_unsmoothP_unsmoothR");
            return ORC_EXIT_PARAM;
        }
        st = orc_aggregate(&L->A, param->aggregation_type, param->strong_coupled,
                           param->max_aggregation, L->agg, &L->nagg);
        if (st == ORC_EXIT_HEC || st == ORC_EXIT_HEM) return st;
        /* amg_setup_ua.c:1325-1328 */
        if (L->nagg * 4 > n) param->strong_coupled /= 2;
        else if (L->nagg * 1.25 < n) param->strong_coupled *= 2;
        if (st < 0) { status = HZ_SUCCESS; break; }
        if (L->nagg < ORC_MIN_CDOF) break;
        if (param->AMG_type == HZ_SA_AMG) {
            dCSRmat N;
            orc_strength(&L->A, L->theta, &N);
            float *om = NULL;
            if (g_orc_omega_fn) {
                om = (float *)xcalloc((size_t)L->nagg + 1, sizeof(float));
                g_orc_omega_fn(lvl, &L->A, L->agg, L->nagg, om);
            }
            orc_smooth_p(&L->A, &N, L->agg, L->nagg, param->tentative_smooth, param->smooth_filter, &L->P, om);
            free(om);
            csr_free(&N);
            csr_transpose(&L->P, &L->R);
            orc_rap(&L->R, &L->A, &L->P, &m->lv[lvl + 1].A);
            ++lvl;
            continue;
        }
        /* R members (dcsr_trans of the tentative P) */
        L->rp = (INT *)xcalloc((size_t)L->nagg + 1, sizeof(INT));
        L->rm = (INT *)xcalloc((size_t)n, sizeof(INT));
        for (INT i = 0; i < n; ++i) if (L->agg[i] > HZ_UNPT) L->rp[L->agg[i] + 1]++;
        for (INT c = 0; c < L->nagg; ++c) L->rp[c + 1] += L->rp[c];
        {
            INT *fill = (INT *)xcalloc((size_t)L->nagg + 1, sizeof(INT));
            memcpy(fill, L->rp, sizeof(INT) * ((size_t)L->nagg + 1));
            for (INT i = 0; i < n; ++i) if (L->agg[i] > HZ_UNPT) L->rm[fill[L->agg[i]]++] = i;
            free(fill);
        }
        orc_rap_agg(&L->A, L->agg, L->nagg, &m->lv[lvl + 1].A);
        ++lvl;
    }
    m->nl = lvl + 1;
    if (param->coarse_solver == HZ_SOLVER_UMFPACK) {
        dCSRmat T;
        csr_transpose(&m->lv[lvl].A, &T);
        csr_free(&m->lv[lvl].A);
        m->lv[lvl].A = T;
        orc_coarse_factor(&m->lv[lvl]);
    }
    m->lv[lvl].theta = param->strong_coupled;
    for (INT l = 0; l < m->nl; ++l) {
        INT mm = m->lv[l].A.row;
        m->lv[l].b = (REAL *)xcalloc((size_t)mm, sizeof(REAL));
        m->lv[l].x = (REAL *)xcalloc((size_t)mm, sizeof(REAL));
        m->lv[l].w = (REAL *)xcalloc((size_t)mm, sizeof(REAL));
    }
    if (param->print_level >= 2) {
        REAL gc = 0.0, oc = 0.0;
        printf("-----------------------------------------------------------\n");
        printf("  Level   Num of rows   Num of nonzeros   Avg. NNZ / row   \n");
        printf("-----------------------------------------------------------\n");
        for (INT l = 0; l < m->nl; ++l) {
            REAL avg = (REAL)m->lv[l].A.nnz / m->lv[l].A.row;
            printf("%5lld %13lld %17lld %14.2f\n", (long long)l, (long long)m->lv[l].A.row,
                   (long long)m->lv[l].A.nnz, avg);
            gc += m->lv[l].A.row; oc += m->lv[l].A.nnz;
        }
        printf("-----------------------------------------------------------\n");
        gc /= m->lv[0].A.row; oc /= m->lv[0].A.nnz;
        printf("  Grid complexity = %.3f  |", gc);
        printf("  Operator complexity = %.3f\n", oc);
        printf("-----------------------------------------------------------\n");
    }
    /* amg_setup_ua.c:1426-1431, 1668-1673 */
    if (param->print_level > 0)
        orc_print_cputime(param->AMG_type == HZ_SA_AMG ? "Smoothed aggregation setup" : "Unsmoothed aggregation setup",
                          orc_now() - setup_start);
    return status;
}

/* ------------------------------------------------------------------ */
/* PCG (dcsr_pcg, krylov.c:122-372 and dcsr_pcg_w_cond_est 400-718)    */
/* ------------------------------------------------------------------ */
typedef void (*orc_prec_fn)(const REAL *r, REAL *z, void *data);

static void orc_itinfo(INT prt, INT stop, INT iter, REAL rel, REAL absr, REAL fac) {
    if (prt < 2) return;
    if (iter > 0) {
        printf("%6lld | %13.6e   | %13.6e  | %10.4f\n", (long long)iter, rel, absr, fac);
    } else {
        printf("-----------------------------------------------------------\n");
        if (stop == HZ_STOP_REL_RES) printf("It Num |   ||r||/||b||   |     ||r||      |  Conv. Factor\n");
        else if (stop == HZ_STOP_REL_PRECRES) printf("It Num | ||r||_B/||b||_B |    ||r||_B     |  Conv. Factor\n");
        else printf("It Num |   ||r||/||x||   |     ||r||      |  Conv. Factor\n");
        printf("-----------------------------------------------------------\n");
        printf("%6lld | %13.6e   | %13.6e  |     -.-- \n", (long long)iter, rel, absr);
    }
}

/* Extreme eigenvalues of the Lanczos tridiagonal by Sturm bisection; returns
 * lambda_max / lambda_min (stands in for LAPACK dsterf_, krylov.c:663-705). */
static INT orc_sturm_count(INT n, const REAL *d, const REAL *e2, REAL x) {
    INT c = 0;
    REAL q = d[0] - x;
    if (q < 0) ++c;
    for (INT i = 1; i < n; ++i) {
        if (q == 0.0) q = 1e-300;
        q = d[i] - x - e2[i - 1] / q;
        if (q < 0) ++c;
    }
    return c;
}
static REAL orc_bisect(INT n, const REAL *d, const REAL *e2, REAL lo, REAL hi, INT k) {
    for (INT it = 0; it < 200; ++it) {
        REAL mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (orc_sturm_count(n, d, e2, mid) > k) hi = mid; else lo = mid;
    }
    return 0.5 * (lo + hi);
}
REAL orc_tridiag_condest(INT m, const REAL *alpha, const REAL *beta) {
    REAL *d0 = (REAL *)xcalloc((size_t)m, sizeof(REAL));
    REAL *e2 = (REAL *)xcalloc((size_t)m, sizeof(REAL));
    REAL *e = (REAL *)xcalloc((size_t)m, sizeof(REAL));
    d0[0] = 1.0 / alpha[0];
    for (INT i = 0; i < m - 1; ++i) {
        d0[i + 1] = 1.0 / alpha[i + 1] + beta[i] / alpha[i];
        e[i] = sqrt(beta[i]) / alpha[i];
        e2[i] = e[i] * e[i];
    }
    REAL lo = d0[0], hi = d0[0];
    for (INT i = 0; i < m; ++i) {
        REAL r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < m - 1 ? fabs(e[i]) : 0.0);
        if (d0[i] - r < lo) lo = d0[i] - r;
        if (d0[i] + r > hi) hi = d0[i] + r;
    }
    REAL pad = 1e-12 * (fabs(lo) + fabs(hi)) + 1e-300;
    REAL lmin = orc_bisect(m, d0, e2, lo - pad, hi + pad, 0);
    REAL lmax = orc_bisect(m, d0, e2, lo - pad, hi + pad, m - 1);
    free(d0); free(e2); free(e);
    return lmax / lmin;
}

static REAL orc_resid_measure(INT stop, INT m, const dCSRmat *A, const REAL *b, const REAL *u,
                              REAL *r, REAL *z, orc_prec_fn pc, void *pcd, REAL normr0,
                              REAL normu, REAL *absres) {
    memcpy(r, b, sizeof(REAL) * (size_t)m);
    orc_dcsr_aAxpy(-1.0, A, u, r);
    if (stop == HZ_STOP_REL_PRECRES) {
        if (pc) pc(r, z, pcd); else memcpy(z, r, sizeof(REAL) * (size_t)m);
        *absres = sqrt(fabs(orc_dot(m, z, r)));
        return *absres / normr0;
    }
    *absres = orc_norm2(m, r);
    return (stop == HZ_STOP_MOD_REL_RES) ? *absres / normu : *absres / normr0;
}

static INT orc_pcg(const dCSRmat *A, const REAL *b, REAL *u, orc_prec_fn pc, void *pcd,
                   REAL tol, INT MaxIt, INT stop, INT prt, int want_cond, REAL *condest,
                   orc_trace *tr);

/* x = a x (array.c:332-341) */
static void orc_ax(INT n, REAL a, REAL *x) {
    if (a != 1.0)
        for (INT i = 0; i < n; ++i) x[i] *= a;
}

/* dcsr_pvgmres (krylov.c:4612-4931) without preconditioner, stop type
 * STOP_REL_RES: GMRES(m) with modified Gram-Schmidt and Givens rotations, the
 * restart length adapted per cycle from the convergence rate cr of the last
 * one (Baker, Jessup and Kolev 2009: back to the maximum above cos 8 deg,
 * kept below cos 80 deg, otherwise shortened by 3 down to 3), and the true
 * residual checked before convergence is accepted.  Returns the iteration
 * count or ERROR_SOLVER_MAXIT. */
INT orc_pvgmres(const dCSRmat *A, const REAL *b, REAL *x, REAL tol, INT MaxIt, INT restart) {
    const INT n = A->row;
    const REAL cr_max = 0.99, cr_min = 0.174;
    const INT d = 3, restart_max = restart, restart_min = 3;
    INT iter = 0, i = 0;
    REAL r_norm, absres0, absres, relres, cr = 1.0, r_norm_old, t, gamma;
    INT Restart = restart;
    const INT R1 = restart + 1;
    REAL *work = (REAL *)xcalloc((size_t)(R1 + 2) * n + (size_t)R1 * restart + 3 * (size_t)R1, sizeof(REAL));
    REAL *r = work, *w = r + n, *P = w + n, *hh = P + (size_t)R1 * n, *c = hh + (size_t)R1 * restart,
         *sn = c + R1, *rs = sn + R1;
#define ORC_P(k) (P + (size_t)(k) * n)
#define ORC_H(a, b) hh[(size_t)(a) * restart + (b)]
    orc_dcsr_residual(A, b, x, ORC_P(0));
    r_norm = orc_norm2(n, ORC_P(0));
    absres0 = r_norm > ORC_SMALLREAL ? r_norm : ORC_SMALLREAL;
    relres = r_norm / absres0;
    if (relres < tol || absres0 < 1e-3 * tol) goto FINISHED;
    while (iter < MaxIt) {
        rs[0] = r_norm_old = r_norm;
        t = 1.0 / r_norm;
        orc_ax(n, t, ORC_P(0));
        if (cr > cr_max || iter == 0) Restart = restart_max;
        else if (cr < cr_min) { /* keep */
        } else if (Restart - d > restart_min) Restart -= d;
        else Restart = restart_max;
        i = 0;
        while (i < Restart && iter < MaxIt) {
            ++i;
            ++iter;
            memcpy(r, ORC_P(i - 1), sizeof(REAL) * (size_t)n);
            orc_dcsr_mxv(A, r, ORC_P(i));
            for (INT j = 0; j < i; ++j) {
                ORC_H(j, i - 1) = orc_dot(n, ORC_P(j), ORC_P(i));
                orc_axpy(n, -ORC_H(j, i - 1), ORC_P(j), ORC_P(i));
            }
            t = orc_norm2(n, ORC_P(i));
            ORC_H(i, i - 1) = t;
            if (t != 0.0) {
                t = 1.0 / t;
                orc_ax(n, t, ORC_P(i));
            }
            for (INT j = 1; j < i; ++j) {
                t = ORC_H(j - 1, i - 1);
                ORC_H(j - 1, i - 1) = sn[j - 1] * ORC_H(j, i - 1) + c[j - 1] * t;
                ORC_H(j, i - 1) = -sn[j - 1] * t + c[j - 1] * ORC_H(j, i - 1);
            }
            t = ORC_H(i, i - 1) * ORC_H(i, i - 1);
            t += ORC_H(i - 1, i - 1) * ORC_H(i - 1, i - 1);
            gamma = sqrt(t);
            if (gamma == 0.0) gamma = ORC_SMALLREAL;
            c[i - 1] = ORC_H(i - 1, i - 1) / gamma;
            sn[i - 1] = ORC_H(i, i - 1) / gamma;
            rs[i] = -sn[i - 1] * rs[i - 1];
            rs[i - 1] = c[i - 1] * rs[i - 1];
            ORC_H(i - 1, i - 1) = sn[i - 1] * ORC_H(i, i - 1) + c[i - 1] * ORC_H(i - 1, i - 1);
            absres = r_norm = fabs(rs[i]);
            relres = absres / absres0;
            if (relres < tol) break;
        }
        /* solution: back substitution, then x += sum_j rs_j p_j */
        rs[i - 1] = rs[i - 1] / ORC_H(i - 1, i - 1);
        for (INT k = i - 2; k >= 0; --k) {
            t = 0.0;
            for (INT j = k + 1; j < i; ++j) t -= ORC_H(k, j) * rs[j];
            t += rs[k];
            rs[k] = t / ORC_H(k, k);
        }
        memcpy(w, ORC_P(i - 1), sizeof(REAL) * (size_t)n);
        orc_ax(n, rs[i - 1], w);
        for (INT j = i - 2; j >= 0; --j) orc_axpy(n, rs[j], ORC_P(j), w);
        memcpy(r, w, sizeof(REAL) * (size_t)n);
        orc_axpy(n, 1.0, r, x);
        if (relres < tol) {
            orc_dcsr_residual(A, b, x, r);
            r_norm = orc_norm2(n, r);
            absres = r_norm;
            relres = absres / absres0;
            if (relres < tol) break;
            memcpy(ORC_P(0), r, sizeof(REAL) * (size_t)n);
            i = 0;
        }
        /* residual vector of the cycle for the next one */
        for (INT j = i; j > 0; --j) {
            rs[j - 1] = -sn[j - 1] * rs[j];
            rs[j] = c[j - 1] * rs[j];
        }
        if (i) orc_axpy(n, rs[i] - 1.0, ORC_P(i), ORC_P(i));
        for (INT j = i - 1; j > 0; --j) orc_axpy(n, rs[j], ORC_P(j), ORC_P(i));
        if (i) {
            orc_axpy(n, rs[0] - 1.0, ORC_P(0), ORC_P(0));
            orc_axpy(n, 1.0, ORC_P(i), ORC_P(0));
        }
        cr = r_norm / r_norm_old;
    }
FINISHED:
#undef ORC_P
#undef ORC_H
    free(work);
    return iter >= MaxIt ? HZ_ERROR_SOLVER_MAXIT : iter;
}

/* dcsr_pcg without preconditioner (krylov.c:122-372 as the coarse solve calls it). */
INT orc_cg(const dCSRmat *A, const REAL *b, REAL *x, REAL tol, INT MaxIt) {
    return orc_pcg(A, b, x, NULL, NULL, tol, MaxIt, HZ_STOP_REL_RES, 0, 0, NULL, NULL);
}

/* coarse_itsolver (mgcycle.c:34-53): CG from x, GMRES(30) from CG's last
 * iterate when CG fails. */
INT orc_coarse_itsolver(const dCSRmat *A, const REAL *b, REAL *x, REAL ctol) {
    const INT n = A->row;
    INT status = orc_pcg(A, b, x, NULL, NULL, ctol, 20 * n, HZ_STOP_REL_RES, 0, 0, NULL, NULL);
    if (status < 0) status = orc_pvgmres(A, b, x, ctol, 20 * n, 30);
    return status;
}

static INT orc_pcg(const dCSRmat *A, const REAL *b, REAL *u, orc_prec_fn pc, void *pcd,
                   REAL tol, INT MaxIt, INT stop, INT prt, int want_cond, REAL *condest,
                   orc_trace *tr) {
    const INT m = A->row;
    const REAL maxdiff = tol * ORC_STAG_RATIO;
    INT iter = 0, stag = 1, more_step = 1, restart_step = 1;
    REAL absres0 = ORC_BIGREAL, absres = ORC_BIGREAL, relres = ORC_BIGREAL;
    REAL normu = ORC_BIGREAL, normr0 = ORC_BIGREAL;
    REAL alpha, beta, temp1, temp2;
    REAL *p = (REAL *)xcalloc((size_t)m * 4, sizeof(REAL));
    REAL *z = p + m, *r = z + m, *t = r + m;
    REAL *alpha_all = want_cond ? (REAL *)xcalloc((size_t)MaxIt + 1, sizeof(REAL)) : NULL;
    REAL *beta_all = want_cond ? (REAL *)xcalloc((size_t)MaxIt + 1, sizeof(REAL)) : NULL;
    INT iter_condest = 0;
    if (tr) tr->relres_count = 0;

    memcpy(r, b, sizeof(REAL) * (size_t)m);
    orc_dcsr_aAxpy(-1.0, A, u, r);
    if (pc) pc(r, z, pcd); else memcpy(z, r, sizeof(REAL) * (size_t)m);
    switch (stop) {
        case HZ_STOP_REL_RES:
            absres0 = orc_norm2(m, r);
            normr0 = absres0 > ORC_SMALLREAL ? absres0 : ORC_SMALLREAL;
            relres = absres0 / normr0;
            break;
        case HZ_STOP_REL_PRECRES:
            absres0 = sqrt(orc_dot(m, r, z));
            normr0 = absres0 > ORC_SMALLREAL ? absres0 : ORC_SMALLREAL;
            relres = absres0 / normr0;
            break;
        case HZ_STOP_MOD_REL_RES:
            absres0 = orc_norm2(m, r);
            normu = orc_norm2(m, u);
            if (normu < ORC_SMALLREAL) normu = ORC_SMALLREAL;
            relres = absres0 / normu;
            break;
        default:
            printf("### ERROR: Unrecognised stopping type for %s!\n", "dcsr_pcg_w_cond_est");
            goto FINISHED;
    }
    if (relres < tol || absres0 < 1e-3 * tol) goto FINISHED;
    orc_itinfo(prt, stop, iter, relres, absres0, 0.0);
    memcpy(p, z, sizeof(REAL) * (size_t)m);
    temp1 = orc_dot(m, z, r);

    while (iter++ < MaxIt) {
        orc_dcsr_mxv(A, p, t);
        temp2 = orc_dot(m, t, p);
        alpha = temp1 / temp2;
        if (want_cond) alpha_all[iter - 1] = alpha;
        orc_axpy(m, alpha, p, u);
        orc_axpy(m, -alpha, t, r);
        if (stop == HZ_STOP_REL_PRECRES) {
            if (pc) pc(r, z, pcd); else memcpy(z, r, sizeof(REAL) * (size_t)m);
            absres = sqrt(fabs(orc_dot(m, z, r)));
            relres = absres / normr0;
        } else {
            absres = orc_norm2(m, r);
            relres = (stop == HZ_STOP_MOD_REL_RES) ? absres / normu : absres / normr0;
        }
        if (tr && tr->relres_count < 4096) tr->relres_hist[tr->relres_count++] = relres;
        orc_itinfo(prt, stop, iter, relres, absres, absres / absres0);
        /* Check I */
        if (orc_norminf(m, u) <= ORC_SMALLREAL) {
            if (prt > 1) printf("### HAZMATH WARNING: Iteration stopped due to the solution is almost zero! %s : %lld\n", "dcsr_pcg_w_cond_est", 0LL);
            iter = HZ_ERROR_SOLVER_SOLSTAG;
            break;
        }
        /* Check II */
        normu = orc_norm2(m, u);
        REAL reldiff = fabs(alpha) * orc_norm2(m, p) / normu;
        if ((stag <= ORC_MAX_STAG) & (reldiff < maxdiff)) {
            if (prt >= 4) {
                printf("||u-u'|| = %e and the comp. rel. res. = %e.\n", reldiff, relres);
                printf("### HAZMATH WARNING: Iteration restarted due to stagnation! %s : %lld\n", "dcsr_pcg_w_cond_est", 0LL);
            }
            relres = orc_resid_measure(stop, m, A, b, u, r, z, pc, pcd, normr0, normu, &absres);
            if (prt >= 4) printf("### HAZMATH WARNING: The actual relative residual = %e!\n", relres);
            if (relres < tol) break;
            if (stag >= ORC_MAX_STAG) {
                if (prt > 1) printf("### HAZMATH WARNING: Iteration stopped due to staggnation! %s : %lld\n", "dcsr_pcg_w_cond_est", 0LL);
                iter = HZ_ERROR_SOLVER_STAG;
                break;
            }
            if (stag == 1 && restart_step == 1) iter_condest = iter;
            memset(p, 0, sizeof(REAL) * (size_t)m);
            ++stag; ++restart_step;
        }
        /* Check III */
        if (relres < tol) {
            REAL computed = relres;
            relres = orc_resid_measure(stop, m, A, b, u, r, z, pc, pcd, normr0, normu, &absres);
            if (relres < tol) break;
            if (prt >= 4) {
                printf("### HAZMATH WARNING: The computed relative residual = %e!\n", computed);
                printf("### HAZMATH WARNING: The actual relative residual = %e!\n", relres);
            }
            if (more_step >= ORC_MAX_RESTART) {
                if (prt > 1) printf("### HAZMATH WARNING: The tolerence might be too small! %s : %lld\n", "dcsr_pcg_w_cond_est", 0LL);
                iter = HZ_ERROR_SOLVER_TOLSMALL;
                break;
            }
            if (stag == 1 && restart_step == 1) iter_condest = iter;
            memset(p, 0, sizeof(REAL) * (size_t)m);
            ++more_step; ++restart_step;
        }
        absres0 = absres;
        if (stop != HZ_STOP_REL_PRECRES) {
            if (pc) pc(r, z, pcd); else memcpy(z, r, sizeof(REAL) * (size_t)m);
        }
        temp2 = orc_dot(m, z, r);
        beta = temp2 / temp1;
        temp1 = temp2;
        if (want_cond) beta_all[iter - 1] = beta;
        orc_axpby(m, 1.0, z, beta, p);
    }

    if (want_cond) {
        if (stag == 1 && restart_step == 1) iter_condest = iter;
        if (iter_condest > MaxIt) iter_condest = MaxIt;   /* reference reads past its arrays here */
        if (iter_condest < 1) {
            printf("### HAZMATH WARNING: %s takes too few iterations for estimating condition number!!!\n",
                   "dcsr_pcg_w_cond_est");
        } else {
            *condest = orc_tridiag_condest(iter_condest, alpha_all, beta_all);
        }
    }
FINISHED:
    if (prt > 0) {
        if (iter > MaxIt) printf("### HAZMATH WARNING: Max iter %lld reached with rel. resid. %e.\n", (long long)MaxIt, relres);
        else if (iter >= 0) printf("Num_iter(krylov.c) = %lld with relative residual %e.\n", (long long)iter, relres);
    }
    if (tr) tr->final_relres = relres;
    free(p); free(alpha_all); free(beta_all);
    return (iter > MaxIt) ? HZ_ERROR_SOLVER_MAXIT : iter;
}

/* ------------------------------------------------------------------ */
/* V/W cycle: mgcycle, mgcycle.c:245-409 (UA transfers, standard       */
/* smoothers), driven by precond_amg, precond.c:164-183                */
/* ------------------------------------------------------------------ */
typedef struct {
    orc_amg *m;
    AMG_param cyc;   /* param_amg_init + param_prec_to_amg (precond.c:173-174) */
    REAL jacobi_w;
    INT maxit;
} orc_prec_data;

/* Chebyshev steps below the finest level (0: the cycle's counts); the
 * library's hz_options.cheb_coarse_steps, no reference counterpart. */
static INT g_orc_cheb_coarse = 0;
void orc_set_cheb_coarse_steps(INT steps) { g_orc_cheb_coarse = steps; }

static void orc_smooth(const orc_prec_data *pd, orc_level *L, INT sweeps, int post) {
    if (pd->cyc.smoother == 91) {   /* Chebyshev (build-owned): bound scale 1 or the level's set scale */
        const INT l = (INT)(L - pd->m->lv);
        if (g_orc_cheb_coarse > 0 && l >= 1) sweeps = g_orc_cheb_coarse;
        orc_cheby_smooth(&L->A, L->b, L->x, sweeps, g_orc_cheb_scale_set ? g_orc_cheb_scale[l] : 1.0);
        return;
    }
    if (orc_is_gs_family(pd->cyc.smoother)) {
        orc_gs_smooth_ordered(&L->A, L->b, L->x, pd->cyc.smoother, post, sweeps, pd->cyc.relaxation, NULL, NULL, 0);
        return;
    }
    for (INT s = 0; s < sweeps; ++s)
        orc_smoother_sweep(&L->A, L->b, L->x, pd->cyc.smoother, pd->jacobi_w);
}

static void orc_mgcycle(orc_prec_data *pd) {
    orc_amg *m = pd->m;
    const AMG_param *p = &pd->cyc;
    const INT nl = m->nl;
    const REAL ctol = p->tol * 1e-2;
    INT num_lvl[HZ_MAX_AMG_LVL] = {0};
    INT l = 0;
    REAL alpha = 1.0;
forward:
    while (l < nl - 1) {
        orc_level *L = &m->lv[l];
        num_lvl[l]++;
        orc_smooth(pd, L, p->presmooth_iter, 0);
        orc_dcsr_residual(&L->A, L->b, L->x, L->w);
        orc_level *C = &m->lv[l + 1];
        if (L->P.IA) {
            orc_dcsr_mxv(&L->R, L->w, C->b);        /* dcsr_mxv (mgcycle.c:302) */
        } else {
            for (INT ic = 0; ic < L->nagg; ++ic) {   /* dcsr_mxv_agg, sparse.c:1663-1679 */
                REAL s = 0.0;
                for (INT q = L->rp[ic]; q < L->rp[ic + 1]; ++q) s += L->w[L->rm[q]];
                C->b[ic] = s;
            }
        }
        ++l;
        memset(m->lv[l].x, 0, sizeof(REAL) * (size_t)m->lv[l].A.row);
    }
    {
        orc_level *C = &m->lv[nl - 1];
        if (p->coarse_solver == HZ_SOLVER_UMFPACK) {
            orc_coarse_solve(C, C->b, C->x);
        } else {
            orc_coarse_itsolver(&C->A, C->b, C->x, ctol);
        }
    }
    while (l > 0) {
        --l;
        orc_level *L = &m->lv[l], *C = &m->lv[l + 1];
        if (p->coarse_scaling == 1) {
            alpha = orc_dot(C->A.row, C->x, C->b) / orc_dcsr_vmv(&C->A, C->x, C->x);
            if (alpha > 2.0) alpha = 2.0;
        }
        for (INT i = 0; i < L->A.row; ++i) {    /* dcsr_aAxpy_agg, sparse.c:1748-1784 / dcsr_aAxpy 1693-1730 */
            REAL t = 0.0;
            if (L->P.IA) {
                for (INT k = L->P.IA[i]; k < L->P.IA[i + 1]; ++k) t += L->P.val[k] * C->x[L->P.JA[k]];
            } else if (L->agg[i] > HZ_UNPT) {
                t += C->x[L->agg[i]];
            }
            if (alpha == 1.0) L->x[i] += t;
            else if (alpha == -1.0) L->x[i] -= t;
            else L->x[i] += t * alpha;
        }
        orc_smooth(pd, L, p->postsmooth_iter, 1);
        if (num_lvl[l] < p->cycle_type) break;
        num_lvl[l] = 0;
    }
    if (l > 0) goto forward;
}

static void orc_precond_amg(const REAL *r, REAL *z, void *data) {
    orc_prec_data *pd = (orc_prec_data *)data;
    orc_level *L0 = &pd->m->lv[0];
    INT n = L0->A.row;
    memcpy(L0->b, r, sizeof(REAL) * (size_t)n);
    memset(L0->x, 0, sizeof(REAL) * (size_t)n);
    for (INT i = 0; i < pd->maxit; ++i) orc_mgcycle(pd);
    memcpy(z, L0->x, sizeof(REAL) * (size_t)n);
}

static REAL g_orc_jacobi_w = 1.0;
void orc_set_jacobi_weight(REAL w) { g_orc_jacobi_w = w; }

/* param_amg_to_prec + param_prec_to_amg (parameters.c:604-675): the cycle runs
 * with param_amg_init defaults overwritten by the caller's cycle fields. */
static void orc_prec_data_init(orc_prec_data *pd, orc_amg *m, const AMG_param *amgparam) {
    pd->m = m;
    orc_param_amg_init(&pd->cyc);
    pd->cyc.AMG_type = amgparam->AMG_type;
    pd->cyc.print_level = amgparam->print_level;
    pd->cyc.cycle_type = amgparam->cycle_type;
    pd->cyc.smoother = amgparam->smoother;
    pd->cyc.presmooth_iter = amgparam->presmooth_iter;
    pd->cyc.postsmooth_iter = amgparam->postsmooth_iter;
    pd->cyc.relaxation = amgparam->relaxation;
    pd->cyc.polynomial_degree = amgparam->polynomial_degree;
    pd->cyc.coarse_solver = amgparam->coarse_solver;
    pd->cyc.coarse_scaling = amgparam->coarse_scaling;
    pd->cyc.amli_degree = amgparam->amli_degree;
    pd->maxit = amgparam->maxit;
    pd->jacobi_w = g_orc_jacobi_w;
}

/* linear_solver_dcsr_krylov_amg, itsolver.c:846-931, with the CG branch of
 * solver_dcsr_linear_itsolver, itsolver.c:135-211. */
INT orc_krylov_amg(dCSRmat *A, dvector *b, dvector *x, linear_itsolver_param *itparam,
                   AMG_param *amgparam, orc_trace *trace) {
    const SHORT prt = itparam->linear_print_level;
    if (trace) memset(trace, 0, sizeof(*trace));
    orc_amg *m = (orc_amg *)xcalloc(1, sizeof(orc_amg));
    double t0 = orc_now();
    /* itsolver.c:866-883 */
    if (prt > 0) printf(amgparam->AMG_type == HZ_SA_AMG ? "\n Calling SA AMG ...\n" : "\n Calling UA AMG ...\n");
    if (amgparam->smoother != HZ_SMOOTHER_JACOBI && amgparam->smoother != HZ_SMOOTHER_L1DIAG &&
        amgparam->smoother != 91 && !orc_is_gs_family(amgparam->smoother)) {
        printf("### ERROR: Wrong smoother type %lld!\n", (long long)amgparam->smoother);
        free(m);
        return HZ_ERROR_AMG_SMOOTH_TYPE;   /* smoothers outside the restated set */
    }
    INT status = orc_setup(m, A, amgparam);
    double t1 = orc_now();
    if (trace) {
        trace->num_levels = m->nl;
        for (INT l = 0; l < m->nl; ++l) {
            trace->level_rows[l] = m->lv[l].A.row;
            trace->level_nnz[l] = m->lv[l].A.nnz;
            trace->theta[l] = m->lv[l].theta;
        }
        trace->setup_s = t1 - t0;
    }
    if (status < 0) { orc_amg_free(m); free(m); return status; }
    orc_prec_data pd;
    orc_prec_data_init(&pd, m, amgparam);
    INT iter;
    REAL condest = 0.0;
    if (itparam->linear_tol < ORC_SMALLREAL)
        printf("### HAZMATH WARNING: Convergence tolerance for iterative solver is too small!\n");
    if (itparam->linear_maxit <= 0)
        printf("### HAZMATH WARNING: Max number of iterations should be a POSITIVE integer!\n");
    if (itparam->linear_itsolver_type != HZ_SOLVER_CG) {
        printf("### ERROR: Unknown itertive solver type %lld!\n", (long long)itparam->linear_itsolver_type);
        iter = HZ_ERROR_SOLVER_TYPE;
    } else {
        if (prt > 0) {
            printf("**********************************************************\n");
            printf(" --> using Conjugate Gradient Method:\n");
        }
        iter = orc_pcg(&m->lv[0].A, b->val, x->val, orc_precond_amg, &pd, itparam->linear_tol,
                       itparam->linear_maxit, itparam->linear_stop_type, prt, 1, &condest, trace);
        printf("Estimated condition number = %13.6e\n", condest);
    }
    double t2 = orc_now();
    /* itsolver.c:201-206, 917-922 */
    if (prt >= 2 && iter >= 0) {
        orc_print_cputime("Iterative method", t2 - t1);
        printf("**********************************************************\n");
    }
    if (prt >= 1) {
        orc_print_cputime("AMG_Krylov method totally", t2 - t0);
        printf("**********************************************************\n");
    }
    if (trace) {
        trace->iterations = iter;
        trace->condest = condest;
        trace->solve_s = t2 - t1;
    }
    fflush(stdout);
    orc_amg_free(m);
    free(m);
    return iter;
}

/* linear_solver_amg, itsolver.c:524-640, with amg_solve, amg_solve.c:109-161:
 * AMG cycles from the caller's x until ||b - A x|| / ||b|| < tol. */
INT orc_linear_solver_amg(dCSRmat *A, dvector *b, dvector *x, AMG_param *param) {
    const SHORT prt = param->print_level;
    if (A->row != A->col || A->nnz <= 0) return -15;   /* ERROR_MAT_SIZE (check_error exits) */
    if (param->smoother != HZ_SMOOTHER_JACOBI && param->smoother != HZ_SMOOTHER_L1DIAG &&
        !orc_is_gs_family(param->smoother))
        return HZ_ERROR_AMG_SMOOTH_TYPE;
    orc_amg *m = (orc_amg *)xcalloc(1, sizeof(orc_amg));
    const double t0 = orc_now();
    /* itsolver.c:545-560 */
    if (prt > 0) printf(param->AMG_type == HZ_SA_AMG ? "\n Calling SA AMG ...\n" : "\n Calling UA AMG ...\n");
    INT status = orc_setup(m, A, param);
    const double t_solve = orc_now();
    if (status < 0) { orc_amg_free(m); free(m); return status; }
    /* amg_solve runs mgcycle with the caller's param directly */
    orc_prec_data pd;
    pd.m = m;
    pd.cyc = *param;
    pd.maxit = 1;
    pd.jacobi_w = g_orc_jacobi_w;
    orc_level *L0 = &m->lv[0];
    const INT n = L0->A.row;
    memcpy(L0->b, b->val, sizeof(REAL) * (size_t)n);
    memcpy(L0->x, x->val, sizeof(REAL) * (size_t)n);
    const INT MaxIt = param->maxit;
    const REAL tol = param->tol;
    const REAL sumb = orc_norm2(n, L0->b);
    REAL relres1 = ORC_BIGREAL, absres0 = sumb, absres, factor;
    INT iter = 0;
    orc_itinfo(prt, HZ_STOP_REL_RES, iter, 1.0, sumb, 0.0);
    while ((++iter <= MaxIt) & (sumb > ORC_SMALLREAL)) {
        orc_mgcycle(&pd);
        orc_dcsr_residual(&L0->A, L0->b, L0->x, L0->w);
        absres = orc_norm2(n, L0->w);
        relres1 = absres / sumb;
        factor = absres / absres0;
        absres0 = absres;
        orc_itinfo(prt, HZ_STOP_REL_RES, iter, relres1, absres, factor);
        if (relres1 < tol) break;
    }
    if (prt > 0) {
        if (iter > MaxIt)
            printf("### HAZMATH WARNING: Max iter %lld reached with rel. resid. %e.\n", (long long)MaxIt, relres1);
        else if (iter >= 0)
            printf("Num_iter(amg_solve.c) = %lld with relative residual %e.\n", (long long)iter, relres1);
        orc_print_cputime("AMG solve", orc_now() - t_solve);   /* amg_solve.c:153-157 */
    }
    memcpy(x->val, L0->x, sizeof(REAL) * (size_t)n);
    if (prt > 0) {   /* itsolver.c:631-636 */
        orc_print_cputime("AMG totally", orc_now() - t0);
        printf("**********************************************************\n");
    }
    fflush(stdout);
    orc_amg_free(m);
    free(m);
    return iter;
}

/* create_precond_amg (helper.c:206-283) / precond_amg (precond.c:164-183). */
void *orc_precond_create(dCSRmat *A, AMG_param *amgparam) {
    if (amgparam->smoother != HZ_SMOOTHER_JACOBI && amgparam->smoother != HZ_SMOOTHER_L1DIAG &&
        amgparam->smoother != 91 && !orc_is_gs_family(amgparam->smoother))
        return NULL;
    orc_amg *m = (orc_amg *)xcalloc(1, sizeof(orc_amg));
    if (orc_setup(m, A, amgparam) < 0) { orc_amg_free(m); free(m); return NULL; }
    orc_prec_data *pd = (orc_prec_data *)xcalloc(1, sizeof(orc_prec_data));
    orc_prec_data_init(pd, m, amgparam);
    return pd;
}

void orc_precond_apply(void *h, const REAL *r, REAL *z) { orc_precond_amg(r, z, h); }

void orc_precond_free(void *h) {
    orc_prec_data *pd = (orc_prec_data *)h;
    if (!pd) return;
    orc_amg_free(pd->m);
    free(pd->m);
    free(pd);
}

INT orc_setup_hierarchy(dCSRmat *A, AMG_param *amgparam, hz_hierarchy *out) {
    orc_amg *m = (orc_amg *)xcalloc(1, sizeof(orc_amg));
    memset(out, 0, sizeof(*out));
    INT st = orc_setup(m, A, amgparam);
    out->num_levels = m->nl;
    for (INT l = 0; l < m->nl; ++l) {
        out->A[l] = m->lv[l].A;
        memset(&m->lv[l].A, 0, sizeof(dCSRmat));
        out->theta[l] = m->lv[l].theta;
        if (l < m->nl - 1) {
            out->aggregates[l] = m->lv[l].agg;
            out->num_aggregates[l] = m->lv[l].nagg;
            m->lv[l].agg = NULL;
            out->P[l] = m->lv[l].P;
            memset(&m->lv[l].P, 0, sizeof(dCSRmat));
        }
    }
    orc_amg_free(m);
    free(m);
    return st;
}

void orc_hierarchy_free(hz_hierarchy *h) {
    for (INT l = 0; l < HZ_MAX_AMG_LVL; ++l) {
        csr_free(&h->A[l]);
        csr_free(&h->P[l]);
        free(h->aggregates[l]);
        h->aggregates[l] = NULL;
    }
    h->num_levels = 0;
}
