// hz_internal.h -- device data structures and kernel entry points (host side).
#pragma once

#include "hz_common.h"

namespace hz {

// CSR on the device (0-based, any column order).  Same arrays as dCSRmat.
struct DevCSR {
    int n = 0, ncol = 0, nnz = 0;
    int *ia = nullptr, *ja = nullptr;
    double *val = nullptr;
};

// SELL-64: rows are grouped in slices of 64 consecutive rows; slice s stores
// width_s columns-major entries (entry k of row 64s+lane at sptr[s]+64k+lane).
// Row order inside each row is the CSR order, so per-row sums are formed in the
// same order as the CSR loops of the reference.
struct DevSELL {
    int n = 0, nslices = 0;
    long long nnz = 0;         // stored CSR entries (algorithmic bytes count these, not the padding)
    long long nent = 0;        // padded entries
    long long *sptr = nullptr; // nslices+1 offsets
    int *len = nullptr;        // row lengths
    int *col = nullptr;
    double *val = nullptr;
    double *dl1 = nullptr;     // row l1 norms (L1DIAG diagonal), formed by sell_from_csr
    double *djac = nullptr;    // last stored diagonal entry of each row (Jacobi diagonal)
    double *dfirst = nullptr;  // first stored diagonal entry of each row (0.0 if none)
    // the columns again as paired 16-bit offsets (col_base), when the FP64
    // cycle's level has them; null otherwise
    unsigned int *pcol = nullptr;
};

// Device scalars of the PCG (one 256-byte block; host copies it back once per
// iteration).
struct PcgScalars {
    double temp1;        // (z, r) of the previous step
    double temp2;        // (A p, p) / (z, r)
    double alpha, beta;
    double rr;           // ||r||^2
    double uinf;         // ||u||_inf
    double uu;           // ||u||^2
    double pp;           // ||p||^2
    double zr;           // (z, r)
    double xb, xax;      // coarse scaling: (x,b) and x^T A x
    double pad[21];
};

// Reduction workspace (partials + tickets), sized for the largest grid used.
// Algorithmic bytes of the solve kernels launched since the counter was
// last cleared (each distinct input array read once, each output written
// once; gathered coarse corrections not counted): the numerator of the
// whole-solve bandwidth.  The PCG clears it when it starts.
extern double g_alg_bytes;
inline void count_bytes(double b) { g_alg_bytes += b; }
// matrix stream of a CSR / SELL launch: values (vb bytes each) and int32
// columns of nnz entries, int32 row pointers of n rows
inline double matrix_bytes(long long nnz, long n, int vb = 8) { return (double)nnz * (vb + 4) + 4.0 * (n + 1); }

struct ReduceWs {
    double *partials = nullptr;   // kMaxBlocks * kSlots
    unsigned int *tickets = nullptr;
    static constexpr int kMaxBlocks = 8192;
    static constexpr int kSlots = 4;
};

// HIP events around the launches of one kernel inside a solve (the stream is
// not synchronised; the events are read once at the end).
struct LaunchTimer {
    std::vector<hipEvent_t> ev;
    int used = 0;
    void begin() { record(); }
    void end() { record(); }
    void record() {
        if (used == (int)ev.size()) {
            hipEvent_t e;
            HZ_CHECK(hipEventCreate(&e));
            ev.push_back(e);
        }
        HZ_CHECK(hipEventRecord(ev[used++], stream()));
    }
    // total milliseconds and launch count
    double total_ms(int *count) {
        HZ_CHECK(hipStreamSynchronize(stream()));
        double t = 0.0;
        for (int i = 0; i + 1 < used; i += 2) {
            float ms = 0.f;
            HZ_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            t += ms;
        }
        if (count) *count = used / 2;
        return t;
    }
    ~LaunchTimer() {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// ---- kernels_solve.hip ----------------------------------------------------
void sell_from_csr(const DevCSR &A, DevSELL &S);
// y = A x ; if dot_out: *dot_out = sum_i y_i * x_i (deterministic).  timer:
// events around the SpMV kernel itself (not the partials' finalize).
// cols_only: the PCG's SpMV ahead of an FP32 cycle, which shares only the
// column indices (those stay in the last-level cache, the values stream).
void spmv(const DevSELL &S, const double *x, double *y, double *dot_out, const ReduceWs &ws,
          LaunchTimer *timer = nullptr, bool cols_only = false);
// w = b - A x (timer: events around the kernel)
void residual(const DevSELL &S, const double *b, const double *x, double *w, LaunchTimer *timer = nullptr);
// One smoothing sweep out of place: xout = sweep(x).  smoother: 1 Jacobi (weight w), 10 L1DIAG.
void smooth_sweep(const DevSELL &S, int smoother, double w, const double *b, const double *x,
                  double *xout);
// Same sweep, also writing sum_i xout_i * b_i (deterministic) to dot_out.
void smooth_sweep_dot(const DevSELL &S, int smoother, double w, const double *b, const double *x,
                      double *xout, double *dot_out, const ReduceWs &ws, LaunchTimer *timer = nullptr);
// First sweep from a zero iterate: x = w*b/d (Jacobi) or b/d_l1 (L1DIAG); d precomputed.
void smooth_from_zero(int n, int smoother, double w, const double *d, const double *b, double *x);
// Smoother diagonal: Jacobi uses the last stored diagonal entry of each row
// (smoother.c:57-60), L1DIAG uses the row's l1 norm (smoother.c:352-357).
void smoother_diag(const DevSELL &S, int smoother, double *d);
// bc = R w (member sums, ascending members);  optional fused zero-iterate sweep of the coarse level.
void restrict_agg(int nc, int n, const int *rptr, const int *rmem, const double *w, double *bc,
                  int smoother, double wgt, const double *dc, double *xc);
// Fused residual + restriction (+ coarse zero-iterate sweep when xc != null),
// one thread per coarse row reading the member rows of the CSR.
void down_fused(int nc, const int *rptr, const int *rmem, const DevCSR &A, const double *b,
                const double *x, double *bc, int smoother, double wgt, const double *dc, double *xc);
// Fused prolongation + one post-smoothing sweep: xout = sweep(x + alpha P xc).
// dot_out (optional): sum_i xout_i * b_i.
void up_fused(const DevSELL &S, int smoother, double w, const int *agg, const double *xc,
              double alpha, const double *alpha_dev, const double *b, const double *x, double *xout,
              double *dot_out, const ReduceWs &ws);
// x += alpha * P xc  (aggregate gather-add)
void prolong_agg(int n, const int *agg, const double *xc, double alpha, const double *alpha_dev,
                 double *x);
// vector helpers
void fill(double *x, double v, long n);
void copy(double *dst, const double *src, long n);
// PCG fused vector kernels (scalars live on the device)
// The preconditioner's first level-0 sweep from a zero iterate, x0 = sweep(0)
// on the updated residual, when pcg_update is to form it (smooth_from_zero's
// arithmetic).
struct FirstSweep {
    int smoother = 0;
    double w = 1.0;
    const double *d = nullptr;
    double *x0 = nullptr;
};
void pcg_update(int n, PcgScalars *s, const double *p, const double *t, double *u, double *r,
                double *alpha_hist, int iter, const ReduceWs &ws, const FirstSweep *first = nullptr);
void dot(int n, const double *x, const double *y, double *out, const ReduceWs &ws);
void norm2sq(int n, const double *x, double *out, const ReduceWs &ws);
// beta = zr/temp1; temp1 = zr; p = z + beta p ; beta_hist[iter-1] = beta
void pcg_direction(int n, PcgScalars *s, const double *z, double *p, double *beta_hist, int iter);
// coarse scaling alpha = min((x,b)/(x^T A x), 2)
void coarse_scaling_alpha(const DevSELL &S, const double *x, const double *b, PcgScalars *s,
                          double *alpha_out, const ReduceWs &ws);

// ---- kernels_setup.hip ----------------------------------------------------
// Strength graph N (construct_strongly_coupled, amg_setup_ua.c:375-452).
// dfirst (optional): the rows' first stored diagonal entries, when already formed.
void strength_graph(const DevCSR &A, double theta, DevCSR &N, int *counts_ws, const double *dfirst = nullptr);
// Exclusive scan of n ints into out (n+1 entries, out[n] = total); returns total.
long long exclusive_scan_int(const int *in, int *out, int n);
// The same scan left on the stream (no host wait); the total stays in out[n].
void exclusive_scan_int_device(const int *in, int *out, int n);
// HEC aggregation (exact restatement of the sequential turn-order algorithm).
// perm: turn order (perm[t] = vertex); returns number of aggregates or a negative code.
int hec_aggregate(const DevCSR &N, const int *perm, int *agg);
// R = P^T: rptr (nagg+1), rmem (members ascending).
void build_restriction(int n, const int *agg, int nagg, int *rptr, int *rmem);
// Galerkin product with unit R/P (dcsr_rap_agg semantics).  Allocates Ac arrays.
void rap_agg(const DevCSR &A, const int *agg, int nagg, const int *rptr, const int *rmem,
             DevCSR &Ac);
// Transpose (sorted columns) of a device CSR (dcsr_trans semantics).
void csr_transpose(const DevCSR &A, DevCSR &AT);

// ---- kernels_dense.hip ----------------------------------------------------
// Explicit inverse of the dense coarsest matrix (block Gauss-Jordan).
// Returns 0 on success.
int coarse_dense_setup(const DevCSR &A, double *Ainv);
// x = Ainv b
void coarse_dense_solve(int n, const double *Ainv, const double *b, double *x);
// In-place inverse of a dense row-major n x n device matrix (block Gauss-Jordan,
// rank-64 updates on the FP64 matrix cores).  Returns nonzero on a zero pivot.
int dense_inverse_inplace(int n, double *M);
bool dense_mfma_active();
// Single-workgroup CG for coarse_solver = 0 (coarse_itsolver, mgcycle.c:34-53).
// coarse_itsolver (mgcycle.c:34-53) in one workgroup: CG, then GMRES(30) from
// CG's last iterate when CG fails.  Optional device outputs: status[0] the
// result code, status[1] CG's; x_cg CG's last iterate.
constexpr int kCoarseGmresRestart = 30;
inline size_t coarse_itsolver_work(int n) { return (size_t)(kCoarseGmresRestart + 3) * (size_t)n + 4; }
void coarse_cg(const DevSELL &S, const double *b, double *x, double tol, int maxit, double *work,
               int *status = nullptr, double *x_cg = nullptr);

// ---- kernels_sa.hip: smoothed aggregation ---------------------------------
// P = (I - omega D^-1 M) P_tent with M = A (filter 0) or the filtered strength
// graph N (filter 1; N's diagonal is modified in place).  omega_agg (optional):
// row i uses omega_agg[agg[i]] instead of omega.
void sa_prolongation(const DevCSR &A, DevCSR &N, const int *agg, int nagg, double omega, int filter,
                     DevCSR &P, const float *omega_agg = nullptr);
// AT = A^T with every row listing its source rows in increasing order (dcsr_trans).
void csr_transpose_gpu(const DevCSR &A, DevCSR &AT);
// C = R A P with the layout and summation order of dcsr_rap.
// narrow_first: try tables of half the width first (rows of the finest
// level's coarse operator are usually narrow); the same matrix either way.
void rap_general(const DevCSR &R, const DevCSR &A, const DevCSR &P, DevCSR &C, bool narrow_first = false);
// hz_options.rap_wave_stage_pairs: (R entry, A entry) pairs the wavefront
// Galerkin kernel stages in LDS at a time; 0 = its capacity, -1 = every row
// through the one-thread-per-row kernel, -2 = the default with a 16-column
// limit.
extern int g_rap_wave_pairs;
// hz_options.hem_pipeline_max_sweeps: sweeps after which HEM's one-launch
// pipeline hands a chunk to the per-round kernel (0 = 8192).
extern int g_hem_pipe_max_sweeps;
// HEM levels whose pipeline gave up (process-wide count).
extern int g_hem_pipe_giveups;
// y = A x (row sums from 0.0, CSR order) and x += alpha P xc (dcsr_aAxpy).
void csr_mxv(const DevCSR &A, const double *x, double *y);
void csr_prolong(const DevCSR &P, const double *xc, double alpha, const double *alpha_dev, double *x);

// ---- kernels_learned.hip: Chebyshev smoother, aggregate features ----------
struct ChebLevel {
    double *diag = nullptr;   // first diagonal entry per row
    double lmax = 0.0, lmin = 0.0;
    double gersh = 0.0;       // max_i sum_j |a_ij| / |a_ii| (lmax = scale * gersh)
};
// diag and lmax = scale * Gershgorin bound of D^-1 A, lmin = lmax / 30
ChebLevel cheby_prepare(const DevCSR &A, double scale);
// `steps` Chebyshev steps from x (x and xt swap roles each step; x ends on the result)
// A (optional): the level's CSR, used on small levels with long rows.
// x_zero: x is zero (its memory need not hold it); the first step then reads no matrix.
void cheby_apply(const DevSELL &S, const ChebLevel &c, int steps, const double *b, double *&x, double *&xt,
                 double *dwork, const DevCSR *A = nullptr, bool x_zero = false);
// network outputs [nagg][2] -> omega[c] = clamp(out[c][0], 0, 1.6); returns
// the level's Chebyshev bound scale clamp(mean_c out[c][1], 0.2, 1.5)
double learned_parameters(int nagg, const float *out, float *omega);
// lambda_max(D^-1 A) as the largest Ritz value of `steps` Lanczos steps on
// D^-1/2 A D^-1/2 (d: the diagonal the Chebyshev smoother divides by); an
// estimate from below.  *steps_done (optional): steps before the Krylov space
// became invariant.
double lanczos_lambda_dinv(const DevCSR &A, const double *d, int steps, int *steps_done);
// The library's second stream (default priority).
hipStream_t side_stream();
// The same estimate (all `steps`, no early stop) queued on a second stream
// (default priority), so that it runs while the host works; finish() waits
// for it and returns the largest Ritz value, drain() only waits.
struct LanczosAsync {
    bool active = false;
    double *ab = nullptr;
    int nsteps = 0;
    void start(const DevCSR &A, const double *d, int steps);
    double finish();
    void drain();
    ~LanczosAsync() {
        if (active) (void)hipStreamSynchronize(side_stream());
    }
};
// feat [nagg][12] (FP32) of level `level`
void aggregate_features(const DevCSR &A, const int *agg, int nagg, const int *rptr, const int *rmem, int level,
                        float *feat);

// ---- kernels_mlp.hip: the per-aggregate parameter network -----------------
constexpr int kMlpWgradBlocks = 256;
struct MlpWork {
    int cap = 0;   // rows (multiple of 64)
    float *X = nullptr, *Z1 = nullptr, *H1 = nullptr, *Z2 = nullptr, *H2 = nullptr, *Y = nullptr;
    float *T = nullptr, *dY = nullptr, *dH = nullptr, *dZ = nullptr, *part = nullptr;
    double *lpart = nullptr;
};
MlpWork mlp_work(int capacity);
// X = rows idx[0..B) (or 0..B) of x [.][12], padded
void mlp_inputs(const float *x, const int *idx, int B, const MlpWork &w);
void mlp_targets(const float *t, const int *idx, int B, const MlpWork &w);
void mlp_forward(const float *theta, const MlpWork &w, int B);
void mlp_outputs(int B, const MlpWork &w, float *y);   // y [B][2]
// gradient of the loss w.r.t. theta (after mlp_forward with the same inputs) and the loss
void mlp_backward(const float *theta, const MlpWork &w, int B, float *grad, float *loss);
void mlp_adam(float *theta, const float *grad, float *m, float *v, float lr, int step);
// Forward pass with device x, y and theta through the static scratch array,
// queued on st; no device memory is allocated.  B <= HZ_MLP_NOALLOC_MAX_ROWS.
int mlp_forward_noalloc(const float *theta, const float *x, int B, float *y, hipStream_t st);
// The HIP runtime is up with a device visible (no allocation; the caller's
// libc random() state is kept).
bool runtime_visible();

// ---- kernels_mixed.hip: the V-cycle in FP32 / BF16 ------------------------
constexpr int kLowF32 = 1, kLowBF16 = 2;
// Low-precision copies of a level's values: SELL-64 order (sharing DevSELL's
// column indices) and CSR order (sharing DevCSR's), FP32 or BF16.
struct LowMatrix {
    int kind = 0;
    void *sell = nullptr;
    void *csr = nullptr;
    // the SELL column indices again as 16-bit offsets from the row (padding
    // kColPad16), two entries of a row to a 32-bit word, slice s from word
    // sptr[s] / 2 + 32 s, when every stored entry's fits; null otherwise
    unsigned int *dcol = nullptr;
};
void low_values(const DevSELL &S, const DevCSR &A, int kind, LowMatrix &out, bool col16 = false);
// queue the conversion of S's columns into paired 16-bit offsets (col_base);
// the device int *misfit is cleared first and set when some offset does not fit
unsigned int *col16_launch(const DevSELL &S, int *misfit);
// HZ_COL16=0 switches the 16-bit column offsets off
bool col16_enabled();
// FP32 copy of a CSR matrix's values (smoothed aggregation's P and R).
float *values_f32(const DevCSR &A);
// `steps` Chebyshev steps in FP32 on the level's low-precision values (x and
// xt swap roles each step); with out64 the last step writes x + d to out64 in
// FP64 instead, with (out64, r64) summed into *dot_out when dot_out is set.
// x_zero: as for cheby_apply; first_done: x and dwork already hold the first
// step from zero (formed where b was).
void cheby_apply_m(const DevSELL &S, const DevCSR &A, const LowMatrix &M, const ChebLevel &c, const float *diag,
                   int steps, const float *b, float *&x, float *&xt, float *dwork, double *out64, const double *r64,
                   double *dot_out, const ReduceWs &ws, bool x_zero = false, bool first_done = false);
// y = A x and x += P xc in FP32 with FP32 values (smoothed aggregation transfers)
void csr_mxv_m(const DevCSR &A, const float *val, const float *x, float *y, const float *diag = nullptr,
               float c2 = 0.0f, float *xfirst = nullptr, float *dfirst = nullptr);
// xfirst (csr_mxv_m): also the next level's first Chebyshev step from zero on y
// (diag, c2 = cheb_first_c2 of that level; dfirst receives its d)
float cheb_first_c2(const ChebLevel &c);
void csr_prolong_m(const DevCSR &P, const float *val, const float *xc, float *x);
void d2f(const double *a, float *b, long n);
void f2d(const float *a, double *b, long n);
void fill_f(float *x, long n);
// b = fl(r); x = first sweep from zero when x != null; with dcheb, the
// Chebyshev smoother's first step from zero instead (d: its diagonal, c2 its
// first coefficient; dcheb receives the step's d)
void enter_m(int n, const double *r, int smoother, float w, const float *d, float *b, float *x, float c2 = 0.0f,
             float *dcheb = nullptr);
// xout = sweep(x [+ P xc when agg != null]); with out64: writes the result in
// FP64 to out64 instead and sum_i out64_i r64_i to dot_out.
void sweep_m(const DevSELL &S, const LowMatrix &M, int smoother, float w, const int *agg, const float *xc,
             const float *b, const float *x, float *xout, double *out64, const double *r64, double *dot_out,
             const ReduceWs &ws);
// A (optional): the level's CSR; small levels with long rows then use it.
void residual_m(const DevSELL &S, const LowMatrix &M, const float *b, const float *x, float *w,
                const DevCSR *A = nullptr);
void restrict_m(int nc, int n, const int *rptr, const int *rmem, const float *w, float *bc, int smoother, float wgt,
                const float *dc, float *xc);
void down_fused_m(int nc, const int *rptr, const int *rmem, const DevCSR &A, const LowMatrix &M, const float *b,
                  const float *x, float *bc, int smoother, float wgt, const float *dc, float *xc);
void prolong_m(int n, const int *agg, const float *xc, float *x);
void coarse_dense_solve_m(int n, const double *Ainv, const float *b, float *x);

// ---- HEC turn order -------------------------------------------------------
// Turn order of HEC from libc rand() (iarray_shuffle, array.c:302-318).
void hec_turn_order(int n, int *perm_dev, int rng_source);

// ---- kernels_gs.hip: Gauss-Seidel family as multicolour sweeps -------------
struct GsLevel {
    int ncolors = 0, nslices = 0;
    std::vector<int> cptr;   // host: colour c holds row-list positions [cptr[c], cptr[c+1])
    std::vector<int> sl0;    // host: colour c holds slices [sl0[c], sl0[c+1])
    int *rows = nullptr;     // colour-major row list (rows ascending inside a colour)
    int *spos = nullptr, *scnt = nullptr;   // per slice: first list position, rows
    long long *sptr = nullptr;               // per slice: entry offset (nslices + 1)
    int *col = nullptr;
    double *val = nullptr;   // off-diagonal entries, slice column-major (col: list positions)
    double *diag = nullptr;  // a_ii per list position
    int *dsl0 = nullptr;     // sl0 on the device
    int n = 0;
    bool in_lds = false;     // small level: a whole smoothing call is one workgroup, x in LDS
    double *xp = nullptr, *bp = nullptr;   // x and b permuted to the list order (other levels)
};
bool is_gs_family(int smoother);
void gs_prepare(const DevCSR &A, GsLevel &G);
// pre (post = false) or post smoothing with smoother 2, 3, 5, 6, 7 or 8, in place
void gs_smooth(const GsLevel &G, int smoother, bool post, int nsweeps, double relax, const double *b, double *x);

// ---- kernels_batch.hip: many systems x kBatchRhs right-hand sides ---------
constexpr int kBatchRhs = 8;
struct BatchStats {
    double setup_ms = 0, solve_ms = 0, op_complexity = 0;
    int num_levels = 0, graph_launches = 0, shape_level0 = 0;
    bool time_kernels = false;   // also time the level-0 SpMM and FP32 sweep
    double spmm_ms = 0, sweep_ms = 0;
    long long rows = 0, nnz = 0, sell_entries = 0;
    int sweep_shape0 = 0, spmm_shape0 = 0;   // level-0 kernel shape codes
    int tuned = 0, tune_cached = 0;          // tuner decisions timed / taken from the cache
    double tune_ms = 0;                      // host wall time of the timed decisions
};
// A0: block-diagonal stack of the systems (device CSR, global column indices),
// off0: nsys + 1 row offsets; b, u: [n][kBatchRhs] FP64 on the device (u: initial
// guess in, solution out); iters: nsys * kBatchRhs results (iterations or an
// error code per column).
int batch_solve(const DevCSR &A0, const std::vector<int> &off0, const double *b, double *u,
                const linear_itsolver_param *itp, AMG_param *amg, double jacobi_w, int *iters, BatchStats *stats);
// strength graph with a per-row squared threshold
void strength_graph_rows(const DevCSR &A, const double *th2_row, DevCSR &N, int *counts_ws);

}  // namespace hz
