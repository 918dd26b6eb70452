// amg.hpp -- device AMG hierarchy, V/W cycle and PCG driver.
#pragma once

#include <vector>

#include "hz_internal.h"

namespace hz {

struct Level {
    DevCSR A;                 // CSR (setup input; coarsest: sorted transpose for coarse_solver 32)
    DevSELL S;                // SELL-64 copy used by every solve kernel
    double *dsm = nullptr;    // smoother diagonal (Jacobi: last diagonal entry; L1DIAG: row l1 norm)
    int *agg = nullptr;       // P: row i -> aggregate (UNPT = -1)
    int nagg = 0;
    int *rptr = nullptr;      // R = P^T rows
    int *rmem = nullptr;
    double *b = nullptr, *x = nullptr, *w = nullptr, *xt = nullptr;
    double theta = 0.0;       // strong_coupled used to aggregate this level
    bool sa = false;          // smoothed aggregation: transfers use P and R below
    DevCSR P, R;
    ChebLevel cheb;           // Chebyshev smoother data
    double *cwork = nullptr;  // Chebyshev direction vector
    GsLevel gs;               // colouring and colour-sliced matrix (GS family)
    double *Ainv = nullptr;   // coarsest, coarse_solver 32
    double *cgwork = nullptr; // coarsest, coarse_solver 0
    // FP32 cycle (CycleParams::precision > 0)
    LowMatrix low;            // values in FP32 or BF16 (non-coarsest levels)
    float *dsmf = nullptr;    // smoother diagonal rounded to FP32
    float *bf = nullptr, *xf = nullptr, *xtf = nullptr, *wf = nullptr;
    float *Pf = nullptr, *Rf = nullptr;           // smoothed aggregation: P and R values in FP32
    float *cdiagf = nullptr, *cworkf = nullptr;   // Chebyshev: diagonal and direction vector in FP32
};

struct CycleParams {
    int cycle_type = 1;
    int smoother = 1;
    int presmooth = 1, postsmooth = 1;
    int coarse_solver = 0;
    int coarse_scaling = 0;
    double coarse_tol = 1e-8;   // param_amg_init tol 1e-6 times 1e-2 (mgcycle.c:255)
    double jacobi_w = 1.0;
    double relax = 1.0;         // SOR-type relaxation (AMG_param relaxation)
    int maxit = 1;              // cycles per preconditioner application
    int precision = 0;          // 0 FP64 cycle; 1 FP32 cycle; 2 FP32 cycle, BF16 values below level 0
    int cheb_coarse = 0;        // Chebyshev steps per smoothing below level 0 (0: presmooth / postsmooth)
};

struct Hierarchy {
    std::vector<Level> lv;
    int nl = 0;
    CycleParams cp;
    ReduceWs ws;
    PcgScalars *sc = nullptr;
    double *alpha_cs = nullptr;   // coarse scaling factor (device)
    double *x0buf = nullptr;      // level-0 ping-pong partner of the preconditioner output
    int hec_rng = 0;              // 0: GPU stream from the libc state (checked), 1: host rand()
    // learned parameters: when set (SA only), each level's per-aggregate damping
    // and Chebyshev bound scale come from the network with these weights
    const float *mlp_theta = nullptr;
    // when set (SA only): level 0's per-aggregate prolongator damping (device,
    // omega_level0_n entries; the setup fails when level 0 has another count)
    const float *omega_level0 = nullptr;
    int omega_level0_n = 0;
    std::vector<double> cheb_scale;   // per level (default 1.0)
    std::vector<double> lambda_ratio; // learned: lambda_est(D^-1 A) / Gershgorin bound per level
    LaunchTimer *spmv_timer = nullptr;  // when set: events around the PCG's fine-level SpMV
    LaunchTimer *res_timer = nullptr;   // ... the cycle's fine-level residual (SELL form)
    LaunchTimer *sweep_timer = nullptr; // ... the cycle's last fine-level sweep (with (z, r))
};

// print_cputime (message.c:73-77): "<message> costs <seconds> seconds."
void print_cputime(const char *message, double seconds);

// Runs the reference's setup loop on the device.  A0 arrays are borrowed.
// Returns SUCCESS, a negative HAZmath code, or a negative exit code
// (-1016 / -1017 / -1013) for the reference's fatal paths.
int amg_setup(Hierarchy &H, const DevCSR &A0, AMG_param *param, bool keep_for_export);
void amg_prepare_cycle(Hierarchy &H, const AMG_param *param);
// z = B(r): cp.maxit cycles from a zero iterate.  Writes (z, r) to zr_out and
// returns true when the last fine sweep produced it.  first_done: the first
// level-0 sweep from zero is already in amg_precond_first_sweep(H, z)->x0.
bool amg_precond(Hierarchy &H, const double *r, double *z, double *zr_out, bool first_done = false);
// Where and how the preconditioner's first level-0 sweep from zero writes, so
// the PCG update can form it; false when the cycle does not start with such a
// sweep (FP32 cycle, one level, Chebyshev or Gauss-Seidel-type smoothers, no
// pre-smoothing).
bool amg_precond_first_sweep(Hierarchy &H, double *z, FirstSweep &out);

struct PcgResult {
    int iter = 0;
    double relres = 0.0;
    double condest = 0.0;
    double alg_bytes = 0.0;   // algorithmic bytes of the solve's kernels (g_alg_bytes)
    std::vector<double> relres_hist;
};

// PCG with condition estimate, dcsr_pcg_w_cond_est (krylov.c:400-718).
PcgResult pcg_solve(Hierarchy &H, const DevSELL &A, const double *b, double *u, double tol,
                    int maxit, int stop_type, int prtlvl);

struct AmgSolveResult {
    int iter = 0;
    double relres = 0.0;
};

// AMG cycles as the solver, amg_solve (amg_solve.c:109-161); x holds the
// initial guess on entry.  Returns the reference's iteration count (MaxIt+1
// when it did not converge).
AmgSolveResult amg_solve(Hierarchy &H, const double *b, double *x, double tol, int MaxIt, int prt);

// (||x_K||_A / ||x_{K/2}||_A)^(1/(K-K/2)) of the cycle iterated on A x = 0 from a random x_0.
double error_contraction(Hierarchy &H, int iters, unsigned long long seed);

// The dense coarsest solve (coarse_solver SOLVER_UMFPACK) inverts the level
// explicitly; larger coarsest levels are refused with HZ_ERROR_ALLOC_MEM
// (32768^2 doubles = 8 GiB, about a second of FP64 MFMA work).
constexpr int kDenseCoarsestMaxRows = 32768;

// Host-staged sequential aggregation for VMB (and HEM on non-symmetric patterns).
int aggregate_host(const DevCSR &N, int type, int max_aggregation, int *agg_dev, int *nagg);

// Levels up to hem_host_max_rows() rows run HEM on the host (set through
// hz_options.hem_host_max_rows -> g_hem_host_max).
extern int g_hem_host_max;
int hem_host_max_rows();

// HEM on the GPU: SUCCESS, -1017, or 1 when the strength pattern is not
// symmetric (then the host sweep is used).  With host_on_giveup (and no sweep
// cap set through hz_options), a pass 1 whose dependency chains need more
// than kPipeProbeSweeps sweeps of one chunk stops early and returns 2: the
// caller then runs the host sweep, which is linear in n whatever the chains
// (a negative hz_options.hem_pipeline_max_sweeps sets that cap).
int hem_aggregate_gpu(const DevCSR &N, int *agg, int *nagg_out, bool host_on_giveup = false);

// HEM aggregation of one level in the form its size calls for (full host
// sweep, preference lists, or the GPU pipeline; hz_options.hem_host_max_rows):
// SUCCESS or -1017.
int hem_level(const DevCSR &N, const AMG_param *param, int *agg, int *nagg);

// HEM pass 1 on the host over GPU-sorted preference lists, pass 2 on the GPU:
// SUCCESS, -1017, or 1 when the pattern needs the full host sweep.
int aggregate_hem_staged(const DevCSR &N, int *agg_dev, int *nagg);

// Pieces shared by the two HEM paths (kernels_hem.hip).
struct HemPrefLists {
    int *info = nullptr, *pp = nullptr, *cols = nullptr, *flags = nullptr;
};
void hem_pref_lists(const DevCSR &N, HemPrefLists &out);
void hem_agg_from_pairs(const int *pairs_host, int nc, int n, int *agg);
int hem_pass2_from_host(const DevCSR &N, int *agg_dev, int nc, const int *list_host, int nl, int *rounds);

}  // namespace hz
