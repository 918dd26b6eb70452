// api.hip -- exported C ABI (include/hazmath_mi355x.h).
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "amg.hpp"
#include "tune_cache.h"

using namespace hz;

namespace {
hz_options g_opt = {1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};

// The library's device memory for one top-level call: the arena is reset on
// entry, and with hz_options.release_after_call its chunks go back to HIP when
// the call returns (the reference frees its work arrays before returning).
struct CallArena {
    CallArena() { arena().reset(); }
    ~CallArena() {
        if (g_opt.release_after_call) {
            (void)hipStreamSynchronize(stream());
            release_library_arena();
        }
    }
    CallArena(const CallArena &) = delete;
    CallArena &operator=(const CallArena &) = delete;
};
hz_solve_info g_info;
}  // namespace

namespace hz {
// check_error (message.c:191ff) exits with the status; return_on_fatal returns it.
int fatal_status(int code) {
    if (g_opt.return_on_fatal) return code;
    std::fflush(stdout);
    std::exit(code);
}
}  // namespace hz

namespace {

// The caller's libc random() generator (HEC draws its turn order from it, as
// the reference does).  Initialising the HIP runtime consumes random() values,
// so the state is saved around that initialisation and put back.
struct RandomSnapshot {
    char *buf = nullptr;      // header word of the active state
    int words = 0;            // header + degree
    int32_t data[64];
};

alignas(8) char g_switch_state[128];

bool save_random(RandomSnapshot &s) {
    static const int kDegrees[5] = {0, 7, 15, 31, 63};
    char *cur = initstate(1, g_switch_state, sizeof(g_switch_state));   // records cur's position
    if (!cur) return false;
    int32_t hdr;
    std::memcpy(&hdr, cur, sizeof(hdr));
    setstate(cur);
    const int type = hdr % 5;
    if (type < 0 || type > 4) return false;
    s.buf = cur;
    s.words = 1 + kDegrees[type];
    std::memcpy(s.data, cur, sizeof(int32_t) * s.words);
    return true;
}

void restore_random(const RandomSnapshot &s) {
    setstate(g_switch_state);
    std::memcpy(s.buf, s.data, sizeof(int32_t) * s.words);
    setstate(s.buf);
}

bool runtime_up() {
    static int ready = -1;
    if (ready < 0) {
        RandomSnapshot snap;
        const bool saved = save_random(snap);
        int n = 0;
        ready = (hipGetDeviceCount(&n) == hipSuccess && n >= 1) ? 1 : 0;
        if (ready) {
            // bring the runtime fully up now: context, the library's stream,
            // device and pinned allocations, a launch
            try {
                (void)hipFree(nullptr);
                void *d = nullptr, *h = nullptr;
                HZ_CHECK(hipMalloc(&d, 256));
                HZ_CHECK(hipHostMalloc(&h, 256, hipHostMallocDefault));
                HZ_CHECK(hipMemsetAsync(d, 0, 256, stream()));
                HZ_CHECK(hipStreamSynchronize(stream()));
                (void)hipFree(d);
                (void)hipHostFree(h);
            } catch (const DeviceError &) {
                ready = 0;
            }
        }
        if (saved) restore_random(snap);
    }
    return ready == 1;
}

}  // namespace

namespace hz {
bool runtime_visible() {
    static int ready = -1;
    if (ready < 0) {
        RandomSnapshot snap;
        const bool saved = save_random(snap);
        int n = 0;
        ready = (hipGetDeviceCount(&n) == hipSuccess && n >= 1) ? 1 : 0;
        if (saved) restore_random(snap);
    }
    return ready == 1;
}
}  // namespace hz

namespace {

bool device_ready() {
    if (!runtime_up()) {
        std::fprintf(stderr, "hazmath_mi355x: no HIP device available\n");
        return false;
    }
    return true;
}

template <class T>
T *upload(const T *h, size_t n) {
    T *d = arena().alloc_n<T>(n + 1);
    if (n) HZ_CHECK(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, stream()));
    return d;
}

DevCSR upload_csr(const dCSRmat *A) {
    DevCSR d;
    d.n = A->row;
    d.ncol = A->col;
    d.nnz = A->IA[A->row];
    d.ia = upload(A->IA, (size_t)A->row + 1);
    d.ja = upload(A->JA, (size_t)d.nnz);
    d.val = upload(A->val, (size_t)d.nnz);
    return d;
}

void download_csr(const DevCSR &d, dCSRmat *A) {
    A->row = d.n;
    A->col = d.ncol;
    A->nnz = d.nnz;
    A->IA = (INT *)std::calloc((size_t)d.n + 1, sizeof(INT));
    A->JA = (INT *)std::calloc((size_t)d.nnz + 1, sizeof(INT));
    A->val = (REAL *)std::calloc((size_t)d.nnz + 1, sizeof(REAL));
    HZ_CHECK(memcpy_sync(A->IA, d.ia, sizeof(int) * ((size_t)d.n + 1), hipMemcpyDeviceToHost));
    if (d.nnz) {
        HZ_CHECK(memcpy_sync(A->JA, d.ja, sizeof(int) * d.nnz, hipMemcpyDeviceToHost));
        HZ_CHECK(memcpy_sync(A->val, d.val, sizeof(double) * d.nnz, hipMemcpyDeviceToHost));
    }
}


int fatal(int code) {
    // reference: exit(16) / exit(17) / check_error -> exit(status)
    if (g_opt.return_on_fatal) return code;
    std::fflush(stdout);
    if (code == -1016) std::exit(16);
    if (code == -1017) std::exit(17);
    std::exit(HZ_ERROR_AMG_AGG_TYPE);
}

// Features of AMG_param this build implements on the GPU.
int check_supported(const linear_itsolver_param *it, const AMG_param *amg) {
    (void)it;
    if (amg->smoother != HZ_SMOOTHER_JACOBI && amg->smoother != HZ_SMOOTHER_L1DIAG &&
        amg->smoother != HZ_SMOOTHER_CHEBY && !is_gs_family(amg->smoother)) {
        std::printf("### ERROR: Wrong smoother type %lld!\n", (long long)amg->smoother);
        return HZ_ERROR_AMG_SMOOTH_TYPE;
    }
    if (is_gs_family(amg->smoother) && g_opt.vcycle_precision > 0) {
        std::printf("### ERROR: Gauss-Seidel-type smoothers need the FP64 cycle (vcycle_precision = 0)\n");
        return HZ_ERROR_INPUT_PAR;
    }
    if (amg->AMG_type != HZ_UA_AMG && amg->AMG_type != HZ_SA_AMG) {
        std::printf("### ERROR: AMG type %lld is not available in this build\n", (long long)amg->AMG_type);
        return HZ_ERROR_INPUT_PAR;
    }
    if (amg->cycle_type != HZ_V_CYCLE && amg->cycle_type != HZ_W_CYCLE) {
        std::printf("### ERROR: cycle type %lld is not available in this build\n", (long long)amg->cycle_type);
        return HZ_ERROR_INPUT_PAR;
    }
    if (g_opt.cheb_coarse_steps < 0) {
        std::printf("### ERROR: cheb_coarse_steps %lld is negative\n", (long long)g_opt.cheb_coarse_steps);
        return HZ_ERROR_INPUT_PAR;
    }
    if (g_opt.vcycle_precision < 0 || g_opt.vcycle_precision > 2) {
        std::printf("### ERROR: vcycle_precision %lld is not 0, 1 or 2\n", (long long)g_opt.vcycle_precision);
        return HZ_ERROR_INPUT_PAR;
    }
    if (g_opt.vcycle_precision > 0 && amg->coarse_scaling == 1) {
        std::printf("### ERROR: coarse scaling needs the FP64 cycle (vcycle_precision = 0)\n");
        return HZ_ERROR_INPUT_PAR;
    }
    if (amg->Schwarz_levels > 0) {
        std::printf("### ERROR: Schwarz smoothing is not available in this build\n");
        return HZ_ERROR_INPUT_PAR;
    }
    return HZ_SUCCESS;
}

// The batched driver: CG + UA-AMG with HEM pairs, V(pre, post) Jacobi / L1DIAG,
// dense coarsest solve, one cycle per application, relative-residual stop.
int batch_check(const linear_itsolver_param *it, const AMG_param *amg) {
    const char *what = nullptr;
    if (it->linear_itsolver_type != HZ_SOLVER_CG) what = "the Krylov method must be CG (1)";
    else if (it->linear_stop_type != HZ_STOP_REL_RES) what = "the stop type must be 1 (relative residual)";
    else if (amg->AMG_type != HZ_UA_AMG) what = "AMG_type must be UA";
    else if (amg->aggregation_type != HZ_HEM) what = "the aggregation must be HEM (5)";
    else if (amg->smoother != HZ_SMOOTHER_JACOBI && amg->smoother != HZ_SMOOTHER_L1DIAG)
        what = "the smoother must be JACOBI (1) or L1DIAG (10)";
    else if (amg->cycle_type != HZ_V_CYCLE) what = "the cycle must be V";
    else if (amg->coarse_solver != HZ_SOLVER_UMFPACK) what = "the coarse solver must be the direct one (32)";
    else if (amg->coarse_scaling != 0) what = "coarse scaling must be off";
    else if (amg->maxit != 1) what = "one cycle per preconditioner application (maxit = 1)";
    else if (amg->Schwarz_levels > 0) what = "Schwarz levels must be 0";
    if (what) {
        std::printf("### ERROR: batched AMG-PCG: %s\n", what);
        return HZ_ERROR_INPUT_PAR;
    }
    return HZ_SUCCESS;
}

hz_batch_stats g_batch_stats{};

INT batch_core(const DevCSR &A, const std::vector<int> &off, const double *db, double *dx,
               linear_itsolver_param *it, AMG_param *amg, INT *iters) {
    const int st = batch_check(it, amg);
    if (st < 0) return st;
    BatchStats stats;
    stats.time_kernels = g_opt.batch_time_kernels != 0;
    std::vector<int> out((off.size() - 1) * kBatchRhs);
    const int r = batch_solve(A, off, db, dx, it, amg, g_opt.jacobi_weight, out.data(), &stats);
    if (r <= -1000) return fatal(r);
    if (r < 0) return r;
    std::memset(&g_info, 0, sizeof(g_info));
    int maxit = 0;
    for (int v : out) maxit = std::max(maxit, v);
    g_info.iterations = maxit;
    g_info.num_levels = stats.num_levels;
    g_info.setup_ms = stats.setup_ms;
    g_info.solve_ms = stats.solve_ms;
    g_info.total_ms = stats.setup_ms + stats.solve_ms;
    g_batch_stats.setup_ms = stats.setup_ms;
    g_batch_stats.solve_ms = stats.solve_ms;
    g_batch_stats.op_complexity = stats.op_complexity;
    g_batch_stats.num_levels = stats.num_levels;
    g_batch_stats.graph_launches = stats.graph_launches;
    g_batch_stats.shape_level0 = stats.shape_level0;
    g_batch_stats.spmm_ms = stats.spmm_ms;
    g_batch_stats.sweep_ms = stats.sweep_ms;
    g_batch_stats.rows = stats.rows;
    g_batch_stats.nnz = stats.nnz;
    g_batch_stats.sell_entries = stats.sell_entries;
    g_batch_stats.sweep_shape0 = stats.sweep_shape0;
    g_batch_stats.spmm_shape0 = stats.spmm_shape0;
    g_batch_stats.tuned = stats.tuned;
    g_batch_stats.tune_cached = stats.tune_cached;
    g_batch_stats.tune_ms = stats.tune_ms;
    if (iters) std::copy(out.begin(), out.end(), iters);
    return HZ_SUCCESS;
}

// Three timing events, destroyed on every path out of core_solve.
struct SolveEvents {
    hipEvent_t e[3] = {nullptr, nullptr, nullptr};
    SolveEvents() {
        for (auto &x : e) HZ_CHECK(hipEventCreate(&x));
    }
    ~SolveEvents() {
        for (auto &x : e)
            if (x) (void)hipEventDestroy(x);
    }
    SolveEvents(const SolveEvents &) = delete;
    SolveEvents &operator=(const SolveEvents &) = delete;
};

// linear_solver_dcsr_krylov_amg (itsolver.c:846-931) on device-resident data.
INT core_solve(const DevCSR &A, const double *db, double *dx, linear_itsolver_param *it,
               AMG_param *amg, const float *mlp_theta = nullptr) {
    std::memset(&g_info, 0, sizeof(g_info));
    const short prt = it->linear_print_level;
    SolveEvents ev;
    hipEvent_t e0 = ev.e[0], e1 = ev.e[1], e2 = ev.e[2];
    HZ_CHECK(hipEventRecord(e0, stream()));
    const double wall0 = wall_ms();
    if (prt > 0) std::printf(amg->AMG_type == HZ_SA_AMG ? "\n Calling SA AMG ...\n" : "\n Calling UA AMG ...\n");
    int st = check_supported(it, amg);
    if (st < 0) return st;
    const int giveups0 = g_hem_pipe_giveups;
    Hierarchy H;
    H.cp.jacobi_w = g_opt.jacobi_weight;
    H.hec_rng = g_opt.hec_rng_source;
    H.mlp_theta = mlp_theta;
    st = amg_setup(H, A, amg, false);
    if (st <= -1000) return fatal(st);
    if (st < 0) return st;
    H.cp.precision = g_opt.vcycle_precision;
    H.cp.cheb_coarse = (int)g_opt.cheb_coarse_steps;
    amg_prepare_cycle(H, amg);
    H.cp.jacobi_w = g_opt.jacobi_weight;
    HZ_CHECK(hipEventRecord(e1, stream()));
    const double wall_solve0 = wall_ms();
    INT iter;
    double condest = 0.0;
    PcgResult pr;
    if (it->linear_tol < kSmallReal)
        std::printf("### HAZMATH WARNING: Convergence tolerance for iterative solver is too small!\n");
    if (it->linear_maxit <= 0)
        std::printf("### HAZMATH WARNING: Max number of iterations should be a POSITIVE integer!\n");
    if (it->linear_itsolver_type != HZ_SOLVER_CG) {
        std::printf("### ERROR: Unknown itertive solver type %lld!\n", (long long)it->linear_itsolver_type);
        iter = HZ_ERROR_SOLVER_TYPE;
    } else {
        if (prt > 0) {
            std::printf("**********************************************************\n");
            std::printf(" --> using Conjugate Gradient Method:\n");
        }
        LaunchTimer spmv_timer, res_timer, sweep_timer;
        if (g_opt.time_pcg_spmv) {
            H.spmv_timer = &spmv_timer;
            H.res_timer = &res_timer;
            H.sweep_timer = &sweep_timer;
        }
        pr = pcg_solve(H, H.lv[0].S, db, dx, it->linear_tol, it->linear_maxit, it->linear_stop_type, prt);
        if (H.spmv_timer) {
            int cnt = 0;
            const double tot = spmv_timer.total_ms(&cnt);
            g_info.pcg_spmv_count = cnt;
            g_info.pcg_spmv_ms = cnt ? tot / cnt : 0.0;
            const double tr = res_timer.total_ms(&cnt);
            g_info.cycle_residual_ms = cnt ? tr / cnt : 0.0;
            const double ts = sweep_timer.total_ms(&cnt);
            g_info.cycle_sweep_ms = cnt ? ts / cnt : 0.0;
            H.spmv_timer = H.res_timer = H.sweep_timer = nullptr;
        } else {
            g_info.pcg_spmv_count = 0;
            g_info.pcg_spmv_ms = 0.0;
        }
        iter = pr.iter;
        condest = pr.condest;
        std::printf("Estimated condition number = %13.6e\n", condest);
    }
    HZ_CHECK(hipEventRecord(e2, stream()));
    HZ_CHECK(hipEventSynchronize(e2));
    // itsolver.c:201-206 and 917-922
    if (prt >= 2 && iter >= 0) {
        print_cputime("Iterative method", (wall_ms() - wall_solve0) * 1e-3);
        std::printf("**********************************************************\n");
    }
    if (prt >= 1) {
        print_cputime("AMG_Krylov method totally", (wall_ms() - wall0) * 1e-3);
        std::printf("**********************************************************\n");
    }
    float ms_setup = 0.f, ms_solve = 0.f;
    HZ_CHECK(hipEventElapsedTime(&ms_setup, e0, e1));
    HZ_CHECK(hipEventElapsedTime(&ms_solve, e1, e2));
    g_info.iterations = iter;
    g_info.hem_pipeline_giveups = g_hem_pipe_giveups - giveups0;
    for (int l = 0; l < H.nl && l < HZ_MAX_AMG_LVL; ++l) {
        if (l < (int)H.cheb_scale.size() && mlp_theta) g_info.cheb_scale[l] = H.cheb_scale[l];
        if (l < (int)H.lambda_ratio.size()) g_info.lambda_ratio[l] = H.lambda_ratio[l];
    }
    g_info.num_levels = H.nl;
    for (int l = 0; l < H.nl && l < HZ_MAX_AMG_LVL; ++l) {
        g_info.level_rows[l] = H.lv[l].A.n;
        g_info.level_nnz[l] = H.lv[l].A.nnz;
    }
    g_info.setup_ms = ms_setup;
    g_info.solve_ms = ms_solve;
    g_info.total_ms = ms_setup + ms_solve;
    g_info.final_relres = pr.relres;
    g_info.solve_alg_bytes = pr.alg_bytes;
    g_info.condest = condest;
    g_info.relres_count = (int)std::min<size_t>(pr.relres_hist.size(), 4096);
    for (int i = 0; i < g_info.relres_count; ++i) g_info.relres_hist[i] = pr.relres_hist[i];
    if (g_opt.verbose_timing)
        std::printf("hazmath_mi355x: setup %.3f ms, solve %.3f ms, %d levels\n", ms_setup, ms_solve, H.nl);
    std::fflush(stdout);
    return iter;
}

// Preconditioner object behind precond.data (create_precond_amg): the
// hierarchy lives in its own arena so later library calls do not reuse its
// memory, plus device staging vectors for host-array applications.
struct PrecondAmg {
    Arena arena{size_t(16) << 20};
    Hierarchy H;
    int n = 0;
    double *r = nullptr, *z = nullptr;
    Arena::Mark mark{0, 0};
};

}  // namespace

extern "C" {

// parameters.c:158-205
void param_amg_init(AMG_param *p) {
    std::memset(p, 0, sizeof(*p));
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
    p->amli_coef = nullptr;
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
    p->HAZDIR = nullptr;
    p->Schwarz_on_blk = nullptr;
    p->Schwarz_patch_type = nullptr;
    p->damping_param = 1.0;
    p->BSR_alpha = -1000.;
    p->BSR_omega = -1000.;
}

// parameters.c:238-251
void param_linear_solver_init(linear_itsolver_param *p) {
    std::memset(p, 0, sizeof(*p));
    p->linear_print_level = 0;
    p->linear_itsolver_type = 4;
    p->linear_precond_type = 0;
    p->linear_stop_type = HZ_STOP_REL_RES;
    p->linear_maxit = 500;
    p->linear_restart = 100;
    p->linear_tol = 1e-6;
    p->HX_smooth_iter = 1;
}

// sparse.c:27-57
dCSRmat dcsr_create(const INT m, const INT n, const INT nnz) {
    dCSRmat A;
    A.IA = (m > 0) ? (INT *)std::calloc((size_t)m + 1, sizeof(INT)) : nullptr;
    A.JA = (n > 0) ? (INT *)std::calloc((size_t)nnz, sizeof(INT)) : nullptr;
    A.val = (nnz > 0) ? (REAL *)std::calloc((size_t)nnz, sizeof(REAL)) : nullptr;
    A.row = m;
    A.col = n;
    A.nnz = nnz;
    return A;
}

// sparse.c:573-592
void dcsr_free(dCSRmat *A) {
    if (!A) return;
    std::free(A->IA);
    std::free(A->JA);
    std::free(A->val);
    A->IA = nullptr;
    A->JA = nullptr;
    A->val = nullptr;
}

// vec.c:68-79
dvector dvec_create(const INT m) {
    dvector u;
    u.row = m;
    u.val = (m > 0) ? (REAL *)std::calloc((size_t)m, sizeof(REAL)) : nullptr;
    return u;
}

// vec.c:157-163
void dvec_free(dvector *u) {
    if (!u) return;
    std::free(u->val);
    u->val = nullptr;
    u->row = 0;
}

// vec.c:280-295
void dvec_set(INT n, dvector *x, const REAL val) {
    if (n > 0) x->row = n;
    else n = x->row;
    for (INT i = 0; i < n; ++i) x->val[i] = val;
}

// io.c:2296-2346 + format.c:785-827
dCSRmat *dcoo_read_dcsr_p(FILE *fp) {
    if (!fp) return nullptr;
    char buf[512];
    long pos;
    while (true) {
        pos = std::ftell(fp);
        if (std::fscanf(fp, "%511s", buf) != 1) return nullptr;
        if (buf[0] == '[' || buf[0] == '%' || buf[0] == '|') {
            if (!std::fgets(buf, sizeof(buf), fp)) return nullptr;
            continue;
        }
        break;
    }
    std::fseek(fp, pos, SEEK_SET);
    long long m, n, nnz;
    if (std::fscanf(fp, "%lld %lld %lld", &m, &n, &nnz) != 3) return nullptr;
    std::vector<INT> ri(nnz), ci(nnz);
    std::vector<REAL> v(nnz);
    for (long long k = 0; k < nnz; ++k) {
        long long i, j;
        double a;
        if (std::fscanf(fp, "%lld %lld %le", &i, &j, &a) != 3) return nullptr;
        ri[k] = (INT)i;
        ci[k] = (INT)j;
        v[k] = a;
    }
    dCSRmat *A = (dCSRmat *)std::calloc(1, sizeof(dCSRmat));
    *A = dcsr_create((INT)m, (INT)n, (INT)nnz);
    std::vector<INT> next((size_t)m + 1, 0);
    for (long long k = 0; k < nnz; ++k) next[ri[k] + 1]++;
    A->IA[0] = 0;
    for (long long i = 1; i <= m; ++i) {
        A->IA[i] = A->IA[i - 1] + next[i];
        next[i] = A->IA[i];
    }
    next[0] = 0;
    for (long long k = 0; k < nnz; ++k) {
        const INT r = ri[k];
        const INT p = next[r]++;
        A->JA[p] = ci[k];
        A->val[p] = v[k];
    }
    return A;
}

// io.c:2427-2445
dvector *dvector_read_p(FILE *fp) {
    if (!fp) return nullptr;
    long long n;
    if (std::fscanf(fp, "%lld", &n) != 1) return nullptr;
    dvector *b = (dvector *)std::calloc(1, sizeof(dvector));
    *b = dvec_create((INT)n);
    for (long long i = 0; i < n; ++i) {
        double v;
        if (std::fscanf(fp, "%lg", &v) != 1) break;
        b->val[i] = v;
    }
    return b;
}

void hz_get_options(hz_options *opt) { *opt = g_opt; }
void hz_set_options(const hz_options *opt) {
    g_opt = *opt;
    g_hem_host_max = opt->hem_host_max_rows;
    g_rap_wave_pairs = opt->rap_wave_stage_pairs;
    g_hem_pipe_max_sweeps = opt->hem_pipeline_max_sweeps;
}
void hz_get_last_info(hz_solve_info *info) { *info = g_info; }

void hz_release_device_memory(void) {
    if (!runtime_up()) return;
    (void)hipStreamSynchronize(stream());
    release_library_arena();
}

long long hz_device_pool_bytes(void) { return (long long)library_arena_bytes(); }

const char *hz_version(void) { return "hazmath_mi355x 0.1 (gfx950)"; }

INT hz_device_count(void) {
    int n = 0;
    if (!runtime_up() || hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

INT hz_krylov_amg_device(INT n, INT nnz, const INT *dIA, const INT *dJA, const REAL *dval,
                         const REAL *db, REAL *dx, linear_itsolver_param *itparam,
                         AMG_param *amgparam) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR A;
        A.n = n;
        A.ncol = n;
        A.nnz = nnz;
        A.ia = const_cast<int *>(dIA);
        A.ja = const_cast<int *>(dJA);
        A.val = const_cast<double *>(dval);
        return core_solve(A, db, dx, itparam, amgparam);
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

// itsolver.c:846-931
INT linear_solver_dcsr_krylov_amg(dCSRmat *A, dvector *b, dvector *x, linear_itsolver_param *itparam,
                                  AMG_param *amgparam) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        const double *db = upload(b->val, (size_t)A->row);
        double *dx = upload(x->val, (size_t)A->row);
        const INT st = core_solve(dA, db, dx, itparam, amgparam);
        HZ_CHECK(hipMemcpyAsync(x->val, dx, sizeof(double) * (size_t)A->row, hipMemcpyDeviceToHost,
                                stream()));
        HZ_CHECK(hipStreamSynchronize(stream()));
        return st;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_setup_hierarchy(dCSRmat *A, AMG_param *amgparam, hz_hierarchy *out) {
    std::memset(out, 0, sizeof(*out));
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        Hierarchy H;
        H.hec_rng = g_opt.hec_rng_source;
        const int giveups0 = g_hem_pipe_giveups;
        int st = amg_setup(H, dA, amgparam, true);
        g_info.hem_pipeline_giveups = g_hem_pipe_giveups - giveups0;
        if (st <= -1000) return fatal(st);
        if (st < 0) return st;
        out->num_levels = H.nl;
        for (int l = 0; l < H.nl; ++l) {
            download_csr(H.lv[l].A, &out->A[l]);
            out->theta[l] = H.lv[l].theta;
            if (l < H.nl - 1) {
                out->num_aggregates[l] = H.lv[l].nagg;
                out->aggregates[l] = (INT *)std::calloc((size_t)H.lv[l].A.n + 1, sizeof(INT));
                HZ_CHECK(memcpy_sync(out->aggregates[l], H.lv[l].agg, sizeof(int) * H.lv[l].A.n,
                                   hipMemcpyDeviceToHost));
                if (H.lv[l].sa) download_csr(H.lv[l].P, &out->P[l]);
            }
        }
        return st;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

void hz_hierarchy_free(hz_hierarchy *h) {
    for (int l = 0; l < HZ_MAX_AMG_LVL; ++l) {
        dcsr_free(&h->A[l]);
        dcsr_free(&h->P[l]);
        std::free(h->aggregates[l]);
        h->aggregates[l] = nullptr;
    }
    h->num_levels = 0;
}

INT hz_dcsr_mxv(dCSRmat *A, const REAL *x, REAL *y) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR dA = upload_csr(A);
        DevSELL S;
        sell_from_csr(dA, S);
        const double *dx = upload(x, (size_t)A->col);
        double *dy = arena().alloc_n<double>((size_t)A->row + 1);
        ReduceWs ws;
        spmv(S, dx, dy, nullptr, ws);
        HZ_CHECK(hipMemcpyAsync(y, dy, sizeof(double) * A->row, hipMemcpyDeviceToHost, stream()));
        HZ_CHECK(hipStreamSynchronize(stream()));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_dcsr_residual(dCSRmat *A, const REAL *b, const REAL *x, REAL *w) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR dA = upload_csr(A);
        DevSELL S;
        sell_from_csr(dA, S);
        const double *db = upload(b, (size_t)A->row);
        const double *dx = upload(x, (size_t)A->col);
        double *dw = arena().alloc_n<double>((size_t)A->row + 1);
        residual(S, db, dx, dw);
        HZ_CHECK(hipMemcpyAsync(w, dw, sizeof(double) * A->row, hipMemcpyDeviceToHost, stream()));
        HZ_CHECK(hipStreamSynchronize(stream()));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_smoother_sweep(dCSRmat *A, const REAL *b, REAL *x, INT smoother, REAL w) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR dA = upload_csr(A);
        DevSELL S;
        sell_from_csr(dA, S);
        const double *db = upload(b, (size_t)A->row);
        double *dx = upload(x, (size_t)A->row);
        double *dout = arena().alloc_n<double>((size_t)A->row + 1);
        smooth_sweep(S, smoother, w, db, dx, dout);
        HZ_CHECK(hipMemcpyAsync(x, dout, sizeof(double) * A->row, hipMemcpyDeviceToHost, stream()));
        HZ_CHECK(hipStreamSynchronize(stream()));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_time_fine_kernels(INT n, INT nnz, const INT *dIA, const INT *dJA, const REAL *dval, INT reps,
                         hz_kernel_times *out) {
    std::memset(out, 0, sizeof(*out));
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR A;
        A.n = n;
        A.ncol = n;
        A.nnz = nnz;
        A.ia = const_cast<int *>(dIA);
        A.ja = const_cast<int *>(dJA);
        A.val = const_cast<double *>(dval);
        DevSELL S;
        sell_from_csr(A, S);
        double *x = arena().alloc_n<double>((size_t)n + 1);
        double *y = arena().alloc_n<double>((size_t)n + 1);
        double *b = arena().alloc_n<double>((size_t)n + 1);
        double *dot_out = arena().alloc_n<double>(1);
        fill(x, 1.0, n);
        fill(b, 0.5, n);
        ReduceWs ws;
        ws.partials = arena().alloc_n<double>((size_t)std::max(8192, div_up(S.nslices, 4) + 1) * 4);
        ws.tickets = arena().alloc_n<unsigned int>(64);
        HZ_CHECK(hipMemsetAsync(ws.tickets, 0, 64 * sizeof(unsigned int), stream()));
        hipEvent_t e0, e1;
        HZ_CHECK(hipEventCreate(&e0));
        HZ_CHECK(hipEventCreate(&e1));
        auto time_it = [&](auto &&launch) -> double {
            launch();   // warm
            HZ_CHECK(hipEventRecord(e0, stream()));
            for (int r = 0; r < reps; ++r) launch();
            HZ_CHECK(hipEventRecord(e1, stream()));
            HZ_CHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HZ_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / reps;
        };
        out->reps = reps;
        out->spmv_ms = time_it([&] { spmv(S, x, y, nullptr, ws); });
        out->residual_ms = time_it([&] { residual(S, b, x, y); });
        out->jacobi_ms = time_it([&] { smooth_sweep(S, HZ_SMOOTHER_JACOBI, 1.0, b, x, y); });
        out->spmv_dot_ms = time_it([&] { spmv(S, x, y, dot_out, ws); });
        out->sell_entries = S.nent;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_rap_agg(dCSRmat *A, const INT *aggregates, INT num_aggregates, dCSRmat *Ac) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR dA = upload_csr(A);
        int *dagg = upload(aggregates, (size_t)A->row);
        int *rptr = arena().alloc_n<int>((size_t)num_aggregates + 1);
        int *rmem = arena().alloc_n<int>((size_t)A->row + 1);
        build_restriction(A->row, dagg, num_aggregates, rptr, rmem);
        DevCSR dC;
        rap_agg(dA, dagg, num_aggregates, rptr, rmem, dC);
        download_csr(dC, Ac);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_dense_inverse(INT n, REAL *M) {
    if (n < 0) return HZ_ERROR_INPUT_PAR;
    if (n == 0) return HZ_SUCCESS;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        double *dM = upload(M, (size_t)n * n);
        const int st = dense_inverse_inplace(n, dM);
        if (st != 0) return HZ_ERROR_SOLVER_MISC;
        HZ_CHECK(memcpy_sync(M, dM, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_dense_mfma_active(void) {
    if (!device_ready()) return 0;
    try {
        return dense_mfma_active() ? 1 : 0;
    } catch (const DeviceError &) {
        return 0;
    }
}

// itsolver.c:524-640 (UA setup + amg_solve).  x holds the initial guess.
INT linear_solver_amg(dCSRmat *A, dvector *b, dvector *x, AMG_param *param) {
    std::memset(&g_info, 0, sizeof(g_info));
    const int prt = param->print_level;
    if (A->row != A->col) {
        std::printf("### ERROR: A is not a square matrix!\n");
        return fatal_status(HZ_ERROR_MAT_SIZE);
    }
    if (A->nnz <= 0) {
        std::printf("### ERROR: A has no nonzero entries!\n");
        return fatal_status(HZ_ERROR_MAT_SIZE);
    }
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const auto t0 = std::chrono::steady_clock::now();
        int st = check_supported(nullptr, param);
        if (st < 0) return st;
        const DevCSR dA = upload_csr(A);
        const double *db = upload(b->val, (size_t)A->row);
        double *dx = upload(x->val, (size_t)A->row);
        if (prt > 0) std::printf(param->AMG_type == HZ_SA_AMG ? "\n Calling SA AMG ...\n" : "\n Calling UA AMG ...\n");
        Hierarchy H;
        H.cp.jacobi_w = g_opt.jacobi_weight;
        H.hec_rng = g_opt.hec_rng_source;
        st = amg_setup(H, dA, param, false);
        if (st <= -1000) return fatal(st);
        if (st < 0) return st;   // the reference falls back to GMRES here; not on this path
        amg_prepare_cycle(H, param);
        H.cp.jacobi_w = g_opt.jacobi_weight;
        // amg_solve hands the caller's param to mgcycle: coarse tolerance tol * 1e-2
        H.cp.coarse_tol = param->tol * 1e-2;
        const AmgSolveResult r = amg_solve(H, db, dx, param->tol, param->maxit, prt);
        HZ_CHECK(memcpy_sync(x->val, dx, sizeof(double) * (size_t)A->row, hipMemcpyDeviceToHost));
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        g_info.iterations = r.iter;
        g_info.num_levels = H.nl;
        for (int l = 0; l < H.nl && l < HZ_MAX_AMG_LVL; ++l) {
            g_info.level_rows[l] = H.lv[l].A.n;
            g_info.level_nnz[l] = H.lv[l].A.nnz;
        }
        g_info.total_ms = ms;
        g_info.final_relres = r.relres;
        if (prt > 0) {
            print_cputime("AMG totally", ms * 1e-3);
            std::printf("**********************************************************\n");
        }
        std::fflush(stdout);
        return r.iter;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

// precond_amg (precond.c:164-183) on host arrays: z = B r.
void hz_precond_amg_apply(REAL *r, REAL *z, void *data) {
    auto *P = static_cast<PrecondAmg *>(data);
    try {
        ArenaScope scope(&P->arena);
        HZ_CHECK(hipMemcpyAsync(P->r, r, sizeof(double) * (size_t)P->n, hipMemcpyHostToDevice, stream()));
        amg_precond(P->H, P->r, P->z, nullptr);
        HZ_CHECK(memcpy_sync(z, P->z, sizeof(double) * (size_t)P->n, hipMemcpyDeviceToHost));
        P->arena.rewind(P->mark);
    } catch (const DeviceError &) {
        std::fprintf(stderr, "hazmath_mi355x: preconditioner application failed\n");
        std::memset(z, 0, sizeof(double) * (size_t)P->n);
    }
}

void hz_precond_amg_apply_device(const REAL *dr, REAL *dz, void *data) {
    auto *P = static_cast<PrecondAmg *>(data);
    try {
        ArenaScope scope(&P->arena);
        amg_precond(P->H, dr, dz, nullptr);
        HZ_CHECK(hipStreamSynchronize(stream()));
        P->arena.rewind(P->mark);
    } catch (const DeviceError &) {
        std::fprintf(stderr, "hazmath_mi355x: preconditioner application failed\n");
    }
}

// helper.c:206-283: UA setup once; pc->fct applies one preconditioner step.
precond *create_precond_amg(dCSRmat *A, AMG_param *amgparam) {
    if (!device_ready()) return nullptr;
    if (check_supported(nullptr, amgparam) < 0) return nullptr;
    auto *P = new PrecondAmg();
    try {
        const auto t0 = std::chrono::steady_clock::now();
        ArenaScope scope(&P->arena);
        const DevCSR dA = upload_csr(A);
        if (amgparam->print_level > 0) std::printf("\n Calling UA AMG ...\n");
        P->H.cp.jacobi_w = g_opt.jacobi_weight;
        P->H.hec_rng = g_opt.hec_rng_source;
        const int st = amg_setup(P->H, dA, amgparam, false);
        if (st < 0) {
            std::printf("\n In function create_precond: AMG data failed to set up \n");
            P->arena.release();
            delete P;
            return nullptr;
        }
        P->H.cp.precision = g_opt.vcycle_precision;
        P->H.cp.cheb_coarse = (int)g_opt.cheb_coarse_steps;
        amg_prepare_cycle(P->H, amgparam);
        P->H.cp.jacobi_w = g_opt.jacobi_weight;
        P->n = A->row;
        P->r = arena().alloc_n<double>((size_t)P->n + 1);
        P->z = arena().alloc_n<double>((size_t)P->n + 1);
        HZ_CHECK(hipStreamSynchronize(stream()));
        P->mark = P->arena.mark();
        auto *pc = static_cast<precond *>(std::calloc(1, sizeof(precond)));
        pc->data = P;
        pc->fct = hz_precond_amg_apply;
        pc->setup_time = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return pc;
    } catch (const DeviceError &) {
        P->arena.release();
        delete P;
        return nullptr;
    }
}

// helper.c:1254-1260
void apply_precond(REAL *r, REAL *z, precond *pc) { pc->fct(r, z, pc->data); }

void hz_precond_free(precond *pc) {
    if (!pc) return;
    auto *P = static_cast<PrecondAmg *>(pc->data);
    if (P) {
        (void)hipStreamSynchronize(stream());
        P->arena.release();
        delete P;
    }
    std::free(pc);
}

void hz_mlp_init(float *theta, unsigned long long seed) {
    std::memset(theta, 0, sizeof(float) * HZ_MLP_PARAMS);
    auto next = [&seed]() {
        unsigned long long z = (seed += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    struct Layer { int off, fin, fout, rows, cols; };
    const Layer layers[3] = {{0, HZ_MLP_INPUTS, 64, 16, 64}, {16 * 64 + 64, 64, 64, 64, 64},
                             {16 * 64 + 64 + 64 * 64 + 64, 64, HZ_MLP_OUTPUTS, 64, 16}};
    for (const Layer &L : layers) {
        const double a = std::sqrt(6.0 / (L.fin + L.fout));
        for (int i = 0; i < L.fin; ++i)
            for (int j = 0; j < L.fout; ++j) theta[L.off + i * L.cols + j] = (float)((2.0 * next() - 1.0) * a);
    }
}

INT hz_mlp_forward(const float *theta, const float *x, INT B, float *y) {
    if (B < 0) return HZ_ERROR_INPUT_PAR;
    if (B == 0) return HZ_SUCCESS;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const float *dth = upload(theta, HZ_MLP_PARAMS);
        const float *dx = upload(x, (size_t)B * HZ_MLP_INPUTS);
        MlpWork w = mlp_work(B);
        float *dy = arena().alloc_n<float>((size_t)B * HZ_MLP_OUTPUTS);
        mlp_inputs(dx, nullptr, B, w);
        mlp_forward(dth, w, B);
        mlp_outputs(B, w, dy);
        HZ_CHECK(memcpy_sync(y, dy, sizeof(float) * B * HZ_MLP_OUTPUTS, hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_mlp_forward_noalloc(const float *theta, const float *x, INT B, float *y, void *stream) {
    if (B < 0 || B > HZ_MLP_NOALLOC_MAX_ROWS) {
        std::fprintf(stderr, "hazmath_mi355x: hz_mlp_forward_noalloc: %lld rows is outside 0..%d\n", (long long)B,
                     HZ_MLP_NOALLOC_MAX_ROWS);
        return HZ_ERROR_INPUT_PAR;
    }
    if (B == 0) return HZ_SUCCESS;
    if (!theta || !x || !y) return HZ_ERROR_INPUT_PAR;
    if (!runtime_visible()) return HZ_ERROR_DEVICE;
    return mlp_forward_noalloc(theta, x, B, y, static_cast<hipStream_t>(stream));
}

INT hz_mlp_loss_grad(const float *theta, const float *x, const float *t, INT B, float *grad, float *loss) {
    if (B <= 0) return HZ_ERROR_INPUT_PAR;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const float *dth = upload(theta, HZ_MLP_PARAMS);
        const float *dx = upload(x, (size_t)B * HZ_MLP_INPUTS);
        const float *dt = upload(t, (size_t)B * HZ_MLP_OUTPUTS);
        MlpWork w = mlp_work(B);
        float *dg = arena().alloc_n<float>(HZ_MLP_PARAMS);
        float *dl = arena().alloc_n<float>(1);
        HZ_CHECK(hipMemsetAsync(dg, 0, sizeof(float) * HZ_MLP_PARAMS, stream()));
        mlp_inputs(dx, nullptr, B, w);
        mlp_targets(dt, nullptr, B, w);
        mlp_forward(dth, w, B);
        mlp_backward(dth, w, B, dg, dl);
        HZ_CHECK(memcpy_sync(grad, dg, sizeof(float) * HZ_MLP_PARAMS, hipMemcpyDeviceToHost));
        HZ_CHECK(memcpy_sync(loss, dl, sizeof(float), hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_mlp_train(float *theta, const float *x, const float *t, INT B, INT epochs, INT batch, float lr,
                 unsigned long long seed, float *loss_hist) {
    if (B <= 0 || epochs < 0 || batch <= 0) return HZ_ERROR_INPUT_PAR;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        float *dth = upload(theta, HZ_MLP_PARAMS);
        const float *dx = upload(x, (size_t)B * HZ_MLP_INPUTS);
        const float *dt = upload(t, (size_t)B * HZ_MLP_OUTPUTS);
        const int bs = std::min<int>(batch, B);
        MlpWork w = mlp_work(std::max<int>(bs, B));
        float *dg = arena().alloc_n<float>(HZ_MLP_PARAMS);
        float *dm = arena().alloc_n<float>(HZ_MLP_PARAMS);
        float *dv = arena().alloc_n<float>(HZ_MLP_PARAMS);
        float *dl = arena().alloc_n<float>(1);
        int *didx = arena().alloc_n<int>((size_t)B);
        HZ_CHECK(hipMemsetAsync(dm, 0, sizeof(float) * HZ_MLP_PARAMS, stream()));
        HZ_CHECK(hipMemsetAsync(dv, 0, sizeof(float) * HZ_MLP_PARAMS, stream()));
        auto full_loss = [&]() {
            mlp_inputs(dx, nullptr, B, w);
            mlp_targets(dt, nullptr, B, w);
            mlp_forward(dth, w, B);
            mlp_backward(dth, w, B, dg, dl);
            float l = 0.0f;
            HZ_CHECK(memcpy_sync(&l, dl, sizeof(float), hipMemcpyDeviceToHost));
            return l;
        };
        std::vector<int> perm(B);
        unsigned long long s = seed;
        auto next = [&s]() {
            unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            return z ^ (z >> 31);
        };
        int step = 0;
        for (int e = 0; e < epochs; ++e) {
            loss_hist[e] = full_loss();
            for (int i = 0; i < B; ++i) perm[i] = i;
            for (int i = B - 1; i > 0; --i) std::swap(perm[i], perm[next() % (unsigned long long)(i + 1)]);
            HZ_CHECK(hipMemcpyAsync(didx, perm.data(), sizeof(int) * B, hipMemcpyHostToDevice, stream()));
            for (int o = 0; o < B; o += bs) {
                const int nb = std::min(bs, B - o);
                mlp_inputs(dx, didx + o, nb, w);
                mlp_targets(dt, didx + o, nb, w);
                mlp_forward(dth, w, nb);
                mlp_backward(dth, w, nb, dg, dl);
                mlp_adam(dth, dg, dm, dv, lr, ++step);
            }
            HZ_CHECK(hipStreamSynchronize(stream()));   // perm is reused next epoch
        }
        loss_hist[epochs] = full_loss();
        HZ_CHECK(memcpy_sync(theta, dth, sizeof(float) * HZ_MLP_PARAMS, hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_krylov_amg_learned(dCSRmat *A, dvector *b, dvector *x, linear_itsolver_param *itparam,
                          AMG_param *amgparam, const float *theta, INT budget, INT cheb_steps) {
    if (itparam->linear_stop_type != HZ_STOP_REL_RES && itparam->linear_stop_type != HZ_STOP_MOD_REL_RES) {
        std::printf("### ERROR: the learned AMG-PCG judges its result by the true residual: stop type 1 or 3\n");
        return HZ_ERROR_INPUT_PAR;
    }
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const size_t n = (size_t)A->row;
        const DevCSR dA = upload_csr(A);
        const double *db = upload(b->val, n);
        double *dx = upload(x->val, n);
        double *dx0 = upload(x->val, n);
        const float *dth = upload(theta, HZ_MLP_PARAMS);
        // learned method: SA with network damping + Chebyshev smoother
        AMG_param lp = *amgparam;
        lp.AMG_type = HZ_SA_AMG;
        lp.aggregation_type = HZ_VMB;   // aggregates of ~3^d keep the smoothed operators sparse
        lp.smoother = HZ_SMOOTHER_CHEBY;
        lp.presmooth_iter = lp.postsmooth_iter = (SHORT)std::max(1, cheb_steps);
        linear_itsolver_param li = *itparam;
        li.linear_maxit = std::min<int>(itparam->linear_maxit, budget);
        // The FP64 true residual is the judge, normalised as the caller's stop
        // type normalises it (krylov.c:441-466): by max(1e-20, ||b - A x0||)
        // for the relative residual, by ||x|| for the modified one.
        DevSELL S;
        sell_from_csr(dA, S);
        double *w = arena().alloc_n<double>(n + 1);
        auto residual_norm = [&](const double *xv) {
            residual(S, db, xv, w);
            std::vector<double> hw(n);
            HZ_CHECK(memcpy_sync(hw.data(), w, sizeof(double) * n, hipMemcpyDeviceToHost));
            double rr = 0.0;
            for (size_t i = 0; i < n; ++i) rr += hw[i] * hw[i];
            return std::sqrt(rr);
        };
        const double normr0 = std::max(kSmallReal, residual_norm(dx0));
        INT st = core_solve(dA, db, dx, &li, &lp, dth);
        bool ok = st >= 0 && st <= budget;
        if (ok) {
            const double rn = residual_norm(dx);
            double denom = normr0;
            if (itparam->linear_stop_type == HZ_STOP_MOD_REL_RES) {
                std::vector<double> hx(n);
                HZ_CHECK(memcpy_sync(hx.data(), dx, sizeof(double) * n, hipMemcpyDeviceToHost));
                double xx = 0.0;
                for (double v : hx) xx += v * v;
                denom = std::max(kSmallReal, std::sqrt(xx));
            }
            ok = rn <= itparam->linear_tol * denom * 1.0000001;
        }
        const hz_solve_info learned_info = g_info;
        const INT learned_st = st;
        if (!ok) {
            HZ_CHECK(hipMemcpyAsync(dx, dx0, sizeof(double) * n, hipMemcpyDeviceToDevice, stream()));
            AMG_param pp = *amgparam;
            st = core_solve(dA, db, dx, itparam, &pp);
            g_info.fallback_used = 1;
        } else {
            g_info = learned_info;
            g_info.fallback_used = 0;
        }
        g_info.learned_iterations = learned_st;
        g_info.learned_total_ms = learned_info.total_ms;
        HZ_CHECK(memcpy_sync(x->val, dx, sizeof(double) * n, hipMemcpyDeviceToHost));
        return st;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_two_grid_rate(dCSRmat *A, AMG_param *amgparam, REAL omega, REAL lscale, INT iters, REAL *rate) {
    *rate = 0.0;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        AMG_param p = *amgparam;
        p.AMG_type = HZ_SA_AMG;
        p.tentative_smooth = omega;
        p.max_levels = 2;
        p.coarse_dof = 1;
        p.coarse_solver = HZ_SOLVER_UMFPACK;
        p.smoother = HZ_SMOOTHER_CHEBY;
        p.cycle_type = HZ_V_CYCLE;
        Hierarchy H;
        H.hec_rng = g_opt.hec_rng_source;
        H.cheb_scale.assign(2, lscale);
        const int st = amg_setup(H, dA, &p, false);
        if (st <= -1000) return fatal(st);
        if (st < 0) return st;
        amg_prepare_cycle(H, &p);
        *rate = error_contraction(H, iters, 0x5eedULL);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_cycle_rate_agg(dCSRmat *A, AMG_param *amgparam, const float *omega_agg, INT nagg, REAL lscale, INT iters,
                      REAL *rate) {
    *rate = 0.0;
    if (!omega_agg || nagg <= 0 || iters < 2) return HZ_ERROR_INPUT_PAR;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        AMG_param p = *amgparam;
        p.AMG_type = HZ_SA_AMG;
        p.smoother = HZ_SMOOTHER_CHEBY;
        p.cycle_type = HZ_V_CYCLE;
        Hierarchy H;
        H.hec_rng = g_opt.hec_rng_source;
        H.cheb_scale.assign(HZ_MAX_AMG_LVL, lscale);
        H.omega_level0 = upload(omega_agg, (size_t)nagg);
        H.omega_level0_n = nagg;
        const int st = amg_setup(H, dA, &p, false);
        if (st <= -1000) return fatal(st);
        if (st < 0) return st;
        amg_prepare_cycle(H, &p);
        *rate = error_contraction(H, iters, 0x5eedULL);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_aggregate_features(dCSRmat *A, const INT *aggregates, INT num_aggregates, INT level, float *feat) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        int *dagg = upload(aggregates, (size_t)A->row);
        int *rp = arena().alloc_n<int>((size_t)num_aggregates + 1);
        int *rm = arena().alloc_n<int>((size_t)A->row + 1);
        build_restriction(A->row, dagg, num_aggregates, rp, rm);
        float *df = arena().alloc_n<float>((size_t)num_aggregates * HZ_MLP_INPUTS + 1);
        aggregate_features(dA, dagg, num_aggregates, rp, rm, level, df);
        HZ_CHECK(memcpy_sync(feat, df, sizeof(float) * num_aggregates * HZ_MLP_INPUTS, hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_lambda_max_dinv(dCSRmat *A, INT steps, REAL *lambda, REAL *gershgorin) {
    *lambda = 0.0;
    *gershgorin = 0.0;
    if (steps <= 0) return HZ_ERROR_INPUT_PAR;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        const ChebLevel g1 = cheby_prepare(dA, 1.0);
        *gershgorin = g1.lmax;
        *lambda = lanczos_lambda_dinv(dA, g1.diag, steps, nullptr);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_cheby_smooth(dCSRmat *A, const REAL *b, REAL *x, INT steps, REAL scale) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        DevSELL S;
        sell_from_csr(dA, S);
        const double *db = upload(b, (size_t)A->row);
        double *dx = upload(x, (size_t)A->row);
        double *dt = arena().alloc_n<double>((size_t)A->row + 1);
        double *dd = arena().alloc_n<double>((size_t)A->row + 1);
        const ChebLevel c = cheby_prepare(dA, scale);
        cheby_apply(S, c, steps, db, dx, dt, dd, &dA);
        HZ_CHECK(memcpy_sync(x, dx, sizeof(double) * A->row, hipMemcpyDeviceToHost));
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_coarse_itsolver(dCSRmat *A, const REAL *b, REAL *x, REAL ctol, INT *cg_status, REAL *x_cg) {
    if (A->row != A->col) return HZ_ERROR_MAT_SIZE;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        DevSELL S;
        sell_from_csr(dA, S);
        const double *db = upload(b, (size_t)A->row);
        double *dx = upload(x, (size_t)A->row);
        double *dxcg = x_cg ? arena().alloc_n<double>((size_t)A->row + 1) : nullptr;
        double *work = arena().alloc_n<double>(coarse_itsolver_work(A->row));
        int *dst = arena().alloc_n<int>(2);
        HZ_CHECK(hipMemsetAsync(dst, 0, 2 * sizeof(int), stream()));
        coarse_cg(S, db, dx, ctol, 20 * A->row, work, dst, dxcg);
        int st[2] = {0, 0};
        HZ_CHECK(memcpy_sync(st, dst, 2 * sizeof(int), hipMemcpyDeviceToHost));
        HZ_CHECK(memcpy_sync(x, dx, sizeof(double) * A->row, hipMemcpyDeviceToHost));
        if (x_cg) HZ_CHECK(memcpy_sync(x_cg, dxcg, sizeof(double) * A->row, hipMemcpyDeviceToHost));
        if (cg_status) *cg_status = st[1];
        return st[0];
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_batch_krylov_amg_device(INT nsys, const INT *row_offsets, INT nnz, const INT *dIA, const INT *dJA,
                               const REAL *dval, const REAL *db, REAL *dx, linear_itsolver_param *itparam,
                               AMG_param *amgparam, INT *iters) {
    if (nsys <= 0 || !row_offsets || row_offsets[0] != 0) return HZ_ERROR_INPUT_PAR;
    if (batch_check(itparam, amgparam) < 0) return HZ_ERROR_INPUT_PAR;
    for (INT s = 0; s < nsys; ++s)
        if (row_offsets[s + 1] <= row_offsets[s]) return HZ_ERROR_MAT_SIZE;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        DevCSR A;
        A.n = A.ncol = row_offsets[nsys];
        A.nnz = nnz;
        A.ia = const_cast<int *>(dIA);
        A.ja = const_cast<int *>(dJA);
        A.val = const_cast<double *>(dval);
        std::vector<int> off(row_offsets, row_offsets + nsys + 1);
        return batch_core(A, off, db, dx, itparam, amgparam, iters);
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_batch_krylov_amg(INT nsys, dCSRmat *const *A, INT nrhs, REAL *const *b, REAL *const *x,
                        linear_itsolver_param *itparam, AMG_param *amgparam, INT *iters) {
    if (nsys <= 0 || nrhs < 1 || nrhs > HZ_BATCH_RHS) return HZ_ERROR_INPUT_PAR;
    if (batch_check(itparam, amgparam) < 0) return HZ_ERROR_INPUT_PAR;
    std::vector<int> off(nsys + 1, 0);
    long long nnz = 0;
    for (INT s = 0; s < nsys; ++s) {
        if (A[s]->row != A[s]->col || A[s]->row <= 0 || A[s]->nnz <= 0) return HZ_ERROR_MAT_SIZE;
        off[s + 1] = off[s] + A[s]->row;
        nnz += A[s]->IA[A[s]->row];
    }
    if (nnz > INT32_MAX) return HZ_ERROR_MAT_SIZE;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const int n = off[nsys];
        // the block-diagonal stack and the interleaved right-hand sides
        std::vector<int> ia(n + 1), ja(nnz);
        std::vector<double> va(nnz), hb((size_t)n * HZ_BATCH_RHS, 0.0), hx((size_t)n * HZ_BATCH_RHS, 0.0);
        long long pos = 0;
        for (INT s = 0; s < nsys; ++s) {
            const dCSRmat *M = A[s];
            const int r0 = off[s];
            for (int i = 0; i < M->row; ++i) {
                ia[r0 + i] = (int)pos;
                for (int k = M->IA[i]; k < M->IA[i + 1]; ++k, ++pos) {
                    ja[pos] = M->JA[k] + r0;
                    va[pos] = M->val[k];
                }
                for (int q = 0; q < nrhs; ++q) {
                    hb[(size_t)(r0 + i) * HZ_BATCH_RHS + q] = b[s][(size_t)q * M->row + i];
                    hx[(size_t)(r0 + i) * HZ_BATCH_RHS + q] = x[s][(size_t)q * M->row + i];
                }
            }
        }
        ia[n] = (int)pos;
        DevCSR dA;
        dA.n = dA.ncol = n;
        dA.nnz = (int)nnz;
        dA.ia = upload(ia.data(), ia.size());
        dA.ja = upload(ja.data(), ja.size());
        dA.val = upload(va.data(), va.size());
        const double *db = upload(hb.data(), hb.size());
        double *dx = upload(hx.data(), hx.size());
        std::vector<int> all((size_t)nsys * HZ_BATCH_RHS);
        const INT st = batch_core(dA, off, db, dx, itparam, amgparam, all.data());
        if (st < 0) return st;
        HZ_CHECK(memcpy_sync(hx.data(), dx, sizeof(double) * hx.size(), hipMemcpyDeviceToHost));
        for (INT s = 0; s < nsys; ++s) {
            const int r0 = off[s], m = A[s]->row;
            for (int q = 0; q < nrhs; ++q) {
                for (int i = 0; i < m; ++i) x[s][(size_t)q * m + i] = hx[(size_t)(r0 + i) * HZ_BATCH_RHS + q];
                if (iters) iters[(size_t)s * nrhs + q] = all[(size_t)s * HZ_BATCH_RHS + q];
            }
        }
        return st;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

void hz_batch_get_stats(hz_batch_stats *out) {
    if (out) *out = g_batch_stats;
}

INT hz_batch_tune_cache(const char *path) { return TuneCache::get().reset(path, false); }

INT hz_gs_smooth(dCSRmat *A, const REAL *b, REAL *x, INT smoother, INT post, INT nsweeps, REAL relax, INT *order,
                 INT *cptr, INT *ncolors) {
    if (!is_gs_family((int)smoother)) return HZ_ERROR_AMG_SMOOTH_TYPE;
    if (A->row != A->col) return HZ_ERROR_MAT_SIZE;
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        const double *db = upload(b, (size_t)A->row);
        double *dx = upload(x, (size_t)A->row);
        GsLevel G;
        gs_prepare(dA, G);
        gs_smooth(G, (int)smoother, post != 0, (int)nsweeps, relax, db, dx);
        HZ_CHECK(memcpy_sync(x, dx, sizeof(double) * A->row, hipMemcpyDeviceToHost));
        if (order && A->row > 0) HZ_CHECK(memcpy_sync(order, G.rows, sizeof(int) * A->row, hipMemcpyDeviceToHost));
        if (cptr)
            for (int c = 0; c <= G.ncolors; ++c) cptr[c] = G.cptr.empty() ? 0 : G.cptr[c];
        if (ncolors) *ncolors = G.ncolors;
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

INT hz_sa_prolongation(dCSRmat *A, const INT *aggregates, INT num_aggregates, REAL theta, REAL omega,
                       const float *omega_agg, INT filter, dCSRmat *P) {
    if (!device_ready()) return HZ_ERROR_DEVICE;
    try {
        CallArena call_arena;
        const DevCSR dA = upload_csr(A);
        int *dagg = upload(aggregates, (size_t)A->row);
        int *counts = arena().alloc_n<int>((size_t)A->row + 1);
        DevCSR N;
        strength_graph(dA, theta, N, counts);
        const float *dom = omega_agg ? upload(omega_agg, (size_t)num_aggregates) : nullptr;
        DevCSR dP;
        sa_prolongation(dA, N, dagg, num_aggregates, omega, filter, dP, dom);
        download_csr(dP, P);
        return HZ_SUCCESS;
    } catch (const DeviceError &) {
        return HZ_ERROR_DEVICE;
    }
}

}  // extern "C"
