// amg.hip -- AMG setup loop, V/W cycle and PCG driver on the device.
//
// Control flow follows the reference line by line:
//   setup  : amg_setup_unsmoothP_unsmoothR  (src/solver/amg_setup_ua.c:1228-1443)
//   cycle  : mgcycle                        (src/solver/mgcycle.c:245-409)
//   precond: precond_amg                    (src/solver/precond.c:164-183)
//   PCG    : dcsr_pcg_w_cond_est            (src/solver/krylov.c:400-718)
// Every vector operation runs on the GPU; the host only holds the scalar
// control flow (one 256-byte device->host copy per PCG iteration).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "amg.hpp"

namespace hz {

// ---------------------------------------------------------------------------
// Arena / stream
// ---------------------------------------------------------------------------
void *Arena::alloc(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    if (bytes == 0) bytes = 256;
    while (cur_ < chunks_.size()) {
        Chunk &c = chunks_[cur_];
        if (c.used + bytes <= c.size) {
            void *p = c.base + c.used;
            c.used += bytes;
            return p;
        }
        ++cur_;
    }
    const size_t sz = std::max(bytes, min_chunk_);
    char *p = nullptr;
    HZ_CHECK(hipMalloc(&p, sz));
    chunks_.push_back({p, sz, bytes});
    cur_ = chunks_.size() - 1;
    return p;
}

void Arena::reset() {
    for (auto &c : chunks_) c.used = 0;
    cur_ = 0;
}

void Arena::release() {
    for (auto &c : chunks_) (void)hipFree(c.base);
    chunks_.clear();
    cur_ = 0;
}

size_t Arena::bytes_reserved() const {
    size_t s = 0;
    for (auto &c : chunks_) s += c.size;
    return s;
}

size_t Arena::bytes_in_use() const {
    size_t s = 0;
    for (auto &c : chunks_) s += c.used;
    return s;
}

namespace {
Arena g_default_arena;
Arena *g_current_arena = &g_default_arena;
}  // namespace

Arena &arena() { return *g_current_arena; }

void release_library_arena() { g_default_arena.release(); }
size_t library_arena_bytes() { return g_default_arena.bytes_reserved(); }

ArenaScope::ArenaScope(Arena *a) : prev_(g_current_arena) { g_current_arena = a; }
ArenaScope::~ArenaScope() { g_current_arena = prev_; }

Arena::Mark Arena::mark() const {
    if (chunks_.empty()) return {0, 0};
    return {cur_, chunks_[cur_].used};
}

void Arena::rewind(const Mark &m) {
    for (size_t c = m.chunk + 1; c < chunks_.size(); ++c) chunks_[c].used = 0;
    if (m.chunk < chunks_.size()) chunks_[m.chunk].used = m.used;
    cur_ = m.chunk;
}

bool trace_enabled() {
    static int on = -1;
    if (on < 0) {
        const char *e = std::getenv("HZ_TRACE");
        on = (e && e[0] && e[0] != '0') ? 1 : 0;
    }
    return on == 1;
}

double wall_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

hipStream_t stream() {
    static hipStream_t s = nullptr;
    if (!s) HZ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
}

// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------
// Levels up to this many rows run HEM's sequential sweep on the host
// (HZ_HEM_HOST_MAX overrides the default).  Measured on the 3-D Kuhn 128^3
// hierarchy with pinned staging (setup ms by threshold): 0: 43.2, 300K: 29.5,
// 600K: 27.9, 1.1M: 31.3 -- the GPU pipeline's dependency wavefront does not
// shrink with the level, while the host sweep is linear in it.
int g_hem_host_max = 0;   // hz_options.hem_host_max_rows: 0 = default, -1 = GPU pipeline,
                          // -2 = full host sweep, -3 = staged host sweep (every level)

int hem_host_max_rows() {
    if (g_hem_host_max < 0) return -1;
    if (g_hem_host_max > 0) return g_hem_host_max;
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("HZ_HEM_HOST_MAX");
        v = e ? std::atoi(e) : 1100000;
    }
    return v;
}

// HEM for one level.  Small levels copy the whole strength graph and sweep it
// on the host (two transfers, nothing else to wait for); mid-size levels copy
// only GPU-sorted preference lists; the largest run the GPU pipeline, whose
// cost follows the length of the sweep's dependency chains rather than n.
namespace {
constexpr int kHemFullHostRows = 32768;
}  // namespace
int hem_level(const DevCSR &N, const AMG_param *param, int *agg, int *nagg) {
    const int n = N.n;
    int st;
    if (g_hem_host_max == -1) st = hem_aggregate_gpu(N, agg, nagg);
    else if (g_hem_host_max == -2) st = 1;
    else if (g_hem_host_max == -3) st = aggregate_hem_staged(N, agg, nagg);
    else if (n > hem_host_max_rows()) st = hem_aggregate_gpu(N, agg, nagg, true);
    else if (n > kHemFullHostRows) st = aggregate_hem_staged(N, agg, nagg);
    else st = 1;
    // the pipeline met chains longer than its cap: the host sweep over preference lists
    if (st == 2) st = aggregate_hem_staged(N, agg, nagg);
    // the full host sweep: small levels, patterns the other forms cannot
    // express, and the reference's message for a failed pass 2
    if (st == 1 || st == -1017) st = aggregate_host(N, HZ_HEM, param->max_aggregation, agg, nagg);
    return st;
}

// print_cputime, message.c:73-77 (wall seconds here).
void print_cputime(const char *message, double seconds) {
    std::printf("%s costs %.4f seconds.\n", message, seconds);
}

int amg_setup(Hierarchy &H, const DevCSR &A0, AMG_param *param, bool) {
    const double setup_start = wall_ms();
    const short min_cdof = (short)std::max(param->coarse_dof, 1);
    int max_levels = std::max<int>(1, param->max_levels);
    if (max_levels > HZ_MAX_AMG_LVL) max_levels = HZ_MAX_AMG_LVL;
    H.lv.assign(max_levels, Level{});
    H.lv[0].A = A0;
    sell_from_csr(H.lv[0].A, H.lv[0].S);
    int *counts = arena().alloc_n<int>((size_t)A0.n + 1);
    int lvl = 0;
    const int type = param->aggregation_type;
    while (H.lv[lvl].A.n > min_cdof && lvl < max_levels - 1) {
        Level &L = H.lv[lvl];
        const int n = L.A.n;
        L.theta = param->strong_coupled;
        if (type != HZ_VMB && type != HZ_HEC && type != HZ_HEM) {
            std::printf("\n!!! ERROR HAZMATH DANGER: in function '%s' -- Unknown AMG aggregation type !!!\n\n",
                        "amg_setup_unsmoothP_unsmoothR");
            return -1013;
        }
        const bool tr = trace_enabled();
        double t0 = 0, t1 = 0, t2 = 0, t3 = 0;
        if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); t0 = wall_ms(); }
        DevCSR N;
        strength_graph(L.A, param->strong_coupled, N, counts, L.S.dfirst);
        // learned SA: the Chebyshev diagonal and Gershgorin bound, and the Lanczos
        // estimate behind the bound's floor, need only A: the estimate runs on the
        // second stream while the host aggregates
        const bool learned_level = param->AMG_type == HZ_SA_AMG && H.mlp_theta && !(lvl == 0 && H.omega_level0);
        ChebLevel g1;
        LanczosAsync lz;
        if (learned_level) {
            g1 = cheby_prepare(L.A, 1.0);
            if (g1.lmax > 0.0) lz.start(L.A, g1.diag, 30);
        }
        if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); t1 = wall_ms(); }
        L.agg = arena().alloc_n<int>((size_t)n + 1);
        int nagg = 0, st = HZ_SUCCESS;
        if (type == HZ_HEC) {
            int *perm = arena().alloc_n<int>((size_t)n + 1);
            hec_turn_order(n, perm, H.hec_rng);
            const int r = hec_aggregate(N, perm, L.agg);
            if (r == -1016) {
                std::fprintf(stderr, "\n\n%%%% ERROR Stop: negative index encountered in %s()\n\n",
                             "aggregation_hec");
                return -1016;
            }
            nagg = r;
        } else if (type == HZ_HEM) {
            // small levels: the sequential sweep on the host beats the GPU's
            // wavefront (whose length does not shrink with the level)
            st = hem_level(N, param, L.agg, &nagg);
            if (st == -1017) return -1017;
        } else {
            st = aggregate_host(N, type, param->max_aggregation, L.agg, &nagg);
        }
        L.nagg = nagg;
        // amg_setup_ua.c:1325-1328
        if (nagg * 4 > n) param->strong_coupled /= 2;
        else if (nagg * 1.25 < n) param->strong_coupled *= 2;
        if (st < 0) break;
        if (nagg < 1) break;
        if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); t2 = wall_ms(); }
        Level &C = H.lv[lvl + 1];
        if (param->AMG_type == HZ_SA_AMG) {
            // amg_setup_smoothP_smoothR: smoothed P, R = P^T, C = R A P
            const float *omega_agg = nullptr;
            if (lvl == 0 && H.omega_level0) {
                if (nagg != H.omega_level0_n) return HZ_ERROR_INPUT_PAR;
                omega_agg = H.omega_level0;
            } else if (H.mlp_theta) {
                // learned parameters: features of the aggregates -> network ->
                // per-aggregate damping and this level's Chebyshev bound scale
                int *rp = arena().alloc_n<int>((size_t)nagg + 1);
                int *rm = arena().alloc_n<int>((size_t)n + 1);
                build_restriction(n, L.agg, nagg, rp, rm);
                float *feat = arena().alloc_n<float>((size_t)nagg * HZ_MLP_INPUTS + 1);
                aggregate_features(L.A, L.agg, nagg, rp, rm, lvl, feat);
                MlpWork w = mlp_work(nagg);
                float *out = arena().alloc_n<float>((size_t)nagg * HZ_MLP_OUTPUTS + 1);
                mlp_inputs(feat, nullptr, nagg, w);
                mlp_forward(H.mlp_theta, w, nagg);
                mlp_outputs(nagg, w, out);
                float *om = arena().alloc_n<float>((size_t)nagg + 1);
                double scale = learned_parameters(nagg, out, om);
                // floor: the Chebyshev interval must reach the level's lambda_max(D^-1 A).  The Lanczos
                // estimate (30 steps, queued before the aggregation) is from below, hence 1.1x; the
                // Gershgorin bound (scale 1) always covers lambda_max, so the floor never exceeds 1 and is
                // not needed (its result not read) when the network asks for 1 or more.
                if ((int)H.cheb_scale.size() < max_levels) H.cheb_scale.assign(max_levels, 1.0);
                if ((int)H.lambda_ratio.size() < max_levels) H.lambda_ratio.assign(max_levels, 0.0);
                if (g1.lmax > 0.0 && scale < 1.0) {
                    const double ratio = lz.finish() / g1.lmax;
                    const double floor = std::min(1.0, 1.1 * ratio);
                    scale = std::min(1.5, std::max(scale, floor));
                    H.lambda_ratio[lvl] = ratio;
                }
                H.cheb_scale[lvl] = scale;
                L.cheb = g1;   // the diagonal and bound, rescaled when the cycle is prepared
                omega_agg = om;
            }
            double ts[4] = {0, 0, 0, 0};
            if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); ts[0] = wall_ms(); }
            sa_prolongation(L.A, N, L.agg, nagg, param->tentative_smooth, param->smooth_filter, L.P, omega_agg);
            if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); ts[1] = wall_ms(); }
            csr_transpose_gpu(L.P, L.R);
            if (tr) { HZ_CHECK(hipStreamSynchronize(stream())); ts[2] = wall_ms(); }
            rap_general(L.R, L.A, L.P, C.A, lvl == 0);
            if (tr) {
                HZ_CHECK(hipStreamSynchronize(stream()));
                ts[3] = wall_ms();
                std::fprintf(stderr, "[hz]   SA level %d: learned parameters %.2f ms, P %.2f ms (nnz %d), R %.2f ms, "
                             "RAP %.2f ms (nnz %d)\n", lvl, ts[0] - t2, ts[1] - ts[0], L.P.nnz, ts[2] - ts[1],
                             ts[3] - ts[2], C.A.nnz);
            }
            L.sa = true;
        } else {
            L.rptr = arena().alloc_n<int>((size_t)nagg + 1);
            L.rmem = arena().alloc_n<int>((size_t)n + 1);
            build_restriction(n, L.agg, nagg, L.rptr, L.rmem);
            rap_agg(L.A, L.agg, nagg, L.rptr, L.rmem, C.A);
        }
        sell_from_csr(C.A, C.S);
        if (tr) {
            HZ_CHECK(hipStreamSynchronize(stream()));
            t3 = wall_ms();
            std::fprintf(stderr, "[hz] level %d n=%d nnz=%d: strength %.2f ms, aggregate %.2f ms (nagg=%d), "
                         "R+RAP+SELL %.2f ms\n", lvl, n, L.A.nnz, t1 - t0, t2 - t1, nagg, t3 - t2);
        }
        ++lvl;
    }
    H.nl = lvl + 1;
    H.lv[lvl].theta = param->strong_coupled;
    Level &Cst = H.lv[lvl];
    if (param->coarse_solver == HZ_SOLVER_UMFPACK) {
        // the dense inverse holds n^2 values: refused beyond kDenseCoarsestMaxRows
        if (Cst.A.n > kDenseCoarsestMaxRows) {
            std::fprintf(stderr, "hazmath_mi355x: the coarsest level has %d rows; the dense coarsest solve takes at "
                         "most %d (raise max_levels or lower coarse_dof)\n", Cst.A.n, kDenseCoarsestMaxRows);
            return HZ_ERROR_ALLOC_MEM;
        }
        // amg_setup_ua.c:1390-1401: the coarsest matrix is replaced by its sorted transpose
        DevCSR T;
        csr_transpose(Cst.A, T);
        Cst.A = T;
        sell_from_csr(Cst.A, Cst.S);
        Cst.Ainv = arena().alloc_n<double>((size_t)Cst.A.n * Cst.A.n + 1);
        if (coarse_dense_setup(Cst.A, Cst.Ainv) != 0)
            std::fprintf(stderr, "hazmath_mi355x: zero pivot in the coarsest-level Gauss-Jordan inverse\n");
    } else {
        Cst.cgwork = arena().alloc_n<double>(coarse_itsolver_work(Cst.A.n));
    }
    // amg_setup_ua.c:1426-1431 / 1668-1673: print_amg_complexity (message.c:105-136)
    // and the setup time
    if (param->print_level > 0) HZ_CHECK(hipStreamSynchronize(stream()));
    if (param->print_level >= 2) {
        double gc = 0.0, oc = 0.0;
        std::printf("-----------------------------------------------------------\n");
        std::printf("  Level   Num of rows   Num of nonzeros   Avg. NNZ / row   \n");
        std::printf("-----------------------------------------------------------\n");
        for (int l = 0; l < H.nl; ++l) {
            const double avg = (double)H.lv[l].A.nnz / H.lv[l].A.n;
            std::printf("%5lld %13lld %17lld %14.2f\n", (long long)l, (long long)H.lv[l].A.n,
                        (long long)H.lv[l].A.nnz, avg);
            gc += H.lv[l].A.n;
            oc += H.lv[l].A.nnz;
        }
        std::printf("-----------------------------------------------------------\n");
        gc /= H.lv[0].A.n;
        oc /= H.lv[0].A.nnz;
        std::printf("  Grid complexity = %.3f  |", gc);
        std::printf("  Operator complexity = %.3f\n", oc);
        std::printf("-----------------------------------------------------------\n");
    }
    if (param->print_level > 0)
        print_cputime(param->AMG_type == HZ_SA_AMG ? "Smoothed aggregation setup" : "Unsmoothed aggregation setup",
                      (wall_ms() - setup_start) * 1e-3);
    return HZ_SUCCESS;
}

void amg_prepare_cycle(Hierarchy &H, const AMG_param *param) {
    // precond_amg rebuilds its parameters from param_amg_init + param_prec_to_amg
    // (precond.c:173-174, parameters.c:649-675): tol is the default 1e-6.
    CycleParams &cp = H.cp;
    cp.cycle_type = param->cycle_type;
    cp.smoother = param->smoother;
    cp.presmooth = param->presmooth_iter;
    cp.postsmooth = param->postsmooth_iter;
    cp.coarse_solver = param->coarse_solver;
    cp.coarse_scaling = param->coarse_scaling;
    cp.relax = param->relaxation;
    cp.coarse_tol = 1e-6 * 1e-2;
    cp.maxit = param->maxit;
    const int n0 = H.lv[0].A.n;
    int maxgrid = std::max(8192, div_up(H.lv[0].S.nslices, 4) + 1);
    H.ws.partials = arena().alloc_n<double>((size_t)maxgrid * ReduceWs::kSlots);
    H.ws.tickets = arena().alloc_n<unsigned int>(64);
    HZ_CHECK(hipMemsetAsync(H.ws.tickets, 0, 64 * sizeof(unsigned int), stream()));
    H.sc = arena().alloc_n<PcgScalars>(1);
    HZ_CHECK(hipMemsetAsync(H.sc, 0, sizeof(PcgScalars), stream()));
    H.alpha_cs = arena().alloc_n<double>(1);
    for (int l = 0; l < H.nl; ++l) {
        Level &L = H.lv[l];
        const size_t n = (size_t)L.A.n + 1;
        if (l > 0) {
            L.b = arena().alloc_n<double>(n);
            L.x = arena().alloc_n<double>(n);
        }
        L.w = arena().alloc_n<double>(n);
        if (l > 0) L.xt = arena().alloc_n<double>(n);
        else H.x0buf = arena().alloc_n<double>(n);
        if (l < H.nl - 1) {
            L.dsm = arena().alloc_n<double>(n);
            smoother_diag(L.S, cp.smoother, L.dsm);
            if (cp.smoother == HZ_SMOOTHER_CHEBY) {
                const double sc = l < (int)H.cheb_scale.size() ? H.cheb_scale[l] : 1.0;
                if (L.cheb.diag) {   // formed during the setup (learned parameters)
                    L.cheb.lmax = sc * L.cheb.gersh;
                    L.cheb.lmin = L.cheb.lmax / 30.0;
                } else {
                    L.cheb = cheby_prepare(L.A, sc);
                }
                L.cwork = arena().alloc_n<double>(n);
            }
            if (is_gs_family(cp.smoother)) gs_prepare(L.A, L.gs);
        }
        if (cp.precision > 0) {
            L.bf = arena().alloc_n<float>(n);
            L.xf = arena().alloc_n<float>(n);
            if (l < H.nl - 1) {
                L.xtf = arena().alloc_n<float>(n);
                L.wf = arena().alloc_n<float>(n);
                L.dsmf = arena().alloc_n<float>(n);
                d2f(L.dsm, L.dsmf, L.A.n);
                low_values(L.S, L.A, (cp.precision == 2 && l > 0) ? kLowBF16 : kLowF32, L.low,
                           cp.smoother == HZ_SMOOTHER_CHEBY);
                // level 0: the PCG's product reads the same offsets (the
                // cycle's last step leaves them in the last-level cache)
                if (l == 0 && cp.smoother == HZ_SMOOTHER_CHEBY) L.S.pcol = L.low.dcol;
                if (L.sa) {
                    L.Pf = values_f32(L.P);
                    L.Rf = values_f32(L.R);
                }
                if (cp.smoother == HZ_SMOOTHER_CHEBY) {
                    L.cdiagf = arena().alloc_n<float>(n);
                    d2f(L.cheb.diag, L.cdiagf, L.A.n);
                    L.cworkf = arena().alloc_n<float>(n);
                }
            }
        }
    }
    // the FP64 Jacobi / L1 cycle and the PCG's product read 16-bit column
    // offsets on levels of 65,536 rows or more where every offset fits: one
    // conversion a level, one wait for all of them
    if (cp.precision == 0 && (cp.smoother == HZ_SMOOTHER_L1DIAG || cp.smoother == HZ_SMOOTHER_JACOBI) &&
        col16_enabled()) {
        std::vector<int> lv;
        for (int l = 0; l < H.nl - 1; ++l)
            if (H.lv[l].S.n >= 65536) lv.push_back(l);
        if (!lv.empty()) {
            int *misfit = arena().alloc_n<int>(lv.size());
            std::vector<unsigned int *> d(lv.size());
            for (size_t i = 0; i < lv.size(); ++i) d[i] = col16_launch(H.lv[lv[i]].S, misfit + i);
            std::vector<int> h(lv.size(), 1);
            HZ_CHECK(hipMemcpyAsync(h.data(), misfit, sizeof(int) * lv.size(), hipMemcpyDeviceToHost, stream()));
            HZ_CHECK(hipStreamSynchronize(stream()));
            for (size_t i = 0; i < lv.size(); ++i) H.lv[lv[i]].S.pcol = h[i] == 0 ? d[i] : nullptr;
        }
    }
    (void)n0;
}

namespace {

struct CycleState {
    int num_lvl[HZ_MAX_AMG_LVL];
    bool zero[HZ_MAX_AMG_LVL];        // x of the level is logically zero
    bool first_done[HZ_MAX_AMG_LVL];  // zero-iterate sweep already fused into restriction
};

// Chebyshev steps of a smoothing on level l: the caller's count, or below
// level 0 the coarse-level count when one is set.
int cheb_steps_at(const CycleParams &cp, int l, int steps) {
    return (cp.cheb_coarse > 0 && l >= 1) ? cp.cheb_coarse : steps;
}

void presmooth(Hierarchy &H, int l, CycleState &cs) {
    Level &L = H.lv[l];
    const CycleParams &cp = H.cp;
    int sweeps = cp.smoother == HZ_SMOOTHER_CHEBY ? cheb_steps_at(cp, l, cp.presmooth) : cp.presmooth;
    if (cp.smoother == HZ_SMOOTHER_CHEBY || is_gs_family(cp.smoother)) {
        // Chebyshev from zero: its first step reads no x (nor the matrix)
        const bool cheb_from_zero = cp.smoother == HZ_SMOOTHER_CHEBY && cs.zero[l] && sweeps > 0;
        if (cs.zero[l] && !cheb_from_zero) fill(L.x, 0.0, L.A.n);
        cs.zero[l] = false;
        cs.first_done[l] = false;
        if (cp.smoother == HZ_SMOOTHER_CHEBY)
            cheby_apply(L.S, L.cheb, sweeps, L.b, L.x, L.xt, L.cwork, &L.A, cheb_from_zero);
        else gs_smooth(L.gs, cp.smoother, false, sweeps, cp.relax, L.b, L.x);
        return;
    }
    if (cs.zero[l]) {
        if (sweeps == 0) {
            fill(L.x, 0.0, L.A.n);
        } else if (cs.first_done[l]) {
            --sweeps;
        } else {
            smooth_from_zero(L.A.n, cp.smoother, cp.jacobi_w, L.dsm, L.b, L.x);
            --sweeps;
        }
        cs.zero[l] = false;
        cs.first_done[l] = false;
    }
    for (int s = 0; s < sweeps; ++s) {
        smooth_sweep(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt);
        std::swap(L.x, L.xt);
    }
}

// Levels up to this many rows are latency-bound: they use the fused
// one-thread-per-coarse-row restriction and the prolongation folded into the
// sweep (HZ_FUSE_DOWN_MAX overrides the default).
int fused_down_max_rows() {
    static int v = -1;
    if (v < 0) {
        const char *e = std::getenv("HZ_FUSE_DOWN_MAX");
        v = e ? std::atoi(e) : 300000;
    }
    return v;
}

// Prolongation x_l += alpha P x_{l+1} followed by the post-smoothing sweeps;
// the prolongation is folded into the first sweep.  dot_out: when non-null
// and the fused sweep is the last one, it also yields sum_i x_i b_i.
void prolong_postsmooth(Hierarchy &H, int l, double alpha, const double *alpha_dev, double *dot_out,
                        bool *dot_done) {
    Level &L = H.lv[l];
    Level &C = H.lv[l + 1];
    const CycleParams &cp = H.cp;
    if (cp.postsmooth <= 0 || cp.smoother == HZ_SMOOTHER_CHEBY || is_gs_family(cp.smoother)) {
        if (L.sa) csr_prolong(L.P, C.x, alpha, alpha_dev, L.x);
        else prolong_agg(L.A.n, L.agg, C.x, alpha, alpha_dev, L.x);
        if (cp.smoother == HZ_SMOOTHER_CHEBY)
            cheby_apply(L.S, L.cheb, cheb_steps_at(cp, l, cp.postsmooth), L.b, L.x, L.xt, L.cwork, &L.A);
        else if (is_gs_family(cp.smoother)) gs_smooth(L.gs, cp.smoother, true, cp.postsmooth, cp.relax, L.b, L.x);
        return;
    }
    const bool want_dot = dot_out && cp.postsmooth == 1;
    if (L.sa) {
        csr_prolong(L.P, C.x, alpha, alpha_dev, L.x);   // dcsr_aAxpy(alpha, P), mgcycle.c:346
        if (want_dot)
            smooth_sweep_dot(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt, dot_out, H.ws);
        else
            smooth_sweep(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt);
    } else if (L.A.n <= fused_down_max_rows()) {
        // latency-bound level: one kernel
        up_fused(L.S, cp.smoother, cp.jacobi_w, L.agg, C.x, alpha, alpha_dev, L.b, L.x, L.xt,
                 want_dot ? dot_out : nullptr, H.ws);
    } else {
        // bandwidth-bound level: gathering P xc per matrix entry costs more than one pass
        prolong_agg(L.A.n, L.agg, C.x, alpha, alpha_dev, L.x);
        if (want_dot)
            smooth_sweep_dot(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt, dot_out, H.ws,
                             l == 0 ? H.sweep_timer : nullptr);
        else
            smooth_sweep(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt);
    }
    if (want_dot && dot_done) *dot_done = true;
    std::swap(L.x, L.xt);
    for (int s = 1; s < cp.postsmooth; ++s) {
        smooth_sweep(L.S, cp.smoother, cp.jacobi_w, L.b, L.x, L.xt);
        std::swap(L.x, L.xt);
    }
}


// mgcycle, mgcycle.c:245-409 (UA transfers, no Schwarz levels)
void mgcycle(Hierarchy &H, CycleState &cs, double *zr_out, bool *zr_done) {
    const CycleParams &cp = H.cp;
    const int nl = H.nl;
    int l = 0;
    double alpha = 1.0;
    bool alpha_on_device = false;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) cs.num_lvl[i] = 0;
forward:
    while (l < nl - 1) {
        Level &L = H.lv[l];
        Level &C = H.lv[l + 1];
        cs.num_lvl[l]++;
        presmooth(H, l, cs);
        // the next level restarts from zero; fuse its first sweep into the restriction
        const bool fuse = (l + 1 < nl - 1) && cp.presmooth > 0 && !L.sa && cp.smoother != HZ_SMOOTHER_CHEBY &&
                          !is_gs_family(cp.smoother);
        if (L.sa) {
            residual(L.S, L.b, L.x, L.w);
            csr_mxv(L.R, L.w, C.b);    // dcsr_mxv(R), mgcycle.c:302
        } else if (L.A.n <= fused_down_max_rows()) {
            down_fused(L.nagg, L.rptr, L.rmem, L.A, L.b, L.x, C.b, cp.smoother, cp.jacobi_w,
                       fuse ? C.dsm : nullptr, fuse ? C.x : nullptr);
        } else {
            residual(L.S, L.b, L.x, L.w, l == 0 ? H.res_timer : nullptr);
            restrict_agg(L.nagg, L.A.n, L.rptr, L.rmem, L.w, C.b, cp.smoother, cp.jacobi_w,
                         fuse ? C.dsm : nullptr, fuse ? C.x : nullptr);
        }
        ++l;
        cs.zero[l] = true;
        cs.first_done[l] = fuse;
    }
    {
        Level &C = H.lv[nl - 1];
        if (cp.coarse_solver == HZ_SOLVER_UMFPACK) {
            coarse_dense_solve(C.A.n, C.Ainv, C.b, C.x);
        } else {
            // CG from the level's iterate: zero below the finest level, the
            // caller's guess when the finest level is also the coarsest
            if (cs.zero[nl - 1]) fill(C.x, 0.0, C.A.n);
            coarse_cg(C.S, C.b, C.x, cp.coarse_tol, 20 * C.A.n, C.cgwork);
        }
        cs.zero[nl - 1] = false;
    }
    while (l > 0) {
        --l;
        Level &L = H.lv[l];
        Level &C = H.lv[l + 1];
        if (cp.coarse_scaling == 1) {
            coarse_scaling_alpha(C.S, C.x, C.b, H.sc, H.alpha_cs, H.ws);
            alpha_on_device = true;
        }
        prolong_postsmooth(H, l, alpha, alpha_on_device ? H.alpha_cs : nullptr,
                           l == 0 ? zr_out : nullptr, l == 0 ? zr_done : nullptr);
        if (cs.num_lvl[l] < cp.cycle_type) break;
        cs.num_lvl[l] = 0;
    }
    if (l > 0) goto forward;
}

}  // namespace

// precond_amg (precond.c:164-183): b0 = r, x0 = 0, maxit cycles, z = x0.
// When zr_out is given and the last level-0 sweep can produce it, (z, r) is
// written there and true is returned.
// ---------------------------------------------------------------------------
// The same cycle on FP32 vectors and low-precision values (kernels_mixed.hip).
// ---------------------------------------------------------------------------
void presmooth_m(Hierarchy &H, int l, CycleState &cs) {
    Level &L = H.lv[l];
    const CycleParams &cp = H.cp;
    const float w = (float)cp.jacobi_w;
    int sweeps = cp.smoother == HZ_SMOOTHER_CHEBY ? cheb_steps_at(cp, l, cp.presmooth) : cp.presmooth;
    if (cp.smoother == HZ_SMOOTHER_CHEBY) {
        // from zero, the first step reads no x (nor the matrix); it may have
        // been formed already, where b was
        const bool formed = cs.zero[l] && cs.first_done[l] && sweeps > 0;
        const bool from_zero = cs.zero[l] && sweeps > 0 && !formed;
        if (cs.zero[l] && sweeps <= 0) fill_f(L.xf, L.A.n);
        cs.zero[l] = false;
        cs.first_done[l] = false;
        cheby_apply_m(L.S, L.A, L.low, L.cheb, L.cdiagf, sweeps, L.bf, L.xf, L.xtf, L.cworkf, nullptr, nullptr,
                      nullptr, H.ws, from_zero, formed);
        return;
    }
    if (cs.zero[l]) {
        if (sweeps == 0) {
            fill_f(L.xf, L.A.n);
        } else if (cs.first_done[l]) {
            --sweeps;
        } else {
            // every level below 0 enters with its first sweep fused into the
            // restriction, and level 0 in enter_m, so this only zeroes
            fill_f(L.xf, L.A.n);
        }
        cs.zero[l] = false;
        cs.first_done[l] = false;
    }
    for (int s = 0; s < sweeps; ++s) {
        sweep_m(L.S, L.low, cp.smoother, w, nullptr, nullptr, L.bf, L.xf, L.xtf, nullptr, nullptr, nullptr, H.ws);
        std::swap(L.xf, L.xtf);
    }
}

// Prolongation + post-smoothing of level l; on the final pass over level 0 the
// last sweep writes z in FP64 (and (z, r) into dot_out).  Returns true when z
// was written that way.
bool prolong_postsmooth_m(Hierarchy &H, int l, bool final, double *z, const double *r, double *dot_out) {
    Level &L = H.lv[l];
    Level &C = H.lv[l + 1];
    const CycleParams &cp = H.cp;
    const float w = (float)cp.jacobi_w;
    const bool cheb = cp.smoother == HZ_SMOOTHER_CHEBY;
    if (cp.postsmooth <= 0 || L.sa || cheb) {
        // separate prolongation: x += P xc (mgcycle.c:344-347)
        if (L.sa) csr_prolong_m(L.P, L.Pf, C.xf, L.xf);
        else prolong_m(L.A.n, L.agg, C.xf, L.xf);
        if (cp.postsmooth <= 0) return false;
        if (cheb) {
            cheby_apply_m(L.S, L.A, L.low, L.cheb, L.cdiagf, cheb_steps_at(cp, l, cp.postsmooth), L.bf, L.xf, L.xtf,
                          L.cworkf,
                          final ? z : nullptr, r, final ? dot_out : nullptr, H.ws);
            return final;
        }
    }
    for (int s = 0; s < cp.postsmooth; ++s) {
        const bool last = final && s == cp.postsmooth - 1;
        const int *agg = nullptr;
        const float *xc = nullptr;
        if (s == 0 && !L.sa) {
            if (L.A.n <= fused_down_max_rows()) {
                agg = L.agg;
                xc = C.xf;
            } else {
                prolong_m(L.A.n, L.agg, C.xf, L.xf);
            }
        }
        sweep_m(L.S, L.low, cp.smoother, w, agg, xc, L.bf, L.xf, L.xtf, last ? z : nullptr, r,
                last ? dot_out : nullptr, H.ws);
        if (last) return true;
        std::swap(L.xf, L.xtf);
    }
    return false;
}

bool mgcycle_m(Hierarchy &H, CycleState &cs, bool final, double *z, const double *r, double *zr_out) {
    const CycleParams &cp = H.cp;
    const int nl = H.nl;
    const float w = (float)cp.jacobi_w;
    bool z_written = false;
    int l = 0;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) cs.num_lvl[i] = 0;
forward:
    while (l < nl - 1) {
        Level &L = H.lv[l];
        Level &C = H.lv[l + 1];
        cs.num_lvl[l]++;
        presmooth_m(H, l, cs);
        // the next level's first sweep from zero rides on the restriction (UA, Jacobi / L1)
        const bool fuse = (l + 1 < nl - 1) && cp.presmooth > 0 && cp.smoother != HZ_SMOOTHER_CHEBY && !L.sa;
        // Chebyshev below: the next level's first step from zero rides on the restriction
        const bool cheb_fuse = L.sa && cp.smoother == HZ_SMOOTHER_CHEBY && (l + 1 < nl - 1) &&
                               cheb_steps_at(cp, l + 1, cp.presmooth) > 0;
        if (L.sa) {
            residual_m(L.S, L.low, L.bf, L.xf, L.wf, &L.A);
            if (cheb_fuse)   // dcsr_mxv(R), mgcycle.c:302
                csr_mxv_m(L.R, L.Rf, L.wf, C.bf, C.cdiagf, cheb_first_c2(C.cheb), C.xf, C.cworkf);
            else
                csr_mxv_m(L.R, L.Rf, L.wf, C.bf);
        } else if (L.A.n <= fused_down_max_rows()) {
            down_fused_m(L.nagg, L.rptr, L.rmem, L.A, L.low, L.bf, L.xf, C.bf, cp.smoother, w,
                         fuse ? C.dsmf : nullptr, fuse ? C.xf : nullptr);
        } else {
            residual_m(L.S, L.low, L.bf, L.xf, L.wf);
            restrict_m(L.nagg, L.A.n, L.rptr, L.rmem, L.wf, C.bf, cp.smoother, w, fuse ? C.dsmf : nullptr,
                       fuse ? C.xf : nullptr);
        }
        ++l;
        cs.zero[l] = true;
        cs.first_done[l] = fuse || cheb_fuse;
    }
    {
        Level &C = H.lv[nl - 1];
        if (cp.coarse_solver == HZ_SOLVER_UMFPACK) {
            coarse_dense_solve_m(C.A.n, C.Ainv, C.bf, C.xf);
        } else {
            f2d(C.bf, C.b, C.A.n);
            fill(C.x, 0.0, C.A.n);
            coarse_cg(C.S, C.b, C.x, cp.coarse_tol, 20 * C.A.n, C.cgwork);
            d2f(C.x, C.xf, C.A.n);
        }
        cs.zero[nl - 1] = false;
    }
    while (l > 0) {
        --l;
        if (prolong_postsmooth_m(H, l, final && l == 0, z, r, zr_out)) z_written = true;
        if (cs.num_lvl[l] < cp.cycle_type) break;
        cs.num_lvl[l] = 0;
    }
    if (l > 0) goto forward;
    return z_written;
}

bool amg_precond_m(Hierarchy &H, const double *r, double *z, double *zr_out) {
    Level &L0 = H.lv[0];
    const CycleParams &cp = H.cp;
    const int n = L0.A.n;
    if (cp.maxit <= 0) {
        fill(z, 0.0, n);
        return false;
    }
    const bool first_fused = cp.presmooth > 0 && cp.smoother != HZ_SMOOTHER_CHEBY;
    // Chebyshev: its first step from zero is formed with b
    const bool cheb_first = cp.smoother == HZ_SMOOTHER_CHEBY && cheb_steps_at(cp, 0, cp.presmooth) > 0;
    if (cheb_first)
        enter_m(n, r, cp.smoother, (float)cp.jacobi_w, L0.cdiagf, L0.bf, L0.xf, cheb_first_c2(L0.cheb), L0.cworkf);
    else
        enter_m(n, r, cp.smoother, (float)cp.jacobi_w, L0.dsmf, L0.bf, first_fused ? L0.xf : nullptr);
    CycleState cs;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) {
        cs.zero[i] = false;
        cs.first_done[i] = false;
    }
    cs.zero[0] = true;
    cs.first_done[0] = first_fused || cheb_first;
    bool z_written = false;
    for (int it = 0; it < cp.maxit; ++it) z_written = mgcycle_m(H, cs, it == cp.maxit - 1, z, r, zr_out);
    if (!z_written) {
        f2d(L0.xf, z, n);
        return false;
    }
    return zr_out != nullptr;
}

namespace {
// Level-0 ping-pong of the preconditioner: the buffer its iterate starts in,
// chosen so that the last sweep of level 0 writes into z.  (Gauss-Seidel-type
// sweeps work in place: no swaps.)
double *precond_x0(Hierarchy &H, double *z) {
    const CycleParams &cp = H.cp;
    const int swaps = is_gs_family(cp.smoother)
                          ? 0
                          : (cp.presmooth > 0 ? cp.presmooth - 1 : 0) + cp.postsmooth +
                                (cp.maxit > 1 ? (cp.maxit - 1) * (cp.presmooth + cp.postsmooth) : 0);
    return swaps % 2 == 1 ? H.x0buf : z;
}
}  // namespace

bool amg_precond_first_sweep(Hierarchy &H, double *z, FirstSweep &out) {
    const CycleParams &cp = H.cp;
    if (H.cp.precision > 0 || H.nl < 2 || cp.maxit < 1 || cp.presmooth < 1) return false;
    if (cp.smoother != HZ_SMOOTHER_JACOBI && cp.smoother != HZ_SMOOTHER_L1DIAG) return false;
    out.smoother = cp.smoother;
    out.w = cp.jacobi_w;
    out.d = H.lv[0].dsm;
    out.x0 = precond_x0(H, z);
    return true;
}

bool amg_precond(Hierarchy &H, const double *r, double *z, double *zr_out, bool first_done) {
    if (H.cp.precision > 0 && H.nl > 1) return amg_precond_m(H, r, z, zr_out);
    Level &L0 = H.lv[0];
    const CycleParams &cp = H.cp;
    const int n = L0.A.n;
    L0.b = const_cast<double *>(r);
    if (H.nl == 1) {
        // a single level is solved directly by the coarsest solver
        if (cp.coarse_solver == HZ_SOLVER_UMFPACK) {
            for (int it = 0; it < cp.maxit; ++it) coarse_dense_solve(n, L0.Ainv, r, z);
        } else {
            fill(z, 0.0, n);
            for (int it = 0; it < cp.maxit; ++it) coarse_cg(L0.S, r, z, cp.coarse_tol, 20 * n, L0.cgwork);
        }
        if (cp.maxit <= 0) fill(z, 0.0, n);
        return false;
    }
    // the ping-pong is arranged so that the last sweep of level 0 writes into z
    L0.x = precond_x0(H, z);
    L0.xt = L0.x == z ? H.x0buf : z;
    CycleState cs;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) {
        cs.zero[i] = false;
        cs.first_done[i] = false;
    }
    cs.zero[0] = true;
    cs.first_done[0] = first_done;
    bool zr_done = false;
    for (int it = 0; it < cp.maxit; ++it)
        mgcycle(H, cs, it == cp.maxit - 1 ? zr_out : nullptr, &zr_done);
    if (cp.maxit <= 0) fill(L0.x, 0.0, n);
    if (L0.x != z) copy(z, L0.x, n);
    L0.x = nullptr;
    L0.xt = nullptr;
    return zr_done;
}

// ---------------------------------------------------------------------------
// Condition number estimate from the Lanczos tridiagonal (krylov.c:663-705):
// extreme eigenvalues by Sturm bisection (stands in for LAPACK dsterf).
// ---------------------------------------------------------------------------
static int sturm_count(int n, const double *d, const double *e2, double x) {
    int c = 0;
    double q = d[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < n; ++i) {
        if (q == 0.0) q = 1e-300;
        q = d[i] - x - e2[i - 1] / q;
        if (q < 0) ++c;
    }
    return c;
}

static double bisect(int n, const double *d, const double *e2, double lo, double hi, int k) {
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (sturm_count(n, d, e2, mid) > k) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

static double tridiag_condest(int m, const double *alpha, const double *beta) {
    std::vector<double> d0(m), e(m, 0.0), e2(m, 0.0);
    d0[0] = 1.0 / alpha[0];
    for (int i = 0; i < m - 1; ++i) {
        d0[i + 1] = 1.0 / alpha[i + 1] + beta[i] / alpha[i];
        e[i] = std::sqrt(beta[i]) / alpha[i];
        e2[i] = e[i] * e[i];
    }
    double lo = d0[0], hi = d0[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i < m - 1 ? std::fabs(e[i]) : 0.0);
        lo = std::min(lo, d0[i] - r);
        hi = std::max(hi, d0[i] + r);
    }
    const double pad = 1e-12 * (std::fabs(lo) + std::fabs(hi)) + 1e-300;
    const double lmin = bisect(m, d0.data(), e2.data(), lo - pad, hi + pad, 0);
    const double lmax = bisect(m, d0.data(), e2.data(), lo - pad, hi + pad, m - 1);
    return lmax / lmin;
}

static void itinfo(int prt, int stop, int iter, double rel, double absr, double fac) {
    if (prt < 2) return;
    if (iter > 0) {
        std::printf("%6lld | %13.6e   | %13.6e  | %10.4f\n", (long long)iter, rel, absr, fac);
    } else {
        std::printf("-----------------------------------------------------------\n");
        if (stop == HZ_STOP_REL_RES) std::printf("It Num |   ||r||/||b||   |     ||r||      |  Conv. Factor\n");
        else if (stop == HZ_STOP_REL_PRECRES) std::printf("It Num | ||r||_B/||b||_B |    ||r||_B     |  Conv. Factor\n");
        else std::printf("It Num |   ||r||/||x||   |     ||r||      |  Conv. Factor\n");
        std::printf("-----------------------------------------------------------\n");
        std::printf("%6lld | %13.6e   | %13.6e  |     -.-- \n", (long long)iter, rel, absr);
    }
}

namespace {
PcgScalars *pinned_scalars() {
    static PcgScalars *h = nullptr;
    if (!h) HZ_CHECK(hipHostMalloc((void **)&h, sizeof(PcgScalars), hipHostMallocDefault));
    return h;
}

hipEvent_t fetch_event() {
    static hipEvent_t e = nullptr;
    if (!e) HZ_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    return e;
}

// Queue a copy of the device scalars into pinned memory (non-blocking).
void fetch_async(const Hierarchy &H) {
    HZ_CHECK(hipMemcpyAsync(pinned_scalars(), H.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream()));
    HZ_CHECK(hipEventRecord(fetch_event(), stream()));
}

void fetch_wait(PcgScalars &out) {
    HZ_CHECK(hipEventSynchronize(fetch_event()));
    out = *pinned_scalars();
}

void fetch(const Hierarchy &H, PcgScalars &out) {
    PcgScalars *h = pinned_scalars();
    HZ_CHECK(hipMemcpyAsync(h, H.sc, sizeof(PcgScalars), hipMemcpyDeviceToHost, stream()));
    HZ_CHECK(hipStreamSynchronize(stream()));
    out = *h;
}
}  // namespace

// dcsr_pcg_w_cond_est, krylov.c:400-718, with the AMG preconditioner.
PcgResult pcg_solve(Hierarchy &H, const DevSELL &A, const double *b, double *u, double tol,
                    int MaxIt, int stop, int prt) {
    PcgResult res;
    g_alg_bytes = 0.0;
    const int m = A.n;
    const double maxdiff = tol * 1e-4;
    int iter = 0, stag = 1, more_step = 1, restart_step = 1;
    double absres0 = kBigReal, absres = kBigReal, relres = kBigReal;
    double normu = kBigReal, normr0 = kBigReal;
    double *p = arena().alloc_n<double>((size_t)m + 1);
    double *z = arena().alloc_n<double>((size_t)m + 1);
    double *r = arena().alloc_n<double>((size_t)m + 1);
    double *t = arena().alloc_n<double>((size_t)m + 1);
    const int histn = std::max(MaxIt, 1) + 1;
    double *alpha_hist = arena().alloc_n<double>((size_t)histn);
    double *beta_hist = arena().alloc_n<double>((size_t)histn);
    int iter_condest = 0;
    PcgScalars hs{};
    double *sc_rr = &H.sc->rr, *sc_zr = &H.sc->zr;

    auto precond_zr = [&](const double *rr_, double *zz_, bool first_done = false) {
        if (!amg_precond(H, rr_, zz_, sc_zr, first_done)) dot(m, zz_, rr_, sc_zr, H.ws);
    };
    // the PCG update forms the preconditioner's first level-0 sweep when the
    // preconditioner is queued right behind it
    FirstSweep first;
    const bool fuse_first = stop != HZ_STOP_REL_PRECRES && amg_precond_first_sweep(H, z, first);
    auto measure = [&](double &absr) -> double {
        // r = b - A u; then the stop-type residual
        residual(A, b, u, r);
        if (stop == HZ_STOP_REL_PRECRES) {
            precond_zr(r, z);
            fetch(H, hs);
            absr = std::sqrt(std::fabs(hs.zr));
            return absr / normr0;
        }
        norm2sq(m, r, sc_rr, H.ws);
        fetch(H, hs);
        absr = std::sqrt(hs.rr);
        return (stop == HZ_STOP_MOD_REL_RES) ? absr / normu : absr / normr0;
    };

    const bool speculate = (stop != HZ_STOP_REL_PRECRES);
    residual(A, b, u, r);
    precond_zr(r, z);
    switch (stop) {
        case HZ_STOP_REL_RES:
            norm2sq(m, r, sc_rr, H.ws);
            fetch(H, hs);
            absres0 = std::sqrt(hs.rr);
            normr0 = std::max(kSmallReal, absres0);
            relres = absres0 / normr0;
            break;
        case HZ_STOP_REL_PRECRES:
            dot(m, r, z, sc_zr, H.ws);
            fetch(H, hs);
            absres0 = std::sqrt(hs.zr);
            normr0 = std::max(kSmallReal, absres0);
            relres = absres0 / normr0;
            break;
        case HZ_STOP_MOD_REL_RES:
            norm2sq(m, r, sc_rr, H.ws);
            fetch(H, hs);
            absres0 = std::sqrt(hs.rr);
            norm2sq(m, u, sc_rr, H.ws);
            fetch(H, hs);
            normu = std::max(kSmallReal, std::sqrt(hs.rr));
            relres = absres0 / normu;
            break;
        default:
            std::printf("### ERROR: Unrecognised stopping type for %s!\n", "dcsr_pcg_w_cond_est");
            goto FINISHED;
    }
    if (relres < tol || absres0 < 1e-3 * tol) goto FINISHED;
    itinfo(prt, stop, iter, relres, absres0, 0.0);
    copy(p, z, m);   // (z, r) already sits in sc->zr: it becomes temp1 at the first update

    // The preconditioner and the direction update of an iteration are queued
    // before the host has read that iteration's scalars (they do not touch u or
    // r); the rare restart / convergence paths below redo or drop that work.
    while (iter++ < MaxIt) {
        spmv(A, p, t, &H.sc->temp2, H.ws, H.spmv_timer, H.cp.precision > 0 && H.nl > 1 && !A.pcol);
        pcg_update(m, H.sc, p, t, u, r, alpha_hist, iter, H.ws, fuse_first ? &first : nullptr);
        fetch_async(H);
        if (speculate) {
            precond_zr(r, z, fuse_first);
            pcg_direction(m, H.sc, z, p, beta_hist, iter);
        }
        fetch_wait(hs);
        const double alpha = hs.alpha;
        bool redo = false;
        if (stop == HZ_STOP_REL_PRECRES) {
            amg_precond(H, r, z, nullptr);
            dot(m, z, r, sc_zr, H.ws);
            fetch(H, hs);
            absres = std::sqrt(std::fabs(hs.zr));
            relres = absres / normr0;
        } else {
            absres = std::sqrt(hs.rr);
            relres = (stop == HZ_STOP_MOD_REL_RES) ? absres / normu : absres / normr0;
        }
        res.relres_hist.push_back(relres);
        itinfo(prt, stop, iter, relres, absres, absres / absres0);
        // Check I
        if (hs.uinf <= kSmallReal) {
            if (prt > 1)
                std::printf("### HAZMATH WARNING: Iteration stopped due to the solution is almost zero! %s : %lld\n",
                            "dcsr_pcg_w_cond_est", 0LL);
            iter = HZ_ERROR_SOLVER_SOLSTAG;
            break;
        }
        // Check II
        normu = std::sqrt(hs.uu);
        const double reldiff = std::fabs(alpha) * std::sqrt(hs.pp) / normu;
        if ((stag <= 20) & (reldiff < maxdiff)) {
            if (prt >= 4) {
                std::printf("||u-u'|| = %e and the comp. rel. res. = %e.\n", reldiff, relres);
                std::printf("### HAZMATH WARNING: Iteration restarted due to stagnation! %s : %lld\n",
                            "dcsr_pcg_w_cond_est", 0LL);
            }
            relres = measure(absres);
            redo = true;
            if (prt >= 4) std::printf("### HAZMATH WARNING: The actual relative residual = %e!\n", relres);
            if (relres < tol) break;
            if (stag >= 20) {
                if (prt > 1)
                    std::printf("### HAZMATH WARNING: Iteration stopped due to staggnation! %s : %lld\n",
                                "dcsr_pcg_w_cond_est", 0LL);
                iter = HZ_ERROR_SOLVER_STAG;
                break;
            }
            if (stag == 1 && restart_step == 1) iter_condest = iter;
            fill(p, 0.0, m);
            ++stag;
            ++restart_step;
        }
        // Check III
        if (relres < tol) {
            const double computed = relres;
            relres = measure(absres);
            redo = true;
            if (relres < tol) break;
            if (prt >= 4) {
                std::printf("### HAZMATH WARNING: The computed relative residual = %e!\n", computed);
                std::printf("### HAZMATH WARNING: The actual relative residual = %e!\n", relres);
            }
            if (more_step >= 20) {
                if (prt > 1)
                    std::printf("### HAZMATH WARNING: The tolerence might be too small! %s : %lld\n",
                                "dcsr_pcg_w_cond_est", 0LL);
                iter = HZ_ERROR_SOLVER_TOLSMALL;
                break;
            }
            if (stag == 1 && restart_step == 1) iter_condest = iter;
            fill(p, 0.0, m);
            ++more_step;
            ++restart_step;
        }
        absres0 = absres;
        if (!speculate) {
            dot(m, z, r, sc_zr, H.ws);
            pcg_direction(m, H.sc, z, p, beta_hist, iter);
        } else if (redo) {
            // r was recomputed (and p reset): redo z = B(r), (z, r) and p
            precond_zr(r, z);
            pcg_direction(m, H.sc, z, p, beta_hist, iter);
        }
    }
    {
        if (stag == 1 && restart_step == 1) iter_condest = iter;
        if (iter_condest > MaxIt) iter_condest = MaxIt;
        if (iter_condest < 1) {
            std::printf("### HAZMATH WARNING: %s takes too few iterations for estimating condition number!!!\n",
                        "dcsr_pcg_w_cond_est");
        } else {
            std::vector<double> ha(iter_condest), hb(iter_condest);
            HZ_CHECK(hipMemcpyAsync(ha.data(), alpha_hist, sizeof(double) * iter_condest,
                                    hipMemcpyDeviceToHost, stream()));
            HZ_CHECK(hipMemcpyAsync(hb.data(), beta_hist, sizeof(double) * iter_condest,
                                    hipMemcpyDeviceToHost, stream()));
            HZ_CHECK(hipStreamSynchronize(stream()));
            res.condest = tridiag_condest(iter_condest, ha.data(), hb.data());
        }
    }
FINISHED:
    if (prt > 0) {
        if (iter > MaxIt)
            std::printf("### HAZMATH WARNING: Max iter %lld reached with rel. resid. %e.\n", (long long)MaxIt,
                        relres);
        else if (iter >= 0)
            std::printf("Num_iter(krylov.c) = %lld with relative residual %e.\n", (long long)iter, relres);
    }
    HZ_CHECK(hipStreamSynchronize(stream()));
    res.alg_bytes = g_alg_bytes;
    res.relres = relres;
    res.iter = (iter > MaxIt) ? HZ_ERROR_SOLVER_MAXIT : iter;
    return res;
}

// ---------------------------------------------------------------------------
// AMG as a stand-alone iterative solver, amg_solve (amg_solve.c:109-161):
// cycles from the caller's x until ||b - A x|| / ||b|| < tol.
// ---------------------------------------------------------------------------
AmgSolveResult amg_solve(Hierarchy &H, const double *b, double *x, double tol, int MaxIt, int prt) {
    AmgSolveResult res;
    const double solve_start = wall_ms();
    Level &L0 = H.lv[0];
    const int n = L0.A.n;
    double *r = arena().alloc_n<double>((size_t)n + 1);
    PcgScalars hs{};
    norm2sq(n, b, &H.sc->rr, H.ws);
    fetch(H, hs);
    const double sumb = std::sqrt(hs.rr);
    double relres1 = kBigReal, absres0 = sumb;
    itinfo(prt, HZ_STOP_REL_RES, 0, 1.0, sumb, 0.0);
    // the cycle ping-pongs level 0 between x and a scratch vector
    L0.b = const_cast<double *>(b);
    L0.x = x;
    L0.xt = H.x0buf;
    CycleState cs;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) {
        cs.zero[i] = false;
        cs.first_done[i] = false;
    }
    int iter = 0;
    // the reference evaluates both operands of '&' before testing
    while ((++iter <= MaxIt) & (sumb > kSmallReal)) {
        bool unused = false;
        if (H.nl == 1) {
            if (H.cp.coarse_solver == HZ_SOLVER_UMFPACK)
                coarse_dense_solve(n, L0.Ainv, b, L0.x);
            else
                coarse_cg(L0.S, b, L0.x, H.cp.coarse_tol, 20 * n, L0.cgwork);
        } else {
            mgcycle(H, cs, nullptr, &unused);
        }
        residual(L0.S, b, L0.x, r);
        norm2sq(n, r, &H.sc->rr, H.ws);
        fetch(H, hs);
        const double absres = std::sqrt(hs.rr);
        relres1 = absres / sumb;
        const double factor = absres / absres0;
        absres0 = absres;
        itinfo(prt, HZ_STOP_REL_RES, iter, relres1, absres, factor);
        if (relres1 < tol) break;
    }
    if (L0.x != x) copy(x, L0.x, n);
    L0.x = nullptr;
    L0.xt = nullptr;
    L0.b = nullptr;
    if (prt > 0) {
        if (iter > MaxIt)
            std::printf("### HAZMATH WARNING: Max iter %lld reached with rel. resid. %e.\n", (long long)MaxIt,
                        relres1);
        else if (iter >= 0)
            std::printf("Num_iter(amg_solve.c) = %lld with relative residual %e.\n", (long long)iter, relres1);
        HZ_CHECK(hipStreamSynchronize(stream()));
        print_cputime("AMG solve", (wall_ms() - solve_start) * 1e-3);
    }
    HZ_CHECK(hipStreamSynchronize(stream()));
    res.iter = iter;
    res.relres = relres1;
    return res;
}

// ---------------------------------------------------------------------------
// Error contraction of the cycle as a stationary iteration: b = 0, x_0 random
// (splitmix64(seed)), rate = (||x_K||_A / ||x_{K/2}||_A)^(1/(K - K/2)), an
// estimate of the asymptotic factor.  Used to label parameter choices by their
// two-grid convergence factor.
// ---------------------------------------------------------------------------
double error_contraction(Hierarchy &H, int iters, unsigned long long seed) {
    Level &L0 = H.lv[0];
    const int n = L0.A.n;
    std::vector<double> hx(n);
    unsigned long long s = seed;
    for (int i = 0; i < n; ++i) {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        hx[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
    double *x = arena().alloc_n<double>((size_t)n + 1);
    double *b = arena().alloc_n<double>((size_t)n + 1);
    double *w = arena().alloc_n<double>((size_t)n + 1);
    HZ_CHECK(memcpy_sync(x, hx.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    fill(b, 0.0, n);
    PcgScalars hs{};
    auto anorm = [&](const double *v) {
        spmv(L0.S, v, w, &H.sc->rr, H.ws);
        fetch(H, hs);
        return std::sqrt(std::fabs(hs.rr));
    };
    const double e0 = anorm(x);
    L0.b = b;
    L0.x = x;
    L0.xt = H.x0buf;
    CycleState cs;
    for (int i = 0; i < HZ_MAX_AMG_LVL; ++i) {
        cs.zero[i] = false;
        cs.first_done[i] = false;
    }
    // the first half of the iterations removes the fast transient; the rate is
    // measured over the second half
    const int half = iters / 2;
    double eh = e0;
    for (int it = 0; it < iters; ++it) {
        bool unused = false;
        mgcycle(H, cs, nullptr, &unused);
        if (it + 1 == half) eh = anorm(L0.x);
    }
    const double ek = anorm(L0.x);
    L0.x = L0.xt = L0.b = nullptr;
    if (eh == 0.0 || iters - half <= 0) return 0.0;
    return std::pow(ek / eh, 1.0 / (iters - half));
}

}  // namespace hz
