// kernels_gs.hip -- Gauss-Seidel-type smoothers (smoother_dcsr_gs / _sgs /
// _sor, src/solver/smoother.c:109-307, dispatched like dcsr_presmoothing /
// dcsr_postsmoothing, src/solver/mgcycle.c:76-223) as multicolour sweeps.
//
// The rows are coloured so that no two rows of one colour are coupled in
// either direction (Jones-Plassmann on the pattern of A + A^T: a row takes the
// smallest colour its coloured neighbours leave free once its hash priority
// beats every uncoloured neighbour's).  A forward sweep visits the colours in
// increasing order, a backward sweep in decreasing order, and the rows of one
// colour update in parallel.  That is Gauss-Seidel in the colour-major row
// order, bit for bit (the rows of a colour read no value written in the same
// colour), not in the natural order the reference sweeps in: results follow
// the reference's by convergence behaviour, not bit for bit.
//
// Layout: per colour, the rows (ascending) are cut into slices of 64; a slice
// stores its rows' off-diagonal entries column-major (entry k of lane t at
// sptr[s] + 64 k + t, CSR order inside the row, zero-padded to the slice
// width), and the diagonal a_ii separately.  One launch per colour; a wave
// handles one slice.
#include <algorithm>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;
static_assert(kSlice == 64, "one wave per slice");
constexpr int kGsMaxPass = 8;               // passes per single-workgroup launch
constexpr int kGsLdsRows = 8192;            // levels up to this size keep x in LDS (64 KiB)
constexpr int kSlicesPerBlock = kBlock / kSlice;

int grid_rows(long n) { return (int)std::max(1L, std::min<long>((n + kBlock - 1) / kBlock, 16384)); }

__device__ __forceinline__ unsigned jp_priority(int i) {
    unsigned h = (unsigned)i * 2654435761u;
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    return h;
}

// Neighbour j of row i beats i: higher priority, ties by index.
__device__ __forceinline__ bool beats(int j, unsigned pj, int i, unsigned pi) {
    return pj > pi || (pj == pi && j > i);
}

// One Jones-Plassmann round over the pattern of A (ia/ja) and of A^T
// (ta/tj).  color_in[i] = -1 while row i is uncoloured.
__global__ void k_jp_round(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                           const int *__restrict__ ta, const int *__restrict__ tj, const int *__restrict__ color_in,
                           int *__restrict__ color_out, int *__restrict__ left) {
    int local = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int c = color_in[i];
        if (c < 0) {
            const unsigned pi = jp_priority(i);
            bool wins = true;
            for (int pass = 0; pass < 2 && wins; ++pass) {
                const int *p = pass ? ta : ia;
                const int *q = pass ? tj : ja;
                for (int k = p[i]; k < p[i + 1]; ++k) {
                    const int j = q[k];
                    if (j != i && color_in[j] < 0 && beats(j, jp_priority(j), i, pi)) {
                        wins = false;
                        break;
                    }
                }
            }
            if (wins) {
                // smallest colour free among the coloured neighbours, 64 at a time
                for (int base = 0;; base += 64) {
                    unsigned long long used = 0ull;
                    for (int pass = 0; pass < 2; ++pass) {
                        const int *p = pass ? ta : ia;
                        const int *q = pass ? tj : ja;
                        for (int k = p[i]; k < p[i + 1]; ++k) {
                            const int j = q[k];
                            const int cj = j != i ? color_in[j] - base : -1;
                            if (cj >= 0 && cj < 64) used |= 1ull << cj;
                        }
                    }
                    if (~used) {
                        c = base + __ffsll((long long)~used) - 1;
                        break;
                    }
                }
            } else {
                ++local;
            }
        }
        color_out[i] = c;
    }
    if (local && left) atomicAdd(left, local);
}

__global__ void k_offdiag_count(int n, const int *__restrict__ ia, const int *__restrict__ ja, int *__restrict__ cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int c = 0;
        for (int k = ia[i]; k < ia[i + 1]; ++k) c += ja[k] != i;
        cnt[i] = c;
    }
}

// Slice fill: one thread per (slice, lane).  Columns are stored as list
// positions (pos_of[j]): the sweeps work on vectors permuted to the
// colour-major order, where the rows a colour's slice reads lie close
// together, as they do in the natural order.
__global__ void k_gs_fill(int nslices, const int *__restrict__ spos, const int *__restrict__ scnt,
                          const long long *__restrict__ sptr, const int *__restrict__ rows,
                          const int *__restrict__ pos_of, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ a, int *__restrict__ col, double *__restrict__ val,
                          double *__restrict__ diag) {
    const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int s = (int)(gid / kSlice), lane = (int)(gid % kSlice);
    if (s >= nslices) return;
    const long long base = sptr[s];
    const int width = (int)((sptr[s + 1] - base) / kSlice);
    const int pos0 = spos[s];
    const int me = pos0 + (lane < scnt[s] ? lane : 0);
    const int i = rows[me];
    int k2 = 0;
    if (lane < scnt[s]) {
        double d = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            const int j = ja[k];
            if (j == i) {
                d = a[k];
            } else {
                col[base + (long long)k2 * kSlice + lane] = pos_of[j];
                val[base + (long long)k2 * kSlice + lane] = a[k];
                ++k2;
            }
        }
        diag[me] = d;
    }
    for (; k2 < width; ++k2) {   // padding: a zero entry on a valid position
        col[base + (long long)k2 * kSlice + lane] = me;
        val[base + (long long)k2 * kSlice + lane] = 0.0;
    }
}

__global__ void k_pos_of(int n, const int *__restrict__ rows, int *__restrict__ pos_of) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) pos_of[rows[p]] = p;
}

// xp[p] = x[rows[p]], bp[p] = b[rows[p]]
__global__ void k_perm_in(int n, const int *__restrict__ rows, const double *__restrict__ x,
                          const double *__restrict__ b, double *__restrict__ xp, double *__restrict__ bp) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        const int i = rows[p];
        xp[p] = x[i];
        bp[p] = b[i];
    }
}

__global__ void k_perm_out(int n, const int *__restrict__ rows, const double *__restrict__ xp,
                           double *__restrict__ x) {
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) x[rows[p]] = xp[p];
}

// The update of one row from t = b_i - sum_{j != i} a_ij x_j (CSR order):
//   mode 0 (GS)  x_i = t * (1 / a_ii)          (smoother.c:140-146)
//   mode 1 (SGS) x_i = t / a_ii                (smoother.c:222-227)
//   mode 2 (SOR) x_i = w (t / a_ii) + (1 - w) x_i  (smoother.c:283-289)
// rows with |a_ii| <= SMALLREAL keep their value.
__device__ __forceinline__ double gs_update(int mode, double t, double d, double xo, double w) {
    if (!(fabs(d) > kSmallReal)) return xo;
    if (mode == 0) return t * (1.0 / d);
    if (mode == 1) return t / d;
    return w * (t / d) + (1 - w) * xo;
}

// t - sum_k val[k] x[col[k]] over one lane's slice entries, subtracted in
// entry order; the loads of 8 entries are issued before their products so
// that the gathers overlap (the order of the subtractions is unchanged).
template <class XT>
__device__ __forceinline__ double slice_row_sub(double t, const int *__restrict__ cp, const double *__restrict__ vp,
                                                int width, const XT *x) {
    int k = 0;
    for (; k + 8 <= width; k += 8) {
        int c[8];
        double v[8], xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            c[u] = ld_nt(cp + (long long)(k + u) * kSlice);
            v[u] = ld_nt(vp + (long long)(k + u) * kSlice);
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = x[c[u]];
#pragma unroll
        for (int u = 0; u < 8; ++u) t -= v[u] * xv[u];
    }
    for (; k < width; ++k) t -= ld_nt(vp + (long long)k * kSlice) * x[ld_nt(cp + (long long)k * kSlice)];
    return t;
}

// One colour on the permuted vectors; a wave per slice.
template <int kMode>
__global__ void __launch_bounds__(kBlock) k_gs_color(int s0, int ns, const int *__restrict__ spos,
                                                     const int *__restrict__ scnt, const long long *__restrict__ sptr,
                                                     const int *__restrict__ col, const double *__restrict__ val,
                                                     const double *__restrict__ diag, const double *__restrict__ bp,
                                                     double *__restrict__ xp, double w) {
    const int sl = blockIdx.x * kSlicesPerBlock + threadIdx.x / kSlice;
    const int lane = threadIdx.x % kSlice;
    if (sl >= ns) return;
    const int s = s0 + sl;
    // slice header loads issued together
    const int cnt = scnt[s], p0 = spos[s];
    const long long base = sptr[s], end = sptr[s + 1];
    if (lane >= cnt) return;
    const int width = (int)((end - base) / kSlice);
    const int pos = p0 + lane;
    const double bi = bp[pos], d = diag[pos];
    const double t = slice_row_sub(bi, col + base + lane, val + base + lane, width, xp);
    xp[pos] = gs_update(kMode, t, d, xp[pos], w);
}

// Whole smoothing call of a small level in one workgroup, x held in LDS
// (permuted): the passes in turn, the colours of a pass in turn, a barrier
// after each colour.  Same arithmetic as k_gs_color.
struct GsProgram {
    int npass;
    int forward[kGsMaxPass], mode[kGsMaxPass], skip[kGsMaxPass];
    double w[kGsMaxPass];
};

__global__ void __launch_bounds__(1024) k_gs_lds(int n, int ncolors, const int *__restrict__ sl0,
                                                 const int *__restrict__ spos, const int *__restrict__ scnt,
                                                 const long long *__restrict__ sptr, const int *__restrict__ rows,
                                                 const int *__restrict__ col, const double *__restrict__ val,
                                                 const double *__restrict__ diag, const double *__restrict__ b,
                                                 double *__restrict__ x, GsProgram prog) {
    extern __shared__ double xs[];
    for (int p = threadIdx.x; p < n; p += blockDim.x) xs[p] = x[rows[p]];
    __syncthreads();
    const int wave = threadIdx.x / kSlice, nwaves = blockDim.x / kSlice;
    const int lane = threadIdx.x % kSlice;
    for (int ps = 0; ps < prog.npass; ++ps) {
        const int count = ncolors - prog.skip[ps];
        const int mode = prog.mode[ps];
        const double w = prog.w[ps];
        for (int q = 0; q < count; ++q) {
            const int c = prog.forward[ps] ? q + prog.skip[ps] : ncolors - 1 - prog.skip[ps] - q;
            for (int s = sl0[c] + wave; s < sl0[c + 1]; s += nwaves) {
                const int cnt = scnt[s], p0 = spos[s];
                const long long base = sptr[s], end = sptr[s + 1];
                if (lane >= cnt) continue;
                const int width = (int)((end - base) / kSlice);
                const int pos = p0 + lane;
                const double bi = b[rows[pos]], d = diag[pos];
                const double t = slice_row_sub(bi, col + base + lane, val + base + lane, width, xs);
                xs[pos] = gs_update(mode, t, d, xs[pos], w);
            }
            __syncthreads();
        }
    }
    for (int p = threadIdx.x; p < n; p += blockDim.x) x[rows[p]] = xs[p];
}

}  // namespace

bool is_gs_family(int smoother) {
    return smoother == HZ_SMOOTHER_GS || smoother == HZ_SMOOTHER_SGS || smoother == HZ_SMOOTHER_SOR ||
           smoother == HZ_SMOOTHER_SSOR || smoother == HZ_SMOOTHER_GSOR || smoother == HZ_SMOOTHER_SGSOR;
}

void gs_prepare(const DevCSR &A, GsLevel &G) {
    G = GsLevel();
    const int n = A.n;
    if (n == 0) return;
    hipStream_t st = stream();
    // colouring on the pattern of A + A^T
    DevCSR AT;
    csr_transpose_gpu(A, AT);
    int *c0 = arena().alloc_n<int>((size_t)n + 1);
    int *c1 = arena().alloc_n<int>((size_t)n + 1);
    int *left = arena().alloc_n<int>(1);
    HZ_CHECK(hipMemsetAsync(c0, 0xff, sizeof(int) * n, st));
    // Rounds after the last one change nothing, so the count of uncoloured
    // rows is read back only every kCheck rounds.  Every round colours at
    // least the uncoloured row of highest priority: n rounds always suffice.
    constexpr int kCheck = 8;
    for (int round = 0; round <= n; round += kCheck) {
        HZ_CHECK(hipMemsetAsync(left, 0, sizeof(int), st));
        for (int r = 0; r < kCheck; ++r) {
            k_jp_round<<<grid_rows(n), kBlock, 0, st>>>(n, A.ia, A.ja, AT.ia, AT.ja, c0, c1,
                                                        r == kCheck - 1 ? left : nullptr);
            HZ_LAUNCH_CHECK();
            std::swap(c0, c1);
        }
        int hl = 0;
        HZ_CHECK(memcpy_sync(&hl, left, sizeof(int), hipMemcpyDeviceToHost));
        if (hl == 0) break;
    }
    std::vector<int> color(n), offd(n);
    k_offdiag_count<<<grid_rows(n), kBlock, 0, st>>>(n, A.ia, A.ja, c1);
    HZ_LAUNCH_CHECK();
    HZ_CHECK(memcpy_sync(color.data(), c0, sizeof(int) * n, hipMemcpyDeviceToHost));
    HZ_CHECK(memcpy_sync(offd.data(), c1, sizeof(int) * n, hipMemcpyDeviceToHost));
    int nc = 0;
    for (int v : color) nc = std::max(nc, v + 1);
    G.ncolors = nc;
    G.cptr.assign(nc + 1, 0);
    for (int v : color) G.cptr[v + 1]++;
    for (int c = 0; c < nc; ++c) G.cptr[c + 1] += G.cptr[c];
    std::vector<int> rows(n), fillp(G.cptr.begin(), G.cptr.end() - 1);
    for (int i = 0; i < n; ++i) rows[fillp[color[i]]++] = i;
    // slices of each colour (diagonal entries are stored apart)
    G.sl0.assign(nc + 1, 0);
    std::vector<int> spos, scnt;
    std::vector<long long> sptr(1, 0);
    for (int c = 0; c < nc; ++c) {
        G.sl0[c] = (int)spos.size();
        for (int p = G.cptr[c]; p < G.cptr[c + 1]; p += kSlice) {
            const int cnt = std::min(kSlice, G.cptr[c + 1] - p);
            int width = 0;
            for (int t = 0; t < cnt; ++t) width = std::max(width, offd[rows[p + t]]);
            spos.push_back(p);
            scnt.push_back(cnt);
            sptr.push_back(sptr.back() + (long long)width * kSlice);
        }
    }
    const int ns = (int)spos.size();
    G.sl0[nc] = ns;
    G.nslices = ns;
    G.n = n;
    G.rows = arena().alloc_n<int>((size_t)n + 1);
    G.spos = arena().alloc_n<int>((size_t)ns + 1);
    G.scnt = arena().alloc_n<int>((size_t)ns + 1);
    G.sptr = arena().alloc_n<long long>((size_t)ns + 1);
    G.diag = arena().alloc_n<double>((size_t)n + 1);
    G.col = arena().alloc_n<int>((size_t)sptr.back() + 1);
    G.val = arena().alloc_n<double>((size_t)sptr.back() + 1);
    HZ_CHECK(memcpy_sync(G.rows, rows.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    HZ_CHECK(memcpy_sync(G.spos, spos.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HZ_CHECK(memcpy_sync(G.scnt, scnt.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HZ_CHECK(memcpy_sync(G.sptr, sptr.data(), sizeof(long long) * (ns + 1), hipMemcpyHostToDevice));
    int *pos_of = c1;   // scratch: the colouring buffers are free now
    k_pos_of<<<grid_rows(n), kBlock, 0, st>>>(n, G.rows, pos_of);
    HZ_LAUNCH_CHECK();
    const long threads = (long)ns * kSlice;
    k_gs_fill<<<div_up(threads, kBlock), kBlock, 0, st>>>(ns, G.spos, G.scnt, G.sptr, G.rows, pos_of, A.ia, A.ja,
                                                          A.val, G.col, G.val, G.diag);
    HZ_LAUNCH_CHECK();
    G.dsl0 = arena().alloc_n<int>((size_t)nc + 1);
    HZ_CHECK(memcpy_sync(G.dsl0, G.sl0.data(), sizeof(int) * (nc + 1), hipMemcpyHostToDevice));
    G.in_lds = n <= kGsLdsRows;
    if (!G.in_lds) {
        G.xp = arena().alloc_n<double>((size_t)n + 1);
        G.bp = arena().alloc_n<double>((size_t)n + 1);
    }
    HZ_CHECK(hipStreamSynchronize(st));
}

namespace {

struct Pass {
    bool forward;
    int mode;      // as in gs_update
    double w;
    bool skip;     // skip the first colour visited (the backward half of SGS starts at n-2, smoother.c:236)
};

// dcsr_presmoothing (mgcycle.c:76-137, rows 0..n-1 step 1) and
// dcsr_postsmoothing (mgcycle.c:162-223, step -1) as a list of passes;
// smoother_dcsr_gs / _sor with L sweeps are L passes in one direction.
std::vector<Pass> gs_program(int smoother, bool post, int L, double relax) {
    std::vector<Pass> prog;
    auto call = [&](bool fwd, int mode, double w) {
        for (int s = 0; s < L; ++s) prog.push_back({fwd, mode, w, false});
    };
    constexpr bool F = true, B = false;
    switch (smoother) {
        case HZ_SMOOTHER_GS: call(post ? B : F, 0, 1.0); break;
        case HZ_SMOOTHER_SGS:   // smoother_dcsr_sgs: L times forward, backward
            for (int s = 0; s < L; ++s) {
                prog.push_back({F, 1, 1.0, false});
                prog.push_back({B, 1, 1.0, true});
            }
            break;
        case HZ_SMOOTHER_SOR: call(post ? B : F, 2, relax); break;
        case HZ_SMOOTHER_SSOR:
            call(F, 2, relax);
            call(B, 2, relax);
            break;
        case HZ_SMOOTHER_GSOR:
            if (!post) {
                call(F, 0, 1.0);
                call(B, 2, relax);
            } else {
                call(F, 2, relax);
                call(B, 0, 1.0);
            }
            break;
        case HZ_SMOOTHER_SGSOR:
            if (!post) {
                call(F, 0, 1.0);
                call(B, 0, 1.0);
                call(F, 2, relax);
                call(B, 2, relax);
            } else {
                call(F, 2, relax);
                call(B, 2, relax);
                call(F, 0, 1.0);
                call(B, 0, 1.0);
            }
            break;
        default: break;
    }
    return prog;
}

void color_launch(const GsLevel &G, int c, int mode, double w) {
    const int s0 = G.sl0[c], ns = G.sl0[c + 1] - G.sl0[c];
    if (ns == 0) return;
    const int grid = div_up(ns, kSlicesPerBlock);
    hipStream_t st = stream();
    if (mode == 0)
        k_gs_color<0><<<grid, kBlock, 0, st>>>(s0, ns, G.spos, G.scnt, G.sptr, G.col, G.val, G.diag, G.bp, G.xp, w);
    else if (mode == 1)
        k_gs_color<1><<<grid, kBlock, 0, st>>>(s0, ns, G.spos, G.scnt, G.sptr, G.col, G.val, G.diag, G.bp, G.xp, w);
    else
        k_gs_color<2><<<grid, kBlock, 0, st>>>(s0, ns, G.spos, G.scnt, G.sptr, G.col, G.val, G.diag, G.bp, G.xp, w);
}

}  // namespace

void gs_smooth(const GsLevel &G, int smoother, bool post, int nsweeps, double relax, const double *b, double *x) {
    const std::vector<Pass> prog = gs_program(smoother, post, nsweeps, relax);
    if (prog.empty() || G.n == 0) return;
    hipStream_t st = stream();
    if (G.in_lds) {
        for (size_t p0 = 0; p0 < prog.size(); p0 += kGsMaxPass) {
            GsProgram gp{};
            gp.npass = (int)std::min<size_t>(kGsMaxPass, prog.size() - p0);
            for (int k = 0; k < gp.npass; ++k) {
                const Pass &P = prog[p0 + k];
                gp.forward[k] = P.forward;
                gp.mode[k] = P.mode;
                gp.skip[k] = (P.skip && G.ncolors > 0) ? 1 : 0;
                gp.w[k] = P.w;
            }
            k_gs_lds<<<1, 1024, sizeof(double) * G.n, st>>>(G.n, G.ncolors, G.dsl0, G.spos, G.scnt, G.sptr, G.rows,
                                                            G.col, G.val, G.diag, b, x, gp);
            HZ_LAUNCH_CHECK();
        }
        return;
    }
    k_perm_in<<<grid_rows(G.n), kBlock, 0, st>>>(G.n, G.rows, x, b, G.xp, G.bp);
    for (const Pass &P : prog) {
        for (int q = P.skip ? 1 : 0; q < G.ncolors; ++q)
            color_launch(G, P.forward ? q : G.ncolors - 1 - q, P.mode, P.w);
    }
    k_perm_out<<<grid_rows(G.n), kBlock, 0, st>>>(G.n, G.rows, G.xp, x);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
