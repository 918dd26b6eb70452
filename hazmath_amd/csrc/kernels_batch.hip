// kernels_batch.hip -- many independent systems, kBatchRhs right-hand sides
// each, through one UA-AMG-PCG: linear_solver_dcsr_krylov_amg
// (itsolver.c:846-931) applied to every (system, right-hand side) pair at once.
//
// Layout.  The systems are stacked into one block-diagonal "super" matrix per
// level: level l holds, one after the other, the level-l matrices of the
// systems whose hierarchies reach level l.  Vectors interleave the right-hand
// sides (entry (i, q) at i * kBatchRhs + q), so one pass over a matrix row
// serves every right-hand side.  Each smoothing sweep, residual, restriction,
// prolongation and coarse solve of a level is one launch covering every
// system at that level.
//
// Setup.  construct_strongly_coupled, aggregation_hem and dcsr_rap_agg act row
// by row or through neighbours only, so on the block-diagonal matrix they
// produce exactly each system's own result (aggregates numbered block after
// block).  What is per system -- the adaptive strength threshold
// (amg_setup_ua.c:1325-1328), the coarsening stop rules (rows <= coarse_dof,
// max_levels, no aggregates) -- is kept per block: a system that stops at
// level l is solved there by its dense inverse, and its rows at level l are
// left alone by the smoothing and transfer kernels ("dead" rows).
//
// Cycle.  V-cycle in FP32 (values and vectors), the smoothers of the FP32
// single-system cycle (Jacobi with weight w / L1DIAG, same arithmetic per
// row), FP64 dense coarsest inverse with FP64 accumulation.
//
// Krylov.  dcsr_pcg_w_cond_est (krylov.c:400-718) per column in FP64: the
// initial exit test, alpha/beta recurrences, relative-residual stop with the
// false-convergence check III (true residual recomputed, restart with p = 0,
// at most MAX_RESTART = 20 restarts).  Checks I and II (zero solution,
// stagnation restart) and the condition number estimate are not restated.
// Converged columns freeze.  One PCG iteration is captured once into a
// hipGraph and replayed; the host polls the count of unfinished columns.
//
// Autotuning.  The 8-right-hand-side kernels come in two shapes: a lane per
// row with all right-hand sides in registers, or two lanes per row with half
// of them each (twice the waves, for latency-bound levels).  The tuner times
// both shapes on each level's own matrix at setup and keeps the faster.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <type_traits>
#include <vector>

#include "amg.hpp"
#include "tune_cache.h"

namespace hz {

namespace {

constexpr int kR = kBatchRhs;          // right-hand sides per system (interleaved)
static_assert(kR == HZ_BATCH_RHS, "C ABI batch width");
static_assert(kR == 8, "the kernels split the right-hand sides into float4 / double2 groups");
constexpr int kBlock = 256;
constexpr int kSlicesPerBlock = kBlock / kWave;
constexpr int kChunkRows = 2048;       // level-0 rows per reduction chunk (never across systems)
constexpr float kSmallRealF = 1e-20f;
constexpr int kMaxRestart = 20;        // MAX_RESTART, include/macro.h:33
constexpr int kCoarseMax = 2048;       // dense coarsest limit (the block's b is staged in 64 KiB of LDS)

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(g, 8192));
}

// ---------------------------------------------------------------------------
// setup kernels
// ---------------------------------------------------------------------------
// blk[i] = the block holding row i (off: nblk + 1 row offsets)
__global__ void k_row_block(int n, int nblk, const int *__restrict__ off, int *__restrict__ blk) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int lo = 0, hi = nblk;   // off[lo] <= i < off[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (off[mid] <= i) lo = mid;
            else hi = mid;
        }
        blk[i] = lo;
    }
}

__global__ void k_gather_i(int n, const int *__restrict__ blk, const int *__restrict__ tab, int *__restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = tab[blk[i]];
}

__global__ void k_fill_i(int n, int v, int *__restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_gather_d(int n, const int *__restrict__ blk, const double *__restrict__ tab,
                           double *__restrict__ out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = tab[blk[i]];
}

// per block: smallest and largest aggregate number of its rows (-1: none)
__global__ void k_block_aggrange(int nblk, const int *__restrict__ off, const int *__restrict__ agg,
                                 int *__restrict__ amin, int *__restrict__ amax) {
    __shared__ int smin[kBlock / kWave], smax[kBlock / kWave];
    const int b = blockIdx.x;
    if (b >= nblk) return;
    int lo = INT_MAX, hi = -1;
    for (int i = off[b] + threadIdx.x; i < off[b + 1]; i += blockDim.x) {
        const int g = agg[i];
        if (g >= 0) {
            lo = min(lo, g);
            hi = max(hi, g);
        }
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) {
        lo = min(lo, __shfl_xor(lo, o, kWave));
        hi = max(hi, __shfl_xor(hi, o, kWave));
    }
    if (threadIdx.x % kWave == 0) {
        smin[threadIdx.x / kWave] = lo;
        smax[threadIdx.x / kWave] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x / kWave); ++w) {
            lo = min(lo, smin[w]);
            hi = max(hi, smax[w]);
        }
        amin[b] = hi >= 0 ? lo : -1;
        amax[b] = hi;
    }
}

// agg[i] += shift[blk[i]] for rows of continuing blocks, -1 for the others
__global__ void k_agg_renumber(int n, const int *__restrict__ blk, const int *__restrict__ keep,
                               const int *__restrict__ shift, int *__restrict__ agg,
                               unsigned char *__restrict__ live) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = blk[i];
        const int g = agg[i];
        agg[i] = (keep[b] && g >= 0) ? g + shift[b] : -1;
        live[i] = (unsigned char)(keep[b] ? 1 : 0);
    }
}

// dense row-major copy of the diagonal block rows/cols [r0, r0+nb)
__global__ void k_dense_block(const int *__restrict__ ia, const int *__restrict__ ja, const double *__restrict__ a,
                              int r0, int nb, double *__restrict__ M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) {
        double *row = M + (size_t)i * nb;
        for (int j = 0; j < nb; ++j) row[j] = 0.0;
        for (int k = ia[r0 + i]; k < ia[r0 + i + 1]; ++k) {
            const int j = ja[k] - r0;
            if (j >= 0 && j < nb) row[j] += a[k];
        }
    }
}

__global__ void k_to_f32(long long n, const double *__restrict__ v, float *__restrict__ o) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        o[i] = (float)v[i];
}

// ---------------------------------------------------------------------------
// FP32 cycle kernels, kR right-hand sides interleaved.  kH lanes per row:
// lane half h handles right-hand sides [h * kR / kH, (h + 1) * kR / kH).
// ---------------------------------------------------------------------------
template <int kN>
__device__ __forceinline__ void ld_vec(const float *p, float *v) {
    if constexpr (kN == 8) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        const float4 b = *reinterpret_cast<const float4 *>(p + 4);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else if constexpr (kN == 4) {
        const float4 a = *reinterpret_cast<const float4 *>(p);
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    } else {
        static_assert(kN == 2, "2, 4 or 8 right-hand sides per lane");
        const float2 a = *reinterpret_cast<const float2 *>(p);
        v[0] = a.x; v[1] = a.y;
    }
}

template <int kN>
__device__ __forceinline__ void st_vec(float *p, const float *v) {
    if constexpr (kN == 8) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4 *>(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
    } else if constexpr (kN == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        *reinterpret_cast<float2 *>(p) = make_float2(v[0], v[1]);
    }
}

// rows of systems whose every right-hand side has finished are skipped:
// act[system] = 0 once all its columns are done; rsys: row -> system
__device__ __forceinline__ bool skip_row(const int *rsys, const int *act, long r) {
    return act != nullptr && act[rsys[r]] == 0;
}

// which slice / row / right-hand-side range this thread covers
template <int kH>
struct LaneMap {
    int s, r, q0;
    __device__ __forceinline__ LaneMap() {
        // XCD-aware: each XCD gets a contiguous range of slices, so the
        // neighbouring rows a slice gathers stay in that XCD's L2
        const int wave = xcd_block_id() * (blockDim.x / kWave) + threadIdx.x / kWave;
        s = wave / kH;
        r = s * kSlice + threadIdx.x % kWave;
        q0 = (wave % kH) * (kR / kH);
    }
};

// One sweep xout = sweep(x) for live rows (Jacobi weight w / L1DIAG as in
// the FP32 single-system cycle); dead rows copy x.
template <int kSm, int kH, int kU>
__global__ __launch_bounds__(kBlock) void k_bsweep(int n, int nslices, const long long *__restrict__ sptr,
                                                   const int *__restrict__ col, const float *__restrict__ val,
                                                   float wgt, const unsigned char *__restrict__ live,
                                                   const float *__restrict__ b, const float *__restrict__ x,
                                                   float *__restrict__ xout, const int *__restrict__ rsys,
                                                   const int *__restrict__ act) {
    constexpr int kN = kR / kH;
    const LaneMap<kH> m;
    if (m.s >= nslices || m.r >= n) return;
    const int r = m.r;
    if (skip_row(rsys, act, r)) return;
    float xi[kN];
    ld_vec<kN>(x + (size_t)r * kR + m.q0, xi);
    if (!live[r]) {
        st_vec<kN>(xout + (size_t)r * kR + m.q0, xi);
        return;
    }
    float t[kN];
    ld_vec<kN>(b + (size_t)r * kR + m.q0, t);
    float d = 0.0f;
    const long long base = sptr[m.s];
    const int wd = (int)((sptr[m.s + 1] - base) / kSlice);
    const int lane = threadIdx.x % kWave;
    const int *c = col + base + lane;
    const float *v = val + base + lane;
    // entries in row order; the loads of kU entries are issued before their use
    auto step = [&](int cj, float a, const float *xj) {
        if (cj < 0) return;
        if (kSm == HZ_SMOOTHER_L1DIAG) {
#pragma unroll
            for (int q = 0; q < kN; ++q) t[q] = t[q] - a * xj[q];
            d = d + fabsf(a);
        } else if (cj == r) {
            d = a;
        } else {
#pragma unroll
            for (int q = 0; q < kN; ++q) t[q] = t[q] - a * xj[q];
        }
    };
    // rounds of kU entries, masked past the slice width (no one-entry tail)
    for (int k = 0; k < wd; k += kU) {
        int cc[kU];
        float aa[kU], xx[kU][kN];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool ok = k + u < wd;
            cc[u] = ok ? ld_nt(c + (long long)(k + u) * kSlice) : -1;
            aa[u] = ok ? ld_nt(v + (long long)(k + u) * kSlice) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (cc[u] >= 0) ld_vec<kN>(x + (size_t)cc[u] * kR + m.q0, xx[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) step(cc[u], aa[u], xx[u]);
    }
    float xn[kN];
#pragma unroll
    for (int q = 0; q < kN; ++q) {
        xn[q] = xi[q];
        if (fabsf(d) > kSmallRealF) {
            if (kSm == HZ_SMOOTHER_L1DIAG) xn[q] = xi[q] + t[q] / d;
            else xn[q] = (1.0f - wgt) * xi[q] + wgt * t[q] / d;
        }
    }
    st_vec<kN>(xout + (size_t)r * kR + m.q0, xn);
}

// w = b - A x for live rows
template <int kH, int kU>
__global__ __launch_bounds__(kBlock) void k_bresidual(int n, int nslices, const long long *__restrict__ sptr,
                                                      const int *__restrict__ col, const float *__restrict__ val,
                                                      const unsigned char *__restrict__ live,
                                                      const float *__restrict__ b, const float *__restrict__ x,
                                                      float *__restrict__ w, const int *__restrict__ rsys,
                                                      const int *__restrict__ act) {
    constexpr int kN = kR / kH;
    const LaneMap<kH> m;
    if (m.s >= nslices || m.r >= n) return;
    const int r = m.r;
    if (!live[r] || skip_row(rsys, act, r)) return;
    float acc[kN];
#pragma unroll
    for (int q = 0; q < kN; ++q) acc[q] = 0.0f;
    const long long base = sptr[m.s];
    const int wd = (int)((sptr[m.s + 1] - base) / kSlice);
    const int lane = threadIdx.x % kWave;
    const int *c = col + base + lane;
    const float *v = val + base + lane;
    // rounds of kU entries, masked past the slice width (no one-entry tail)
    for (int k = 0; k < wd; k += kU) {
        int cc[kU];
        float aa[kU], xx[kU][kN];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool ok = k + u < wd;
            cc[u] = ok ? ld_nt(c + (long long)(k + u) * kSlice) : -1;
            aa[u] = ok ? ld_nt(v + (long long)(k + u) * kSlice) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (cc[u] >= 0) ld_vec<kN>(x + (size_t)cc[u] * kR + m.q0, xx[u]);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (cc[u] < 0) continue;
#pragma unroll
            for (int q = 0; q < kN; ++q) acc[q] = acc[q] + aa[u] * xx[u][q];
        }
    }
    float bi[kN];
    ld_vec<kN>(b + (size_t)r * kR + m.q0, bi);
#pragma unroll
    for (int q = 0; q < kN; ++q) acc[q] = bi[q] - acc[q];
    st_vec<kN>(w + (size_t)r * kR + m.q0, acc);
}

// bc = R w (member sums in ascending member order), one thread per coarse row
__global__ void k_brestrict(int nc, const int *__restrict__ rptr, const int *__restrict__ rmem,
                            const float *__restrict__ w, float *__restrict__ bc, const int *__restrict__ crsys,
                            const int *__restrict__ act) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        if (skip_row(crsys, act, c)) continue;
        float s[kR];
#pragma unroll
        for (int q = 0; q < kR; ++q) s[q] = 0.0f;
        for (int p = rptr[c]; p < rptr[c + 1]; ++p) {
            float v[kR];
            ld_vec<kR>(w + (size_t)rmem[p] * kR, v);
#pragma unroll
            for (int q = 0; q < kR; ++q) s[q] = s[q] + v[q];
        }
        st_vec<kR>(bc + (size_t)c * kR, s);
    }
}

// x = first sweep from zero for live rows (Jacobi: w b / d, L1DIAG: b / d with d precomputed)
__global__ void k_bzero_sweep(int n, int smoother, float wgt, const float *__restrict__ d,
                              const unsigned char *__restrict__ live, const float *__restrict__ b,
                              float *__restrict__ x, const int *__restrict__ rsys, const int *__restrict__ act) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!live[i] || skip_row(rsys, act, i)) continue;
        const float di = d[i];
        float bi[kR], v[kR];
        ld_vec<kR>(b + (size_t)i * kR, bi);
#pragma unroll
        for (int q = 0; q < kR; ++q) {
            v[q] = 0.0f;
            if (fabsf(di) > kSmallRealF) v[q] = smoother == HZ_SMOOTHER_L1DIAG ? bi[q] / di : wgt * bi[q] / di;
        }
        st_vec<kR>(x + (size_t)i * kR, v);
    }
}

// x += P xc
__global__ void k_bprolong(int n, const int *__restrict__ agg, const float *__restrict__ xc,
                           float *__restrict__ x, const int *__restrict__ rsys, const int *__restrict__ act) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int g = agg[i];
        if (g <= HZ_UNPT || skip_row(rsys, act, i)) continue;
        float xi[kR], c[kR];
        ld_vec<kR>(x + (size_t)i * kR, xi);
        ld_vec<kR>(xc + (size_t)g * kR, c);
#pragma unroll
        for (int q = 0; q < kR; ++q) xi[q] = xi[q] + c[q];
        st_vec<kR>(x + (size_t)i * kR, xi);
    }
}

// Systems whose hierarchy ends at this level: x = Ainv b per block (FP64
// inverse and accumulation, sum over j ascending).  One workgroup per block,
// a thread per row with all right-hand sides; the block's b staged in LDS;
// the inverse stored transposed (entry (i, j) at io + j nb + i) so a
// workgroup's loads of one j are contiguous.  tab[4 k .. 4 k + 3] = row
// offset, size, inverse offset (in doubles), system.
__global__ __launch_bounds__(256) void k_bcoarse(const long long *__restrict__ tab, const double *__restrict__ AinvT,
                                                 const float *__restrict__ b, float *__restrict__ x,
                                                 const int *__restrict__ act) {
    extern __shared__ float bs[];
    const long long r0 = tab[4 * blockIdx.x], nb = tab[4 * blockIdx.x + 1], io = tab[4 * blockIdx.x + 2];
    if (act && act[tab[4 * blockIdx.x + 3]] == 0) return;
    for (long long e = threadIdx.x; e < nb * kR; e += blockDim.x) bs[e] = b[r0 * kR + e];
    __syncthreads();
    for (long long i = threadIdx.x; i < nb; i += blockDim.x) {
        double acc[kR];
#pragma unroll
        for (int q = 0; q < kR; ++q) acc[q] = 0.0;
        const double *Mi = AinvT + io + i;
        for (long long j = 0; j < nb; ++j) {
            const double m = Mi[j * nb];
            float bj[kR];
            ld_vec<kR>(bs + j * kR, bj);
#pragma unroll
            for (int q = 0; q < kR; ++q) acc[q] += m * (double)bj[q];
        }
        float o[kR];
#pragma unroll
        for (int q = 0; q < kR; ++q) o[q] = (float)acc[q];
        st_vec<kR>(x + (r0 + i) * kR, o);
    }
}

// Explicit inverses of many small dense matrices at once: workgroup k inverts
// block k (size tab[4k+1], stored row-major at W + woff[k]) in place by
// Gauss-Jordan elimination without pivoting (the coarsest Galerkin matrices are
// SPD), the elimination of the single-system coarsest inverse applied to the
// whole block.  Writes the transpose of the inverse to T + tab[4k+2].  A zero
// pivot sets status[k] = 1 and leaves the block.
__global__ __launch_bounds__(1024) void k_batch_gj_inverse(const long long *__restrict__ tab,
                                                           const long long *__restrict__ woff,
                                                           double *__restrict__ W, double *__restrict__ T,
                                                           int *__restrict__ status) {
    __shared__ double colk[kCoarseMax];
    __shared__ double rowk[kCoarseMax];
    const int blk = blockIdx.x;
    const int nb = (int)tab[4 * blk + 1];
    double *M = W + woff[blk];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 0; k < nb; ++k) {
        const double piv = M[(size_t)k * nb + k];
        if (piv == 0.0) {
            if (tid == 0) status[blk] = 1;
            return;   // uniform: every thread read the same pivot
        }
        for (int i = tid; i < nb; i += nt) colk[i] = M[(size_t)i * nb + k];
        __syncthreads();
        for (int j = tid; j < nb; j += nt) {
            const double v = (j == k) ? 1.0 / piv : M[(size_t)k * nb + j] / piv;
            rowk[j] = v;
            M[(size_t)k * nb + j] = v;
        }
        __syncthreads();
        const long total = (long)nb * nb;
        for (long e = tid; e < total; e += nt) {
            const int i = (int)(e / nb), j = (int)(e % nb);
            if (i == k) continue;
            if (j != k) M[e] -= colk[i] * rowk[j];
            else M[e] = -colk[i] * rowk[k];
        }
        __syncthreads();
    }
    double *Tb = T + tab[4 * blk + 2];
    const long total = (long)nb * nb;
    for (long e = tid; e < total; e += nt) {
        const long i = e / nb, j = e % nb;
        Tb[j * nb + i] = M[e];
    }
}

__global__ void k_enter_b(int n, const double *__restrict__ r, float *__restrict__ b,
                          const int *__restrict__ rsys, const int *__restrict__ act) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (skip_row(rsys, act, i)) continue;
        const double2 *rp = reinterpret_cast<const double2 *>(r + (size_t)i * kR);
        float v[kR];
#pragma unroll
        for (int h = 0; h < kR / 2; ++h) {
            const double2 t = rp[h];
            v[2 * h] = (float)t.x;
            v[2 * h + 1] = (float)t.y;
        }
        st_vec<kR>(b + (size_t)i * kR, v);
    }
}

__global__ void k_exit_b(int n, const float *__restrict__ x, double *__restrict__ z,
                         const int *__restrict__ rsys, const int *__restrict__ act) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (skip_row(rsys, act, i)) continue;
        float v[kR];
        ld_vec<kR>(x + (size_t)i * kR, v);
        double2 *zp = reinterpret_cast<double2 *>(z + (size_t)i * kR);
#pragma unroll
        for (int h = 0; h < kR / 2; ++h) zp[h] = make_double2((double)v[2 * h], (double)v[2 * h + 1]);
    }
}

// ---------------------------------------------------------------------------
// FP64 PCG kernels on level 0 (kR right-hand sides interleaved)
// ---------------------------------------------------------------------------
// y = A x (kSub = 1: y = A x; kSub = 2: y = b - A x); two lanes per row
// (four right-hand sides each)
template <int kSub, int kH, int kU>
__global__ __launch_bounds__(kBlock) void k_spmm64(int n, int nslices, const long long *__restrict__ sptr,
                                                   const int *__restrict__ col, const double *__restrict__ val,
                                                   const double *__restrict__ x, const double *__restrict__ b,
                                                   double *__restrict__ y, const int *__restrict__ gate,
                                                   const int *__restrict__ rsys, const int *__restrict__ act) {
    if (gate && *gate == 0) return;
    constexpr int kN = kR / kH;      // right-hand sides per lane
    constexpr int kV = kN / 2;       // double2 loads per gathered row
    const LaneMap<kH> m;
    if (m.s >= nslices || m.r >= n) return;
    if (skip_row(rsys, act, m.r)) return;
    double acc[kN];
#pragma unroll
    for (int q = 0; q < kN; ++q) acc[q] = 0.0;
    const long long base = sptr[m.s];
    const int wd = (int)((sptr[m.s + 1] - base) / kSlice);
    const int lane = threadIdx.x % kWave;
    const int *c = col + base + lane;
    const double *v = val + base + lane;
    auto madd = [&](double a, const double2 *xv) {
#pragma unroll
        for (int h = 0; h < kV; ++h) {
            acc[2 * h] += a * xv[h].x;
            acc[2 * h + 1] += a * xv[h].y;
        }
    };
    // rounds of kU entries, masked past the slice width (no one-entry tail)
    for (int k = 0; k < wd; k += kU) {
        int cc[kU];
        double aa[kU];
        double2 xx[kU][kV];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool ok = k + u < wd;
            cc[u] = ok ? ld_nt(c + (long long)(k + u) * kSlice) : -1;
            aa[u] = ok ? ld_nt(v + (long long)(k + u) * kSlice) : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            if (cc[u] >= 0) {
                const double2 *xp = reinterpret_cast<const double2 *>(x + (size_t)cc[u] * kR + m.q0);
#pragma unroll
                for (int h = 0; h < kV; ++h) xx[u][h] = xp[h];
            }
        }
#pragma unroll
        for (int u = 0; u < kU; ++u)
            if (cc[u] >= 0) madd(aa[u], xx[u]);
    }
    const size_t o = (size_t)m.r * kR + m.q0;
#pragma unroll
    for (int q = 0; q < kN; ++q) y[o + q] = kSub == 2 ? b[o + q] - acc[q] : acc[q];
}

// chunk c: rows [crow[c], crow[c+1]) of one system; partial sums per right-hand side
struct ChunkDot {
    __device__ static void store(double *part, int c, double v, double *sh) {
        // threads t and t + 8k share right-hand side t % 8: reduce per rhs
        sh[threadIdx.x] = v;
        __syncthreads();
        if (threadIdx.x < kR) {
            double s = 0.0;
            for (int t = threadIdx.x; t < (int)blockDim.x; t += kR) s += sh[t];
            part[(size_t)c * kR + threadIdx.x] = s;
        }
    }
};

// partial (u, v) per chunk and rhs
__global__ __launch_bounds__(kBlock) void k_cdot(const int *__restrict__ crow, const double *__restrict__ u,
                                                 const double *__restrict__ v, double *__restrict__ part,
                                                 const int *__restrict__ csys, const int *__restrict__ act) {
    __shared__ double sh[kBlock];
    const int c = blockIdx.x;
    if (act && act[csys[c]] == 0) return;   // whole workgroup
    const long e0 = (long)crow[c] * kR, e1 = (long)crow[c + 1] * kR;
    double s = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += kBlock) s += u[e] * v[e];
    ChunkDot::store(part, c, s, sh);
}

// Per-column PCG state (column = system * kR + rhs).
struct ColState {
    double normr0, temp1, alpha, beta, relres;
    int done, cand, restart, more, iters, pad;
};

// u += alpha p, r -= alpha t for unfinished columns; partial ||r||^2
__global__ __launch_bounds__(kBlock) void k_cupdate(const int *__restrict__ crow, const int *__restrict__ csys,
                                                    const ColState *__restrict__ st, const double *__restrict__ p,
                                                    const double *__restrict__ t, double *__restrict__ u,
                                                    double *__restrict__ r, double *__restrict__ part,
                                                    const int *__restrict__ act) {
    __shared__ double sh[kBlock];
    const int c = blockIdx.x;
    if (act && act[csys[c]] == 0) return;
    const int q = threadIdx.x % kR;
    const ColState &cs = st[csys[c] * kR + q];
    const bool active = !cs.done;
    const double al = cs.alpha;
    const long e0 = (long)crow[c] * kR, e1 = (long)crow[c + 1] * kR;
    double s = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += kBlock) {
        double re = r[e];
        if (active) {
            u[e] += al * p[e];
            re -= al * t[e];
            r[e] = re;
        }
        s += re * re;
    }
    ChunkDot::store(part, c, s, sh);
}

// r = t (= b - A u) for candidate columns; partial ||r||^2 of those columns
__global__ __launch_bounds__(kBlock) void k_creplace(const int *__restrict__ crow, const int *__restrict__ csys,
                                                     const ColState *__restrict__ st, const double *__restrict__ t,
                                                     double *__restrict__ r, double *__restrict__ part,
                                                     const int *__restrict__ gate, const int *__restrict__ act) {
    if (*gate == 0) return;
    __shared__ double sh[kBlock];
    const int c = blockIdx.x;
    if (act && act[csys[c]] == 0) return;
    const int q = threadIdx.x % kR;
    const bool cand = st[csys[c] * kR + q].cand != 0;
    const long e0 = (long)crow[c] * kR, e1 = (long)crow[c + 1] * kR;
    double s = 0.0;
    for (long e = e0 + threadIdx.x; e < e1; e += kBlock) {
        if (cand) {
            const double re = t[e];
            r[e] = re;
            s += re * re;
        }
    }
    ChunkDot::store(part, c, s, sh);
}

// p = z + beta p for unfinished columns
__global__ __launch_bounds__(kBlock) void k_cdirection(const int *__restrict__ crow, const int *__restrict__ csys,
                                                       const ColState *__restrict__ st, const double *__restrict__ z,
                                                       double *__restrict__ p, const int *__restrict__ act) {
    const int c = blockIdx.x;
    if (act && act[csys[c]] == 0) return;
    const int q = threadIdx.x % kR;
    const ColState &cs = st[csys[c] * kR + q];
    if (cs.done) return;
    const double be = cs.beta;
    const long e0 = (long)crow[c] * kR, e1 = (long)crow[c + 1] * kR;
    for (long e = e0 + threadIdx.x; e < e1; e += kBlock) p[e] = z[e] + be * p[e];
}

// column sums of the chunk partials of its system, in chunk order
__device__ __forceinline__ double col_sum(const double *part, const int *schunk, int sys, int q) {
    double s = 0.0;
    for (int c = schunk[sys]; c < schunk[sys + 1]; ++c) s += part[(size_t)c * kR + q];
    return s;
}

// Finalize kernels: one thread per column, single workgroup of 1024 threads.
// kind 0: initial residual (absres0, normr0, exit test);  1: alpha = temp1 / (t, p);
// 2: relres after the update, candidates for check III;  3: check III;  4: beta.
struct PcgCtl {
    int it;          // iterations done
    int maxit;
    int anycand;     // some column reached relres < tol this iteration
    int active;      // unfinished columns
    double tol;
};

template <int kKind>
__global__ __launch_bounds__(1024) void k_cfinal(int ncols, const double *__restrict__ part,
                                                 const int *__restrict__ schunk, ColState *__restrict__ st,
                                                 PcgCtl *__restrict__ ctl, int *__restrict__ sysact) {
    __shared__ int sh[1024 / kWave];
    int flag = 0;
    const double tol = ctl->tol;
    const int it = ctl->it + (kKind == 2 ? 1 : 0);
    // uniform trip count: the 8 columns of a system sit in 8 neighbouring lanes
    for (int base = 0; base < ncols; base += blockDim.x) {
        const int col = base + threadIdx.x;
        const bool valid = col < ncols;
        int open = 0;   // this column still iterates
        if (valid) {
            ColState &cs = st[col];
            const int sys = col / kR, q = col % kR;
            if (kKind == 0) {
                const double absres0 = sqrt(col_sum(part, schunk, sys, q));
                cs.normr0 = fmax(kSmallReal, absres0);
                cs.relres = absres0 / cs.normr0;
                cs.done = (cs.relres < tol || absres0 < 1e-3 * tol) ? 1 : 0;
                cs.iters = 0;
                cs.cand = cs.restart = cs.more = 0;
                cs.alpha = cs.beta = 0.0;
                flag += !cs.done;
            } else if (kKind == 1) {
                if (!cs.done) cs.alpha = cs.temp1 / col_sum(part, schunk, sys, q);
            } else if (kKind == 2) {
                cs.cand = 0;
                if (!cs.done) {
                    cs.relres = sqrt(col_sum(part, schunk, sys, q)) / cs.normr0;
                    if (cs.relres < tol) {
                        cs.cand = 1;
                        flag = 1;
                    }
                }
            } else if (kKind == 3) {
                if (cs.cand) {
                    cs.relres = sqrt(col_sum(part, schunk, sys, q)) / cs.normr0;
                    if (cs.relres < tol) {
                        cs.done = 1;
                        cs.iters = it;
                    } else if (cs.more >= kMaxRestart) {
                        cs.done = 1;
                        cs.iters = HZ_ERROR_SOLVER_TOLSMALL;
                    } else {
                        cs.restart = 1;
                        ++cs.more;
                    }
                    cs.cand = 0;
                }
                if (!cs.done && it >= ctl->maxit) {
                    cs.done = 1;
                    cs.iters = HZ_ERROR_SOLVER_MAXIT;
                }
                flag += !cs.done;
            } else {   // kKind == 4: beta and the new (z, r); a restart drops p
                if (!cs.done) {
                    const double zr = col_sum(part, schunk, sys, q);
                    cs.beta = cs.restart ? 0.0 : zr / cs.temp1;
                    cs.temp1 = zr;
                    cs.restart = 0;
                }
            }
            open = kKind == 2 ? cs.cand : !cs.done;
        }
        if (kKind == 0 || kKind == 2 || kKind == 3) {
            // per system: active while any of its kR columns iterates (kinds 0, 3),
            // or holding a check-III candidate (kind 2)
#pragma unroll
            for (int o = 1; o < kR; o <<= 1) open |= __shfl_xor(open, o, kWave);
            if (valid && col % kR == 0) sysact[col / kR] = open;
        }
    }
    if (kKind == 0 || kKind == 2 || kKind == 3) {
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) flag += __shfl_xor(flag, o, kWave);
        if (threadIdx.x % kWave == 0) sh[threadIdx.x / kWave] = flag;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < (int)(blockDim.x / kWave); ++w) s += sh[w];
            if (kKind == 2) {
                ctl->anycand = s > 0 ? 1 : 0;
                ctl->it = it;
            } else {
                ctl->active = s;
            }
        }
    }
}

// initial (z, r) -> temp1, p = z handled by k_cdirection with beta = 0
__global__ __launch_bounds__(1024) void k_cfinal_temp1(int ncols, const double *__restrict__ part,
                                                       const int *__restrict__ schunk, ColState *__restrict__ st) {
    for (int col = threadIdx.x; col < ncols; col += blockDim.x) {
        ColState &cs = st[col];
        cs.temp1 = col_sum(part, schunk, col / kR, col % kR);
        cs.beta = 0.0;
    }
}

__global__ void k_fill_d(long n, double v, double *__restrict__ x) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] = v;
}

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct BatchLevel {
    DevCSR A;
    DevSELL S;
    float *valf = nullptr;           // FP32 SELL values
    float *dsm = nullptr;            // smoother diagonal, FP32
    int *blk = nullptr;              // row -> block of this level
    int *rsys = nullptr;             // row -> system
    std::vector<int> sys, off;       // blocks: system id, row offsets (nblk + 1)
    unsigned char *live = nullptr;   // rows of systems that continue below
    int *agg = nullptr, nagg = 0, *rptr = nullptr, *rmem = nullptr;
    int ncoarse = 0;                 // systems solved at this level
    int cmax = 0;                    // their largest size
    long long *ctab = nullptr;       // (row offset, size, inverse offset, system) per such system
    double *Ainv = nullptr;
    float *b = nullptr, *x = nullptr, *xt = nullptr, *w = nullptr;
    int shape = 104;                 // FP32 cycle kernels' shape chosen by the tuner (shape code)
    int spmm_shape = 204;            // level 0: shape of the FP64 PCG SpMM (tuned)
};

struct BatchHierarchy {
    std::vector<BatchLevel> lv;
    int nl = 0, nsys = 0;
    int smoother = HZ_SMOOTHER_L1DIAG, pre = 1, post = 1;
    float wgt = 1.0f;
    std::vector<int> levels_of;      // per system: number of levels
    int *act = nullptr;              // per system: 1 while any right-hand side iterates (null: all)
};

namespace {

// Kernel shapes: code = 100 x (lanes per row, each holding 8 / lanes of the
// right-hand sides) + (entries whose loads are issued together per round).
constexpr int kShapes[] = {104, 108, 204, 208, 404, 408};

template <int V>
using IC = std::integral_constant<int, V>;

// f(IC<lanes>, IC<round>) for a shape code
template <class F>
void with_shape(int code, F &&f) {
    switch (code) {
        case 104: f(IC<1>{}, IC<4>{}); break;
        case 108: f(IC<1>{}, IC<8>{}); break;
        case 204: f(IC<2>{}, IC<4>{}); break;
        case 208: f(IC<2>{}, IC<8>{}); break;
        case 404: f(IC<4>{}, IC<4>{}); break;
        case 408: f(IC<4>{}, IC<8>{}); break;
        default: throw DeviceError{hipErrorInvalidValue, "batched AMG: unknown kernel shape"};
    }
}

bool valid_shape(int code) {
    for (int c : kShapes)
        if (c == code) return true;
    return false;
}

void bsweep(const BatchLevel &L, int sm, float w, const float *b, const float *x, float *xo,
            const int *act = nullptr) {
    if (L.S.n == 0) return;
    with_shape(L.shape, [&](auto H, auto U) {
        constexpr int kH = decltype(H)::value, kU = decltype(U)::value;
        const int grid = div_up(L.S.nslices * kH, kSlicesPerBlock);
        if (sm == HZ_SMOOTHER_L1DIAG)
            k_bsweep<HZ_SMOOTHER_L1DIAG, kH, kU><<<grid, kBlock, 0, stream()>>>(
                L.S.n, L.S.nslices, L.S.sptr, L.S.col, L.valf, w, L.live, b, x, xo, L.rsys, act);
        else
            k_bsweep<HZ_SMOOTHER_JACOBI, kH, kU><<<grid, kBlock, 0, stream()>>>(
                L.S.n, L.S.nslices, L.S.sptr, L.S.col, L.valf, w, L.live, b, x, xo, L.rsys, act);
    });
    HZ_LAUNCH_CHECK();
}

void bresidual(const BatchLevel &L, const int *act) {
    if (L.S.n == 0) return;
    with_shape(L.shape, [&](auto H, auto U) {
        constexpr int kH = decltype(H)::value, kU = decltype(U)::value;
        const int grid = div_up(L.S.nslices * kH, kSlicesPerBlock);
        k_bresidual<kH, kU><<<grid, kBlock, 0, stream()>>>(L.S.n, L.S.nslices, L.S.sptr, L.S.col, L.valf, L.live,
                                                          L.b, L.x, L.w, L.rsys, act);
    });
    HZ_LAUNCH_CHECK();
}

// HIP-event time of `reps` calls of fn
template <class F>
float time_ms(F fn, int reps) {
    hipEvent_t e0, e1;
    HZ_CHECK(hipEventCreate(&e0));
    HZ_CHECK(hipEventCreate(&e1));
    fn();
    HZ_CHECK(hipEventRecord(e0, stream()));
    for (int i = 0; i < reps; ++i) fn();
    HZ_CHECK(hipEventRecord(e1, stream()));
    HZ_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HZ_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

// The tuner's bookkeeping of the current batched call (reported in hz_batch_stats).
struct TuneStats {
    int timed = 0, cached = 0;
    double ms = 0.0;
};
TuneStats g_tune;

// Cache key of a decision: device, kernel, and the matrix's class (rounded
// log2 of the rows, 4 x the mean entries per row).
std::string tune_key(const char *kernel, const BatchLevel &L) {
    static std::map<int, std::string> names;   // per device ordinal
    int d = 0;
    HZ_CHECK(hipGetDevice(&d));
    std::string &dev = names[d];
    if (dev.empty()) {
        hipDeviceProp_t prop{};
        HZ_CHECK(hipGetDeviceProperties(&prop, d));
        std::string arch = prop.gcnArchName;
        arch = arch.substr(0, arch.find(':'));
        dev = arch + "-" + std::to_string(prop.multiProcessorCount) + "cu";
    }
    const int nclass = (int)std::lround(std::log2(std::max(1, L.S.n)));
    const int wclass = L.S.n ? (int)std::lround(4.0 * (double)L.A.nnz / L.S.n) : 0;
    return dev + "/" + kernel + "/" + std::to_string(nclass) + "/" + std::to_string(wclass);
}

// The shape for one kernel on one matrix: an override from the environment,
// else the cached decision, else every shape timed with HIP events (3 runs
// after one warm-up) and the fastest kept, in memory and in the cache file.
// All shapes add every row's entries in stored order, so they give
// bit-identical results.
template <class Run>
int tune(const char *kernel, const char *env_name, const BatchLevel &L, int &field, Run run) {
    if (const char *env = std::getenv(env_name)) {
        const int forced = std::atoi(env);
        if (valid_shape(forced)) return forced;
    }
    const std::string key = tune_key(kernel, L);
    const int hit = TuneCache::get().lookup(key);
    if (valid_shape(hit)) {
        ++g_tune.cached;
        return hit;
    }
    const double t0 = wall_ms();
    int best = kShapes[0];
    float tbest = 1e30f;
    for (int code : kShapes) {
        field = code;
        const float t = time_ms(run, 3);
        if (t < tbest) {
            tbest = t;
            best = code;
        }
    }
    ++g_tune.timed;
    g_tune.ms += wall_ms() - t0;
    TuneCache::get().store(key, best);
    return best;
}

int tune_shape(BatchLevel &L, int sm, float w) {
    return tune("bsweep", "HZ_BATCH_SHAPE", L, L.shape, [&] { bsweep(L, sm, w, L.b, L.x, L.xt); });
}

}  // namespace

// Setup of the stacked hierarchies.  A0: the level-0 super matrix (device,
// borrowed), off0: nsys + 1 row offsets.  theta: per system, the caller's
// strong_coupled (updated per level per system, as amg_setup_ua.c:1325-1328
// does to the caller's struct).  Returns SUCCESS or a negative code.
int batch_setup(BatchHierarchy &H, const DevCSR &A0, const std::vector<int> &off0, const AMG_param *param,
                double jacobi_w) {
    const int nsys = (int)off0.size() - 1;
    H.nsys = nsys;
    H.smoother = param->smoother;
    H.pre = param->presmooth_iter;
    H.post = param->postsmooth_iter;
    H.wgt = (float)jacobi_w;
    const int min_cdof = std::max(param->coarse_dof, 1);
    int max_levels = std::max<int>(1, param->max_levels);
    if (max_levels > HZ_MAX_AMG_LVL) max_levels = HZ_MAX_AMG_LVL;
    std::vector<double> theta(nsys, param->strong_coupled);
    H.levels_of.assign(nsys, 0);
    H.lv.clear();
    H.lv.emplace_back();
    H.lv[0].A = A0;
    H.lv[0].off = off0;
    H.lv[0].sys.resize(nsys);
    for (int s = 0; s < nsys; ++s) H.lv[0].sys[s] = s;
    hipStream_t st = stream();
    for (int lvl = 0;; ++lvl) {
        BatchLevel &L = H.lv[lvl];
        const int n = L.A.n, nblk = (int)L.sys.size();
        sell_from_csr(L.A, L.S);
        int *doff = arena().alloc_n<int>((size_t)nblk + 1);
        HZ_CHECK(memcpy_sync(doff, L.off.data(), sizeof(int) * (nblk + 1), hipMemcpyHostToDevice));
        L.blk = arena().alloc_n<int>((size_t)n + 1);
        k_row_block<<<grid_for(n), kBlock, 0, st>>>(n, nblk, doff, L.blk);
        int *dsys = arena().alloc_n<int>((size_t)nblk + 1);
        HZ_CHECK(memcpy_sync(dsys, L.sys.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
        L.rsys = arena().alloc_n<int>((size_t)n + 1);
        k_gather_i<<<grid_for(n), kBlock, 0, st>>>(n, L.blk, dsys, L.rsys);
        HZ_LAUNCH_CHECK();
        // which blocks try to coarsen (amg_setup_ua.c:1285)
        std::vector<int> keep(nblk, 0);
        for (int k = 0; k < nblk; ++k)
            keep[k] = (L.off[k + 1] - L.off[k] > min_cdof && lvl < max_levels - 1) ? 1 : 0;
        bool any = false;
        for (int k = 0; k < nblk; ++k) any = any || keep[k];
        std::vector<int> nagg_b(nblk, 0), amin(nblk, -1);
        if (any) {
            std::vector<double> th2(nblk);
            for (int k = 0; k < nblk; ++k) th2[k] = theta[L.sys[k]] * theta[L.sys[k]];
            double *dth2b = arena().alloc_n<double>((size_t)nblk + 1);
            double *dth2 = arena().alloc_n<double>((size_t)n + 1);
            HZ_CHECK(memcpy_sync(dth2b, th2.data(), sizeof(double) * nblk, hipMemcpyHostToDevice));
            k_gather_d<<<grid_for(n), kBlock, 0, st>>>(n, L.blk, dth2b, dth2);
            HZ_LAUNCH_CHECK();
            int *counts = arena().alloc_n<int>((size_t)n + 1);
            DevCSR N;
            strength_graph_rows(L.A, dth2, N, counts);
            L.agg = arena().alloc_n<int>((size_t)n + 1);
            int nagg = 0;
            const bool tr = trace_enabled();
            double t0 = 0;
            if (tr) {
                HZ_CHECK(hipStreamSynchronize(st));
                t0 = wall_ms();
            }
            int sth = hem_level(N, param, L.agg, &nagg);
            if (tr) {
                HZ_CHECK(hipStreamSynchronize(st));
                std::fprintf(stderr, "[hz-batch] level %d n=%d nnz=%d blocks=%d: aggregate %.2f ms\n", lvl, n,
                             L.A.nnz, nblk, wall_ms() - t0);
            }
            if (sth == -1017) return -1017;
            int *damin = arena().alloc_n<int>((size_t)nblk + 1);
            int *damax = arena().alloc_n<int>((size_t)nblk + 1);
            k_block_aggrange<<<nblk, kBlock, 0, st>>>(nblk, doff, L.agg, damin, damax);
            HZ_LAUNCH_CHECK();
            std::vector<int> amax(nblk);
            HZ_CHECK(memcpy_sync(amin.data(), damin, sizeof(int) * nblk, hipMemcpyDeviceToHost));
            HZ_CHECK(memcpy_sync(amax.data(), damax, sizeof(int) * nblk, hipMemcpyDeviceToHost));
            for (int k = 0; k < nblk; ++k) {
                if (!keep[k]) continue;
                const int nb = L.off[k + 1] - L.off[k];
                nagg_b[k] = amax[k] >= 0 ? amax[k] - amin[k] + 1 : 0;
                double &th = theta[L.sys[k]];
                if (nagg_b[k] * 4 > nb) th /= 2;
                else if (nagg_b[k] * 1.25 < nb) th *= 2;
                if (nagg_b[k] < 1) keep[k] = 0;   // P.col < MIN_CDOF (amg_setup_ua.c:1340)
            }
        }
        // systems solved at this level: dense FP64 inverse of their block
        std::vector<long long> ctab;
        long long inv_total = 0;
        for (int k = 0; k < nblk; ++k) {
            if (keep[k]) continue;
            const long long nb = L.off[k + 1] - L.off[k];
            if (nb > kCoarseMax) {
                std::printf("### ERROR: batched AMG: a coarsest level of %lld rows exceeds the dense limit %d\n",
                            nb, kCoarseMax);
                return HZ_ERROR_INPUT_PAR;
            }
            L.cmax = std::max<int>(L.cmax, (int)nb);
            ctab.push_back(L.off[k]);
            ctab.push_back(nb);
            ctab.push_back(inv_total);
            ctab.push_back(L.sys[k]);
            inv_total += nb * nb;
            H.levels_of[L.sys[k]] = lvl + 1;
        }
        L.ncoarse = (int)(ctab.size() / 4);
        if (L.ncoarse) {
            L.Ainv = arena().alloc_n<double>((size_t)inv_total + 1);
            L.ctab = arena().alloc_n<long long>(ctab.size() + 1);
            HZ_CHECK(memcpy_sync(L.ctab, ctab.data(), sizeof(long long) * ctab.size(), hipMemcpyHostToDevice));
            // dense copies of the blocks (work area at the inverses' offsets),
            // then all inverses in one launch
            double *W = arena().alloc_n<double>((size_t)inv_total + 1);
            std::vector<long long> woff(L.ncoarse);
            for (int c = 0; c < L.ncoarse; ++c) {
                const int r0 = (int)ctab[4 * c], nb = (int)ctab[4 * c + 1];
                woff[c] = ctab[4 * c + 2];
                k_dense_block<<<div_up(nb, 64), 64, 0, st>>>(L.A.ia, L.A.ja, L.A.val, r0, nb, W + woff[c]);
            }
            HZ_LAUNCH_CHECK();
            long long *dwoff = arena().alloc_n<long long>((size_t)L.ncoarse + 1);
            int *dstat = arena().alloc_n<int>((size_t)L.ncoarse + 1);
            HZ_CHECK(memcpy_sync(dwoff, woff.data(), sizeof(long long) * L.ncoarse, hipMemcpyHostToDevice));
            HZ_CHECK(hipMemsetAsync(dstat, 0, sizeof(int) * L.ncoarse, st));
            k_batch_gj_inverse<<<L.ncoarse, 1024, 0, st>>>(L.ctab, dwoff, W, L.Ainv, dstat);
            HZ_LAUNCH_CHECK();
            std::vector<int> hstat(L.ncoarse);
            HZ_CHECK(memcpy_sync(hstat.data(), dstat, sizeof(int) * L.ncoarse, hipMemcpyDeviceToHost));
            for (int v : hstat)
                if (v) return HZ_ERROR_SOLVER_MISC;
        }
        // live rows, renumbered aggregates, next level
        L.live = arena().alloc_n<unsigned char>((size_t)n + 1);
        int nc = 0;
        std::vector<int> shift(nblk, 0), noff(1, 0), nsys_next;
        for (int k = 0; k < nblk; ++k) {
            if (!keep[k]) continue;
            shift[k] = nc - amin[k];
            nc += nagg_b[k];
            noff.push_back(nc);
            nsys_next.push_back(L.sys[k]);
        }
        if (!L.agg) L.agg = arena().alloc_n<int>((size_t)n + 1);
        {
            int *dkeep = arena().alloc_n<int>((size_t)nblk + 1);
            int *dshift = arena().alloc_n<int>((size_t)nblk + 1);
            HZ_CHECK(memcpy_sync(dkeep, keep.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
            HZ_CHECK(memcpy_sync(dshift, shift.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
            k_agg_renumber<<<grid_for(n), kBlock, 0, st>>>(n, L.blk, dkeep, dshift, L.agg, L.live);
            HZ_LAUNCH_CHECK();
        }
        L.nagg = nc;
        // cycle storage of this level
        L.valf = arena().alloc_n<float>((size_t)L.S.nent + 1);
        if (L.S.nent)
            k_to_f32<<<grid_for(L.S.nent), kBlock, 0, st>>>(L.S.nent, L.S.val, L.valf);
        double *dsm64 = arena().alloc_n<double>((size_t)n + 1);
        smoother_diag(L.S, H.smoother, dsm64);
        L.dsm = arena().alloc_n<float>((size_t)n + 1);
        d2f(dsm64, L.dsm, n);
        const size_t nv = (size_t)n * kR + kR;
        L.b = arena().alloc_n<float>(nv);
        L.x = arena().alloc_n<float>(nv);
        L.xt = arena().alloc_n<float>(nv);
        L.w = arena().alloc_n<float>(nv);
        HZ_CHECK(hipMemsetAsync(L.x, 0, sizeof(float) * nv, st));
        HZ_CHECK(hipMemsetAsync(L.b, 0, sizeof(float) * nv, st));
        if (nc == 0) break;
        L.rptr = arena().alloc_n<int>((size_t)nc + 1);
        L.rmem = arena().alloc_n<int>((size_t)n + 1);
        build_restriction(n, L.agg, nc, L.rptr, L.rmem);
        BatchLevel C;
        rap_agg(L.A, L.agg, nc, L.rptr, L.rmem, C.A);
        C.off = noff;
        C.sys = nsys_next;
        H.lv.push_back(C);
    }
    H.nl = (int)H.lv.size();
    H.act = arena().alloc_n<int>((size_t)nsys + 1);
    k_fill_i<<<grid_for(nsys), kBlock, 0, st>>>(nsys, 1, H.act);
    HZ_LAUNCH_CHECK();
    HZ_CHECK(hipStreamSynchronize(st));
    for (auto &L : H.lv) L.shape = tune_shape(L, H.smoother, H.wgt);
    return HZ_SUCCESS;
}

// z = B(r) for all columns: one V-cycle from zero (precond_amg with maxit 1,
// cycle_type V), FP32.  r, z: level-0 interleaved FP64.
void batch_vcycle(BatchHierarchy &H, const double *r, double *z) {
    const int nl = H.nl;
    hipStream_t st = stream();
    BatchLevel &L0 = H.lv[0];
    const int *act = H.act;
    k_enter_b<<<grid_for(L0.A.n), kBlock, 0, st>>>(L0.A.n, r, L0.b, L0.rsys, act);
    HZ_LAUNCH_CHECK();
    // descend
    for (int l = 0; l < nl; ++l) {
        BatchLevel &L = H.lv[l];
        const int n = L.A.n;
        if (L.ncoarse)
            k_bcoarse<<<L.ncoarse, 256, sizeof(float) * kR * L.cmax, st>>>(L.ctab, L.Ainv, L.b, L.x, act);
        if (l == nl - 1) break;
        if (H.pre > 0) {
            k_bzero_sweep<<<grid_for(n), kBlock, 0, st>>>(n, H.smoother, H.wgt, L.dsm, L.live, L.b, L.x,
                                                          L.rsys, act);
            for (int s = 1; s < H.pre; ++s) {
                bsweep(L, H.smoother, H.wgt, L.b, L.x, L.xt, act);
                std::swap(L.x, L.xt);
            }
        } else {
            // x = 0 on live rows: the dead rows hold the coarse solution
            k_bzero_sweep<<<grid_for(n), kBlock, 0, st>>>(n, HZ_SMOOTHER_JACOBI, 0.0f, L.dsm, L.live, L.b, L.x,
                                                          L.rsys, act);
        }
        bresidual(L, act);
        BatchLevel &C = H.lv[l + 1];
        const long totc = (long)C.A.n * kR;
        k_brestrict<<<grid_for(C.A.n), kBlock, 0, st>>>(C.A.n, L.rptr, L.rmem, L.w, C.b, C.rsys, act);
        HZ_LAUNCH_CHECK();
    }
    // ascend
    for (int l = nl - 2; l >= 0; --l) {
        BatchLevel &L = H.lv[l];
        BatchLevel &C = H.lv[l + 1];
        k_bprolong<<<grid_for(L.A.n), kBlock, 0, st>>>(L.A.n, L.agg, C.x, L.x, L.rsys, act);
        for (int s = 0; s < H.post; ++s) {
            bsweep(L, H.smoother, H.wgt, L.b, L.x, L.xt, act);
            std::swap(L.x, L.xt);
        }
    }
    k_exit_b<<<grid_for(L0.A.n), kBlock, 0, st>>>(L0.A.n, L0.x, z, L0.rsys, act);
    HZ_LAUNCH_CHECK();
}

struct BatchPcg {
    int nchunks = 0, ncols = 0;
    int *crow = nullptr, *csys = nullptr, *schunk = nullptr;
    double *part = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *t = nullptr;
    const double *b = nullptr;
    ColState *cs = nullptr;
    PcgCtl *ctl = nullptr;
    int *scand = nullptr;   // per system: some column is a check-III candidate this iteration
    hipGraphExec_t exec = nullptr;
};

namespace {

void pcg_alloc(BatchPcg &P, const BatchHierarchy &H, const std::vector<int> &off0) {
    const int nsys = H.nsys;
    std::vector<int> crow, csys, schunk(1, 0);
    for (int s = 0; s < nsys; ++s) {
        for (int r0 = off0[s]; r0 < off0[s + 1]; r0 += kChunkRows) {
            crow.push_back(r0);
            csys.push_back(s);
        }
        schunk.push_back((int)csys.size());
    }
    crow.push_back(off0[nsys]);
    P.nchunks = (int)csys.size();
    P.ncols = nsys * kR;
    P.crow = arena().alloc_n<int>(crow.size() + 1);
    P.csys = arena().alloc_n<int>(csys.size() + 1);
    P.schunk = arena().alloc_n<int>(schunk.size() + 1);
    HZ_CHECK(memcpy_sync(P.crow, crow.data(), sizeof(int) * crow.size(), hipMemcpyHostToDevice));
    if (!csys.empty())
        HZ_CHECK(memcpy_sync(P.csys, csys.data(), sizeof(int) * csys.size(), hipMemcpyHostToDevice));
    HZ_CHECK(memcpy_sync(P.schunk, schunk.data(), sizeof(int) * schunk.size(), hipMemcpyHostToDevice));
    P.part = arena().alloc_n<double>((size_t)P.nchunks * kR + kR);
    const size_t nv = (size_t)off0[nsys] * kR + kR;
    P.r = arena().alloc_n<double>(nv);
    P.z = arena().alloc_n<double>(nv);
    P.p = arena().alloc_n<double>(nv);
    P.t = arena().alloc_n<double>(nv);
    P.cs = arena().alloc_n<ColState>((size_t)P.ncols + 1);
    P.scand = arena().alloc_n<int>((size_t)nsys + 1);
    P.ctl = arena().alloc_n<PcgCtl>(1);
}

// y = A x (sub = 0) or y = b - A x (sub = 1) on level 0 in FP64, in the
// shape the tuner chose for this matrix
void spmm(const BatchLevel &L0, const double *x, const double *b, double *y, int sub, const int *gate,
          const int *act) {
    with_shape(L0.spmm_shape, [&](auto H, auto U) {
        constexpr int kH = decltype(H)::value, kU = decltype(U)::value;
        const int grid = std::max(1, div_up(L0.S.nslices * kH, kSlicesPerBlock));
        if (sub)
            k_spmm64<2, kH, kU><<<grid, kBlock, 0, stream()>>>(L0.S.n, L0.S.nslices, L0.S.sptr, L0.S.col, L0.S.val,
                                                              x, b, y, gate, L0.rsys, act);
        else
            k_spmm64<1, kH, kU><<<grid, kBlock, 0, stream()>>>(L0.S.n, L0.S.nslices, L0.S.sptr, L0.S.col, L0.S.val,
                                                              x, b, y, gate, L0.rsys, act);
    });
    HZ_LAUNCH_CHECK();
}

// The tuner's choice of the FP64 SpMM shape for this level-0 matrix, timed on
// y = A x (HZ_BATCH_SPMM_SHAPE overrides).
int tune_spmm_shape(BatchLevel &L0, const double *x, double *y) {
    return tune("spmm64", "HZ_BATCH_SPMM_SHAPE", L0, L0.spmm_shape,
                [&] { spmm(L0, x, nullptr, y, 0, nullptr, nullptr); });
}

// one PCG iteration (graph body)
void pcg_iteration(BatchHierarchy &H, BatchPcg &P, double *u) {
    const BatchLevel &L0 = H.lv[0];
    hipStream_t st = stream();
    const int nch = P.nchunks, nc = P.ncols;
    const int *act = H.act;
    int *sact = H.act;
    spmm(L0, P.p, nullptr, P.t, 0, nullptr, act);                                            // t = A p
    k_cdot<<<nch, kBlock, 0, st>>>(P.crow, P.p, P.t, P.part, P.csys, act);                   // (t, p)
    k_cfinal<1><<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs, P.ctl, sact);              // alpha
    k_cupdate<<<nch, kBlock, 0, st>>>(P.crow, P.csys, P.cs, P.p, P.t, u, P.r, P.part, act);  // u, r, ||r||
    k_cfinal<2><<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs, P.ctl, P.scand);           // relres, candidates
    // check III: b - A u, only on the systems holding a candidate column
    spmm(L0, u, P.b, P.t, 1, &P.ctl->anycand, P.scand);
    k_creplace<<<nch, kBlock, 0, st>>>(P.crow, P.csys, P.cs, P.t, P.r, P.part, &P.ctl->anycand, P.scand);
    k_cfinal<3><<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs, P.ctl, sact);              // check III, maxit
    HZ_LAUNCH_CHECK();
    batch_vcycle(H, P.r, P.z);                                                               // z = B r
    k_cdot<<<nch, kBlock, 0, st>>>(P.crow, P.z, P.r, P.part, P.csys, act);                   // (z, r)
    k_cfinal<4><<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs, P.ctl, sact);              // beta
    k_cdirection<<<nch, kBlock, 0, st>>>(P.crow, P.csys, P.cs, P.z, P.p, act);              // p = z + beta p
    HZ_LAUNCH_CHECK();
}

}  // namespace

int batch_solve(const DevCSR &A0, const std::vector<int> &off0, const double *b, double *u,
                const linear_itsolver_param *itp, AMG_param *amg, double jacobi_w, int *iters, BatchStats *stats) {
    hipStream_t st = stream();
    const double t0 = wall_ms();
    g_tune = TuneStats{};
    BatchHierarchy H;
    int status = batch_setup(H, A0, off0, amg, jacobi_w);
    if (status < 0) return status;
    HZ_CHECK(hipStreamSynchronize(st));
    const double t1 = wall_ms();
    BatchPcg P;
    pcg_alloc(P, H, off0);
    P.b = b;
    const BatchLevel &L0 = H.lv[0];
    const int nch = P.nchunks, nc = P.ncols;
    const long tot0 = (long)L0.A.n * kR;
    PcgCtl hctl{};
    hctl.maxit = itp->linear_maxit;
    hctl.tol = itp->linear_tol;
    HZ_CHECK(memcpy_sync(P.ctl, &hctl, sizeof(PcgCtl), hipMemcpyHostToDevice));
    // r = b - A u, z = B r, initial exit test, temp1 = (z, r), p = z (krylov.c:432-470)
    // (every system active here: H.act was set to 1 at setup)
    spmm(L0, u, b, P.r, 1, nullptr, nullptr);
    k_fill_d<<<grid_for(tot0), kBlock, 0, st>>>(tot0 + kR, 0.0, P.p);
    H.lv[0].spmm_shape = tune_spmm_shape(H.lv[0], P.p, P.t);
    batch_vcycle(H, P.r, P.z);
    k_cdot<<<nch, kBlock, 0, st>>>(P.crow, P.r, P.r, P.part, P.csys, nullptr);
    k_cfinal<0><<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs, P.ctl, H.act);
    k_cdot<<<nch, kBlock, 0, st>>>(P.crow, P.z, P.r, P.part, P.csys, nullptr);
    k_cfinal_temp1<<<1, 1024, 0, st>>>(nc, P.part, P.schunk, P.cs);
    k_cdirection<<<nch, kBlock, 0, st>>>(P.crow, P.csys, P.cs, P.z, P.p, nullptr);
    HZ_LAUNCH_CHECK();
    // one iteration captured once, replayed until every column is finished
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    HZ_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    pcg_iteration(H, P, u);
    HZ_CHECK(hipStreamEndCapture(st, &graph));
    HZ_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    PcgCtl *pin = nullptr;
    HZ_CHECK(hipHostMalloc((void **)&pin, sizeof(PcgCtl), hipHostMallocDefault));
    HZ_CHECK(hipMemcpyAsync(pin, P.ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    int launched = 0;
    while (pin->active > 0 && launched < hctl.maxit) {
        HZ_CHECK(hipGraphLaunch(exec, st));
        ++launched;
        HZ_CHECK(hipMemcpyAsync(pin, P.ctl, sizeof(PcgCtl), hipMemcpyDeviceToHost, st));
        HZ_CHECK(hipStreamSynchronize(st));
    }
    (void)hipHostFree(pin);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    std::vector<ColState> cs(nc);
    HZ_CHECK(memcpy_sync(cs.data(), P.cs, sizeof(ColState) * nc, hipMemcpyDeviceToHost));
    for (int c = 0; c < nc; ++c) iters[c] = cs[c].done ? cs[c].iters : HZ_ERROR_SOLVER_MAXIT;
    const double t2 = wall_ms();
    if (stats && stats->time_kernels) {
        // level-0 kernels on the final state: FP64 SpMM t = A u and one FP32 sweep
        stats->spmm_ms = time_ms([&] { spmm(L0, u, nullptr, P.t, 0, nullptr, nullptr); }, 10);
        BatchLevel &M0 = H.lv[0];
        stats->sweep_ms = time_ms([&] { bsweep(M0, H.smoother, H.wgt, M0.b, M0.x, M0.xt); }, 10);
        stats->rows = L0.A.n;
        stats->nnz = L0.A.nnz;
        stats->sell_entries = L0.S.nent;
    }
    if (stats) {
        stats->setup_ms = t1 - t0;
        stats->solve_ms = t2 - t1;
        stats->num_levels = H.nl;
        stats->graph_launches = launched;
        long long nnz = 0;
        for (const auto &L : H.lv) nnz += L.A.nnz;
        stats->op_complexity = L0.A.nnz ? (double)nnz / L0.A.nnz : 0.0;
        stats->shape_level0 = (L0.shape / 100) * 10 + L0.spmm_shape / 100;
        stats->sweep_shape0 = L0.shape;
        stats->spmm_shape0 = L0.spmm_shape;
        stats->tuned = g_tune.timed;
        stats->tune_cached = g_tune.cached;
        stats->tune_ms = g_tune.ms;
    }
    return HZ_SUCCESS;
}

}  // namespace hz
