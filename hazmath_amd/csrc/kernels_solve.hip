// kernels_solve.hip -- solve-phase kernels of the UA-AMG-PCG path on gfx950.
//
// Matrix kernels use SELL-64: one wavefront owns 64 consecutive rows, lane l
// owns row 64*s+l and walks that row's entries in CSR order, so every per-row
// sum is formed in the same order as the reference's CSR loops
// (src/utilities/sparse.c:1456-1730, src/solver/smoother.c:35-92,324-381) and,
// built with -ffp-contract=off, rounds identically.  Loads of a slice are
// column-major (64 lanes x 8 B contiguous per step): fully coalesced, and for
// stencil matrices the x gathers of a step are contiguous as well.
#include <cstring>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;               // 4 wavefronts
constexpr int kSlicesPerBlock = kBlock / kWave;
constexpr int kStreamGrid = 2048;         // fixed grid of grid-stride vector kernels
constexpr double kCacheResidentBytes = 192e6;   // three quarters of the last-level cache

__global__ void k_slice_width(int n, const int *__restrict__ ia, long long *__restrict__ wcount,
                              int nslices) {
    const int s = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    int len = (r < n) ? ia[r + 1] - ia[r] : 0;
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off, kWave));
    if (lane == 0) wcount[s] = (long long)len * kSlice;
}

// One wavefront per slice.  The slice's rows are contiguous in the CSR arrays:
// up to kSellStage of their entries are first loaded coalesced into LDS, then
// each lane writes its row down the slice's columns (col -1 / 0.0 padding)
// and forms its row's diagonals on the way: the l1 norm (sum of |a| from 0.0
// in stored order, L1DIAG) and the last stored diagonal entry (Jacobi and the
// rest) -- what k_smoother_diag computes from the SELL copy -- and the first
// stored diagonal entry (dcsr_getdiag, which the strength graph reads).
constexpr int kSellStage = 1024;
__global__ __launch_bounds__(kBlock) void k_sell_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const double *__restrict__ val,
                                                      const long long *__restrict__ sptr, int *__restrict__ scol,
                                                      double *__restrict__ sval, int nslices, double *__restrict__ dl1,
                                                      double *__restrict__ djac, double *__restrict__ dfirst) {
    __shared__ int s_col[kSlicesPerBlock][kSellStage];
    __shared__ double s_val[kSlicesPerBlock][kSellStage];
    const int wv = threadIdx.x / kWave;
    const int s = blockIdx.x * kSlicesPerBlock + wv;
    const int lane = threadIdx.x % kWave;
    const bool active = s < nslices;
    const int r = s * kSlice + lane;
    int e0 = 0, e1 = 0;
    if (active) {
        e0 = ia[s * kSlice];
        e1 = ia[min(n, (s + 1) * kSlice)];
    }
    const bool staged = active && e1 - e0 <= kSellStage;
    if (staged)
        for (int e = e0 + lane; e < e1; e += kWave) {
            s_col[wv][e - e0] = ja[e];
            s_val[wv][e - e0] = val[e];
        }
    __syncthreads();
    if (!active) return;
    const long long base = sptr[s];
    const int w = (int)((sptr[s + 1] - base) / kSlice);
    const int beg = (r < n) ? ia[r] : 0;
    const int len = (r < n) ? ia[r + 1] - beg : 0;
    double l1 = 0.0, dj = 0.0, df = 0.0;
    bool seen = false;
    for (int k = 0; k < w; ++k) {
        const long long e = base + (long long)k * kSlice + lane;
        if (k < len) {
            const int q = beg + k;
            const int j = staged ? s_col[wv][q - e0] : ja[q];
            const double a = staged ? s_val[wv][q - e0] : val[q];
            scol[e] = j;
            sval[e] = a;
            l1 = l1 + fabs(a);
            if (j == r) {
                dj = a;
                if (!seen) df = a;
                seen = true;
            }
        } else {
            scol[e] = -1;
            sval[e] = 0.0;
        }
    }
    if (r < n) {
        dl1[r] = l1;
        djac[r] = dj;
        dfirst[r] = df;
    }
}

// Row sum from 0.0 in stored order over one lane's SELL entries (padding col
// -1 skipped).  Entries go eight at a time: their column and value loads
// (non-temporal) are issued first, then the eight gathers, then the adds in
// order, so a row costs a few memory round trips instead of one per entry.
// kKeep 1: the second half of each round uses ordinary loads, whose lines
// stay in the last-level cache for the next kernel that streams the same
// matrix; kKeep 2: every load is ordinary; kKeep 4: the column indices only.
template <bool kNtTail, bool kWide = false, int kKeep = 0, class CT = int>
__device__ __forceinline__ double sell_row_dot(const CT *c, const double *v, int w, const double *x, int r) {
    // kWide: rounds of 16 masked past the slice width (a 15-entry row in one round)
    constexpr int U = kWide ? 16 : 8;
    static_assert(kWide || std::is_same_v<CT, int>, "paired offsets: masked rounds only");
    constexpr int kKeepFrom = (kKeep == 2 || kKeep == 4) ? 0 : (kKeep == 1 ? U / 2 : U);
    double acc = 0.0;
    int k = 0;
    for (; kWide ? (k < w) : (k + U <= w); k += U) {
        int cc[U];
        double vv[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = !kWide || k + u < w;
            const bool keep = kKeep == 2 || (kKeep == 1 && u >= U / 2);
            vv[u] = ok ? (keep ? v[(k + u) * kSlice] : ld_nt(v + (k + u) * kSlice)) : 0.0;
        }
        ld_cols<U, kKeepFrom>(c, k, w, r, cc);
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (cc[u] >= 0) acc = acc + vv[u] * xx[u];
    }
    for (; k < w; ++k) {
        const int c0 = kNtTail ? ld_nt(c + k * kSlice) : c[k * kSlice];
        if (c0 >= 0) acc = acc + (kNtTail ? ld_nt(v + k * kSlice) : v[k * kSlice]) * x[c0];
    }
    return acc;
}

// y = A x (row sums from 0.0 in CSR order); optional deterministic (y, x).
template <bool kDot, bool kWide = false, int kKeep = 0, class CT = int>
__global__ __launch_bounds__(kBlock) void k_spmv(int n, int nslices, const long long *__restrict__ sptr,
                                                 const CT *__restrict__ col,
                                                 const double *__restrict__ val,
                                                 const double *__restrict__ x, double *__restrict__ y,
                                                 double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    const int blk = xcd_block_id();
    const int s = blk * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = s * kSlice + lane;
    double acc = 0.0;
    double prod = 0.0;
    if (s < nslices) {
        const long long base = sptr[s];
        const int w = (int)((sptr[s + 1] - base) / kSlice);
        const CT *c = col + col_base<CT>(base, s) + lane;
        const double *v = val + base + lane;
        acc = sell_row_dot<true, kWide, kKeep>(c, v, w, x, r);
        if (r < n) {
            y[r] = acc;
            if (kDot) prod = acc * x[r];
        }
    }
    if (kDot) {
        const double bs = block_sum<kBlock>(prod, sh);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

// w = b - A x  (array_cp + dcsr_aAxpy(-1.0): row sum t from 0.0, then b - t)
template <bool kWide = false, int kKeep = 0, class CT = int>
__global__ __launch_bounds__(kBlock) void k_residual(int n, int nslices, const long long *__restrict__ sptr,
                                                     const CT *__restrict__ col,
                                                     const double *__restrict__ val,
                                                     const double *__restrict__ b,
                                                     const double *__restrict__ x,
                                                     double *__restrict__ w) {
    const int s = xcd_block_id() * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    const long long base = sptr[s];
    const int wd = (int)((sptr[s + 1] - base) / kSlice);
    const CT *c = col + col_base<CT>(base, s) + lane;
    const double *v = val + base + lane;
    const double acc = sell_row_dot<false, kWide, kKeep>(c, v, wd, x, r);
    if (r < n) w[r] = b[r] - acc;
}

// x_j + alpha * (P xc)_j formed exactly like dcsr_aAxpy_agg (sparse.c:1748-1784):
// t = 0.0 (+ xc[g] when row j of P is not empty), then x + t (alpha = 1),
// x - t (alpha = -1) or x + t * alpha.
__device__ __forceinline__ double prolong_value(double xj, int g, double xcg, double a) {
    double t = 0.0;
    if (g > HZ_UNPT) t = t + xcg;
    if (a == 1.0) return xj + t;
    if (a == -1.0) return xj - t;
    return xj + t * a;
}

// One smoothing sweep out of place, optionally with the prolongation of the
// level below folded in (every x value read is x + alpha P xc) and optionally
// with per-block partials of sum_i xout_i b_i.
// Jacobi (smoother.c:35-92): t = b - sum_{j!=i} a_ij x_j, d = last diagonal
// entry; x_i <- (1-w) x_i + w t / d when |d| > SMALLREAL.
// L1DIAG (smoother.c:324-381): t = b - sum_j a_ij x_j, d = sum_j |a_ij|; x_i += t/d.
// Entries are consumed four at a time with all their loads issued first, so a
// row costs a few dependent memory round trips instead of one per entry.
template <int kSmoother, bool kProlong, bool kDot, bool kWide = false, class CT = int>
__global__ __launch_bounds__(kBlock) void k_sweep(int n, int nslices, const long long *__restrict__ sptr,
                                                  const CT *__restrict__ col,
                                                  const double *__restrict__ val, double wgt,
                                                  const int *__restrict__ agg,
                                                  const double *__restrict__ xc, double alpha,
                                                  const double *__restrict__ alpha_dev,
                                                  const double *__restrict__ b,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ xout,
                                                  double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    const int blk = xcd_block_id();
    const int s = blk * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = s * kSlice + lane;
    const double al = kProlong ? (alpha_dev ? *alpha_dev : alpha) : 1.0;
    double prod = 0.0;
    if (s < nslices) {
        const long long base = sptr[s];
        const int wd = (int)((sptr[s + 1] - base) / kSlice);
        const CT *c = col + col_base<CT>(base, s) + lane;
        const double *v = val + base + lane;
        double t = (r < n) ? b[r] : 0.0;
        double d = 0.0;
        // entries per round: loads first (matrix non-temporal), then the math in
        // order.  With the prolongation folded in (the latency-bound small
        // levels) rounds are 16 wide and masked past the slice width, so a
        // 15-entry row takes one round; otherwise 8 wide plus a tail.
        constexpr bool kMask = kProlong || kWide;
        constexpr int U = kMask ? 16 : 8;
        static_assert(kMask || std::is_same_v<CT, int>, "paired offsets: masked rounds only");
        int k = 0;
        for (; kMask ? (k < wd) : (k + U <= wd); k += U) {
            int cc[U];
            double vv[U], xx[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool ok = !kMask || k + u < wd;
                // the fine level's last sweep (kDot) precedes the PCG SpMV: the
                // second half of each round uses ordinary loads, which stay in the
                // last-level cache for that SpMV
                const bool keep = kWide && kDot && u >= 8;
                vv[u] = ok ? (keep ? v[(k + u) * kSlice] : ld_nt(v + (k + u) * kSlice)) : 0.0;
            }
            ld_cols<U, (kWide && kDot) ? 8 : U>(c, k, wd, r, cc);
#pragma unroll
            for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
            if (kProlong) {
                int gg[U];
                double xg[U];
#pragma unroll
                for (int u = 0; u < U; ++u) gg[u] = cc[u] >= 0 ? agg[cc[u]] : HZ_UNPT;
#pragma unroll
                for (int u = 0; u < U; ++u) xg[u] = gg[u] > HZ_UNPT ? xc[gg[u]] : 0.0;
#pragma unroll
                for (int u = 0; u < U; ++u) xx[u] = prolong_value(xx[u], gg[u], xg[u], al);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kSmoother == HZ_SMOOTHER_L1DIAG) {
                    if (cc[u] >= 0) {
                        t = t - vv[u] * xx[u];
                        d = d + fabs(vv[u]);
                    }
                } else {
                    if (cc[u] == r) d = vv[u];
                    else if (cc[u] >= 0) t = t - vv[u] * xx[u];
                }
            }
        }
        for (; k < wd; ++k) {
            const int cj = c[k * kSlice];
            const double a = v[k * kSlice];
            double xj = cj >= 0 ? x[cj] : 0.0;
            if (kProlong && cj >= 0) {
                const int g = agg[cj];
                xj = prolong_value(xj, g, g > HZ_UNPT ? xc[g] : 0.0, al);
            }
            if (kSmoother == HZ_SMOOTHER_L1DIAG) {
                if (cj >= 0) {
                    t = t - a * xj;
                    d = d + fabs(a);
                }
            } else {
                if (cj == r) d = a;
                else if (cj >= 0) t = t - a * xj;
            }
        }
        if (r < n) {
            double xi = x[r];
            if (kProlong) {
                const int g = agg[r];
                xi = prolong_value(xi, g, g > HZ_UNPT ? xc[g] : 0.0, al);
            }
            double xn = xi;
            if (fabs(d) > kSmallReal) {
                if (kSmoother == HZ_SMOOTHER_L1DIAG) xn = xi + t / d;
                else xn = (1 - wgt) * xi + wgt * t / d;
            }
            xout[r] = xn;
            if (kDot) prod = xn * b[r];
        }
    }
    if (kDot) {
        const double bs = block_sum<kBlock>(prod, sh);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

template <int kSmoother>
__global__ void k_smoother_diag(int n, int nslices, const long long *__restrict__ sptr,
                                const int *__restrict__ col, const double *__restrict__ val,
                                double *__restrict__ d) {
    const int s = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    const long long base = sptr[s];
    const int wd = (int)((sptr[s + 1] - base) / kSlice);
    double dd = 0.0;
    for (int k = 0; k < wd; ++k) {
        const int cj = col[base + (long long)k * kSlice + lane];
        const double a = val[base + (long long)k * kSlice + lane];
        if (kSmoother == HZ_SMOOTHER_L1DIAG) {
            if (cj >= 0) dd = dd + fabs(a);
        } else if (cj == r) {
            dd = a;
        }
    }
    if (r < n) d[r] = dd;
}

// First sweep from x = 0: t = b exactly (up to the sign of zero), so
// Jacobi gives (1-w)*0 + w*b/d and L1DIAG gives 0 + b/d.
// One sweep from a zero iterate at one row (Jacobi / L1DIAG with x = 0).
__device__ __forceinline__ double sweep_from_zero(double bi, double di, int smoother, double wgt) {
    double xi = 0.0;
    if (fabs(di) > kSmallReal) {
        if (smoother == HZ_SMOOTHER_L1DIAG) xi = 0.0 + bi / di;
        else xi = (1 - wgt) * 0.0 + wgt * bi / di;
    }
    return xi;
}

__global__ void k_smooth_from_zero(int n, int smoother, double wgt, const double *__restrict__ d,
                                   const double *__restrict__ b, double *__restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        x[i] = sweep_from_zero(b[i], d[i], smoother, wgt);
}

// bc[c] = sum over members (ascending) of w, from 0.0 (dcsr_mxv_agg, sparse.c:1663-1679)
__global__ void k_restrict(int nc, const int *__restrict__ rptr, const int *__restrict__ rmem,
                           const double *__restrict__ w, double *__restrict__ bc, int smoother,
                           double wgt, const double *__restrict__ dc, double *__restrict__ xc) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        double s = 0.0;
        const int e = rptr[c + 1];
        for (int q = rptr[c]; q < e; ++q) s = s + w[rmem[q]];
        bc[c] = s;
        if (xc) {
            const double di = dc[c];
            double xi = 0.0;
            if (fabs(di) > kSmallReal) {
                if (smoother == HZ_SMOOTHER_L1DIAG) xi = 0.0 + s / di;
                else xi = (1 - wgt) * 0.0 + wgt * s / di;
            }
            xc[c] = xi;
        }
    }
}

// Residual, restriction and the coarse level's first (zero-iterate) sweep in
// one pass, one thread per coarse row: w_m = b_m - A_m x for each member m
// (row sum from 0.0 in stored order), bc = 0.0 + w_m0 + w_m1 + ... in member
// order, then xc from bc.  Same operations and order as residual + restrict.
// Latency-bound small levels: kLanes lanes share a coarse row, lane j taking
// members j, j + kLanes, ...; each member row's sum is formed in CSR order with
// its loads issued 16 at a time (masked past the row end), and lane 0 adds the
// members' residuals in member order -- the same operations and order as
// residual + restrict.
template <int kLanes>
__global__ void k_down_fused(int nc, const int *__restrict__ rptr, const int *__restrict__ rmem,
                             const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ a, const double *__restrict__ b,
                             const double *__restrict__ x, double *__restrict__ bc, int smoother,
                             double wgt, const double *__restrict__ dc, double *__restrict__ xc) {
    constexpr int U = 16;
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const int lane = threadIdx.x % kWave, gbase = lane - j;
    const bool valid = c < nc;
    const int q0 = valid ? rptr[c] : 0, q1 = valid ? rptr[c + 1] : 0;
    double s = 0.0;
    for (int qb = q0; qb < q1; qb += kLanes) {   // trip count uniform within the lane group
        const int q = qb + j;
        double wv = 0.0;
        if (q < q1) {
            const int m = rmem[q];
            const int k0 = ia[m], k1 = ia[m + 1];
            const double bm = b[m];
            double t = 0.0;
            for (int k = k0; k < k1; k += U) {
                int jj[U];
                double aa[U], xx[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool ok = k + u < k1;
                    jj[u] = ok ? ja[k + u] : -1;
                    aa[u] = ok ? a[k + u] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xx[u] = jj[u] >= 0 ? x[jj[u]] : 0.0;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (jj[u] >= 0) t = t + aa[u] * xx[u];
            }
            wv = bm - t;
        }
#pragma unroll
        for (int u = 0; u < kLanes; ++u) {
            const double v = __shfl(wv, gbase + u, kWave);
            if (j == 0 && qb + u < q1) s = s + v;
        }
    }
    if (valid && j == 0) {
        bc[c] = s;
        if (xc) {
            const double di = dc[c];
            double xi = 0.0;
            if (fabs(di) > kSmallReal) {
                if (smoother == HZ_SMOOTHER_L1DIAG) xi = 0.0 + s / di;
                else xi = (1 - wgt) * 0.0 + wgt * s / di;
            }
            xc[c] = xi;
        }
    }
}

// x += alpha * P xc (dcsr_aAxpy_agg, sparse.c:1748-1784)
__global__ void k_prolong(int n, const int *__restrict__ agg, const double *__restrict__ xc,
                          double alpha, const double *__restrict__ alpha_dev,
                          double *__restrict__ x) {
    const double a = alpha_dev ? *alpha_dev : alpha;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int g = ld_nt(agg + i);
        double t = 0.0;
        if (g > HZ_UNPT) t = t + xc[g];
        const double xi = ld_nt(x + i);
        if (a == 1.0) x[i] = xi + t;
        else if (a == -1.0) x[i] = xi - t;
        else x[i] = xi + t * a;
    }
}

__global__ void k_fill(double *__restrict__ x, double v, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = v;
}

__global__ void k_copy(double *__restrict__ d, const double *__restrict__ s, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        d[i] = s[i];
}

// u += alpha p ; r -= alpha t ; then per-block partials of ||r||^2, ||u||_inf,
// ||u||^2, ||p||^2 (finalized by k_final_update).  alpha = zr / temp2.
// kFirst: also x0 = the preconditioner's first level-0 sweep from zero on the
// new r (the V-cycle then starts after it).  The streamed vectors are read
// with non-temporal loads, as in k_pcg_direction and k_prolong (a 1 GiB
// stream reads at 6.6-6.8 TB/s that way against 4.4-4.8 TB/s with ordinary
// loads).
template <bool kFirst>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int n, const PcgScalars *__restrict__ s,
                                                       const double *__restrict__ p,
                                                       const double *__restrict__ t,
                                                       double *__restrict__ u,
                                                       double *__restrict__ r,
                                                       double *__restrict__ partials, int smoother,
                                                       double wgt, const double *__restrict__ d,
                                                       double *__restrict__ x0) {
    __shared__ double sh[kBlock / kWave];
    const double alpha = s->zr / s->temp2;
    const double malpha = -alpha;
    double rr = 0.0, ui = 0.0, uu = 0.0, pp = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        const double pi = ld_nt(p + i);
        double ui_ = ld_nt(u + i);
        double ri = ld_nt(r + i);
        if (alpha == 1.0) ui_ = ui_ + pi;
        else if (alpha == -1.0) ui_ = ui_ - pi;
        else ui_ = ui_ + alpha * pi;
        const double ti = ld_nt(t + i);
        if (malpha == 1.0) ri = ri + ti;
        else if (malpha == -1.0) ri = ri - ti;
        else ri = ri + malpha * ti;
        u[i] = ui_;
        r[i] = ri;
        if (kFirst) x0[i] = sweep_from_zero(ri, ld_nt(d + i), smoother, wgt);
        rr += ri * ri;
        ui = fmax(ui, fabs(ui_));
        uu += ui_ * ui_;
        pp += pi * pi;
    }
    const int G = gridDim.x;
    double v;
    v = block_sum<kBlock>(rr, sh);
    if (threadIdx.x == 0) partials[0 * G + blockIdx.x] = v;
    v = block_max<kBlock>(ui, sh);
    if (threadIdx.x == 0) partials[1 * G + blockIdx.x] = v;
    v = block_sum<kBlock>(uu, sh);
    if (threadIdx.x == 0) partials[2 * G + blockIdx.x] = v;
    v = block_sum<kBlock>(pp, sh);
    if (threadIdx.x == 0) partials[3 * G + blockIdx.x] = v;
}

constexpr int kFinBlock = 1024;

// Single-block finalizers: the partials are summed in block-index order, so the
// totals depend only on the grid size.
__global__ __launch_bounds__(kFinBlock) void k_final_sum(const double *__restrict__ partials, int G,
                                                         double *__restrict__ out) {
    __shared__ double sh[kFinBlock / kWave];
    const double t = final_sum<kFinBlock>(partials, G, sh);
    if (threadIdx.x == 0) *out = t;
}

__global__ __launch_bounds__(kFinBlock) void k_final_update(PcgScalars *__restrict__ s,
                                                            const double *__restrict__ partials, int G,
                                                            double *__restrict__ alpha_hist, int iter) {
    __shared__ double sh[kFinBlock / kWave];
    const double a0 = final_sum<kFinBlock>(partials + 0 * G, G, sh);
    const double a1 = final_max<kFinBlock>(partials + 1 * G, G, sh);
    const double a2 = final_sum<kFinBlock>(partials + 2 * G, G, sh);
    const double a3 = final_sum<kFinBlock>(partials + 3 * G, G, sh);
    if (threadIdx.x == 0) {
        const double alpha = s->zr / s->temp2;
        s->rr = a0;
        s->uinf = a1;
        s->uu = a2;
        s->pp = a3;
        s->alpha = alpha;
        s->temp1 = s->zr;
        if (alpha_hist) alpha_hist[iter - 1] = alpha;
    }
}

__global__ __launch_bounds__(kBlock) void k_dot(int n, const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    double acc = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc += x[i] * y[i];
    const double v = block_sum<kBlock>(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

// beta = zr / temp1 ; p = 1.0*z + beta*p (array_axpby, array.c:423-433)
__global__ void k_pcg_direction(int n, PcgScalars *__restrict__ s, const double *__restrict__ z,
                                double *__restrict__ p, double *__restrict__ beta_hist, int iter) {
    const double beta = s->zr / s->temp1;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        p[i] = 1.0 * ld_nt(z + i) + beta * ld_nt(p + i);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        s->beta = beta;
        if (beta_hist) beta_hist[iter - 1] = beta;
    }
}

// Coarse scaling factor (mgcycle.c:337-341): alpha = min((x,b) / x^T A x, 2).
__global__ __launch_bounds__(kBlock) void k_coarse_scaling(int n, int nslices,
                                                           const long long *__restrict__ sptr,
                                                           const int *__restrict__ col,
                                                           const double *__restrict__ val,
                                                           const double *__restrict__ x,
                                                           const double *__restrict__ b,
                                                           double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    const int s = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = s * kSlice + lane;
    double xb = 0.0, xax = 0.0;
    if (s < nslices) {
        const long long base = sptr[s];
        const int wd = (int)((sptr[s + 1] - base) / kSlice);
        double acc = 0.0;
        for (int k = 0; k < wd; ++k) {
            const int cj = col[base + (long long)k * kSlice + lane];
            if (cj >= 0) acc = acc + val[base + (long long)k * kSlice + lane] * x[cj];
        }
        if (r < n) {
            xb = x[r] * b[r];
            xax = x[r] * acc;
        }
    }
    const int G = gridDim.x;
    double v = block_sum<kBlock>(xb, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = v;
    v = block_sum<kBlock>(xax, sh);
    if (threadIdx.x == 0) partials[G + blockIdx.x] = v;
}

__global__ __launch_bounds__(kFinBlock) void k_final_cs(const double *__restrict__ partials, int G,
                                                        double *__restrict__ alpha_out) {
    __shared__ double sh[kFinBlock / kWave];
    const double a = final_sum<kFinBlock>(partials, G, sh);
    const double c = final_sum<kFinBlock>(partials + G, G, sh);
    if (threadIdx.x == 0) {
        const double al = a / c;
        *alpha_out = al < 2.0 ? al : 2.0;
    }
}

inline int stream_grid(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > kStreamGrid) g = kStreamGrid;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

double g_alg_bytes = 0.0;


void sell_from_csr(const DevCSR &A, DevSELL &S) {
    S.n = A.n;
    S.nnz = A.nnz;
    S.nslices = div_up(A.n, kSlice);
    const int ns = S.nslices;
    long long *wcount = arena().alloc_n<long long>((size_t)ns + 1);
    S.sptr = arena().alloc_n<long long>((size_t)ns + 1);
    hipStream_t st = stream();
    if (ns > 0) {
        k_slice_width<<<div_up(ns, kSlicesPerBlock), kBlock, 0, st>>>(A.n, A.ia, wcount, ns);
        HZ_LAUNCH_CHECK();
    }
    // exclusive scan of slice sizes (padded entries)
    size_t tmp_bytes = 0;
    HZ_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, wcount, S.sptr, 0LL, (size_t)ns + 1,
                                     rocprim::plus<long long>(), st));
    HZ_CHECK(hipMemsetAsync(wcount + ns, 0, sizeof(long long), st));
    void *tmp = arena().alloc(tmp_bytes);
    HZ_CHECK(rocprim::exclusive_scan(tmp, tmp_bytes, wcount, S.sptr, 0LL, (size_t)ns + 1,
                                     rocprim::plus<long long>(), st));
    long long nent = 0;
    HZ_CHECK(hipMemcpyAsync(&nent, S.sptr + ns, sizeof(long long), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    S.nent = nent;
    S.col = arena().alloc_n<int>((size_t)nent + 1);
    S.val = arena().alloc_n<double>((size_t)nent + 1);
    S.dl1 = arena().alloc_n<double>((size_t)A.n + 1);
    S.djac = arena().alloc_n<double>((size_t)A.n + 1);
    S.dfirst = arena().alloc_n<double>((size_t)A.n + 1);
    if (ns > 0) {
        k_sell_fill<<<div_up(ns, kSlicesPerBlock), kBlock, 0, st>>>(A.n, A.ia, A.ja, A.val, S.sptr, S.col, S.val,
                                                                  ns, S.dl1, S.djac, S.dfirst);
        HZ_LAUNCH_CHECK();
    }
}

// matrix stream of an FP64 SELL launch: 2-byte columns where it reads the
// paired offsets
static double sell_bytes(const DevSELL &S, bool paired) {
    return paired ? (double)S.nnz * 10 + 4.0 * (S.n + 1) : matrix_bytes(S.nnz, S.n);
}

void spmv(const DevSELL &S, const double *x, double *y, double *dot_out, const ReduceWs &ws,
          LaunchTimer *timer, bool cols_only) {
    if (S.n == 0) return;
    // the paired offsets in the PCG's product, unless it leaves the int32
    // columns in the last-level cache for an FP32 cycle
    const bool paired = dot_out && S.pcol && !cols_only;
    count_bytes(sell_bytes(S, paired) + 16.0 * S.n);
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    if (dot_out) {
        if (timer) timer->begin();
        // the PCG's SpMV: rounds of 16, the second half of each with ordinary
        // loads that stay in the last-level cache for the cycle's residual
        // (ahead of an FP32 cycle: every column index, no value)
        if (paired)
            k_spmv<true, true, 1, unsigned int><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.pcol,
                                                                                  S.val, x, y, ws.partials);
        else if (cols_only)
            k_spmv<true, true, 4><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, x, y,
                                                                    ws.partials);
        else
            k_spmv<true, true, 1><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, x, y,
                                                                    ws.partials);
        if (timer) timer->end();
        HZ_LAUNCH_CHECK();
        k_final_sum<<<1, kFinBlock, 0, stream()>>>(ws.partials, grid, dot_out);
    } else {
        k_spmv<false><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, x, y,
                                                     nullptr);
    }
    HZ_LAUNCH_CHECK();
}

void residual(const DevSELL &S, const double *b, const double *x, double *w, LaunchTimer *timer) {
    if (S.n == 0) return;
    count_bytes(sell_bytes(S, S.pcol) + 24.0 * S.n);
    // rounds of 16 masked loads: a 15-entry row in one round of loads
    if (timer) timer->begin();
    const int g = div_up(S.nslices, kSlicesPerBlock);
    // a matrix that fits the 256 MB last-level cache with room to spare is
    // loaded with ordinary loads: the same level's post-sweep, after the
    // coarser levels' work, still finds part of it there
    const bool resident = matrix_bytes(S.nnz, S.n) <= kCacheResidentBytes;
    if (S.pcol && resident)
        k_residual<true, 2, unsigned int><<<g, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.pcol, S.val, b, x, w);
    else if (S.pcol)
        k_residual<true, 0, unsigned int><<<g, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.pcol, S.val, b, x, w);
    else if (resident)
        k_residual<true, 2><<<g, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, b, x, w);
    else
        k_residual<true><<<g, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, b, x, w);
    if (timer) timer->end();
    HZ_LAUNCH_CHECK();
}

namespace {
template <bool kProlong, bool kDot>
void launch_sweep(const DevSELL &S, int smoother, double w, const int *agg, const double *xc,
                  double alpha, const double *alpha_dev, const double *b, const double *x, double *xout,
                  double *partials) {
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    const bool wide = !kProlong;   // rounds of 16 masked loads (the prolongation form has them anyway)
    const bool paired = wide && S.pcol;
    count_bytes(sell_bytes(S, paired) + 24.0 * S.n + (kProlong ? 4.0 * S.n : 0.0));
    if (paired) {
        if (smoother == HZ_SMOOTHER_L1DIAG)
            k_sweep<HZ_SMOOTHER_L1DIAG, kProlong, kDot, true, unsigned int><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.pcol, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
        else
            k_sweep<HZ_SMOOTHER_JACOBI, kProlong, kDot, true, unsigned int><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.pcol, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
    } else if (smoother == HZ_SMOOTHER_L1DIAG) {
        if (wide)
            k_sweep<HZ_SMOOTHER_L1DIAG, kProlong, kDot, true><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
        else
            k_sweep<HZ_SMOOTHER_L1DIAG, kProlong, kDot><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
    } else {
        if (wide)
            k_sweep<HZ_SMOOTHER_JACOBI, kProlong, kDot, true><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
        else
            k_sweep<HZ_SMOOTHER_JACOBI, kProlong, kDot><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, S.val, w, agg, xc, alpha, alpha_dev, b, x, xout, partials);
    }
    HZ_LAUNCH_CHECK();
}
}  // namespace

void smooth_sweep(const DevSELL &S, int smoother, double w, const double *b, const double *x,
                  double *xout) {
    if (S.n == 0) return;
    launch_sweep<false, false>(S, smoother, w, nullptr, nullptr, 1.0, nullptr, b, x, xout, nullptr);
}

void smooth_sweep_dot(const DevSELL &S, int smoother, double w, const double *b, const double *x,
                      double *xout, double *dot_out, const ReduceWs &ws, LaunchTimer *timer) {
    if (S.n == 0) return;
    if (timer) timer->begin();
    launch_sweep<false, true>(S, smoother, w, nullptr, nullptr, 1.0, nullptr, b, x, xout, ws.partials);
    if (timer) timer->end();
    k_final_sum<<<1, kFinBlock, 0, stream()>>>(ws.partials, div_up(S.nslices, kSlicesPerBlock), dot_out);
    HZ_LAUNCH_CHECK();
}

void smoother_diag(const DevSELL &S, int smoother, double *d) {
    if (S.n == 0) return;
    if (S.dl1) {   // formed with the SELL copy
        HZ_CHECK(hipMemcpyAsync(d, smoother == HZ_SMOOTHER_L1DIAG ? S.dl1 : S.djac, sizeof(double) * S.n,
                                hipMemcpyDeviceToDevice, stream()));
        return;
    }
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    if (smoother == HZ_SMOOTHER_L1DIAG)
        k_smoother_diag<HZ_SMOOTHER_L1DIAG><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr,
                                                                           S.col, S.val, d);
    else
        k_smoother_diag<HZ_SMOOTHER_JACOBI><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr,
                                                                           S.col, S.val, d);
    HZ_LAUNCH_CHECK();
}

void smooth_from_zero(int n, int smoother, double w, const double *d, const double *b, double *x) {
    if (n == 0) return;
    count_bytes(24.0 * n);
    k_smooth_from_zero<<<stream_grid(n), kBlock, 0, stream()>>>(n, smoother, w, d, b, x);
    HZ_LAUNCH_CHECK();
}

void restrict_agg(int nc, int n, const int *rptr, const int *rmem, const double *w, double *bc,
                  int smoother, double wgt, const double *dc, double *xc) {
    if (nc == 0) return;
    count_bytes(12.0 * n + 4.0 * (nc + 1) + 8.0 * nc + (xc ? 16.0 * nc : 0.0));
    k_restrict<<<stream_grid(nc), kBlock, 0, stream()>>>(nc, rptr, rmem, w, bc, smoother, wgt, dc, xc);
    HZ_LAUNCH_CHECK();
}

void down_fused(int nc, const int *rptr, const int *rmem, const DevCSR &A, const double *b,
                const double *x, double *bc, int smoother, double wgt, const double *dc, double *xc) {
    if (nc == 0) return;
    count_bytes(matrix_bytes(A.nnz, A.n) + 16.0 * A.n + 4.0 * A.n + 4.0 * (nc + 1) + 8.0 * nc +
                (xc ? 16.0 * nc : 0.0));
    constexpr int kLanes = 4;
    k_down_fused<kLanes><<<div_up((long)nc * kLanes, kBlock), kBlock, 0, stream()>>>(
        nc, rptr, rmem, A.ia, A.ja, A.val, b, x, bc, smoother, wgt, dc, xc);
    HZ_LAUNCH_CHECK();
}

void up_fused(const DevSELL &S, int smoother, double w, const int *agg, const double *xc,
              double alpha, const double *alpha_dev, const double *b, const double *x, double *xout,
              double *dot_out, const ReduceWs &ws) {
    if (S.n == 0) return;
    if (dot_out) {
        launch_sweep<true, true>(S, smoother, w, agg, xc, alpha, alpha_dev, b, x, xout, ws.partials);
        k_final_sum<<<1, kFinBlock, 0, stream()>>>(ws.partials, div_up(S.nslices, kSlicesPerBlock), dot_out);
        HZ_LAUNCH_CHECK();
    } else {
        launch_sweep<true, false>(S, smoother, w, agg, xc, alpha, alpha_dev, b, x, xout, nullptr);
    }
}

void prolong_agg(int n, const int *agg, const double *xc, double alpha, const double *alpha_dev,
                 double *x) {
    if (n == 0) return;
    count_bytes(20.0 * n);
    k_prolong<<<stream_grid(n), kBlock, 0, stream()>>>(n, agg, xc, alpha, alpha_dev, x);
    HZ_LAUNCH_CHECK();
}

void fill(double *x, double v, long n) {
    if (n == 0) return;
    count_bytes(8.0 * n);
    k_fill<<<stream_grid(n), kBlock, 0, stream()>>>(x, v, n);
    HZ_LAUNCH_CHECK();
}

void copy(double *dst, const double *src, long n) {
    if (n == 0) return;
    count_bytes(16.0 * n);
    k_copy<<<stream_grid(n), kBlock, 0, stream()>>>(dst, src, n);
    HZ_LAUNCH_CHECK();
}

void pcg_update(int n, PcgScalars *s, const double *p, const double *t, double *u, double *r,
                double *alpha_hist, int iter, const ReduceWs &ws, const FirstSweep *first) {
    const int G = stream_grid(n);
    if (first) {
        count_bytes(64.0 * n);
        k_pcg_update<true><<<G, kBlock, 0, stream()>>>(n, s, p, t, u, r, ws.partials, first->smoother, first->w,
                                                       first->d, first->x0);
    } else {
        count_bytes(48.0 * n);
        k_pcg_update<false><<<G, kBlock, 0, stream()>>>(n, s, p, t, u, r, ws.partials, 0, 1.0, nullptr, nullptr);
    }
    HZ_LAUNCH_CHECK();
    k_final_update<<<1, kFinBlock, 0, stream()>>>(s, ws.partials, G, alpha_hist, iter);
    HZ_LAUNCH_CHECK();
}

void dot(int n, const double *x, const double *y, double *out, const ReduceWs &ws) {
    const int G = stream_grid(n);
    count_bytes(x == y ? 8.0 * n : 16.0 * n);
    k_dot<<<G, kBlock, 0, stream()>>>(n, x, y, ws.partials);
    HZ_LAUNCH_CHECK();
    k_final_sum<<<1, kFinBlock, 0, stream()>>>(ws.partials, G, out);
    HZ_LAUNCH_CHECK();
}

void norm2sq(int n, const double *x, double *out, const ReduceWs &ws) { dot(n, x, x, out, ws); }

void pcg_direction(int n, PcgScalars *s, const double *z, double *p, double *beta_hist, int iter) {
    count_bytes(24.0 * n);
    k_pcg_direction<<<stream_grid(n), kBlock, 0, stream()>>>(n, s, z, p, beta_hist, iter);
    HZ_LAUNCH_CHECK();
}

void coarse_scaling_alpha(const DevSELL &S, const double *x, const double *b, PcgScalars *,
                          double *alpha_out, const ReduceWs &ws) {
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    k_coarse_scaling<<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, x, b,
                                                    ws.partials);
    HZ_LAUNCH_CHECK();
    k_final_cs<<<1, kFinBlock, 0, stream()>>>(ws.partials, grid, alpha_out);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
