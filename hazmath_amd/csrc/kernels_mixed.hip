// kernels_mixed.hip -- the V-cycle in FP32 (matrix values FP32, or BF16 on
// coarse levels) under an FP64 PCG.
//
// The hierarchy is the FP64 one (same aggregates, same Galerkin products);
// only the values the cycle streams are stored again in low precision, sharing
// the SELL-64 column indices.  Vectors of the cycle are FP32 and every sum is
// accumulated in FP32.  The preconditioner's input r and output z stay FP64:
// the first kernel of a cycle rounds r to FP32 and the last post-smoothing
// sweep of level 0 writes z in FP64 together with the partials of (z, r), so
// the PCG sees the same interface as with the FP64 cycle.  The coarsest solve
// uses the FP64 inverse (FP32 in and out, FP64 accumulation).
#include <algorithm>
#include <cstdlib>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;
constexpr int kSlicesPerBlock = kBlock / kWave;
constexpr int kFinBlock = 1024;
constexpr int kStreamGrid = 2048;
constexpr float kSmallRealF = 1e-20f;

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(g, kStreamGrid));
}

// bf16 stored as the upper half of the FP32 bit pattern
// matrix values are streamed once per launch: non-temporal loads
__device__ __forceinline__ float ldv(const float *p) { return ld_nt(p); }
__device__ __forceinline__ float ldv(const unsigned short *p) {
    return __uint_as_float(((unsigned int)ld_nt(p)) << 16);
}

__device__ __forceinline__ unsigned short f2bf(float f) {
    const unsigned int u = __float_as_uint(f);
    const unsigned int r = u + 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (unsigned short)(r >> 16);
}

__global__ void k_vals_f32(long long n, const double *__restrict__ v, float *__restrict__ o) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        o[i] = (float)v[i];
}

__global__ void k_vals_bf16(long long n, const double *__restrict__ v, unsigned short *__restrict__ o) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        o[i] = f2bf((float)v[i]);
}

__global__ void k_d2f(long n, const double *__restrict__ a, float *__restrict__ b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

__global__ void k_f2d(long n, const float *__restrict__ a, double *__restrict__ b) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

__global__ void k_fill_f(float *__restrict__ x, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        x[i] = 0.0f;
}

__device__ __forceinline__ float zero_sweep(int smoother, float w, float b, float d) {
    if (fabsf(d) > kSmallRealF) return smoother == HZ_SMOOTHER_L1DIAG ? b / d : w * b / d;
    return 0.0f;
}

// b = fl(r); optionally x = first sweep from zero
// 16-bit column offsets of a SELL matrix in the paired form ld_cols reads
// (one wave per slice); any stored entry whose offset from its row needs more
// sets *misfit
__global__ __launch_bounds__(kBlock) void k_col16(int nslices, const long long *__restrict__ sptr,
                                                  const int *__restrict__ col, unsigned int *__restrict__ d,
                                                  int *__restrict__ misfit) {
    const int s = blockIdx.x * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    const long long base = sptr[s];
    const int wd = (int)((sptr[s + 1] - base) / kSlice);
    bool bad = false;
    for (int k = 0; k < wd; k += 2) {
        unsigned int w = 0;
        for (int h = 0; h < 2; ++h) {
            const int c = k + h < wd ? col[base + (long long)(k + h) * kSlice + lane] : -1;
            const int off = c - r;
            const bool fits = off >= -32767 && off <= 32767;
            bad = bad || (c >= 0 && !fits);
            const unsigned int half = c < 0 ? 0x8000u : (fits ? (unsigned int)(unsigned short)(short)off : 0u);
            w |= half << (16 * h);
        }
        d[col_base<unsigned int>(base, s) + (long long)(k / 2) * kSlice + lane] = w;
    }
    if (bad) *misfit = 1;
}

__global__ void k_enter(int n, const double *__restrict__ r, int smoother, float w, const float *__restrict__ d,
                        float *__restrict__ b, float *__restrict__ x, float c2, float *__restrict__ dcheb) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float bi = (float)r[i];
        b[i] = bi;
        if (dcheb) {   // Chebyshev's first step from zero (k_cheb_first_m's values)
            const float di = d[i];
            const float dn = 0.0f + (di != 0.0f ? c2 * bi / di : 0.0f);
            dcheb[i] = dn;
            x[i] = 0.0f + dn;
        } else if (x) {
            x[i] = zero_sweep(smoother, w, bi, d[i]);
        }
    }
}

// One sweep (Jacobi with weight w / L1DIAG) in FP32, optionally with the
// prolongation x + P xc folded into every x read, optionally writing the
// result in FP64 (out64) with block partials of sum_i out64_i r64_i.
template <int kSmoother, bool kProlong, bool kOut64, class MT, bool kKeepCol = false>
__global__ __launch_bounds__(kBlock) void k_sweep_m(int n, int nslices, const long long *__restrict__ sptr,
                                                    const int *__restrict__ col, const MT *__restrict__ val,
                                                    float wgt, const int *__restrict__ agg,
                                                    const float *__restrict__ xc, const float *__restrict__ b,
                                                    const float *__restrict__ x, float *__restrict__ xout,
                                                    double *__restrict__ out64, const double *__restrict__ r64,
                                                    double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    const int blk = xcd_block_id();
    const int s = blk * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = s * kSlice + lane;
    double prod = 0.0;
    if (s < nslices) {
        const long long base = sptr[s];
        const int wd = (int)((sptr[s + 1] - base) / kSlice);
        const int *c = col + base + lane;
        const MT *v = val + base + lane;
        float t = (r < n) ? b[r] : 0.0f;
        float d = 0.0f;
        // rounds of 16 entries masked past the slice width (a 15-entry row in one
        // round): all loads of a round issued before its math
        constexpr bool kMask = true;
        constexpr int U = 16;
        int k = 0;
        for (; kMask ? (k < wd) : (k + U <= wd); k += U) {
            int cc[U];
            float vv[U], xx[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const bool ok = !kMask || k + u < wd;
                cc[u] = ok ? (kKeepCol ? c[(k + u) * kSlice] : ld_nt(c + (k + u) * kSlice)) : -1;
                vv[u] = ok ? ldv(v + (k + u) * kSlice) : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0f;
            if (kProlong) {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int g = cc[u] >= 0 ? agg[cc[u]] : HZ_UNPT;
                    if (g > HZ_UNPT) xx[u] = xx[u] + xc[g];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (kSmoother == HZ_SMOOTHER_L1DIAG) {
                    if (cc[u] >= 0) {
                        t = t - vv[u] * xx[u];
                        d = d + fabsf(vv[u]);
                    }
                } else {
                    if (cc[u] == r) d = vv[u];
                    else if (cc[u] >= 0) t = t - vv[u] * xx[u];
                }
            }
        }
        for (; k < wd; ++k) {
            const int cj = ld_nt(c + k * kSlice);
            const float a = ldv(v + k * kSlice);
            float xj = cj >= 0 ? x[cj] : 0.0f;
            if (kProlong && cj >= 0) {
                const int g = agg[cj];
                if (g > HZ_UNPT) xj = xj + xc[g];
            }
            if (kSmoother == HZ_SMOOTHER_L1DIAG) {
                if (cj >= 0) {
                    t = t - a * xj;
                    d = d + fabsf(a);
                }
            } else {
                if (cj == r) d = a;
                else if (cj >= 0) t = t - a * xj;
            }
        }
        if (r < n) {
            float xi = x[r];
            if (kProlong) {
                const int g = agg[r];
                if (g > HZ_UNPT) xi = xi + xc[g];
            }
            float xn = xi;
            if (fabsf(d) > kSmallRealF) {
                if (kSmoother == HZ_SMOOTHER_L1DIAG) xn = xi + t / d;
                else xn = (1.0f - wgt) * xi + wgt * t / d;
            }
            if (kOut64) {
                const double z = (double)xn;
                out64[r] = z;
                prod = z * r64[r];
            } else {
                xout[r] = xn;
            }
        }
    }
    if (kOut64) {
        const double bs = block_sum<kBlock>(prod, sh);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

template <class MT, class CT>
__global__ __launch_bounds__(kBlock) void k_residual_m(int n, int nslices, const long long *__restrict__ sptr,
                                                       const CT *__restrict__ col, const MT *__restrict__ val,
                                                       const float *__restrict__ b, const float *__restrict__ x,
                                                       float *__restrict__ w) {
    const int s = xcd_block_id() * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    const long long base = sptr[s];
    const int wd = (int)((sptr[s + 1] - base) / kSlice);
    const CT *c = col + col_base<CT>(base, s) + lane;
    const MT *v = val + base + lane;
    float acc = 0.0f;
    constexpr int U = 16;   // rounds of 16 masked past the slice width
    for (int k = 0; k < wd; k += U) {
        int cc[U];
        float vv[U], xx[U];
        ld_cols<U>(c, k, wd, r, cc);
#pragma unroll
        for (int u = 0; u < U; ++u) vv[u] = k + u < wd ? ldv(v + (k + u) * kSlice) : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0f;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (cc[u] >= 0) acc = acc + vv[u] * xx[u];
    }
    if (r < n) w[r] = b[r] - acc;
}

__global__ void k_restrict_m(int nc, const int *__restrict__ rptr, const int *__restrict__ rmem,
                             const float *__restrict__ w, float *__restrict__ bc, int smoother, float wgt,
                             const float *__restrict__ dc, float *__restrict__ xc) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        float s = 0.0f;
        for (int q = rptr[c]; q < rptr[c + 1]; ++q) s = s + w[rmem[q]];
        bc[c] = s;
        if (xc) xc[c] = zero_sweep(smoother, wgt, s, dc[c]);
    }
}

// residual + restriction (+ coarse zero-iterate sweep), one thread per coarse
// row, for latency-bound levels (values read from the CSR copy).
// kLanes lanes share a coarse row (members j, j + kLanes, ...), member rows
// summed in CSR order with their loads issued 16 at a time, lane 0 adding the
// members' residuals in member order (as k_down_fused in FP64).
template <class MT, int kLanes>
__global__ void k_down_fused_m(int nc, const int *__restrict__ rptr, const int *__restrict__ rmem,
                               const int *__restrict__ ia, const int *__restrict__ ja, const MT *__restrict__ a,
                               const float *__restrict__ b, const float *__restrict__ x, float *__restrict__ bc,
                               int smoother, float wgt, const float *__restrict__ dc, float *__restrict__ xc) {
    constexpr int U = 16;
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const int lane = threadIdx.x % kWave, gbase = lane - j;
    const bool valid = c < nc;
    const int q0 = valid ? rptr[c] : 0, q1 = valid ? rptr[c + 1] : 0;
    float s = 0.0f;
    for (int qb = q0; qb < q1; qb += kLanes) {   // trip count uniform within the lane group
        const int q = qb + j;
        float wv = 0.0f;
        if (q < q1) {
            const int m = rmem[q];
            const int k0 = ia[m], k1 = ia[m + 1];
            const float bm = b[m];
            float t = 0.0f;
            for (int k = k0; k < k1; k += U) {
                int jj[U];
                float aa[U], xx[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const bool ok = k + u < k1;
                    jj[u] = ok ? ja[k + u] : -1;
                    aa[u] = ok ? ldv(a + k + u) : 0.0f;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xx[u] = jj[u] >= 0 ? x[jj[u]] : 0.0f;
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if (jj[u] >= 0) t = t + aa[u] * xx[u];
            }
            wv = bm - t;
        }
#pragma unroll
        for (int u = 0; u < kLanes; ++u) {
            const float v = __shfl(wv, gbase + u, kWave);
            if (j == 0 && qb + u < q1) s = s + v;
        }
    }
    if (valid && j == 0) {
        bc[c] = s;
        if (xc) xc[c] = zero_sweep(smoother, wgt, s, dc[c]);
    }
}

__global__ void k_prolong_m(int n, const int *__restrict__ agg, const float *__restrict__ xc,
                            float *__restrict__ x) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int g = agg[i];
        if (g > HZ_UNPT) x[i] = x[i] + xc[g];
    }
}

// One Chebyshev step (the FP64 k_cheb_step's recurrence) in FP32:
// dnew = c1 d + c2 (b - A x) / diag, x + dnew to xout, or on the last step of
// the cycle to out64 (FP64) with block partials of sum_i out64_i r64_i.
template <bool kOut64, class MT, class CT = int>
__global__ __launch_bounds__(kBlock) void k_cheb_step_m(int n, int nslices, const long long *__restrict__ sptr,
                                                        const CT *__restrict__ col, const MT *__restrict__ val,
                                                        const float *__restrict__ diag, float c1, float c2,
                                                        const float *__restrict__ b, const float *__restrict__ x,
                                                        float *__restrict__ xout, float *__restrict__ d,
                                                        double *__restrict__ out64, const double *__restrict__ r64,
                                                        double *__restrict__ partials) {
    __shared__ double sh[kBlock / kWave];
    const int blk = xcd_block_id();
    const int s = blk * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int r = s * kSlice + lane;
    double prod = 0.0;
    if (s < nslices) {
        const long long base = sptr[s];
        const int wd = (int)((sptr[s + 1] - base) / kSlice);
        const CT *c = col + col_base<CT>(base, s) + lane;
        const MT *v = val + base + lane;
        float t = (r < n) ? b[r] : 0.0f;
        constexpr int U = 16;   // rounds of 16 masked past the slice width
        for (int k = 0; k < wd; k += U) {
            int cc[U];
            float vv[U], xx[U];
            // the last step's column loads stay in the last-level cache for
            // the PCG SpMV that follows (it shares them)
            ld_cols<U, kOut64 ? 0 : U>(c, k, wd, r, cc);
#pragma unroll
            for (int u = 0; u < U; ++u) vv[u] = k + u < wd ? ldv(v + (k + u) * kSlice) : 0.0f;
#pragma unroll
            for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0f;
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (cc[u] >= 0) t = t - vv[u] * xx[u];
        }
        if (r < n) {
            const float di = diag[r];
            const float dn = (c1 == 0.0f ? 0.0f : c1 * d[r]) + (di != 0.0f ? c2 * t / di : 0.0f);
            d[r] = dn;
            const float xn = x[r] + dn;
            if (kOut64) {
                const double z = (double)xn;
                out64[r] = z;
                prod = z * r64[r];
            } else {
                xout[r] = xn;
            }
        }
    }
    if (kOut64) {
        const double bs = block_sum<kBlock>(prod, sh);
        if (threadIdx.x == 0) partials[blk] = bs;
    }
}

// Row r's sum_k a_rk x_k over the CSR copy with kLanes lanes per row: lane j
// takes entries j, j + kLanes, ...; each pass issues the column and value
// loads of up to kPass of them, then their gathers, then adds them in that
// order; the lanes' sums are combined by a butterfly (every lane gets it).
template <int kLanes, class MT>
__device__ inline float grp_row_dot(int k0, int k1, int j, const int *__restrict__ ja, const MT *__restrict__ a,
                                    const float *__restrict__ x) {
    constexpr int kPass = 64 / kLanes;
    float acc = 0.0f;
    for (int kb = k0 + j; kb < k1; kb += kLanes * kPass) {
        float av[kPass];
        int cj[kPass];
#pragma unroll
        for (int u = 0; u < kPass; ++u) {
            const int k = kb + u * kLanes;
            av[u] = k < k1 ? ldv(a + k) : 0.0f;
            cj[u] = k < k1 ? ja[k] : -1;
        }
#pragma unroll
        for (int u = 0; u < kPass; ++u)
            if (cj[u] >= 0) acc = acc + av[u] * x[cj[u]];
    }
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    return acc;
}

// w = b - A x over the CSR copy with kLanes lanes per row (small levels with
// long rows, where one lane per row is latency-bound).
template <int kLanes, class MT>
__global__ void k_residual_grp_m(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                 const MT *__restrict__ a, const float *__restrict__ b, const float *__restrict__ x,
                                 float *__restrict__ w) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const bool valid = r < n;
    const int k0 = valid ? ia[r] : 0, k1 = valid ? ia[r + 1] : 0;
    const float acc = grp_row_dot<kLanes>(k0, k1, j, ja, a, x);
    if (valid && j == 0) w[r] = b[r] - acc;
}

// The first Chebyshev step from x = 0: b - A x is b, so d = 0 + c2 b / diag
// and xout = 0 + d, the values the full step stores when it reads x = 0.
__global__ void k_cheb_first_m(int n, const float *__restrict__ diag, float c2, const float *__restrict__ b,
                               float *__restrict__ xout, float *__restrict__ d) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const float di = diag[r];
        const float dn = 0.0f + (di != 0.0f ? c2 * b[r] / di : 0.0f);
        d[r] = dn;
        xout[r] = 0.0f + dn;
    }
}

// The Chebyshev step over the CSR copy with kLanes lanes per row (small levels
// with long rows, latency-bound).
template <int kLanes, class MT>
__global__ void k_cheb_step_grp_m(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                  const MT *__restrict__ a, const float *__restrict__ diag, float c1, float c2,
                                  const float *__restrict__ b, const float *__restrict__ x, float *__restrict__ xout,
                                  float *__restrict__ d) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const bool valid = r < n;
    const int k0 = valid ? ia[r] : 0, k1 = valid ? ia[r + 1] : 0;
    const float acc = grp_row_dot<kLanes>(k0, k1, j, ja, a, x);
    if (valid && j == 0) {
        const float t = b[r] - acc;
        const float di = diag[r];
        const float dn = (c1 == 0.0f ? 0.0f : c1 * d[r]) + (di != 0.0f ? c2 * t / di : 0.0f);
        d[r] = dn;
        xout[r] = x[r] + dn;
    }
}

// y = A x in FP32 over a CSR matrix with FP32 values (smoothed aggregation's
// restriction with R), kLanes lanes per row combined by a butterfly.  Lane j
// takes entries j, j + kLanes, ... of its row; each pass issues the column and
// value loads of up to kPass of them, then their gathers, then adds them in
// that order.
template <int kLanes>
__global__ void k_csr_mxv_m(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const float *__restrict__ a, const float *__restrict__ x, float *__restrict__ y,
                            const float *__restrict__ diag, float c2, float *__restrict__ xfirst,
                            float *__restrict__ dfirst) {
    constexpr int kPass = 4;
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const bool valid = r < n;
    const int k0 = valid ? ia[r] : 0, k1 = valid ? ia[r + 1] : 0;
    float acc = 0.0f;
    for (int kb = k0 + j; kb < k1; kb += kLanes * kPass) {
        float av[kPass];
        int cj[kPass];
#pragma unroll
        for (int u = 0; u < kPass; ++u) {
            const int k = kb + u * kLanes;
            av[u] = k < k1 ? ld_nt(a + k) : 0.0f;
            cj[u] = k < k1 ? ja[k] : -1;
        }
#pragma unroll
        for (int u = 0; u < kPass; ++u)
            if (cj[u] >= 0) acc = acc + av[u] * x[cj[u]];
    }
#pragma unroll
    for (int o = kLanes / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, kWave);
    if (valid && j == 0) {
        y[r] = acc;
        if (xfirst) {   // the coarse level's first Chebyshev step from zero (k_cheb_first_m's values)
            const float di = diag[r];
            const float dn = 0.0f + (di != 0.0f ? c2 * acc / di : 0.0f);
            dfirst[r] = dn;
            xfirst[r] = 0.0f + dn;
        }
    }
}

// x += P xc in FP32 (smoothed aggregation's prolongation), one thread per row;
// each pass issues the loads of up to 8 entries, then their gathers, then adds
// them in CSR order.
__global__ void k_csr_prolong_m(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                const float *__restrict__ a, const float *__restrict__ xc, float *__restrict__ x) {
    constexpr int kPass = 8;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k0 = ia[i], k1 = ia[i + 1];
        float s = 0.0f;
        for (int kb = k0; kb < k1; kb += kPass) {
            float av[kPass];
            int cj[kPass];
#pragma unroll
            for (int u = 0; u < kPass; ++u) {
                av[u] = kb + u < k1 ? ld_nt(a + kb + u) : 0.0f;
                cj[u] = kb + u < k1 ? ja[kb + u] : -1;
            }
#pragma unroll
            for (int u = 0; u < kPass; ++u)
                if (cj[u] >= 0) s = s + av[u] * xc[cj[u]];
        }
        x[i] = x[i] + s;
    }
}

// x = Ainv b with FP64 Ainv and accumulation, FP32 b and x; a wavefront per row
__global__ __launch_bounds__(256) void k_gemv_m(int n, const double *__restrict__ Ainv,
                                                const float *__restrict__ b, float *__restrict__ x) {
    const int row = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= n) return;
    double s = 0.0;
    const double *a = Ainv + (size_t)row * n;
    constexpr int U = 16;   // a lane's loads issued together: entries lane, lane + 64, ... in that order
    for (int kb = lane; kb < n; kb += kWave * U) {
        double av[U], bv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kb + u * kWave;
            av[u] = k < n ? a[k] : 0.0;
            bv[u] = k < n ? (double)b[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (kb + u * kWave < n) s += av[u] * bv[u];
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) x[row] = (float)s;
}

__global__ __launch_bounds__(kFinBlock) void k_final_sum_m(const double *__restrict__ partials, int G,
                                                           double *__restrict__ out) {
    __shared__ double sh[kFinBlock / kWave];
    const double t = final_sum<kFinBlock>(partials, G, sh);
    if (threadIdx.x == 0) *out = t;
}

template <int kSm, bool kPro, bool kOut, class MT>
void launch_sweep_m(const DevSELL &S, const MT *val, float w, const int *agg, const float *xc, const float *b,
                    const float *x, float *xout, double *out64, const double *r64, double *partials) {
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    // the last level-0 sweep (z in FP64) precedes the FP64 PCG SpMV, which
    // shares the column indices: they are loaded with ordinary loads and stay
    // in the last-level cache for it (PCG SpMV 81 -> 69 us, FP32-cycle PCG
    // 20.8 -> 20.4 ms at 128^3)
    k_sweep_m<kSm, kPro, kOut, MT, kOut><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, val, w, agg,
                                                                         xc, b, x, xout, out64, r64, partials);
    HZ_LAUNCH_CHECK();
}

template <int kSm, class MT>
void dispatch_sweep_m(const DevSELL &S, const MT *val, float w, const int *agg, const float *xc,
                      const float *b, const float *x, float *xout, double *out64, const double *r64,
                      double *partials) {
    if (agg) {
        if (out64) launch_sweep_m<kSm, true, true>(S, val, w, agg, xc, b, x, xout, out64, r64, partials);
        else launch_sweep_m<kSm, true, false>(S, val, w, agg, xc, b, x, xout, out64, r64, partials);
    } else {
        if (out64) launch_sweep_m<kSm, false, true>(S, val, w, agg, xc, b, x, xout, out64, r64, partials);
        else launch_sweep_m<kSm, false, false>(S, val, w, agg, xc, b, x, xout, out64, r64, partials);
    }
}

}  // namespace

// Small levels with long rows (at most 65,536 rows, 16 or more entries a row on
// average) run their FP32 residual and Chebyshev steps over the CSR copy with
// 16 lanes a row.  On larger levels the SELL kernels are faster: on the 174,080-row
// second level of the smoothed-aggregation hierarchy of the 128^3 anisotropic
// problem (38 entries a row) the 16-lane form took the solve from 85.3 to 93.1 ms.
static bool grp_level(const DevSELL &S, const DevCSR *A) {
    return A && A->ia && A->n == S.n && S.n <= 65536 && (long long)A->nnz >= 16LL * A->n;
}

// matrix stream of a cycle kernel on the low-precision values: 2-byte column
// offsets where the level has them
static double low_matrix_bytes(const DevSELL &S, const LowMatrix &M) {
    const int vb = M.kind == kLowF32 ? 4 : 2;
    return M.dcol ? (double)S.nnz * (vb + 2) + 4.0 * (S.n + 1) : matrix_bytes(S.nnz, S.n, vb);
}

void low_values(const DevSELL &S, const DevCSR &A, int kind, LowMatrix &out, bool col16) {
    out.kind = kind;
    const long long ne = S.nent, nz = A.nnz;
    const int gs = (int)std::max(1LL, std::min(8192LL, (ne + kBlock - 1) / kBlock));
    const int gc = (int)std::max(1LL, std::min(8192LL, (nz + kBlock - 1) / kBlock));
    if (kind == kLowF32) {
        float *s = arena().alloc_n<float>((size_t)ne + 1);
        float *c = arena().alloc_n<float>((size_t)nz + 1);
        if (ne) k_vals_f32<<<gs, kBlock, 0, stream()>>>(ne, S.val, s);
        if (nz) k_vals_f32<<<gc, kBlock, 0, stream()>>>(nz, A.val, c);
        out.sell = s;
        out.csr = c;
    } else {
        unsigned short *s = arena().alloc_n<unsigned short>((size_t)ne + 1);
        unsigned short *c = arena().alloc_n<unsigned short>((size_t)nz + 1);
        if (ne) k_vals_bf16<<<gs, kBlock, 0, stream()>>>(ne, S.val, s);
        if (nz) k_vals_bf16<<<gc, kBlock, 0, stream()>>>(nz, A.val, c);
        out.sell = s;
        out.csr = c;
    }
    HZ_LAUNCH_CHECK();
    // col16: the Chebyshev steps and residuals of the cycle read 2-byte column
    // offsets instead of the int32 columns where every offset fits (learned
    // FP32 solve at 128^3 anisotropic 67.3 -> 62.0 ms; HZ_COL16=0 keeps the
    // int32 ones).  The Jacobi / L1 cycle reads its level-0 columns from the
    // last-level cache and its smaller levels through the fused restriction:
    // building them there cost 0.8 ms of setup and saved nothing.
    out.dcol = nullptr;
    if (!col16 || !col16_enabled() || ne == 0 || grp_level(S, &A)) return;
    int *misfit = arena().alloc_n<int>(1);
    unsigned int *d = col16_launch(S, misfit);
    int h = 1;
    HZ_CHECK(hipMemcpyAsync(&h, misfit, sizeof(int), hipMemcpyDeviceToHost, stream()));
    HZ_CHECK(hipStreamSynchronize(stream()));
    if (h == 0) out.dcol = d;
}

bool col16_enabled() {
    const char *e = std::getenv("HZ_COL16");
    return !(e && std::atoi(e) == 0);
}

unsigned int *col16_launch(const DevSELL &S, int *misfit) {
    unsigned int *d = arena().alloc_n<unsigned int>((size_t)(S.nent / 2 + 32LL * S.nslices) + 1);
    HZ_CHECK(hipMemsetAsync(misfit, 0, sizeof(int), stream()));
    k_col16<<<div_up(S.nslices, kSlicesPerBlock), kBlock, 0, stream()>>>(S.nslices, S.sptr, S.col, d, misfit);
    HZ_LAUNCH_CHECK();
    return d;
}

void d2f(const double *a, float *b, long n) {
    if (n <= 0) return;
    k_d2f<<<grid_for(n), kBlock, 0, stream()>>>(n, a, b);
    HZ_LAUNCH_CHECK();
}

void f2d(const float *a, double *b, long n) {
    if (n <= 0) return;
    k_f2d<<<grid_for(n), kBlock, 0, stream()>>>(n, a, b);
    HZ_LAUNCH_CHECK();
}

void fill_f(float *x, long n) {
    if (n <= 0) return;
    k_fill_f<<<grid_for(n), kBlock, 0, stream()>>>(x, n);
    HZ_LAUNCH_CHECK();
}

void enter_m(int n, const double *r, int smoother, float w, const float *d, float *b, float *x, float c2,
             float *dcheb) {
    if (n <= 0) return;
    count_bytes(8.0 * n + 4.0 * n + (x ? 8.0 * n : 0.0));
    k_enter<<<grid_for(n), kBlock, 0, stream()>>>(n, r, smoother, w, d, b, x, c2, dcheb);
    HZ_LAUNCH_CHECK();
}

void sweep_m(const DevSELL &S, const LowMatrix &M, int smoother, float w, const int *agg, const float *xc,
             const float *b, const float *x, float *xout, double *out64, const double *r64, double *dot_out,
             const ReduceWs &ws) {
    if (S.n == 0) return;
    count_bytes(matrix_bytes(S.nnz, S.n, M.kind == kLowF32 ? 4 : 2) + 12.0 * S.n + (agg ? 4.0 * S.n : 0.0) +
                (out64 ? 12.0 * S.n : 0.0));
    double *partials = out64 ? ws.partials : nullptr;
    if (M.kind == kLowF32) {
        const float *v = static_cast<const float *>(M.sell);
        if (smoother == HZ_SMOOTHER_L1DIAG)
            dispatch_sweep_m<HZ_SMOOTHER_L1DIAG>(S, v, w, agg, xc, b, x, xout, out64, r64, partials);
        else
            dispatch_sweep_m<HZ_SMOOTHER_JACOBI>(S, v, w, agg, xc, b, x, xout, out64, r64, partials);
    } else {
        const unsigned short *v = static_cast<const unsigned short *>(M.sell);
        if (smoother == HZ_SMOOTHER_L1DIAG)
            dispatch_sweep_m<HZ_SMOOTHER_L1DIAG>(S, v, w, agg, xc, b, x, xout, out64, r64, partials);
        else
            dispatch_sweep_m<HZ_SMOOTHER_JACOBI>(S, v, w, agg, xc, b, x, xout, out64, r64, partials);
    }
    if (out64 && dot_out) {
        k_final_sum_m<<<1, kFinBlock, 0, stream()>>>(ws.partials, div_up(S.nslices, kSlicesPerBlock), dot_out);
        HZ_LAUNCH_CHECK();
    }
}

void residual_m(const DevSELL &S, const LowMatrix &M, const float *b, const float *x, float *w, const DevCSR *A) {
    if (S.n == 0) return;
    const unsigned int *dcol = M.dcol;
    count_bytes(low_matrix_bytes(S, M) + 12.0 * S.n);
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    if (grp_level(S, A)) {
        constexpr int G = 16;
        const int g = div_up((long)S.n * G, kBlock);
        if (M.kind == kLowF32)
            k_residual_grp_m<G, float><<<g, kBlock, 0, stream()>>>(S.n, A->ia, A->ja,
                                                                   static_cast<const float *>(M.csr), b, x, w);
        else
            k_residual_grp_m<G, unsigned short><<<g, kBlock, 0, stream()>>>(
                S.n, A->ia, A->ja, static_cast<const unsigned short *>(M.csr), b, x, w);
    } else if (M.kind == kLowF32) {
        const float *v = static_cast<const float *>(M.sell);
        if (dcol)
            k_residual_m<float, unsigned int><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, dcol, v,
                                                                             b, x, w);
        else
            k_residual_m<float, int><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, v, b, x,
                                                                    w);
    } else {
        const unsigned short *v = static_cast<const unsigned short *>(M.sell);
        if (dcol)
            k_residual_m<unsigned short, unsigned int><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, dcol, v,
                                                                                      b, x, w);
        else
            k_residual_m<unsigned short, int><<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col,
                                                                              v, b, x, w);
    }
    HZ_LAUNCH_CHECK();
}

void restrict_m(int nc, int n, const int *rptr, const int *rmem, const float *w, float *bc, int smoother, float wgt,
                const float *dc, float *xc) {
    if (nc <= 0) return;
    count_bytes(8.0 * n + 4.0 * (nc + 1) + 4.0 * nc + (xc ? 8.0 * nc : 0.0));
    k_restrict_m<<<grid_for(nc), kBlock, 0, stream()>>>(nc, rptr, rmem, w, bc, smoother, wgt, dc, xc);
    HZ_LAUNCH_CHECK();
}

void down_fused_m(int nc, const int *rptr, const int *rmem, const DevCSR &A, const LowMatrix &M, const float *b,
                  const float *x, float *bc, int smoother, float wgt, const float *dc, float *xc) {
    if (nc <= 0) return;
    count_bytes(matrix_bytes(A.nnz, A.n, M.kind == kLowF32 ? 4 : 2) + 8.0 * A.n + 4.0 * A.n + 4.0 * (nc + 1) +
                4.0 * nc + (xc ? 8.0 * nc : 0.0));
    if (M.kind == kLowF32)
        k_down_fused_m<float, 4><<<div_up((long)nc * 4, kBlock), kBlock, 0, stream()>>>(
            nc, rptr, rmem, A.ia, A.ja, static_cast<const float *>(M.csr), b, x, bc, smoother, wgt, dc, xc);
    else
        k_down_fused_m<unsigned short, 4><<<div_up((long)nc * 4, kBlock), kBlock, 0, stream()>>>(
            nc, rptr, rmem, A.ia, A.ja, static_cast<const unsigned short *>(M.csr), b, x, bc, smoother, wgt, dc,
            xc);
    HZ_LAUNCH_CHECK();
}

void prolong_m(int n, const int *agg, const float *xc, float *x) {
    if (n <= 0) return;
    count_bytes(12.0 * n);
    k_prolong_m<<<grid_for(n), kBlock, 0, stream()>>>(n, agg, xc, x);
    HZ_LAUNCH_CHECK();
}

void cheby_apply_m(const DevSELL &S, const DevCSR &A, const LowMatrix &M, const ChebLevel &c, const float *diag,
                   int steps, const float *b, float *&x, float *&xt, float *dwork, double *out64, const double *r64,
                   double *dot_out, const ReduceWs &ws, bool x_zero, bool first_done) {
    if (S.n == 0 || steps <= 0) return;
    const double theta = 0.5 * (c.lmax + c.lmin), delta = 0.5 * (c.lmax - c.lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    constexpr int G = 16;
    const bool grp = grp_level(S, &A);
    for (int k = 0; k < steps; ++k) {
        double c1 = 0.0, c2 = 1.0 / theta;
        if (k > 0) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            c1 = rn * rho;
            c2 = 2.0 * rn / delta;
            rho = rn;
        }
        const bool last64 = out64 && k == steps - 1;
        if (k == 0 && first_done && !last64) continue;   // x and dwork hold the first step already
        if (k == 0 && x_zero && !last64) {
            count_bytes(16.0 * S.n);
            k_cheb_first_m<<<div_up(S.n, kBlock), kBlock, 0, stream()>>>(S.n, diag, (float)c2, b, xt, dwork);
            HZ_LAUNCH_CHECK();
            std::swap(x, xt);
            continue;
        }
        // the last step reads the columns the PCG's product that follows reads
        // (it finds them in the last-level cache): the offsets where the level
        // gave them to the FP64 matrix as well, else the int32 ones.  The first step to read the matrix after the PCG's
        // product took 42.4 us on the int32 columns that product leaves in the
        // last-level cache, 40.0 us on the offsets (128^3 level 0).
        const unsigned int *dcol = (last64 && !S.pcol) ? nullptr : M.dcol;
        count_bytes((dcol ? low_matrix_bytes(S, M) : matrix_bytes(S.nnz, S.n, M.kind == kLowF32 ? 4 : 2)) +
                    24.0 * S.n + (last64 ? 16.0 * S.n : 0.0));
        if (last64 && dcol) {
            if (M.kind == kLowF32)
                k_cheb_step_m<true, float, unsigned int><<<grid, kBlock, 0, stream()>>>(
                    S.n, S.nslices, S.sptr, dcol, static_cast<const float *>(M.sell), diag, (float)c1, (float)c2, b,
                    x, nullptr, dwork, out64, r64, ws.partials);
            else
                k_cheb_step_m<true, unsigned short, unsigned int><<<grid, kBlock, 0, stream()>>>(
                    S.n, S.nslices, S.sptr, dcol, static_cast<const unsigned short *>(M.sell), diag, (float)c1,
                    (float)c2, b, x, nullptr, dwork, out64, r64, ws.partials);
            HZ_LAUNCH_CHECK();
            if (dot_out) {
                k_final_sum_m<<<1, kFinBlock, 0, stream()>>>(ws.partials, grid, dot_out);
                HZ_LAUNCH_CHECK();
            }
            return;
        }
        if (last64) {
            if (M.kind == kLowF32)
                k_cheb_step_m<true, float><<<grid, kBlock, 0, stream()>>>(
                    S.n, S.nslices, S.sptr, S.col, static_cast<const float *>(M.sell), diag, (float)c1, (float)c2, b,
                    x, nullptr, dwork, out64, r64, ws.partials);
            else
                k_cheb_step_m<true, unsigned short><<<grid, kBlock, 0, stream()>>>(
                    S.n, S.nslices, S.sptr, S.col, static_cast<const unsigned short *>(M.sell), diag, (float)c1,
                    (float)c2, b, x, nullptr, dwork, out64, r64, ws.partials);
            HZ_LAUNCH_CHECK();
            if (dot_out) {
                k_final_sum_m<<<1, kFinBlock, 0, stream()>>>(ws.partials, grid, dot_out);
                HZ_LAUNCH_CHECK();
            }
            return;
        }
        if (grp) {
            if (M.kind == kLowF32)
                k_cheb_step_grp_m<G, float><<<div_up((long)S.n * G, kBlock), kBlock, 0, stream()>>>(
                    S.n, A.ia, A.ja, static_cast<const float *>(M.csr), diag, (float)c1, (float)c2, b, x, xt, dwork);
            else
                k_cheb_step_grp_m<G, unsigned short><<<div_up((long)S.n * G, kBlock), kBlock, 0, stream()>>>(
                    S.n, A.ia, A.ja, static_cast<const unsigned short *>(M.csr), diag, (float)c1, (float)c2, b, x, xt,
                    dwork);
        } else if (M.kind == kLowF32 && dcol) {
            k_cheb_step_m<false, float, unsigned int><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, dcol, static_cast<const float *>(M.sell), diag, (float)c1, (float)c2, b, x,
                xt, dwork, nullptr, nullptr, nullptr);
        } else if (M.kind == kLowF32) {
            k_cheb_step_m<false, float><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, static_cast<const float *>(M.sell), diag, (float)c1, (float)c2, b, x,
                xt, dwork, nullptr, nullptr, nullptr);
        } else if (dcol) {
            k_cheb_step_m<false, unsigned short, unsigned int><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, dcol, static_cast<const unsigned short *>(M.sell), diag, (float)c1,
                (float)c2, b, x, xt, dwork, nullptr, nullptr, nullptr);
        } else {
            k_cheb_step_m<false, unsigned short><<<grid, kBlock, 0, stream()>>>(
                S.n, S.nslices, S.sptr, S.col, static_cast<const unsigned short *>(M.sell), diag, (float)c1,
                (float)c2, b, x, xt, dwork, nullptr, nullptr, nullptr);
        }
        HZ_LAUNCH_CHECK();
        std::swap(x, xt);
    }
}

void csr_mxv_m(const DevCSR &A, const float *val, const float *x, float *y, const float *diag, float c2,
               float *xfirst, float *dfirst) {
    if (A.n <= 0) return;
    count_bytes(matrix_bytes(A.nnz, A.n, 4) + 4.0 * A.ncol + 4.0 * A.n + (xfirst ? 12.0 * A.n : 0.0));
    constexpr int G = 16;
    k_csr_mxv_m<G><<<div_up((long)A.n * G, kBlock), kBlock, 0, stream()>>>(A.n, A.ia, A.ja, val, x, y, diag, c2,
                                                                            xfirst, dfirst);
    HZ_LAUNCH_CHECK();
}

float cheb_first_c2(const ChebLevel &c) {
    const double theta = 0.5 * (c.lmax + c.lmin);
    return (float)(1.0 / theta);
}

void csr_prolong_m(const DevCSR &P, const float *val, const float *xc, float *x) {
    if (P.n <= 0) return;
    count_bytes(matrix_bytes(P.nnz, P.n, 4) + 4.0 * P.ncol + 8.0 * P.n);
    k_csr_prolong_m<<<div_up(P.n, kBlock), kBlock, 0, stream()>>>(P.n, P.ia, P.ja, val, xc, x);
    HZ_LAUNCH_CHECK();
}

float *values_f32(const DevCSR &A) {
    float *v = arena().alloc_n<float>((size_t)A.nnz + 1);
    if (A.nnz) {
        const int g = (int)std::max(1LL, std::min(8192LL, ((long long)A.nnz + kBlock - 1) / kBlock));
        k_vals_f32<<<g, kBlock, 0, stream()>>>(A.nnz, A.val, v);
        HZ_LAUNCH_CHECK();
    }
    return v;
}

void coarse_dense_solve_m(int n, const double *Ainv, const float *b, float *x) {
    if (n <= 0) return;
    count_bytes(8.0 * n * (double)n + 8.0 * n);
    k_gemv_m<<<div_up(n, 4), 256, 0, stream()>>>(n, Ainv, b, x);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
