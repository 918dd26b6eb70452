// kernels_learned.hip -- pieces of the learned-parameter AMG that are not in
// the reference: the Chebyshev smoother (HZ_SMOOTHER_CHEBY) and the
// per-aggregate features the parameter network reads.
//
// Chebyshev smoother: k steps of the Chebyshev iteration for A x = b with the
// Jacobi preconditioner D = diag(A) on the interval [lmin, lmax] of D^-1 A
// (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 12.1):
//   theta = (lmax+lmin)/2, delta = (lmax-lmin)/2, sigma = theta/delta,
//   rho_0 = 1/sigma, d_0 = D^-1 r_0 / theta,
//   rho_k = 1/(2 sigma - rho_{k-1}), d_k = rho_k rho_{k-1} d_{k-1} + (2 rho_k/delta) D^-1 r_k,
//   x_{k+1} = x_k + d_k,  r = b - A x recomputed every step.
// lmax = scale * (Gershgorin bound max_i sum_j |a_ij| / |a_ii|), lmin = lmax / 30.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;
constexpr int kSlicesPerBlock = kBlock / kWave;

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(g, 4096));
}

// d_i = first diagonal entry; partials of max_i sum_j |a_ij| / |d_i|
__global__ __launch_bounds__(kBlock) void k_cheb_diag(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const double *__restrict__ a, double *__restrict__ d,
                                                      double *__restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    constexpr int U = 16;   // entries whose loads are issued together
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double di = 0.0, s = 0.0;
        bool found = false;
        const int k1 = ia[i + 1];
        for (int kb = ia[i]; kb < k1; kb += U) {
            int jj[U];
            double vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                jj[u] = kb + u < k1 ? ja[kb + u] : -1;
                vv[u] = kb + u < k1 ? a[kb + u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (jj[u] < 0) continue;
                s += fabs(vv[u]);
                if (!found && jj[u] == i) {
                    di = vv[u];
                    found = true;
                }
            }
        }
        d[i] = di;
        if (di != 0.0) m = fmax(m, s / fabs(di));
    }
    const double bm = block_max<kBlock>(m, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bm;
}

__global__ void k_max_final(int G, const double *__restrict__ part, double *__restrict__ out) {
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int g = 0; g < G; ++g) m = fmax(m, part[g]);
        *out = m;
    }
}

// one Chebyshev step: xout = x + dnew, dnew = c1 d + c2 (b - A x) / diag
__global__ __launch_bounds__(kBlock) void k_cheb_step(int n, int nslices, const long long *__restrict__ sptr,
                                                      const int *__restrict__ col, const double *__restrict__ val,
                                                      const double *__restrict__ diag, double c1, double c2,
                                                      const double *__restrict__ b, const double *__restrict__ x,
                                                      double *__restrict__ xout, double *__restrict__ d) {
    const int s = xcd_block_id() * kSlicesPerBlock + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (s >= nslices) return;
    const int r = s * kSlice + lane;
    const long long base = sptr[s];
    const int wd = (int)((sptr[s + 1] - base) / kSlice);
    const int *c = col + base + lane;
    const double *v = val + base + lane;
    double t = (r < n) ? b[r] : 0.0;
    // rounds of 16 entries masked past the slice width: loads first, then the
    // products in stored order
    constexpr int U = 16;
    for (int k = 0; k < wd; k += U) {
        int cc[U];
        double vv[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool ok = k + u < wd;
            cc[u] = ok ? c[(k + u) * kSlice] : -1;
            vv[u] = ok ? v[(k + u) * kSlice] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = cc[u] >= 0 ? x[cc[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (cc[u] >= 0) t = t - vv[u] * xx[u];
    }
    if (r < n) {
        const double di = diag[r];
        const double dn = (c1 == 0.0 ? 0.0 : c1 * d[r]) + (di != 0.0 ? c2 * t / di : 0.0);
        d[r] = dn;
        xout[r] = x[r] + dn;
    }
}

// The first step from x = 0: b - A x is b, so d = 0 + c2 b / diag and
// xout = 0 + d, the values the full step stores when it reads x = 0 (the
// leading 0 + turns a -0 into +0 whatever the products were).
__global__ void k_cheb_first(int n, const double *__restrict__ diag, double c2, const double *__restrict__ b,
                             double *__restrict__ xout, double *__restrict__ d) {
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
        const double di = diag[r];
        const double dn = 0.0 + (di != 0.0 ? c2 * b[r] / di : 0.0);
        d[r] = dn;
        xout[r] = 0.0 + dn;
    }
}

// The same step over CSR with kLanes lanes per row, for small levels with long
// rows: each round the lanes load kLanes consecutive entries and form the
// products a*x, and the group's first lane subtracts them from t = b in CSR
// (= stored) order -- the same operations and order as one thread per row.
template <int kLanes>
__global__ void k_cheb_step_grp(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                const double *__restrict__ a, const double *__restrict__ diag, double c1, double c2,
                                const double *__restrict__ b, const double *__restrict__ x,
                                double *__restrict__ xout, double *__restrict__ d) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const int gbase = threadIdx.x % kWave - j;
    const bool valid = r < n;
    const int k0 = valid ? ia[r] : 0, k1 = valid ? ia[r + 1] : 0;
    double t = valid ? b[r] : 0.0;
    for (int kb = k0; kb < k1; kb += kLanes) {   // trip count uniform within the lane group
        const int k = kb + j;
        const double pr = k < k1 ? a[k] * x[ja[k]] : 0.0;
#pragma unroll
        for (int u = 0; u < kLanes; ++u) {
            const double v = __shfl(pr, gbase + u, kWave);
            if (j == 0 && kb + u < k1) t = t - v;
        }
    }
    if (valid && j == 0) {
        const double di = diag[r];
        const double dn = (c1 == 0.0 ? 0.0 : c1 * d[r]) + (di != 0.0 ? c2 * t / di : 0.0);
        d[r] = dn;
        xout[r] = x[r] + dn;
    }
}

// ---------------------------------------------------------------------------
// Per-aggregate features (12, FP32), averaged over the aggregate's members
// (R rows).  Per row i with a_ii the first diagonal entry, s = sum_{j!=i}|a_ij|,
// q = sum_{j!=i} a_ij, mx / mn the largest / smallest nonzero |a_ij| (j != i),
// pos = sum of positive off-diagonal entries, in = sum_{j!=i, agg j = agg i}|a_ij|:
//  0 log2(size)      1 mean s/a_ii     2 mean q/a_ii    3 mean mx/s
//  4 mean mn/mx      5 mean in/s       6 mean pos/s     7 mean row length / 16
//  8 level / 10      9 std of log10|a_ii|  10 max mx/s  11 min s/a_ii
// (ratios with a zero denominator count as 0).
// ---------------------------------------------------------------------------
__global__ void k_row_stats(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const double *__restrict__ a, const int *__restrict__ agg, double *__restrict__ rs) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        constexpr int U = 16;   // entries whose loads are issued together
        double di = 0.0, s = 0.0, q = 0.0, mx = 0.0, mn = 0.0, pos = 0.0, in = 0.0;
        bool found = false;
        const int gi = agg[i];
        const int k1 = ia[i + 1];
        for (int kb = ia[i]; kb < k1; kb += U) {
            int jj[U], gj[U];
            double vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                jj[u] = kb + u < k1 ? ja[kb + u] : -1;
                vv[u] = kb + u < k1 ? a[kb + u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) gj[u] = (jj[u] >= 0 && gi > HZ_UNPT) ? agg[jj[u]] : HZ_UNPT - 1;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int j = jj[u];
                if (j < 0) continue;
                const double v = vv[u];
                if (j == i) {
                    if (!found) {
                        di = v;
                        found = true;
                    }
                    continue;
                }
                const double av = fabs(v);
                s += av;
                q += v;
                if (v > 0.0) pos += v;
                if (av > 0.0) {
                    mx = fmax(mx, av);
                    mn = (mn == 0.0) ? av : fmin(mn, av);
                }
                if (gi > HZ_UNPT && gj[u] == gi) in += av;
            }
        }
        double *o = rs + (size_t)i * 9;
        const double ad = fabs(di);
        o[0] = ad > 0.0 ? s / ad : 0.0;
        o[1] = ad > 0.0 ? q / ad : 0.0;
        o[2] = s > 0.0 ? mx / s : 0.0;
        o[3] = mx > 0.0 ? mn / mx : 0.0;
        o[4] = s > 0.0 ? in / s : 0.0;
        o[5] = s > 0.0 ? pos / s : 0.0;
        o[6] = (double)(ia[i + 1] - ia[i]) / 16.0;
        o[7] = ad > 0.0 ? log10(ad) : 0.0;
        o[8] = 0.0;
    }
}

__global__ void k_agg_features(int nagg, const int *__restrict__ rptr, const int *__restrict__ rmem,
                               const double *__restrict__ rs, int level, float *__restrict__ feat) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nagg; c += gridDim.x * blockDim.x) {
        const int b = rptr[c], e = rptr[c + 1], m = e - b;
        double sum[7] = {0, 0, 0, 0, 0, 0, 0}, lg = 0.0, lg2 = 0.0, mxs = 0.0, mins = 0.0;
        for (int q = b; q < e; ++q) {
            const double *o = rs + (size_t)rmem[q] * 9;
            for (int t = 0; t < 7; ++t) sum[t] += o[t];
            lg += o[7];
            lg2 += o[7] * o[7];
            mxs = fmax(mxs, o[2]);
            mins = (q == b) ? o[0] : fmin(mins, o[0]);
        }
        float *f = feat + (size_t)c * HZ_MLP_INPUTS;
        const double inv = m > 0 ? 1.0 / m : 0.0;
        f[0] = (float)(m > 0 ? log2((double)m) : 0.0);
        for (int t = 0; t < 7; ++t) f[1 + t] = (float)(sum[t] * inv);
        f[8] = (float)(level / 10.0);
        const double mu = lg * inv;
        f[9] = (float)(m > 0 ? sqrt(fmax(lg2 * inv - mu * mu, 0.0)) : 0.0);
        f[10] = (float)mxs;
        f[11] = (float)mins;
    }
}

}  // namespace

ChebLevel cheby_prepare(const DevCSR &A, double scale) {
    ChebLevel c;
    c.diag = arena().alloc_n<double>((size_t)A.n + 1);
    const int g = grid_for(A.n);
    double *part = arena().alloc_n<double>((size_t)g + 1);
    double *rho = arena().alloc_n<double>(1);
    k_cheb_diag<<<g, kBlock, 0, stream()>>>(A.n, A.ia, A.ja, A.val, c.diag, part);
    k_max_final<<<1, 64, 0, stream()>>>(g, part, rho);
    HZ_LAUNCH_CHECK();
    double hr = 0.0;
    HZ_CHECK(memcpy_sync(&hr, rho, sizeof(double), hipMemcpyDeviceToHost));
    c.gersh = hr;
    c.lmax = scale * hr;
    c.lmin = c.lmax / 30.0;
    return c;
}

void cheby_apply(const DevSELL &S, const ChebLevel &c, int steps, const double *b, double *&x, double *&xt,
                 double *dwork, const DevCSR *A, bool x_zero) {
    if (S.n == 0 || steps <= 0) return;
    const double theta = 0.5 * (c.lmax + c.lmin), delta = 0.5 * (c.lmax - c.lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    const int grid = div_up(S.nslices, kSlicesPerBlock);
    // small levels with long rows (latency-bound): 16 lanes per CSR row
    constexpr int G = 16;
    const bool grp = A && A->ia && A->n == S.n && S.n <= 65536 && (long long)A->nnz >= 16LL * A->n;
    for (int k = 0; k < steps; ++k) {
        double c1 = 0.0, c2 = 1.0 / theta;
        if (k > 0) {
            const double rn = 1.0 / (2.0 * sigma - rho);
            c1 = rn * rho;
            c2 = 2.0 * rn / delta;
            rho = rn;
        }
        if (k == 0 && x_zero) {
            count_bytes(32.0 * S.n);
            k_cheb_first<<<div_up(S.n, kBlock), kBlock, 0, stream()>>>(S.n, c.diag, c2, b, xt, dwork);
            HZ_LAUNCH_CHECK();
            std::swap(x, xt);
            continue;
        }
        count_bytes(matrix_bytes(S.nnz, S.n) + 48.0 * S.n);
        if (grp)
            k_cheb_step_grp<G><<<div_up((long)S.n * G, kBlock), kBlock, 0, stream()>>>(S.n, A->ia, A->ja, A->val,
                                                                                       c.diag, c1, c2, b, x, xt,
                                                                                       dwork);
        else
            k_cheb_step<<<grid, kBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, c.diag, c1, c2, b, x,
                                                       xt, dwork);
        HZ_LAUNCH_CHECK();
        std::swap(x, xt);
    }
}

void aggregate_features(const DevCSR &A, const int *agg, int nagg, const int *rptr, const int *rmem, int level,
                        float *feat) {
    double *rs = arena().alloc_n<double>((size_t)A.n * 9 + 1);
    k_row_stats<<<div_up(A.n, kBlock), kBlock, 0, stream()>>>(A.n, A.ia, A.ja, A.val, agg, rs);
    k_agg_features<<<grid_for(nagg), kBlock, 0, stream()>>>(nagg, rptr, rmem, rs, level, feat);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz

namespace hz {

namespace {

__global__ __launch_bounds__(kBlock) void k_learned_params(int nagg, const float *__restrict__ out,
                                                           float *__restrict__ omega, double *__restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    double s = 0.0;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nagg; c += gridDim.x * blockDim.x) {
        omega[c] = fminf(fmaxf(out[2 * c], 0.0f), 1.6f);
        s += (double)out[2 * c + 1];
    }
    const double bs = block_sum<kBlock>(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

}  // namespace

namespace {

// Lanczos on B = D^-1/2 A D^-1/2 (similar to D^-1 A, symmetric for SPD A;
// d: first diagonal entries, rows with d <= 0 drop out).  The scalars of the
// recurrence stay on the device (alpha[k], beta[k]); the host reads them once.
// Every kernel runs a grid of at most kLzGrid blocks (grid-stride loops), so
// each reduction has at most kLzGrid partials.  4 lanes share a row of A, each
// issuing the loads of up to kLzPass of its entries (lane j: entries j, j + 4,
// ...) before their gathers, and combine their products by a butterfly: an
// estimate, so the summation order is free.  (16 lanes a row, one entry each,
// on at most 2,048 blocks: 175 us a step on the 128^3 anisotropic level.)
constexpr int kLzLanes = 4;
constexpr int kLzPass = 4;
constexpr int kLzGrid = 8192;

// v_k = s wprev (s = 1 / *norm; written out) and y = B v_k, block partials of
// (y, v_k); v_k's gathered entries are formed from wprev on the fly
__global__ __launch_bounds__(kBlock) void k_lz_apply(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                     const double *__restrict__ a, const double *__restrict__ isd,
                                                     const double *__restrict__ wprev,
                                                     const double *__restrict__ norm, double *__restrict__ v,
                                                     double *__restrict__ y, double *__restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    const double nv = *norm;
    const double sc = nv > 0.0 ? 1.0 / nv : 0.0;
    const int groups = gridDim.x * (kBlock / kLzLanes);
    const int g0 = blockIdx.x * (kBlock / kLzLanes) + threadIdx.x / kLzLanes;
    const int j = threadIdx.x % kLzLanes;
    double q = 0.0;
    for (int i = g0; i < n; i += groups) {   // trip count uniform within the lane group
        const int k0 = ia[i], k1 = ia[i + 1];
        double t = 0.0;
        for (int kb = k0 + j; kb < k1; kb += kLzLanes * kLzPass) {
            int cc[kLzPass];
            double av[kLzPass];
#pragma unroll
            for (int u = 0; u < kLzPass; ++u) {
                const int k = kb + u * kLzLanes;
                cc[u] = k < k1 ? ja[k] : -1;
                av[u] = k < k1 ? a[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kLzPass; ++u)
                if (cc[u] >= 0) t += av[u] * (isd[cc[u]] * (sc * wprev[cc[u]]));
        }
#pragma unroll
        for (int o = kLzLanes / 2; o > 0; o >>= 1) t += __shfl_xor(t, o, kWave);
        if (j == 0) {
            const double vi = sc * wprev[i];
            const double yi = isd[i] * t;
            v[i] = vi;
            y[i] = yi;
            q += yi * vi;
        }
    }
    const double bs = block_sum<kBlock>(q, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// y -= alpha_k v + beta_{k-1} vprev; block partials of |y|^2
__global__ __launch_bounds__(kBlock) void k_lz_orth(int n, const double *__restrict__ alpha,
                                                    const double *__restrict__ beta_prev,
                                                    const double *__restrict__ v, const double *__restrict__ vprev,
                                                    double *__restrict__ y, double *__restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    const double al = *alpha, be = beta_prev ? *beta_prev : 0.0;
    double q = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double yi = y[i] - al * v[i] - (beta_prev ? be * vprev[i] : 0.0);
        y[i] = yi;
        q += yi * yi;
    }
    const double bs = block_sum<kBlock>(q, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// the block partials summed in one block (fixed order); sqrt when asked
__global__ __launch_bounds__(kBlock) void k_sum_partials(const double *__restrict__ part, int g, int take_sqrt,
                                                         double *__restrict__ out) {
    __shared__ double sh[kBlock / kWave];
    double s = 0.0;
    for (int i = threadIdx.x; i < g; i += kBlock) s += part[i];
    const double t = block_sum<kBlock>(s, sh);
    if (threadIdx.x == 0) *out = take_sqrt ? sqrt(fmax(t, 0.0)) : t;
}

// 1/sqrt(d_i) (0 where d_i <= 0), a fixed start vector with every entry
// nonzero and no smooth structure, and block partials of its |.|^2
__global__ __launch_bounds__(kBlock) void k_lz_start(int n, const double *__restrict__ d, double *__restrict__ isd,
                                                     double *__restrict__ w, double *__restrict__ part) {
    __shared__ double sh[kBlock / kWave];
    double q = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        isd[i] = d[i] > 0.0 ? 1.0 / sqrt(d[i]) : 0.0;
        const unsigned h = (unsigned)i * 2654435761u;
        const double wi = 0.25 + (double)((h >> 8) & 0xffffu) / 65536.0;
        w[i] = wi;
        q += wi * wi;
    }
    const double bs = block_sum<kBlock>(q, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = bs;
}

// Number of eigenvalues of the symmetric tridiagonal (diag a, off-diagonal b)
// below x (Sturm sequence).
int sturm_below(int m, const double *a, const double *b, double x) {
    int c = 0;
    double q = a[0] - x;
    if (q < 0.0) ++c;
    for (int i = 1; i < m; ++i) {
        const double qq = (q != 0.0) ? q : 1e-300;
        q = a[i] - x - b[i - 1] * b[i - 1] / qq;
        if (q < 0.0) ++c;
    }
    return c;
}

}  // namespace

// Largest eigenvalue of the Lanczos tridiagonal (diagonal ha[0..m), off-diagonal
// hb[0..m-1)), cut at the first (near-)zero off-diagonal entry, where the
// Krylov space is invariant; *used: the order kept.
static double ritz_max(const double *ha, const double *hb, int m, int *used) {
    double scale = 0.0;
    for (int k = 0; k < m; ++k) scale = std::max(scale, std::fabs(ha[k]));
    for (int k = 0; k + 1 < m; ++k)
        if (hb[k] <= 1e-14 * std::max(scale, 1e-300)) {
            m = k + 1;
            break;
        }
    if (used) *used = m;
    // bisection on the Gershgorin interval of the tridiagonal
    double lo = ha[0], hi = ha[0];
    for (int i = 0; i < m; ++i) {
        const double r = (i > 0 ? std::fabs(hb[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(hb[i]) : 0.0);
        lo = std::min(lo, ha[i] - r);
        hi = std::max(hi, ha[i] + r);
    }
    for (int it = 0; it < 200 && hi - lo > 1e-15 * std::max(std::fabs(hi), 1e-300); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (sturm_below(m, ha, hb, mid) >= m) hi = mid;   // every eigenvalue below mid
        else lo = mid;
    }
    return hi;
}

namespace {
// The estimate's device state: the recurrence scalars alpha[0..steps),
// beta[0..steps) and |w0| at ab.
struct LanczosRun {
    double *ab = nullptr;
    int steps = 0;
};

// Queues the start vector and `steps` Lanczos steps on st.
LanczosRun lanczos_enqueue(const DevCSR &A, const double *d, int steps, hipStream_t st) {
    const int n = A.n;
    LanczosRun run;
    run.steps = steps;
    const int G = std::max(1, std::min(kLzGrid, div_up((long)n * kLzLanes, kBlock)));
    double *isd = arena().alloc_n<double>((size_t)n + 1);
    double *vb[2] = {arena().alloc_n<double>((size_t)n + 1), arena().alloc_n<double>((size_t)n + 1)};
    double *yb[2] = {arena().alloc_n<double>((size_t)n + 1), arena().alloc_n<double>((size_t)n + 1)};
    double *part = arena().alloc_n<double>((size_t)G + 1);
    run.ab = arena().alloc_n<double>(2 * (size_t)steps + 2);   // alpha[0..steps), beta[0..steps), |w0|
    double *alpha = run.ab, *beta = run.ab + steps, *nrm0 = run.ab + 2 * steps;
    // y[1] holds the start vector: step 0 reads it as "wprev" with norm |w0|
    k_lz_start<<<G, kBlock, 0, st>>>(n, d, isd, yb[1], part);
    k_sum_partials<<<1, kBlock, 0, st>>>(part, G, 1, nrm0);
    HZ_LAUNCH_CHECK();
    for (int k = 0; k < steps; ++k) {
        double *v = vb[k & 1], *vprev = vb[(k + 1) & 1];
        double *y = yb[k & 1];
        const double *wprev = yb[(k + 1) & 1];
        k_lz_apply<<<G, kBlock, 0, st>>>(n, A.ia, A.ja, A.val, isd, wprev, k > 0 ? beta + k - 1 : nrm0, v, y, part);
        k_sum_partials<<<1, kBlock, 0, st>>>(part, G, 0, alpha + k);
        k_lz_orth<<<G, kBlock, 0, st>>>(n, alpha + k, k > 0 ? beta + k - 1 : nullptr, v, vprev, y, part);
        k_sum_partials<<<1, kBlock, 0, st>>>(part, G, 1, beta + k);
        HZ_LAUNCH_CHECK();
    }
    return run;
}

// The largest Ritz value of a queued run (waits for st).
double lanczos_finish(const LanczosRun &run, hipStream_t st, int *steps_done) {
    const int steps = run.steps;
    std::vector<double> h(2 * (size_t)steps);
    HZ_CHECK(hipMemcpyAsync(h.data(), run.ab, sizeof(double) * 2 * steps, hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    int used = 0;
    const double theta = ritz_max(h.data(), h.data() + steps, steps, &used);
    if (steps_done) *steps_done = used;
    return theta;
}
}  // namespace

double lanczos_lambda_dinv(const DevCSR &A, const double *d, int steps, int *steps_done) {
    if (steps_done) *steps_done = 0;
    if (A.n == 0 || steps <= 0) return 0.0;
    return lanczos_finish(lanczos_enqueue(A, d, steps, stream()), stream(), steps_done);
}

hipStream_t side_stream() {
    static hipStream_t s = nullptr;
    if (!s) HZ_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));   // default priority
    return s;
}

void LanczosAsync::start(const DevCSR &A, const double *d, int steps) {
    static hipEvent_t ready = nullptr;
    if (!ready) HZ_CHECK(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    active = A.n > 0 && steps > 0;
    if (!active) return;
    // the side stream starts once the main stream has formed A and d
    HZ_CHECK(hipEventRecord(ready, stream()));
    HZ_CHECK(hipStreamWaitEvent(side_stream(), ready, 0));
    const LanczosRun run = lanczos_enqueue(A, d, steps, side_stream());
    ab = run.ab;
    nsteps = run.steps;
}

double LanczosAsync::finish() {
    if (!active) return 0.0;
    active = false;
    LanczosRun run;
    run.ab = ab;
    run.steps = nsteps;
    return lanczos_finish(run, side_stream(), nullptr);
}

void LanczosAsync::drain() {
    if (!active) return;
    active = false;
    HZ_CHECK(hipStreamSynchronize(side_stream()));
}

double learned_parameters(int nagg, const float *out, float *omega) {
    if (nagg <= 0) return 1.0;
    const int g = grid_for(nagg);
    double *part = arena().alloc_n<double>((size_t)g + 1);
    k_learned_params<<<g, kBlock, 0, stream()>>>(nagg, out, omega, part);
    HZ_LAUNCH_CHECK();
    std::vector<double> h(g);
    HZ_CHECK(memcpy_sync(h.data(), part, sizeof(double) * g, hipMemcpyDeviceToHost));
    double s = 0.0;
    for (double v : h) s += v;
    return std::min(1.5, std::max(0.2, s / nagg));
}

}  // namespace hz
