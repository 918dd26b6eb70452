// kernels_dense.hip -- coarsest-level solvers on gfx950.
//
// coarse_solver = 32 (SOLVER_UMFPACK): the reference factors the coarsest
// matrix once with its sparse U^T D U (src/solver/hazmath_direct.c:633-854)
// and solves with it every cycle.  Here the explicit inverse of the dense
// coarsest matrix is formed once by block Gauss-Jordan elimination (no
// pivoting, like the reference's factorization; rank-64 updates on the FP64
// matrix cores), and each cycle costs one dense matrix-vector product spread
// over many workgroups instead of two latency-bound triangular sweeps.
// Results equal the reference's to rounding.
//
// coarse_solver = 0: unpreconditioned CG to the cycle tolerance
// (coarse_itsolver, src/solver/mgcycle.c:34-53 -> dcsr_pcg, krylov.c:122-372),
// run inside one workgroup so the whole coarse solve is a single launch.
#include <algorithm>
#include <cmath>
#include <vector>

#include "hz_internal.h"

namespace hz {

namespace {

__global__ void k_dense_from_csr(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                 const double *__restrict__ val, double *__restrict__ M) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = ia[i]; k < ia[i + 1]; ++k) M[(size_t)i * n + ja[k]] += val[k];
}

// ---------------------------------------------------------------------------
// Blocked Gauss-Jordan inversion for larger coarsest levels.  Per 64-column
// block K (P = A_KK):  R = P^{-1} A_K,: ;  A_i,: -= A_i,K R (i not in K);
// A_i,K = -A_i,K P^{-1};  A_K,: = R with A_KK = P^{-1}.  The rank-64 updates
// are C = beta C + alpha A B products on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64): a 256-thread workgroup owns a 64x64 tile of C, each
// wavefront a 32x32 quarter as 2x2 MFMA 16x16 blocks, operands staged in LDS.
// ---------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int kTile = 64;

template <bool kMfma>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const double *__restrict__ A, int lda,
                                              const double *__restrict__ B, int ldb, double *__restrict__ C,
                                              int ldc, double alpha, double beta) {
    __shared__ double As[kTile][kTile + 1];
    __shared__ double Bs[kTile][kTile + 1];
    const int tm = blockIdx.y * kTile, tn = blockIdx.x * kTile;
    const int tid = threadIdx.x;
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        const int r = idx / kTile, k = idx % kTile;
        As[r][k] = (tm + r < M && k < K) ? A[(size_t)(tm + r) * lda + k] : 0.0;
        const int kk = idx / kTile, c = idx % kTile;
        Bs[kk][c] = (kk < K && tn + c < N) ? B[(size_t)kk * ldb + tn + c] : 0.0;
    }
    __syncthreads();
    const int w = tid / kWave, lane = tid % kWave;
    const int wm = (w / 2) * 32, wn = (w % 2) * 32;
    if (kMfma) {
        d4 acc[2][2];
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = d4{0.0, 0.0, 0.0, 0.0};
        // A operand: lane holds A[row = lane % 16][k = lane / 16]; B operand:
        // B[k = lane / 16][col = lane % 16]; accumulator element r of lane holds
        // D[row = lane / 16 + 4 r][col = lane % 16].
        for (int ks = 0; ks < kTile; ks += 4) {
            const int kq = ks + lane / 16;
            const double a0 = As[wm + 0 + lane % 16][kq], a1 = As[wm + 16 + lane % 16][kq];
            const double b0 = Bs[kq][wn + 0 + lane % 16], b1 = Bs[kq][wn + 16 + lane % 16];
            acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
            acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
            acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
            acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
        }
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni)
                for (int r = 0; r < 4; ++r) {
                    const int row = tm + wm + mi * 16 + lane / 16 + 4 * r;
                    const int col = tn + wn + ni * 16 + lane % 16;
                    if (row < M && col < N) {
                        double *c = &C[(size_t)row * ldc + col];
                        *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * acc[mi][ni][r];
                    }
                }
    } else {
        // plain FMA path: each thread 16 outputs of the tile
        for (int e = tid; e < kTile * kTile; e += 256) {
            const int r = e / kTile, cidx = e % kTile;
            double s = 0.0;
            for (int k = 0; k < kTile; ++k) s += As[r][k] * Bs[k][cidx];
            const int row = tm + r, col = tn + cidx;
            if (row < M && col < N) {
                double *c = &C[(size_t)row * ldc + col];
                *c = (beta == 0.0 ? 0.0 : beta * *c) + alpha * s;
            }
        }
    }
}

// In-place Gauss-Jordan inverse of the b x b block at A[k0][k0] (ld n) into P.
// 1024 threads: thread (i, t) owns row i, columns t, t+16, t+32, t+48.  Per
// pivot k: the column is saved, the pivot row scaled (s[k][k] = 1/piv), then
// the rank-1 update of the other rows together with column k.
__global__ __launch_bounds__(1024) void k_block_inverse(int n, int k0, int b, const double *__restrict__ A,
                                                        double *__restrict__ P, int *__restrict__ status) {
    static_assert(kTile == 64, "thread layout assumes 64 x 64 blocks");
    constexpr int kCols = 16;
    __shared__ double s[kTile][kTile + 1];
    __shared__ double colk[kTile];
    const int tid = threadIdx.x;
    const int ti = tid / kCols, tj = tid % kCols;
    if (ti < b)
        for (int q = 0; q < kTile / kCols; ++q) {
            const int j = tj + kCols * q;
            if (j < b) s[ti][j] = A[(size_t)(k0 + ti) * n + k0 + j];
        }
    __syncthreads();
    for (int k = 0; k < b; ++k) {
        const double piv = s[k][k];
        if (piv == 0.0) {   // the same value for every thread: all leave together
            if (tid == 0) status[0] = 1;
            return;
        }
        if (tid < b) colk[tid] = s[tid][k];
        __syncthreads();
        if (tid < b) s[k][tid] = tid == k ? 1.0 / piv : s[k][tid] / piv;
        __syncthreads();
        if (ti < b && ti != k) {
            const double ci = colk[ti];
            for (int q = 0; q < kTile / kCols; ++q) {
                const int j = tj + kCols * q;
                if (j < b && j != k) s[ti][j] -= ci * s[k][j];
            }
            if (tj == 0) s[ti][k] = -ci * s[k][k];
        }
        __syncthreads();
    }
    if (ti < b)
        for (int q = 0; q < kTile / kCols; ++q) {
            const int j = tj + kCols * q;
            if (j < b) P[ti * b + j] = s[ti][j];
        }
}

// One 64 x 64 product on the FP64 matrix cores, operands already in LDS:
// acc (per wavefront a 32 x 32 quarter as 2 x 2 blocks of 16 x 16) = X Y over
// k = 0..63 in steps of 4, the same order as k_gemm.
__device__ inline void tile_mfma(const double (*X)[kTile + 1], const double (*Y)[kTile + 1], int wm, int wn,
                                 int lane, d4 (&acc)[2][2]) {
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < 2; ++ni) acc[mi][ni] = d4{0.0, 0.0, 0.0, 0.0};
    for (int ks = 0; ks < kTile; ks += 4) {
        const int kq = ks + lane / 16;
        const double a0 = X[wm + 0 + lane % 16][kq], a1 = X[wm + 16 + lane % 16][kq];
        const double b0 = Y[kq][wn + 0 + lane % 16], b1 = Y[kq][wn + 16 + lane % 16];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
}

// The rest of a Gauss-Jordan panel step in one launch, after the pivot block
// K = [k0, k0 + b) has been inverted into Pinv (b x b).  Workgroup (tm, tn)
// owns a 64 x 64 tile:
//   tn = K:            Ck[tm rows] = -(A_tm,K Pinv)               (new column block)
//   tn != K, tm = K:   R[:, tn] = Pinv A_K,tn                     (new row block)
//   tn != K, tm != K:  A_tm,tn -= A_tm,K (Pinv A_K,tn)            (in place)
// Column block K and row block K are only read here (their new values go to
// Ck and R), so no workgroup reads what another writes.  Each value is formed
// as the separate products formed it: R = 0 + 1 (Pinv A_K,:), then
// C = 1 C + (-1) acc, Ck = 0 + (-1) acc.
template <bool kRReady>
__global__ __launch_bounds__(256) void k_gj_panel(int n, int k0, int b, double *__restrict__ M,
                                                  const double *__restrict__ Pinv, double *__restrict__ R,
                                                  double *__restrict__ Ck) {
    __shared__ double S1[kTile][kTile + 1];
    __shared__ double S2[kTile][kTile + 1];
    const int tm = blockIdx.y * kTile, tn = blockIdx.x * kTile;
    const bool colK = tn == k0, rowK = tm == k0;
    if (colK && rowK) return;                // the pivot block: Pinv itself
    if (kRReady && rowK) return;             // R was formed by its own product
    const int tid = threadIdx.x;
    const int w = tid / kWave, lane = tid % kWave;
    const int wm = (w / 2) * 32, wn = (w % 2) * 32;
    d4 acc[2][2];
    if (colK) {
        // S1 = A_tm,K (rows tm.., columns K), S2 = Pinv
        for (int idx = tid; idx < kTile * kTile; idx += 256) {
            const int r = idx / kTile, c = idx % kTile;
            S1[r][c] = (tm + r < n && c < b) ? M[(size_t)(tm + r) * n + k0 + c] : 0.0;
            S2[r][c] = (r < b && c < b) ? Pinv[r * b + c] : 0.0;
        }
        __syncthreads();
        tile_mfma(S1, S2, wm, wn, lane, acc);
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni)
                for (int q = 0; q < 4; ++q) {
                    const int row = tm + wm + mi * 16 + lane / 16 + 4 * q;
                    const int col = wn + ni * 16 + lane % 16;
                    if (row < n && col < b) Ck[(size_t)row * b + col] = 0.0 + -1.0 * acc[mi][ni][q];
                }
        return;
    }
    if (kRReady) {
        // S2 = the R tile, S1 = A_tm,K
        for (int idx = tid; idx < kTile * kTile; idx += 256) {
            const int r = idx / kTile, c = idx % kTile;
            S2[r][c] = (r < b && tn + c < n) ? R[(size_t)r * n + tn + c] : 0.0;
            S1[r][c] = (tm + r < n && c < b) ? M[(size_t)(tm + r) * n + k0 + c] : 0.0;
        }
        __syncthreads();
        tile_mfma(S1, S2, wm, wn, lane, acc);
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni)
                for (int q = 0; q < 4; ++q) {
                    const int row = tm + wm + mi * 16 + lane / 16 + 4 * q;
                    const int col = tn + wn + ni * 16 + lane % 16;
                    if (row < n && col < n) {
                        double *c = &M[(size_t)row * n + col];
                        *c = 1.0 * *c + -1.0 * acc[mi][ni][q];
                    }
                }
        return;
    }
    // S1 = Pinv, S2 = A_K,tn
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        const int r = idx / kTile, c = idx % kTile;
        S1[r][c] = (r < b && c < b) ? Pinv[r * b + c] : 0.0;
        S2[r][c] = (r < b && tn + c < n) ? M[(size_t)(k0 + r) * n + tn + c] : 0.0;
    }
    __syncthreads();
    tile_mfma(S1, S2, wm, wn, lane, acc);
    if (rowK) {
        for (int mi = 0; mi < 2; ++mi)
            for (int ni = 0; ni < 2; ++ni)
                for (int q = 0; q < 4; ++q) {
                    const int row = wm + mi * 16 + lane / 16 + 4 * q;
                    const int col = tn + wn + ni * 16 + lane % 16;
                    if (row < b && col < n) R[(size_t)row * n + col] = 0.0 + 1.0 * acc[mi][ni][q];
                }
        return;
    }
    __syncthreads();   // every wavefront is done with S1 and S2
    // S2 = the R tile (rows k < b, zero beyond), S1 = A_tm,K
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < 2; ++ni)
            for (int q = 0; q < 4; ++q) {
                const int row = wm + mi * 16 + lane / 16 + 4 * q;
                const int col = wn + ni * 16 + lane % 16;
                S2[row][col] = row < b ? 0.0 + 1.0 * acc[mi][ni][q] : 0.0;
            }
    for (int idx = tid; idx < kTile * kTile; idx += 256) {
        const int r = idx / kTile, c = idx % kTile;
        S1[r][c] = (tm + r < n && c < b) ? M[(size_t)(tm + r) * n + k0 + c] : 0.0;
    }
    __syncthreads();
    tile_mfma(S1, S2, wm, wn, lane, acc);
    for (int mi = 0; mi < 2; ++mi)
        for (int ni = 0; ni < 2; ++ni)
            for (int q = 0; q < 4; ++q) {
                const int row = tm + wm + mi * 16 + lane / 16 + 4 * q;
                const int col = tn + wn + ni * 16 + lane % 16;
                if (row < n && col < n) {
                    double *c = &M[(size_t)row * n + col];
                    *c = 1.0 * *c + -1.0 * acc[mi][ni][q];
                }
            }
}

// Row and column block K after a panel step: M[k0 + r][c] = R[r][c] outside
// K and Pinv inside it; M[i][k0 + c] = Ck[i][c] for the rows outside K.
__global__ void k_put_pivot_blocks(int n, int k0, int b, const double *__restrict__ R,
                                   const double *__restrict__ Pinv, const double *__restrict__ Ck,
                                   double *__restrict__ M) {
    const long rows_part = (long)b * n, total = 2 * rows_part;
    for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        if (e < rows_part) {
            const int r = (int)(e / n), c = (int)(e % n);
            const int cc = c - k0;
            M[(size_t)(k0 + r) * n + c] = (cc >= 0 && cc < b) ? Pinv[r * b + cc] : R[(size_t)r * n + c];
        } else {
            const long f = e - rows_part;
            const int i = (int)(f / b), c = (int)(f % b);
            if (i < k0 || i >= k0 + b) M[(size_t)i * n + k0 + c] = Ck[(size_t)i * b + c];
        }
    }
}

// x = Ainv b, one wavefront per row block of 4 rows.
__global__ __launch_bounds__(256) void k_gemv(int n, const double *__restrict__ Ainv,
                                              const double *__restrict__ b, double *__restrict__ x) {
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
            bv[u] = k < n ? b[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (kb + u * kWave < n) s += av[u] * bv[u];
    }
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, kWave);
    if (lane == 0) x[row] = s;
}

// ---------------------------------------------------------------------------
// The coarsest level's iterative solve in one workgroup: CG without
// preconditioner, stop type 1 (||r||/||r0||), with the reference's safeguards
// (dcsr_pcg, krylov.c:122-372), and GMRES(30) when it fails.  x holds the
// initial guess (zero in the cycle).  work: coarse_itsolver_work(n) doubles.
// ---------------------------------------------------------------------------
constexpr int kCgBlock = 1024;

__device__ double cg_block_sum(double v, double *sh) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kCgBlock / kWave; ++i) r += sh[i];
    return r;   // every thread gets the same value
}

__device__ double cg_block_max(double v, double *sh) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    const int lane = threadIdx.x % kWave, w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < kCgBlock / kWave; ++i) r = fmax(r, sh[i]);
    return r;
}

__device__ void cg_matvec(int n, int nslices, const long long *sptr, const int *col, const double *val,
                          const double *x, double *y, double alpha_mode, const double *b) {
    // alpha_mode 0: y = A x ; 1: y = b - A x
    for (int s = threadIdx.x / kWave; s < nslices; s += kCgBlock / kWave) {
        const int lane = threadIdx.x % kWave;
        const int r = s * kSlice + lane;
        const long long base = sptr[s];
        const int wd = (int)((sptr[s + 1] - base) / kSlice);
        double acc = 0.0;
        for (int k = 0; k < wd; ++k) {
            const int c = col[base + (long long)k * kSlice + lane];
            if (c >= 0) acc = acc + val[base + (long long)k * kSlice + lane] * x[c];
        }
        if (r < n) y[r] = (alpha_mode == 0.0) ? acc : b[r] - acc;
    }
}

// dcsr_pcg's result: iterations, or ERROR_SOLVER_SOLSTAG / _STAG / _TOLSMALL /
// _MAXIT (krylov.c:122-372 as coarse_itsolver calls it: no preconditioner).
__device__ int coarse_cg_body(int n, int nslices, const long long *__restrict__ sptr, const int *__restrict__ col,
                              const double *__restrict__ val, const double *__restrict__ b, double *__restrict__ u,
                              double tol, int maxit, double *__restrict__ work, double *sh) {
    double *p = work, *r = work + n, *t = work + 2 * n;
    const int tid = threadIdx.x;
    const double maxdiff = tol * 1e-4;
    int iter = 0, stag = 1, more_step = 1;
    cg_matvec(n, nslices, sptr, col, val, u, r, 1.0, b);
    __syncthreads();
    double acc = 0.0;
    for (int i = tid; i < n; i += kCgBlock) acc += r[i] * r[i];
    double absres0 = sqrt(cg_block_sum(acc, sh));
    const double normr0 = absres0 > kSmallReal ? absres0 : kSmallReal;
    double relres = absres0 / normr0, absres = absres0;
    if (relres < tol || absres0 < 1e-3 * tol) return 0;
    for (int i = tid; i < n; i += kCgBlock) p[i] = r[i];
    acc = 0.0;
    for (int i = tid; i < n; i += kCgBlock) acc += r[i] * r[i];
    double temp1 = cg_block_sum(acc, sh);
    while (iter++ < maxit) {
        __syncthreads();
        cg_matvec(n, nslices, sptr, col, val, p, t, 0.0, nullptr);
        __syncthreads();
        acc = 0.0;
        for (int i = tid; i < n; i += kCgBlock) acc += t[i] * p[i];
        const double temp2 = cg_block_sum(acc, sh);
        const double alpha = temp1 / temp2;
        double rr = 0.0, ui = 0.0, uu = 0.0, pp = 0.0;
        for (int i = tid; i < n; i += kCgBlock) {
            const double ui_ = u[i] + alpha * p[i];
            const double ri = r[i] + (-alpha) * t[i];
            u[i] = ui_;
            r[i] = ri;
            rr += ri * ri;
            ui = fmax(ui, fabs(ui_));
            uu += ui_ * ui_;
            pp += p[i] * p[i];
        }
        absres = sqrt(cg_block_sum(rr, sh));
        relres = absres / normr0;
        const double infnormu = cg_block_max(ui, sh);
        if (infnormu <= kSmallReal) return HZ_ERROR_SOLVER_SOLSTAG;   // Check I
        const double normu = sqrt(cg_block_sum(uu, sh));
        const double reldiff = fabs(alpha) * sqrt(cg_block_sum(pp, sh)) / normu;
        if ((stag <= 20) & (reldiff < maxdiff)) {                  // Check II
            __syncthreads();
            cg_matvec(n, nslices, sptr, col, val, u, r, 1.0, b);
            __syncthreads();
            acc = 0.0;
            for (int i = tid; i < n; i += kCgBlock) acc += r[i] * r[i];
            absres = sqrt(cg_block_sum(acc, sh));
            relres = absres / normr0;
            if (relres < tol) return iter;
            if (stag >= 20) return HZ_ERROR_SOLVER_STAG;
            for (int i = tid; i < n; i += kCgBlock) p[i] = 0.0;
            ++stag;
        }
        if (relres < tol) {                                        // Check III
            __syncthreads();
            cg_matvec(n, nslices, sptr, col, val, u, r, 1.0, b);
            __syncthreads();
            acc = 0.0;
            for (int i = tid; i < n; i += kCgBlock) acc += r[i] * r[i];
            absres = sqrt(cg_block_sum(acc, sh));
            relres = absres / normr0;
            if (relres < tol) return iter;
            if (more_step >= 20) return HZ_ERROR_SOLVER_TOLSMALL;
            for (int i = tid; i < n; i += kCgBlock) p[i] = 0.0;
            ++more_step;
        }
        absres0 = absres;
        __syncthreads();
        acc = 0.0;
        for (int i = tid; i < n; i += kCgBlock) acc += r[i] * r[i];
        const double zr = cg_block_sum(acc, sh);
        const double beta = zr / temp1;
        temp1 = zr;
        for (int i = tid; i < n; i += kCgBlock) p[i] = 1.0 * r[i] + beta * p[i];
    }
    return HZ_ERROR_SOLVER_MAXIT;
}


// Block-wide vector helpers of the coarse GMRES: element i always belongs to
// thread i mod kCgBlock, so consecutive element-wise steps need no barrier.
__device__ double blk_dot(int n, const double *x, const double *y, double *sh) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += kCgBlock) acc += x[i] * y[i];
    return cg_block_sum(acc, sh);
}
__device__ void blk_axpy(int n, double a, const double *x, double *y) {   // array_axpy (array.c:357-373)
    for (int i = threadIdx.x; i < n; i += kCgBlock) {
        if (a == 1.0) y[i] += x[i];
        else if (a == -1.0) y[i] -= x[i];
        else y[i] += a * x[i];
    }
}
__device__ void blk_ax(int n, double a, double *x) {   // array_ax (array.c:332-341)
    if (a != 1.0)
        for (int i = threadIdx.x; i < n; i += kCgBlock) x[i] *= a;
}

// dcsr_pvgmres (krylov.c:4612-4931) without preconditioner, stop type 1, from
// the iterate in x: GMRES(restart) with modified Gram-Schmidt and Givens
// rotations, the restart length adapted per cycle from the last cycle's
// convergence rate, the true residual checked before convergence is accepted.
// The Hessenberg matrix and the rotations live in LDS and are updated by
// thread 0; every branch depends only on block-wide sums or LDS values read
// after a barrier, so all threads take it together.  work: (restart + 3) n.
__device__ int coarse_gmres_body(int n, int nslices, const long long *__restrict__ sptr, const int *__restrict__ col,
                                 const double *__restrict__ val, const double *__restrict__ b, double *__restrict__ x,
                                 double tol, int MaxIt, double *__restrict__ work, double *sh, double *hh, double *cs,
                                 double *sn, double *rs) {
    constexpr int R = kCoarseGmresRestart;
    const int tid = threadIdx.x;
    double *r = work, *w = work + n, *P = work + 2 * (size_t)n;
    auto Pk = [&](int k) { return P + (size_t)k * n; };
    const double cr_max = 0.99, cr_min = 0.174;
    const int d = 3, restart_max = R, restart_min = 3;
    int iter = 0, i = 0, Restart = R;
    double cr = 1.0, r_norm, r_norm_old, absres0, absres, relres;
    __syncthreads();
    cg_matvec(n, nslices, sptr, col, val, x, Pk(0), 1.0, b);
    __syncthreads();
    r_norm = sqrt(blk_dot(n, Pk(0), Pk(0), sh));
    absres0 = r_norm > kSmallReal ? r_norm : kSmallReal;
    relres = r_norm / absres0;
    if (relres < tol || absres0 < 1e-3 * tol) return 0;
    while (iter < MaxIt) {
        r_norm_old = r_norm;
        if (tid == 0) rs[0] = r_norm;
        blk_ax(n, 1.0 / r_norm, Pk(0));
        if (cr > cr_max || iter == 0) Restart = restart_max;
        else if (cr < cr_min) {
        } else if (Restart - d > restart_min) Restart -= d;
        else Restart = restart_max;
        i = 0;
        while (i < Restart && iter < MaxIt) {
            ++i;
            ++iter;
            __syncthreads();
            cg_matvec(n, nslices, sptr, col, val, Pk(i - 1), Pk(i), 0.0, nullptr);
            __syncthreads();
            for (int j = 0; j < i; ++j) {
                const double h = blk_dot(n, Pk(j), Pk(i), sh);
                if (tid == 0) hh[j * R + i - 1] = h;
                blk_axpy(n, -h, Pk(j), Pk(i));
            }
            double t = sqrt(blk_dot(n, Pk(i), Pk(i), sh));
            if (tid == 0) hh[i * R + i - 1] = t;
            if (t != 0.0) blk_ax(n, 1.0 / t, Pk(i));
            __syncthreads();
            if (tid == 0) {
                for (int j = 1; j < i; ++j) {
                    t = hh[(j - 1) * R + i - 1];
                    hh[(j - 1) * R + i - 1] = sn[j - 1] * hh[j * R + i - 1] + cs[j - 1] * t;
                    hh[j * R + i - 1] = -sn[j - 1] * t + cs[j - 1] * hh[j * R + i - 1];
                }
                t = hh[i * R + i - 1] * hh[i * R + i - 1];
                t += hh[(i - 1) * R + i - 1] * hh[(i - 1) * R + i - 1];
                double gamma = sqrt(t);
                if (gamma == 0.0) gamma = kSmallReal;
                cs[i - 1] = hh[(i - 1) * R + i - 1] / gamma;
                sn[i - 1] = hh[i * R + i - 1] / gamma;
                rs[i] = -sn[i - 1] * rs[i - 1];
                rs[i - 1] = cs[i - 1] * rs[i - 1];
                hh[(i - 1) * R + i - 1] = sn[i - 1] * hh[i * R + i - 1] + cs[i - 1] * hh[(i - 1) * R + i - 1];
            }
            __syncthreads();
            absres = r_norm = fabs(rs[i]);
            relres = absres / absres0;
            if (relres < tol) break;
        }
        // back substitution, then x += w with w = sum_j rs_j p_j (j = i-1 first)
        __syncthreads();
        if (tid == 0) {
            rs[i - 1] = rs[i - 1] / hh[(i - 1) * R + i - 1];
            for (int k = i - 2; k >= 0; --k) {
                double t = 0.0;
                for (int j = k + 1; j < i; ++j) t -= hh[k * R + j] * rs[j];
                t += rs[k];
                rs[k] = t / hh[k * R + k];
            }
        }
        __syncthreads();
        for (int e = tid; e < n; e += kCgBlock) w[e] = Pk(i - 1)[e];
        blk_ax(n, rs[i - 1], w);
        for (int j = i - 2; j >= 0; --j) blk_axpy(n, rs[j], Pk(j), w);
        blk_axpy(n, 1.0, w, x);
        if (relres < tol) {
            __syncthreads();
            cg_matvec(n, nslices, sptr, col, val, x, r, 1.0, b);
            __syncthreads();
            r_norm = sqrt(blk_dot(n, r, r, sh));
            absres = r_norm;
            relres = absres / absres0;
            if (relres < tol) break;
            for (int e = tid; e < n; e += kCgBlock) Pk(0)[e] = r[e];
            i = 0;
        }
        // the cycle's residual vector, the next cycle's first basis vector
        __syncthreads();
        if (tid == 0)
            for (int j = i; j > 0; --j) {
                rs[j - 1] = -sn[j - 1] * rs[j];
                rs[j] = cs[j - 1] * rs[j];
            }
        __syncthreads();
        if (i) blk_axpy(n, rs[i] - 1.0, Pk(i), Pk(i));
        for (int j = i - 1; j > 0; --j) blk_axpy(n, rs[j], Pk(j), Pk(i));
        if (i) {
            blk_axpy(n, rs[0] - 1.0, Pk(0), Pk(0));
            blk_axpy(n, 1.0, Pk(i), Pk(0));
        }
        cr = r_norm / r_norm_old;
    }
    return iter >= MaxIt ? HZ_ERROR_SOLVER_MAXIT : iter;
}

// coarse_itsolver (mgcycle.c:34-53): CG from x, and GMRES(30) from CG's last
// iterate when CG fails.  status (optional): [0] the final result code, [1]
// CG's; u_cg (optional): CG's last iterate.
__global__ __launch_bounds__(kCgBlock) void k_coarse_itsolver(int n, int nslices, const long long *__restrict__ sptr,
                                                              const int *__restrict__ col,
                                                              const double *__restrict__ val,
                                                              const double *__restrict__ b, double *__restrict__ u,
                                                              double tol, int maxit, double *__restrict__ work,
                                                              int *__restrict__ status, double *__restrict__ u_cg) {
    __shared__ double sh[kCgBlock / kWave];
    __shared__ double hh[(kCoarseGmresRestart + 1) * kCoarseGmresRestart];
    __shared__ double cs[kCoarseGmresRestart + 1], sn[kCoarseGmresRestart + 1], rs[kCoarseGmresRestart + 1];
    const int st_cg = coarse_cg_body(n, nslices, sptr, col, val, b, u, tol, maxit, work, sh);
    if (u_cg)
        for (int i = threadIdx.x; i < n; i += kCgBlock) u_cg[i] = u[i];
    int st = st_cg;
    if (st < 0) st = coarse_gmres_body(n, nslices, sptr, col, val, b, u, tol, maxit, work, sh, hh, cs, sn, rs);
    if (status && threadIdx.x == 0) {
        status[0] = st;
        status[1] = st_cg;
    }
}

}  // namespace

namespace {

void launch_gemm(bool mfma, int M, int N, int K, const double *A, int lda, const double *B, int ldb, double *C,
                 int ldc, double alpha, double beta) {
    if (M <= 0 || N <= 0) return;
    dim3 grid(div_up(N, kTile), div_up(M, kTile));
    if (mfma)
        k_gemm<true><<<grid, 256, 0, stream()>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
    else
        k_gemm<false><<<grid, 256, 0, stream()>>>(M, N, K, A, lda, B, ldb, C, ldc, alpha, beta);
    HZ_LAUNCH_CHECK();
}

// The MFMA operand/accumulator lane layout is checked once against the FMA
// kernel on a 64x64x64 product; a mismatch keeps the FMA kernel.
bool mfma_layout_ok() {
    static int ok = -1;
    if (ok >= 0) return ok == 1;
    const int n = kTile;
    std::vector<double> ha(n * n), hb(n * n);
    for (int i = 0; i < n * n; ++i) {
        ha[i] = std::sin(0.37 * i + 0.1) + (i % 7) * 0.125;
        hb[i] = std::cos(0.23 * i - 0.4) - (i % 5) * 0.25;
    }
    double *d = arena().alloc_n<double>(4 * (size_t)n * n);
    HZ_CHECK(memcpy_sync(d, ha.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
    HZ_CHECK(memcpy_sync(d + n * n, hb.data(), sizeof(double) * n * n, hipMemcpyHostToDevice));
    launch_gemm(true, n, n, n, d, n, d + n * n, n, d + 2 * n * n, n, 1.0, 0.0);
    launch_gemm(false, n, n, n, d, n, d + n * n, n, d + 3 * n * n, n, 1.0, 0.0);
    std::vector<double> r1(n * n), r2(n * n);
    HZ_CHECK(memcpy_sync(r1.data(), d + 2 * n * n, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    HZ_CHECK(memcpy_sync(r2.data(), d + 3 * n * n, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    double err = 0.0, mag = 0.0;
    for (int i = 0; i < n * n; ++i) {
        err = std::max(err, std::fabs(r1[i] - r2[i]));
        mag = std::max(mag, std::fabs(r2[i]));
    }
    ok = (err <= 1e-12 * mag) ? 1 : 0;
    return ok == 1;
}

// The coarsest inverse runs on the FP64 matrix cores only: a failed lane-layout
// self-check is a device error, never a silent switch to another kernel.
void require_mfma() {
    if (!mfma_layout_ok()) {
        std::fprintf(stderr, "hazmath_mi355x: the FP64 MFMA GEMM disagrees with the FMA GEMM on its "
                             "self-check; refusing to build the coarsest inverse\n");
        throw DeviceError{hipErrorUnknown, "MFMA lane-layout self-check"};
    }
}

// Above this many rows R = P^-1 A_K,: is formed by its own product before the
// panel kernel (which then only reads it): recomputing it in every row of tiles
// doubles the matrix-core work.  Measured inverse times (ms), fused / separate:
// 992 rows 2.00 / 2.15, 2,000 6.41 / 6.34, 4,000 27.8 / 24.7, 7,896 157.9 / 138.2.
constexpr int kGjSeparateRRows = 1536;

// In-place inverse of the dense n x n matrix M by 64-column block Gauss-Jordan.
int blocked_inverse(int n, double *M) {
    require_mfma();
    double *Pinv = arena().alloc_n<double>((size_t)kTile * kTile);
    double *R = arena().alloc_n<double>((size_t)kTile * n);
    double *Ck = arena().alloc_n<double>((size_t)n * kTile);
    int *status = arena().alloc_n<int>(1);
    HZ_CHECK(hipMemsetAsync(status, 0, sizeof(int), stream()));
    for (int k0 = 0; k0 < n; k0 += kTile) {
        const int b = std::min(kTile, n - k0);
        k_block_inverse<<<1, 1024, 0, stream()>>>(n, k0, b, M, Pinv, status);
        HZ_LAUNCH_CHECK();
        // R = P^-1 A_K,:, A_i,: -= A_i,K R and A_i,K = -A_i,K P^-1 in one launch, then rows and columns K;
        // on large matrices R is formed once by its own product instead of by every row of tiles
        const dim3 grid(div_up(n, kTile), div_up(n, kTile));
        if (n > kGjSeparateRRows) {
            launch_gemm(true, b, n, b, Pinv, b, M + (size_t)k0 * n, n, R, n, 1.0, 0.0);
            k_gj_panel<true><<<grid, 256, 0, stream()>>>(n, k0, b, M, Pinv, R, Ck);
        } else {
            k_gj_panel<false><<<grid, 256, 0, stream()>>>(n, k0, b, M, Pinv, R, Ck);
        }
        k_put_pivot_blocks<<<std::min(4096, div_up(2L * b * n, 256)), 256, 0, stream()>>>(n, k0, b, R, Pinv, Ck, M);
        HZ_LAUNCH_CHECK();
    }
    int hs = 0;
    HZ_CHECK(memcpy_sync(&hs, status, sizeof(int), hipMemcpyDeviceToHost));
    return hs;
}

}  // namespace

int dense_inverse_inplace(int n, double *M) { return blocked_inverse(n, M); }

bool dense_mfma_active() { return mfma_layout_ok(); }

int coarse_dense_setup(const DevCSR &A, double *Ainv) {
    const int n = A.n;
    if (n == 0) return 0;
    hipStream_t st = stream();
    HZ_CHECK(hipMemsetAsync(Ainv, 0, sizeof(double) * (size_t)n * n, st));
    k_dense_from_csr<<<div_up(n, 256), 256, 0, st>>>(n, A.ia, A.ja, A.val, Ainv);
    HZ_LAUNCH_CHECK();
    return blocked_inverse(n, Ainv);
}

void coarse_dense_solve(int n, const double *Ainv, const double *b, double *x) {
    if (n == 0) return;
    count_bytes(8.0 * n * (double)n + 16.0 * n);
    k_gemv<<<div_up(n, 4), 256, 0, stream()>>>(n, Ainv, b, x);
    HZ_LAUNCH_CHECK();
}

void coarse_cg(const DevSELL &S, const double *b, double *x, double tol, int maxit, double *work, int *status,
               double *x_cg) {
    if (S.n == 0) return;
    k_coarse_itsolver<<<1, kCgBlock, 0, stream()>>>(S.n, S.nslices, S.sptr, S.col, S.val, b, x, tol, maxit, work,
                                                    status, x_cg);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
