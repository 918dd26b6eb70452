// kernels_sa.hip -- smoothed aggregation (amg_setup_smoothP_smoothR,
// src/solver/amg_setup_ua.c:1461-1687) on the GPU: the smoothed prolongation
// P = (I - omega D^-1 M) P_tent, R = P^T, the general Galerkin product R A P,
// and the CSR transfers the cycle uses with them (mgcycle.c:302,346).
//
// Every sum is formed in the reference's order (row order for the smoother
// rows, traversal order (R row, A row, P row) for the Galerkin product), so the
// results are bit-identical to the sequential code.
#include <algorithm>
#include <climits>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    return (int)std::max(1L, std::min<long>(g, 8192));
}

// Filtered strength graph (smooth_aggregation_p, filter ON): the diagonal
// entries of row i of N absorb (sum of A's off-diagonal entries) - (sum of
// N's off-diagonal entries), each sum from 0.0 in row order.
__global__ void k_sa_filter(int n, const int *__restrict__ aia, const int *__restrict__ aja,
                            const double *__restrict__ aval, const int *__restrict__ nia,
                            const int *__restrict__ nja, double *__restrict__ nval) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double sa = 0.0, sn = 0.0;
        for (int k = aia[i]; k < aia[i + 1]; ++k)
            if (aja[k] != i) sa += aval[k];
        for (int k = nia[i]; k < nia[i + 1]; ++k)
            if (nja[k] != i) sn += nval[k];
        for (int k = nia[i]; k < nia[i + 1]; ++k)
            if (nja[k] == i) nval[k] += sa - sn;
    }
}

// d_i: first diagonal entry (dcsr_getdiag), 1.0 when |d_i| < 1e-6
__global__ void k_sa_diag(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ val, double *__restrict__ d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double di = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] == i) {
                di = val[k];
                break;
            }
        d[i] = fabs(di) < 1e-6 ? 1.0 : di;
    }
}

// true when agg[ja[q]] has not occurred earlier in the row
__device__ __forceinline__ bool first_in_row(const int *ja, const int *agg, int b, int q, int c) {
    for (int p = b; p < q; ++p)
        if (agg[ja[p]] == c) return false;
    return true;
}

__global__ void k_sa_count(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                           const int *__restrict__ agg, int *__restrict__ cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = ia[i], e = ia[i + 1];
        int c0 = 0;
        for (int q = b; q < e; ++q) {
            const int c = agg[ja[q]];
            if (c > HZ_UNPT && first_in_row(ja, agg, b, q, c)) ++c0;
        }
        cnt[i] = c0;
    }
}

// P row i (dcsr_mxm of S and the tentative P, sparse.c:1832-1918): the
// aggregates of row i's columns in order of first occurrence, each valued
// 0.0 + sum over its entries, in row order, of S_ik * 1.0 with
// S_ik = 1 - omega M_ik / d_i (k == i) or -omega M_ik / d_i.
__global__ void k_sa_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ mval, const int *__restrict__ agg,
                          const double *__restrict__ d, double omega_all, const float *__restrict__ omega_agg,
                          const int *__restrict__ pia, int *__restrict__ pja, double *__restrict__ pval) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = ia[i], e = ia[i + 1];
        const double di = d[i];
        // per-aggregate damping: row i uses its own aggregate's value
        const double omega = (omega_agg && agg[i] > HZ_UNPT) ? (double)omega_agg[agg[i]] : omega_all;
        int pos = pia[i];
        for (int q = b; q < e; ++q) {
            const int c = agg[ja[q]];
            if (c <= HZ_UNPT || !first_in_row(ja, agg, b, q, c)) continue;
            double s = 0.0;
            for (int p = q; p < e; ++p) {
                const int k = ja[p];
                if (agg[k] != c) continue;
                const double sv = (k == i) ? 1 - omega * mval[p] / di : -omega * mval[p] / di;
                s += sv * 1.0;
            }
            pja[pos] = c;
            pval[pos] = s;
            ++pos;
        }
    }
}

__global__ void k_row_of_entry(int n, const int *__restrict__ ia, int *__restrict__ row) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        for (int k = ia[i]; k < ia[i + 1]; ++k) row[k] = i;
}

__global__ void k_col_keys(int nnz, const int *__restrict__ ja, unsigned int *__restrict__ keys,
                           int *__restrict__ idx, int *__restrict__ cnt) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) {
        keys[k] = (unsigned int)ja[k];
        idx[k] = k;
        atomicAdd(&cnt[ja[k]], 1);
    }
}

__global__ void k_trans_fill(int nnz, const int *__restrict__ order, const int *__restrict__ row,
                             const double *__restrict__ val, int *__restrict__ tja, double *__restrict__ tval) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += gridDim.x * blockDim.x) {
        const int e = order[k];
        tja[k] = row[e];
        tval[k] = val[e];
    }
}

// ---------------------------------------------------------------------------
// General Galerkin product (dcsr_rap, sparse.c:2401-2538), one thread per
// coarse row with an open-addressing table (coarse column -> output slot) in
// global scratch, sized per row from the number of candidate products.
// ---------------------------------------------------------------------------
__global__ void k_rap_candidates(int nc, const int *__restrict__ ria, const int *__restrict__ rja,
                                 const int *__restrict__ aia, const int *__restrict__ aja,
                                 const int *__restrict__ pia, long long *__restrict__ tsize) {
    for (int ic = blockIdx.x * blockDim.x + threadIdx.x; ic < nc; ic += gridDim.x * blockDim.x) {
        long long L = 1;
        for (int q1 = ria[ic]; q1 < ria[ic + 1]; ++q1) {
            const int i1 = rja[q1];
            for (int q2 = aia[i1]; q2 < aia[i1 + 1]; ++q2) {
                const int i2 = aja[q2];
                L += pia[i2 + 1] - pia[i2];
            }
        }
        const long long bound = L < (long long)nc ? L : (long long)nc;
        long long h = 16;
        while (h < 2 * bound) h <<= 1;
        tsize[ic] = h;
    }
}

__device__ __forceinline__ unsigned long long slot_hash(int key, unsigned long long mask) {
    unsigned long long h = (unsigned long long)(unsigned)key * 0x9E3779B97F4A7C15ull;
    return (h >> 29) & mask;
}

template <bool kFill>
__global__ void k_rap_rows(int r0, int r1, const long long *__restrict__ toff, long long tbase,
                           int *__restrict__ tkey, int *__restrict__ tpos, const int *__restrict__ ria,
                           const int *__restrict__ rja, const double *__restrict__ rval,
                           const int *__restrict__ aia, const int *__restrict__ aja,
                           const double *__restrict__ aval, const int *__restrict__ pia,
                           const int *__restrict__ pja, const double *__restrict__ pval,
                           int *__restrict__ rowcount, const int *__restrict__ cia, int *__restrict__ cja,
                           double *__restrict__ cval) {
    for (int ic = r0 + blockIdx.x * blockDim.x + threadIdx.x; ic < r1; ic += gridDim.x * blockDim.x) {
        const long long t0 = toff[ic] - tbase;
        const unsigned long long mask = (unsigned long long)(toff[ic + 1] - toff[ic]) - 1ull;
        int *key = tkey + t0;
        int *pos = tpos + t0;
        for (unsigned long long t = 0; t <= mask; ++t) key[t] = -1;
        int next = kFill ? cia[ic] : 0;
        // the diagonal comes first
        {
            unsigned long long s = slot_hash(ic, mask);
            key[s] = ic;
            pos[s] = next;
            if (kFill) {
                cja[next] = ic;
                cval[next] = 0.0;
            }
            ++next;
        }
        for (int q1 = ria[ic]; q1 < ria[ic + 1]; ++q1) {
            const double r = kFill ? rval[q1] : 0.0;
            const int i1 = rja[q1];
            for (int q2 = aia[i1]; q2 < aia[i1 + 1]; ++q2) {
                const double ra = kFill ? r * aval[q2] : 0.0;
                const int i2 = aja[q2];
                for (int q3 = pia[i2]; q3 < pia[i2 + 1]; ++q3) {
                    const int i3 = pja[q3];
                    unsigned long long s = slot_hash(i3, mask);
                    while (key[s] != -1 && key[s] != i3) s = (s + 1) & mask;
                    if (key[s] == -1) {
                        key[s] = i3;
                        pos[s] = next;
                        if (kFill) {
                            cja[next] = i3;
                            cval[next] = ra * pval[q3];
                        }
                        ++next;
                    } else if (kFill) {
                        cval[pos[s]] += ra * pval[q3];
                    }
                }
            }
        }
        if (!kFill) rowcount[ic] = next;
    }
}

// y = A x, row sums from 0.0 in CSR order (dcsr_mxv)
__global__ void k_csr_mxv(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; ++k) t += val[k] * x[ja[k]];
        y[i] = t;
    }
}

// x += alpha P xc (dcsr_aAxpy, sparse.c:1693-1730)
__global__ void k_csr_aAxpy(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const double *__restrict__ val, const double *__restrict__ xc, double alpha,
                            const double *__restrict__ alpha_dev, double *__restrict__ x) {
    const double a = alpha_dev ? *alpha_dev : alpha;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double t = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; ++k) t += val[k] * xc[ja[k]];
        if (a == 1.0) x[i] += t;
        else if (a == -1.0) x[i] -= t;
        else x[i] += t * a;
    }
}

}  // namespace

void sa_prolongation(const DevCSR &A, DevCSR &N, const int *agg, int nagg, double omega, int filter,
                     DevCSR &P, const float *omega_agg) {
    hipStream_t st = stream();
    const int n = A.n;
    const DevCSR &M = (filter == 1) ? N : A;
    if (filter == 1) {
        k_sa_filter<<<grid_for(n), kBlock, 0, st>>>(n, A.ia, A.ja, A.val, N.ia, N.ja, N.val);
        HZ_LAUNCH_CHECK();
    }
    double *d = arena().alloc_n<double>((size_t)n + 1);
    int *cnt = arena().alloc_n<int>((size_t)n + 1);
    k_sa_diag<<<grid_for(n), kBlock, 0, st>>>(n, M.ia, M.ja, M.val, d);
    k_sa_count<<<grid_for(n), kBlock, 0, st>>>(n, M.ia, M.ja, agg, cnt);
    HZ_LAUNCH_CHECK();
    P.n = n;
    P.ncol = nagg;
    P.ia = arena().alloc_n<int>((size_t)n + 1);
    P.nnz = (int)exclusive_scan_int(cnt, P.ia, n);
    P.ja = arena().alloc_n<int>((size_t)P.nnz + 1);
    P.val = arena().alloc_n<double>((size_t)P.nnz + 1);
    k_sa_fill<<<grid_for(n), kBlock, 0, st>>>(n, M.ia, M.ja, M.val, agg, d, omega, omega_agg, P.ia, P.ja,
                                              P.val);
    HZ_LAUNCH_CHECK();
}

void csr_transpose_gpu(const DevCSR &A, DevCSR &AT) {
    hipStream_t st = stream();
    const int n = A.n, m = A.ncol, nnz = A.nnz;
    AT.n = m;
    AT.ncol = n;
    AT.nnz = nnz;
    AT.ia = arena().alloc_n<int>((size_t)m + 1);
    AT.ja = arena().alloc_n<int>((size_t)nnz + 1);
    AT.val = arena().alloc_n<double>((size_t)nnz + 1);
    int *cnt = arena().alloc_n<int>((size_t)m + 1);
    HZ_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)m + 1), st));
    if (nnz == 0) {
        exclusive_scan_int(cnt, AT.ia, m);
        return;
    }
    int *row = arena().alloc_n<int>((size_t)nnz + 1);
    unsigned int *keys = arena().alloc_n<unsigned int>((size_t)nnz + 1);
    unsigned int *keys_out = arena().alloc_n<unsigned int>((size_t)nnz + 1);
    int *idx = arena().alloc_n<int>((size_t)nnz + 1);
    int *idx_out = arena().alloc_n<int>((size_t)nnz + 1);
    k_row_of_entry<<<grid_for(n), kBlock, 0, st>>>(n, A.ia, row);
    k_col_keys<<<grid_for(nnz), kBlock, 0, st>>>(nnz, A.ja, keys, idx, cnt);
    HZ_LAUNCH_CHECK();
    exclusive_scan_int(cnt, AT.ia, m);
    int bits = 1;
    while ((1u << bits) <= (unsigned)m && bits < 32) ++bits;
    size_t tmp_bytes = 0;
    // radix sort is stable: entries of a column keep increasing source rows
    HZ_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys_out, idx, idx_out, (size_t)nnz, 0, bits,
                                       st));
    void *tmp = arena().alloc(tmp_bytes);
    HZ_CHECK(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys_out, idx, idx_out, (size_t)nnz, 0, bits, st));
    k_trans_fill<<<grid_for(nnz), kBlock, 0, st>>>(nnz, idx_out, row, A.val, AT.ja, AT.val);
    HZ_LAUNCH_CHECK();
}

void rap_general(const DevCSR &R, const DevCSR &A, const DevCSR &P, DevCSR &C) {
    hipStream_t st = stream();
    const int nc = R.n;
    C.n = nc;
    C.ncol = P.ncol;
    C.ia = arena().alloc_n<int>((size_t)nc + 1);
    if (nc == 0) {
        C.nnz = 0;
        HZ_CHECK(hipMemsetAsync(C.ia, 0, sizeof(int), st));
        C.ja = arena().alloc_n<int>(1);
        C.val = arena().alloc_n<double>(1);
        return;
    }
    long long *tsize = arena().alloc_n<long long>((size_t)nc + 1);
    long long *toff = arena().alloc_n<long long>((size_t)nc + 1);
    k_rap_candidates<<<grid_for(nc), kBlock, 0, st>>>(nc, R.ia, R.ja, A.ia, A.ja, P.ia, tsize);
    HZ_LAUNCH_CHECK();
    {
        size_t tmp_bytes = 0;
        HZ_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, tsize, toff, 0LL, (size_t)nc + 1,
                                         rocprim::plus<long long>(), st));
        void *tmp = arena().alloc(tmp_bytes);
        HZ_CHECK(hipMemsetAsync(tsize + nc, 0, sizeof(long long), st));
        HZ_CHECK(rocprim::exclusive_scan(tmp, tmp_bytes, tsize, toff, 0LL, (size_t)nc + 1,
                                         rocprim::plus<long long>(), st));
    }
    std::vector<long long> hoff((size_t)nc + 1);
    HZ_CHECK(memcpy_sync(hoff.data(), toff, sizeof(long long) * ((size_t)nc + 1), hipMemcpyDeviceToHost));
    // rows are processed in chunks whose tables fit the scratch budget
    const long long budget = std::max<long long>(1LL << 26, [&] {
        long long mx = 0;
        for (int i = 0; i < nc; ++i) mx = std::max(mx, hoff[i + 1] - hoff[i]);
        return mx;
    }());
    const long long tables = std::min<long long>(budget, hoff[nc]);
    int *tkey = arena().alloc_n<int>((size_t)tables + 1);
    int *tpos = arena().alloc_n<int>((size_t)tables + 1);
    std::vector<int> cuts{0};
    for (int i = 0; i < nc; ++i)
        if (hoff[i + 1] - hoff[cuts.back()] > tables) cuts.push_back(i);
    cuts.push_back(nc);
    int *rowcount = arena().alloc_n<int>((size_t)nc + 1);
    for (size_t c = 0; c + 1 < cuts.size(); ++c) {
        const int r0 = cuts[c], r1 = cuts[c + 1];
        k_rap_rows<false><<<grid_for(r1 - r0), kBlock, 0, st>>>(r0, r1, toff, hoff[r0], tkey, tpos, R.ia, R.ja,
                                                                R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val,
                                                                rowcount, nullptr, nullptr, nullptr);
        HZ_LAUNCH_CHECK();
    }
    C.nnz = (int)exclusive_scan_int(rowcount, C.ia, nc);
    C.ja = arena().alloc_n<int>((size_t)C.nnz + 1);
    C.val = arena().alloc_n<double>((size_t)C.nnz + 1);
    for (size_t c = 0; c + 1 < cuts.size(); ++c) {
        const int r0 = cuts[c], r1 = cuts[c + 1];
        k_rap_rows<true><<<grid_for(r1 - r0), kBlock, 0, st>>>(r0, r1, toff, hoff[r0], tkey, tpos, R.ia, R.ja,
                                                               R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val,
                                                               nullptr, C.ia, C.ja, C.val);
        HZ_LAUNCH_CHECK();
    }
}

void csr_mxv(const DevCSR &A, const double *x, double *y) {
    if (A.n == 0) return;
    k_csr_mxv<<<grid_for(A.n), kBlock, 0, stream()>>>(A.n, A.ia, A.ja, A.val, x, y);
    HZ_LAUNCH_CHECK();
}

void csr_prolong(const DevCSR &P, const double *xc, double alpha, const double *alpha_dev, double *x) {
    if (P.n == 0) return;
    k_csr_aAxpy<<<grid_for(P.n), kBlock, 0, stream()>>>(P.n, P.ia, P.ja, P.val, xc, alpha, alpha_dev, x);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
