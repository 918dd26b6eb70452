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
    constexpr int U = 16;   // entries whose loads are issued together
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double sa = 0.0, sn = 0.0;
        const int a1 = aia[i + 1], n0 = nia[i], n1 = nia[i + 1];
        for (int kb = aia[i]; kb < a1; kb += U) {
            int jj[U];
            double vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                jj[u] = kb + u < a1 ? aja[kb + u] : -1;
                vv[u] = kb + u < a1 ? aval[kb + u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (jj[u] >= 0 && jj[u] != i) sa += vv[u];
        }
        for (int kb = n0; kb < n1; kb += U) {
            int jj[U];
            double vv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                jj[u] = kb + u < n1 ? nja[kb + u] : -1;
                vv[u] = kb + u < n1 ? nval[kb + u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
                if (jj[u] >= 0 && jj[u] != i) sn += vv[u];
        }
        for (int k = n0; k < n1; ++k)
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

// Rows of at most kRowReg entries are held in registers by k_sa_count and
// k_sa_fill (their loads issued together); longer rows take the loops over
// memory.  Both forms make the same decisions and sums in the same order.
constexpr int kRowReg = 16;

__global__ void k_sa_count(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                           const int *__restrict__ agg, int *__restrict__ cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = ia[i], e = ia[i + 1];
        if (e - b <= kRowReg) {
            int cc[kRowReg];
#pragma unroll
            for (int u = 0; u < kRowReg; ++u) cc[u] = b + u < e ? ja[b + u] : -1;
#pragma unroll
            for (int u = 0; u < kRowReg; ++u) cc[u] = cc[u] >= 0 ? agg[cc[u]] : HZ_UNPT;
            int c0 = 0;
#pragma unroll
            for (int q = 0; q < kRowReg; ++q) {
                bool first = cc[q] > HZ_UNPT;
#pragma unroll
                for (int p = 0; p < q; ++p) first = first && cc[p] != cc[q];
                c0 += first;
            }
            cnt[i] = c0;
            continue;
        }
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
        if (e - b <= kRowReg) {
            int jj[kRowReg], cc[kRowReg];
            double sv[kRowReg];
#pragma unroll
            for (int u = 0; u < kRowReg; ++u) {
                jj[u] = b + u < e ? ja[b + u] : -1;
                sv[u] = b + u < e ? mval[b + u] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kRowReg; ++u) {
                cc[u] = jj[u] >= 0 ? agg[jj[u]] : HZ_UNPT;
                sv[u] = (jj[u] == i) ? 1 - omega * sv[u] / di : -omega * sv[u] / di;
            }
#pragma unroll
            for (int q = 0; q < kRowReg; ++q) {
                bool first = cc[q] > HZ_UNPT;
#pragma unroll
                for (int p = 0; p < q; ++p) first = first && cc[p] != cc[q];
                if (!first) continue;
                double s = 0.0;
#pragma unroll
                for (int p = q; p < kRowReg; ++p)
                    if (cc[p] == cc[q]) s += sv[p] * 1.0;
                pja[pos] = cc[q];
                pval[pos] = s;
                ++pos;
            }
            continue;
        }
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
// rows (optional): the coarse rows to take, entry t of the list at tsize[t].
__global__ void k_rap_candidates(int nrows, const int *__restrict__ rows, int nc, const int *__restrict__ ria,
                                 const int *__restrict__ rja, const int *__restrict__ aia,
                                 const int *__restrict__ aja, const int *__restrict__ pia,
                                 long long *__restrict__ tsize) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nrows; t += gridDim.x * blockDim.x) {
        const int ic = rows ? rows[t] : t;
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
        tsize[t] = h;
    }
}

__device__ __forceinline__ unsigned long long slot_hash(int key, unsigned long long mask) {
    unsigned long long h = (unsigned long long)(unsigned)key * 0x9E3779B97F4A7C15ull;
    return (h >> 29) & mask;
}

template <bool kFill>
__global__ void k_rap_rows(int r0, int r1, const int *__restrict__ rows, const long long *__restrict__ toff,
                           long long tbase,
                           int *__restrict__ tkey, int *__restrict__ tpos, const int *__restrict__ ria,
                           const int *__restrict__ rja, const double *__restrict__ rval,
                           const int *__restrict__ aia, const int *__restrict__ aja,
                           const double *__restrict__ aval, const int *__restrict__ pia,
                           const int *__restrict__ pja, const double *__restrict__ pval,
                           int *__restrict__ rowcount, const int *__restrict__ cia, int *__restrict__ cja,
                           double *__restrict__ cval) {
    for (int t = r0 + blockIdx.x * blockDim.x + threadIdx.x; t < r1; t += gridDim.x * blockDim.x) {
        const int ic = rows ? rows[t] : t;
        const long long t0 = toff[t] - tbase;
        const unsigned long long mask = (unsigned long long)(toff[t + 1] - toff[t]) - 1ull;
        int *key = tkey + t0;
        int *pos = tpos + t0;
        for (unsigned long long u = 0; u <= mask; ++u) key[u] = -1;
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

// The same product with one wavefront per coarse row.  The row's (R entry,
// A entry) pairs are staged in LDS in traversal order, whole A rows at a
// time and at most max_pairs pairs per stage, with the prefix of their P-row
// lengths; each stage's candidates (q1, q2, q3) are then taken 64 at a time
// in that order.  The lanes of a round are grouped by column (one ballot per
// distinct column); a column's first lane folds the round's values of that
// column into the column's running sum in lane order, so every sum keeps the
// sequential order; new columns take output slots in order of first
// appearance, after the diagonal (slot 0, summed from 0.0).  The column table
// and the sums persist across stages.  The count pass marks with rowcount -1
// the rows it cannot take (more than max_cols columns, or one A row longer
// than a stage); the per-thread kernel takes those and the fill pass skips
// them.
// 256 pairs and 256 columns keep the LDS at 12.8 KB: twelve single-wavefront
// workgroups per CU hide the rounds' gathers from P (with 512 / 512, six).
constexpr int kRapwPairs = 256;
constexpr double kRapSlotBytes = 4e9;   // slots of the one-pass form: at most 4 GB
constexpr int kRapwCols = 256;
// the finest level's product first tries tables of this many columns
constexpr int kRapwNarrowCols = 128;

template <int kHash>
__device__ __forceinline__ unsigned rapw_hash(int key) {
    unsigned h = (unsigned)key * 2654435761u;
    return (h ^ (h >> 15)) & (kHash - 1u);
}

__device__ __forceinline__ int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// kCols: the table's columns and the stage's pairs (LDS 50 kCols bytes)
template <bool kFill, bool kSlots = false, int kCols = kRapwCols>
__global__ __launch_bounds__(kWave) void k_rapw(int nc, const int *__restrict__ ria, const int *__restrict__ rja,
                                                const double *__restrict__ rval, const int *__restrict__ aia,
                                                const int *__restrict__ aja, const double *__restrict__ aval,
                                                const int *__restrict__ pia, const int *__restrict__ pja,
                                                const double *__restrict__ pval, int max_pairs, int max_cols,
                                                int *__restrict__ rowcount, const int *__restrict__ cia,
                                                int *__restrict__ cja, double *__restrict__ cval) {
    constexpr int kPairs = kCols, kHash = 2 * kCols;   // at most kCols + 64 keys: probing ends
    __shared__ int s_pst[kPairs];              // P-row start of each pair
    __shared__ int s_off[kPairs];          // its P-row length, then its first candidate
    __shared__ double s_ra[kFill ? kPairs : 1];
    __shared__ int s_key[kHash], s_pos[kHash];
    __shared__ double s_val[kFill ? kHash : 1];
    __shared__ double s_rv[kFill ? kWave : 1];
    const int lane = threadIdx.x;
    const int ic = blockIdx.x;
    if (ic >= nc) return;   // the whole (single-wavefront) block
    if (kFill && !kSlots && rowcount[ic] < 0) return;
    // column table: the diagonal first, from 0.0
    for (int t = lane; t < kHash; t += kWave) {
        s_key[t] = -1;
        s_pos[t] = -1;
    }
    __syncthreads();
    if (lane == 0) {
        const unsigned d = rapw_hash<kHash>(ic);
        s_key[d] = ic;
        s_pos[d] = 0;
        if (kFill) s_val[d] = 0.0;
    }
    int next = 1;
    bool big = false;
    const int r1 = ria[ic + 1];
    int q1 = ria[ic];
    while (q1 < r1 && !big) {   // uniform
        __syncthreads();
        // 1) a stage of (R entry, A entry) pairs: whole A rows while they fit
        int np = 0;
        while (q1 < r1) {
            const int q = q1 + lane;
            int a0 = 0, alen = 0;
            double r = 0.0;
            if (q < r1) {
                const int i1 = rja[q];
                a0 = aia[i1];
                alen = aia[i1 + 1] - a0;
                if (kFill) r = rval[q];
            }
            const int incl = wave_incl_scan(alen, lane);
            // a prefix of the lanes: incl does not decrease with the lane
            const int nfit = __popcll(__ballot(q < r1 && incl <= max_pairs - np));
            if (nfit == 0) {
                if (np == 0) big = true;   // one A row longer than a stage
                break;
            }
            if (lane < nfit) {
                const int first = np + incl - alen;
                for (int k = 0; k < alen; ++k) {
                    const int q2 = a0 + k;
                    const int i2 = aja[q2];
                    const int ps = pia[i2];
                    s_pst[first + k] = ps;
                    s_off[first + k] = pia[i2 + 1] - ps;
                    if (kFill) s_ra[first + k] = r * aval[q2];
                }
            }
            np += __shfl(incl, nfit - 1, kWave);
            q1 += nfit;
            if (nfit < kWave) break;   // the stage is full or the row is done
        }
        if (big) break;
        __syncthreads();
        // prefix of the P-row lengths: s_off[i] = first candidate of pair i
        int L = 0;
        for (int b = 0; b < np; b += kWave) {
            const int idx = b + lane;
            const int len = idx < np ? s_off[idx] : 0;
            const int incl = wave_incl_scan(len, lane);
            if (idx < np) s_off[idx] = L + incl - len;
            L += __shfl(incl, kWave - 1, kWave);
        }
        __syncthreads();
        // 2) the stage's candidates in rounds of a wavefront
        for (int base = 0; !big && base < L; base += kWave) {   // uniform
            const int p = base + lane;
            const bool has = p < L;
            int col = -1;
            double v = 0.0;
            if (has) {
                int lo = 0, hi = np - 1;   // last pair whose first candidate is <= p
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= p) lo = mid;
                    else hi = mid - 1;
                }
                const int q3 = s_pst[lo] + (p - s_off[lo]);
                col = pja[q3];
                if (kFill) v = s_ra[lo] * pval[q3];
            }
            if (kFill) s_rv[lane] = v;
            // the lanes of each distinct column (one ballot per column)
            unsigned long long mine = 0;
            for (unsigned long long todo = __ballot(has); todo;) {   // uniform
                const int l = __builtin_ctzll(todo);
                const int c = __builtin_amdgcn_readlane(col, l);
                const unsigned long long m = __ballot(has && col == c);
                if (has && col == c) mine = m;
                todo &= ~m;
            }
            const bool lead = has && __builtin_ctzll(mine) == lane;
            unsigned slot = 0;
            if (lead) {   // distinct columns: only the first lane of each probes
                slot = rapw_hash<kHash>(col);
                while (true) {
                    const int prev = atomicCAS(&s_key[slot], -1, col);
                    if (prev == -1 || prev == col) break;
                    slot = (slot + 1) & (kHash - 1u);
                }
            }
            const bool fresh = lead && s_pos[slot] < 0;
            const unsigned long long mf = __ballot(fresh);
            if (fresh) s_pos[slot] = next + __popcll(mf & ((1ull << lane) - 1ull));
            next += __popcll(mf);
            __syncthreads();
            if (kFill && lead) {
                double acc = fresh ? v : s_val[slot] + v;
                for (unsigned long long rest = mine & (mine - 1); rest; rest &= rest - 1)
                    acc = acc + s_rv[__builtin_ctzll(rest)];
                s_val[slot] = acc;
            }
            if (next > max_cols) big = true;   // uniform; at most max_cols + 64 keys so far
            __syncthreads();
        }
    }
    if (!kFill || kSlots) {
        if (lane == 0) rowcount[ic] = big ? -1 : next;
        if (!kFill || big) return;
    }
    __syncthreads();
    const long long o = kSlots ? (long long)ic * max_cols : (long long)cia[ic];
    for (int t = lane; t < kHash; t += kWave)
        if (s_key[t] >= 0) {
            cja[o + s_pos[t]] = s_key[t];
            cval[o + s_pos[t]] = s_val[t];
        }
}

// Rows the one-pass form left in slots of max_cols (count >= 0) move to their
// CSR place, one wavefront per row.
__global__ __launch_bounds__(kBlock) void k_rap_compact_slots(int nc, const int *__restrict__ slotcount, int max_cols,
                                                             const int *__restrict__ sja,
                                                             const double *__restrict__ sval,
                                                             const int *__restrict__ cia, int *__restrict__ cja,
                                                             double *__restrict__ cval) {
    const int row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if (row >= nc) return;
    const int cnt = slotcount[row];
    if (cnt < 0) return;
    const long long src = (long long)row * max_cols;
    const int dst = cia[row];
    for (int k = lane; k < cnt; k += kWave) {
        cja[dst + k] = sja[src + k];
        cval[dst + k] = sval[src + k];
    }
}

__global__ void k_rapw_big(int nc, const int *__restrict__ rowcount, int *__restrict__ list, int *__restrict__ nlist) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        if (rowcount[c] < 0) list[atomicAdd(nlist, 1)] = c;
}

// y = A x, row sums from 0.0 in CSR order (dcsr_mxv, sparse.c:1456-1590).
// kLanes lanes share a row: each round they load kLanes consecutive entries
// and form the products val*x, and the group's first lane adds them to the
// row sum in CSR order -- the same operations and order as one thread per row.
template <int kLanes>
__global__ void k_csr_mxv(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ val, const double *__restrict__ x, double *__restrict__ y) {
    constexpr int kPass = 4;   // rounds whose loads are issued together
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(gt / kLanes), j = (int)(gt % kLanes);
    const int gbase = threadIdx.x % kWave - j;
    const bool valid = i < n;
    const int k0 = valid ? ia[i] : 0, k1 = valid ? ia[i + 1] : 0;
    double t = 0.0;
    for (int kp = k0; kp < k1; kp += kLanes * kPass) {   // trip count uniform within the lane group
        double av[kPass];
        int cj[kPass];
#pragma unroll
        for (int q = 0; q < kPass; ++q) {
            const int k = kp + q * kLanes + j;
            av[q] = k < k1 ? val[k] : 0.0;
            cj[q] = k < k1 ? ja[k] : -1;
        }
        double pr[kPass];
#pragma unroll
        for (int q = 0; q < kPass; ++q) pr[q] = cj[q] >= 0 ? av[q] * x[cj[q]] : 0.0;
#pragma unroll
        for (int q = 0; q < kPass; ++q) {
            const int kb = kp + q * kLanes;
            if (kb >= k1) break;   // uniform within the lane group
#pragma unroll
            for (int u = 0; u < kLanes; ++u) {
                const double v = __shfl(pr[q], gbase + u, kWave);
                if (j == 0 && kb + u < k1) t = t + v;
            }
        }
    }
    if (valid && j == 0) y[i] = t;
}

// x += alpha P xc (dcsr_aAxpy, sparse.c:1693-1730)
__global__ void k_csr_aAxpy(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const double *__restrict__ val, const double *__restrict__ xc, double alpha,
                            const double *__restrict__ alpha_dev, double *__restrict__ x) {
    constexpr int kPass = 8;   // entries whose loads are issued together
    const double a = alpha_dev ? *alpha_dev : alpha;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int k0 = ia[i], k1 = ia[i + 1];
        double t = 0.0;
        for (int kb = k0; kb < k1; kb += kPass) {
            double av[kPass];
            int cj[kPass];
#pragma unroll
            for (int u = 0; u < kPass; ++u) {
                av[u] = kb + u < k1 ? val[kb + u] : 0.0;
                cj[u] = kb + u < k1 ? ja[kb + u] : -1;
            }
#pragma unroll
            for (int u = 0; u < kPass; ++u)
                if (cj[u] >= 0) t += av[u] * xc[cj[u]];
        }
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

int g_rap_wave_pairs = 0;

void rap_general(const DevCSR &R, const DevCSR &A, const DevCSR &P, DevCSR &C, bool narrow_first) {
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
    int *rowcount = arena().alloc_n<int>((size_t)nc + 1);
    const bool wave = g_rap_wave_pairs >= 0 || g_rap_wave_pairs == -2;
    int max_pairs = g_rap_wave_pairs > 0 ? std::min(g_rap_wave_pairs, kRapwPairs) : kRapwPairs;
    int max_cols = g_rap_wave_pairs == -2 ? 16 : kRapwCols;
    // one pass into fixed slots of max_cols per row (then compacted) while the
    // slots fit kRapSlotBytes; otherwise a count pass, then the fill
    const bool slots = wave && (double)nc * max_cols * 12.0 <= kRapSlotBytes;
    int *slotcount = nullptr, *sja = nullptr;
    double *sval = nullptr;
    int *list = arena().alloc_n<int>((size_t)nc + 1);
    int *nlist = arena().alloc_n<int>(1);
    bool narrow_done = false;
    if (slots && narrow_first && g_rap_wave_pairs == 0) {
        // half-size tables first (twice the workgroups per CU); when a row
        // passes them, the whole product runs again with the full tables
        slotcount = arena().alloc_n<int>((size_t)nc + 1);
        sja = arena().alloc_n<int>((size_t)nc * kRapwNarrowCols);
        sval = arena().alloc_n<double>((size_t)nc * kRapwNarrowCols);
        k_rapw<true, true, kRapwNarrowCols><<<nc, kWave, 0, st>>>(nc, R.ia, R.ja, R.val, A.ia, A.ja, A.val, P.ia,
                                                                  P.ja, P.val, kRapwNarrowCols, kRapwNarrowCols,
                                                                  slotcount, nullptr, sja, sval);
        HZ_CHECK(hipMemsetAsync(nlist, 0, sizeof(int), st));
        k_rapw_big<<<grid_for(nc), kBlock, 0, st>>>(nc, slotcount, list, nlist);
        HZ_LAUNCH_CHECK();
        int over = 0;
        HZ_CHECK(memcpy_sync(&over, nlist, sizeof(int), hipMemcpyDeviceToHost));
        narrow_done = over == 0;
        if (narrow_done) {
            max_pairs = max_cols = kRapwNarrowCols;
            HZ_CHECK(hipMemcpyAsync(rowcount, slotcount, sizeof(int) * (size_t)nc, hipMemcpyDeviceToDevice, st));
        }
    }
    if (slots && !narrow_done) {
        slotcount = arena().alloc_n<int>((size_t)nc + 1);
        sja = arena().alloc_n<int>((size_t)nc * max_cols);
        sval = arena().alloc_n<double>((size_t)nc * max_cols);
        k_rapw<true, true><<<nc, kWave, 0, st>>>(nc, R.ia, R.ja, R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val,
                                                 max_pairs, max_cols, slotcount, nullptr, sja, sval);
        HZ_CHECK(hipMemcpyAsync(rowcount, slotcount, sizeof(int) * (size_t)nc, hipMemcpyDeviceToDevice, st));
    } else if (wave && !slots) {
        k_rapw<false><<<nc, kWave, 0, st>>>(nc, R.ia, R.ja, R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val, max_pairs,
                                            max_cols, rowcount, nullptr, nullptr, nullptr);
    } else if (!wave) {
        HZ_CHECK(hipMemsetAsync(rowcount, 0xff, sizeof(int) * (size_t)nc, st));   // every row -1
    }
    HZ_LAUNCH_CHECK();
    // rows beyond the wavefront kernel's LDS: one thread per row, global tables
    HZ_CHECK(hipMemsetAsync(nlist, 0, sizeof(int), st));
    k_rapw_big<<<grid_for(nc), kBlock, 0, st>>>(nc, rowcount, list, nlist);
    HZ_LAUNCH_CHECK();
    int nbig = 0;
    HZ_CHECK(memcpy_sync(&nbig, nlist, sizeof(int), hipMemcpyDeviceToHost));
    std::vector<long long> hoff;
    std::vector<int> cuts;
    long long *toff = nullptr;
    int *tkey = nullptr, *tpos = nullptr;
    if (nbig > 0) {
        long long *tsize = arena().alloc_n<long long>((size_t)nbig + 1);
        toff = arena().alloc_n<long long>((size_t)nbig + 1);
        k_rap_candidates<<<grid_for(nbig), kBlock, 0, st>>>(nbig, list, nc, R.ia, R.ja, A.ia, A.ja, P.ia, tsize);
        HZ_LAUNCH_CHECK();
        size_t tmp_bytes = 0;
        HZ_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, tsize, toff, 0LL, (size_t)nbig + 1,
                                         rocprim::plus<long long>(), st));
        void *tmp = arena().alloc(tmp_bytes);
        HZ_CHECK(hipMemsetAsync(tsize + nbig, 0, sizeof(long long), st));
        HZ_CHECK(rocprim::exclusive_scan(tmp, tmp_bytes, tsize, toff, 0LL, (size_t)nbig + 1,
                                         rocprim::plus<long long>(), st));
        hoff.resize((size_t)nbig + 1);
        HZ_CHECK(memcpy_sync(hoff.data(), toff, sizeof(long long) * ((size_t)nbig + 1), hipMemcpyDeviceToHost));
        // the rows go in chunks whose tables fit the scratch budget
        long long mx = 0;
        for (int i = 0; i < nbig; ++i) mx = std::max(mx, hoff[i + 1] - hoff[i]);
        const long long tables = std::min<long long>(std::max<long long>(1LL << 26, mx), hoff[nbig]);
        tkey = arena().alloc_n<int>((size_t)tables + 1);
        tpos = arena().alloc_n<int>((size_t)tables + 1);
        cuts.push_back(0);
        for (int i = 0; i < nbig; ++i)
            if (hoff[i + 1] - hoff[cuts.back()] > tables) cuts.push_back(i);
        cuts.push_back(nbig);
        for (size_t c = 0; c + 1 < cuts.size(); ++c) {
            const int r0 = cuts[c], r1 = cuts[c + 1];
            k_rap_rows<false><<<grid_for(r1 - r0), kBlock, 0, st>>>(r0, r1, list, toff, hoff[r0], tkey, tpos, R.ia,
                                                                    R.ja, R.val, A.ia, A.ja, A.val, P.ia, P.ja,
                                                                    P.val, rowcount, nullptr, nullptr, nullptr);
            HZ_LAUNCH_CHECK();
        }
    }
    C.nnz = (int)exclusive_scan_int(rowcount, C.ia, nc);
    C.ja = arena().alloc_n<int>((size_t)C.nnz + 1);
    C.val = arena().alloc_n<double>((size_t)C.nnz + 1);
    if (slots) {
        k_rap_compact_slots<<<div_up(nc, kBlock / kWave), kBlock, 0, st>>>(nc, slotcount, max_cols, sja, sval, C.ia,
                                                                           C.ja, C.val);
        HZ_LAUNCH_CHECK();
    } else if (wave) {
        k_rapw<true><<<nc, kWave, 0, st>>>(nc, R.ia, R.ja, R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val, max_pairs,
                                           max_cols, rowcount, C.ia, C.ja, C.val);
        HZ_LAUNCH_CHECK();
    }
    for (size_t c = 0; c + 1 < cuts.size(); ++c) {
        const int r0 = cuts[c], r1 = cuts[c + 1];
        k_rap_rows<true><<<grid_for(r1 - r0), kBlock, 0, st>>>(r0, r1, list, toff, hoff[r0], tkey, tpos, R.ia, R.ja,
                                                               R.val, A.ia, A.ja, A.val, P.ia, P.ja, P.val,
                                                               nullptr, C.ia, C.ja, C.val);
        HZ_LAUNCH_CHECK();
    }
}

void csr_mxv(const DevCSR &A, const double *x, double *y) {
    count_bytes(matrix_bytes(A.nnz, A.n) + 8.0 * A.ncol + 8.0 * A.n);
    if (A.n == 0) return;
    k_csr_mxv<16><<<div_up((long)A.n * 16, kBlock), kBlock, 0, stream()>>>(A.n, A.ia, A.ja, A.val, x, y);
    HZ_LAUNCH_CHECK();
}

void csr_prolong(const DevCSR &P, const double *xc, double alpha, const double *alpha_dev, double *x) {
    count_bytes(matrix_bytes(P.nnz, P.n) + 8.0 * P.ncol + 16.0 * P.n);
    if (P.n == 0) return;
    k_csr_aAxpy<<<div_up(P.n, kBlock), kBlock, 0, stream()>>>(P.n, P.ia, P.ja, P.val, xc, alpha, alpha_dev, x);
    HZ_LAUNCH_CHECK();
}

}  // namespace hz
