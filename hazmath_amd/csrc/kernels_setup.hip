// kernels_setup.hip -- AMG setup kernels on gfx950: strength graph, HEC
// aggregation, restriction lists and the Galerkin product with unit R/P.
//
// All results are bit-identical to the sequential reference
// (src/solver/amg_setup_ua.c, src/utilities/sparse.c):
//  * the strength graph keeps each row's CSR order;
//  * HEC is resolved exactly from the turn order (see hec_aggregate);
//  * R lists each aggregate's members in ascending fine index (stable sort);
//  * the coarse row layout (diagonal first, then columns in first-discovery
//    order) and every coarse value (a sum of fine entries in discovery order)
//    are reproduced, so coarser levels aggregate identically.
#include <climits>

#include <cstring>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_radix_sort.hpp>

#include "hz_internal.h"

namespace hz {

namespace {

constexpr int kBlock = 256;

// ---------------------------------------------------------------------------
// Strength graph: N = {j == i} U {a_ij < 0 and a_ij^2 >= theta^2 |a_ii a_jj|}
// with a_ii the FIRST stored diagonal entry (dcsr_getdiag, sparse.c:2168-2189).
// ---------------------------------------------------------------------------
__global__ void k_first_diag(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ val, double *__restrict__ d) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double di = 0.0;
        for (int k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] == i) { di = val[k]; break; }
        d[i] = di;
    }
}

// Rows by groups of kStrLanes lanes: a round's entries sit on consecutive
// lanes (coalesced loads), kept entries are ranked with a ballot and written
// in stored order.  count pass (kFill false), then fill at nia.
constexpr int kStrLanes = 16;
template <bool kFill>
__global__ __launch_bounds__(kBlock) void k_strength_grp(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                         const double *__restrict__ val, const double *__restrict__ d,
                                                         double th2c, int *__restrict__ count,
                                                         const int *__restrict__ nia, int *__restrict__ nja,
                                                         double *__restrict__ nval, const double *__restrict__ th2_row) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(gt / kStrLanes), j = (int)(gt % kStrLanes);
    const int lane = threadIdx.x % kWave, gbase = lane - j;
    if (i >= n) return;   // the whole group
    const double di = d[i];
    const double th2 = th2_row ? th2_row[i] : th2c;
    const int k0 = ia[i], k1 = ia[i + 1];
    const unsigned long long gmask = ((1ull << kStrLanes) - 1ull) << gbase;
    const unsigned long long below = (1ull << lane) - 1ull;
    int pos = kFill ? nia[i] : 0;
    int cnt = 0;
    for (int base = k0; base < k1; base += kStrLanes) {   // trip count uniform within the group
        const int k = base + j;
        bool keep = false;
        int col = -1;
        double a = 0.0;
        if (k < k1) {
            col = ja[k];
            a = val[k];
            const double dj = d[col];
            keep = col == i || ((a * a) >= th2 * fabs(di * dj) && a < 0e0);
        }
        const unsigned long long m = __ballot(keep) & gmask;
        if (kFill && keep) {
            const int r = __popcll(m & below);
            nja[pos + r] = col;
            nval[pos + r] = a;
        }
        const int c = __popcll(m);
        pos += c;
        cnt += c;
    }
    if (!kFill && j == 0) count[i] = cnt;
}

// ---------------------------------------------------------------------------
// HEC (aggregation_hec, amg_setup_ua.c:752-845) resolved in parallel.
//
// In the sequential loop, vertex i (turn t(i)) with strongest neighbour j:
//  - t(j) < t(i): i copies j's final value                       (FOLLOW j)
//  - t(j) > t(i) and i is the first vertex (in turn order) whose target is j:
//    i opens a new aggregate for {i, j}                          (CREATOR)
//  - t(j) > t(i) otherwise: j still holds the id opened by that first
//    vertex m0, and i takes it                                    (JOIN m0)
// A vertex's own turn always overwrites its value, so FOLLOW chains end at
// a CREATOR, a JOIN or an isolated vertex, and ids are the creators' ranks in
// turn order.
// ---------------------------------------------------------------------------
enum : int { kStUnpt = 0, kStCreator = 1, kStJoin = 2, kStFollow = 3 };

__global__ void k_hec_target(int n, const int *__restrict__ nia, const int *__restrict__ nja,
                             const double *__restrict__ nval, int *__restrict__ target,
                             int *__restrict__ err) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int b = nia[i], e = nia[i + 1];
        if (e - b == 1) { target[i] = -1; continue; }
        int best = -1;
        double m = 0.0;
        for (int k = b; k < e; ++k)
            if (nja[k] != i && fabs(nval[k]) > m) { best = k; m = fabs(nval[k]); }
        if (best < 0) {
            m = -1e0;
            for (int k = b; k < e; ++k)
                if (nja[k] != i && fabs(nval[k]) > m) { best = k; m = fabs(nval[k]); }
        }
        if (best < 0) {
            atomicExch(err, 1);
            target[i] = -2;
            continue;
        }
        target[i] = nja[best];
    }
}

__global__ void k_turn_pos(int n, const int *__restrict__ perm, int *__restrict__ tpos) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x)
        tpos[perm[t]] = t;
}

__global__ void k_fill_int(int *__restrict__ x, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

__global__ void k_hec_first(int n, const int *__restrict__ target, const int *__restrict__ tpos,
                            int *__restrict__ first) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int j = target[i];
        if (j >= 0) atomicMin(&first[j], tpos[i]);
    }
}

__global__ void k_hec_classify(int n, const int *__restrict__ target, const int *__restrict__ tpos,
                               const int *__restrict__ first, int *__restrict__ state,
                               int *__restrict__ cflag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int j = target[i];
        int st;
        if (j < 0) st = kStUnpt;
        else if (tpos[j] < tpos[i]) st = kStFollow;
        else if (first[j] == tpos[i]) st = kStCreator;
        else st = kStJoin;
        state[i] = st;
        cflag[tpos[i]] = (st == kStCreator) ? 1 : 0;
    }
}

__global__ void k_hec_resolve(int n, const int *__restrict__ target, const int *__restrict__ tpos,
                              const int *__restrict__ first, const int *__restrict__ state,
                              const int *__restrict__ cid, int *__restrict__ agg) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int v = i;
        while (state[v] == kStFollow) v = target[v];
        const int st = state[v];
        int a;
        if (st == kStUnpt) a = HZ_UNPT;
        else if (st == kStCreator) a = cid[tpos[v]];
        else a = cid[first[target[v]]];
        agg[i] = a;
    }
}

// ---------------------------------------------------------------------------
// R = P^T
// ---------------------------------------------------------------------------
__global__ void k_agg_keys(int n, const int *__restrict__ agg, int nagg, unsigned int *__restrict__ keys,
                           int *__restrict__ vals) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int a = agg[i];
        keys[i] = (a > HZ_UNPT) ? (unsigned int)a : (unsigned int)nagg;
        vals[i] = i;
    }
}

__global__ void k_agg_count(int n, const int *__restrict__ agg, int *__restrict__ cnt) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int a = agg[i];
        if (a > HZ_UNPT) atomicAdd(&cnt[a], 1);
    }
}

// ---------------------------------------------------------------------------
// Galerkin product with unit R/P (dcsr_rap_agg, sparse.c:2557-2683).
// One wavefront per coarse row c.  The row's candidate sequence is the scan
//   for member i1 of c (ascending), for k in A row i1 (stored order):
//       (agg[A.ja[k]], A.val[k])       -- skipped when agg == UNPT
// Output: slot 0 = (c, 0.0 + sum of candidates with column c, in order),
// then one slot per other column in order of first appearance, whose value is
// the in-order sum of its candidates starting from the first one.
// ---------------------------------------------------------------------------
// Rows up to kRapSmall candidates (HEM pairs of 15-point rows have ~30) are
// built in one pass, two per wavefront, into fixed slots (k_rap_pair) and
// moved to their CSR places after the scan; longer rows up to kRapCap run in
// a large-LDS wavefront per row (count, then fill), longer ones in global
// scratch.  The hash tables have twice the slots of the candidate capacity.
constexpr int kRapSmall = 64;
constexpr int kRapCap = 512;

__device__ __forceinline__ unsigned rap_hash(int key, unsigned mask) {
    unsigned h = (unsigned)key * 2654435761u;
    return (h ^ (h >> 15)) & mask;
}

// Slot of `key` (inserted before) in an open-addressing table.
__device__ __forceinline__ unsigned rap_find(const int *keys, int key, unsigned mask) {
    unsigned slot = rap_hash(key, mask);
    while (keys[slot] != key) slot = (slot + 1) & mask;
    return slot;
}

template <bool kFill, bool kGlobal, int kCap>
__global__ __launch_bounds__(kWave) void k_rap_row(
    int nrows, const int *__restrict__ rowlist, const int *__restrict__ rptr,
    const int *__restrict__ rmem, const int *__restrict__ aia, const int *__restrict__ aja,
    const double *__restrict__ aval, const int *__restrict__ agg, int *__restrict__ rowcount,
    int *__restrict__ rowlen, const int *__restrict__ cia, int *__restrict__ cja,
    double *__restrict__ cval, int *__restrict__ gcol, double *__restrict__ gval,
    const long long *__restrict__ goff, int *__restrict__ gkeys, int *__restrict__ gminp,
    const long long *__restrict__ hoff) {
    constexpr int kHash = 2 * kCap;
    __shared__ int s_col[kGlobal ? 1 : kCap];
    __shared__ double s_val[(kGlobal || !kFill) ? 1 : kCap];
    __shared__ int s_keys[kGlobal ? 1 : kHash];
    __shared__ int s_minp[kGlobal ? 1 : kHash];
    __shared__ int s_maxp[kGlobal ? 1 : kHash];
    __shared__ int s_diag;
    const int lane = threadIdx.x;
    const int idx = blockIdx.x;
    if (idx >= nrows) return;
    const int c = rowlist ? rowlist[idx] : idx;
    int *col = kGlobal ? gcol + goff[idx] : s_col;
    double *val = kGlobal ? gval + goff[idx] : s_val;
    // 1) enumerate candidates (members ascending x stored row order)
    int L = 0;
    const int q1 = rptr[c + 1];
    for (int q = rptr[c]; q < q1; ++q) {
        const int i1 = rmem[q];
        const int b = aia[i1], len = aia[i1 + 1] - b;
        for (int k = lane; k < len; k += kWave) {
            const int p = L + k;
            if (kGlobal || p < kCap) {
                col[p] = agg[aja[b + k]];
                if (kFill) val[p] = aval[b + k];
            }
        }
        L += len;
    }
    if (!kGlobal && L > kCap) {   // handled by a larger variant
        if (!kFill && lane == 0) { rowcount[c] = -1; rowlen[c] = L; }
        return;
    }
    const long long obase = kFill ? (long long)cia[c] : 0;
    // 2) hash table of distinct coarse columns -> first position (LDS
    //    variants: also the last position, and the diagonal column c)
    unsigned mask;
    int *keys, *minp;
    if (kGlobal) {
        const long long h0 = hoff[idx];
        mask = (unsigned)(hoff[idx + 1] - h0) - 1u;
        keys = gkeys + h0;
        minp = gminp + h0;
    } else {
        mask = kHash - 1u;
        keys = s_keys;
        minp = s_minp;
    }
    for (unsigned t = lane; t <= mask; t += kWave) {
        keys[t] = -1;
        minp[t] = INT_MAX;
        if (!kGlobal) s_maxp[t] = -1;
    }
    if (lane == 0) s_diag = 0;
    __syncthreads();
    for (int p = lane; p < L; p += kWave) {
        const int key = col[p];
        if (key <= HZ_UNPT || (kGlobal && key == c)) continue;
        unsigned slot = rap_hash(key, mask);
        while (true) {
            const int prev = atomicCAS(&keys[slot], -1, key);
            if (prev == -1 || prev == key) break;
            slot = (slot + 1) & mask;
        }
        atomicMin(&minp[slot], p);
        if (!kGlobal) atomicMax(&s_maxp[slot], p);
    }
    __syncthreads();
    // 3) first appearances, their ranks and (fill) in-order sums; a sum
    //    scans from the column's first to its last position
    int nfirst_before = 0;
    for (int base = 0; base < L; base += kWave) {
        const int p = base + lane;
        bool first = false, dfirst = false;
        int cp = -1, last = L - 1;
        if (p < L) {
            cp = col[p];
            if (cp > HZ_UNPT && (!kGlobal || cp != c)) {
                const unsigned slot = rap_find(keys, cp, mask);
                const bool f = minp[slot] == p;
                if (!kGlobal) last = s_maxp[slot];
                first = f && cp != c;
                dfirst = f && cp == c;
            }
        }
        const unsigned long long m = __ballot(first);
        if (kFill && (first || dfirst)) {
            // the diagonal's sum starts from 0.0, the others from the first value
            double v = dfirst ? 0.0 + val[p] : val[p];
            for (int q = p + 1; q <= last; ++q)
                if (col[q] == cp) v = v + val[q];
            const long long pos = dfirst ? obase : obase + 1 + nfirst_before + __popcll(m & ((1ull << lane) - 1ull));
            cja[pos] = cp;
            cval[pos] = v;
            if (dfirst) s_diag = 1;
        }
        nfirst_before += __popcll(m);
    }
    if (kFill && !kGlobal) __syncthreads();
    if (lane == 0) {
        if (kFill && kGlobal) {
            double v = 0.0;
            for (int q = 0; q < L; ++q)
                if (col[q] == c) v = v + val[q];
            cja[obase] = c;
            cval[obase] = v;
        } else if (kFill && !s_diag) {   // no candidate in column c
            cja[obase] = c;
            cval[obase] = 0.0;
        }
        if (!kFill) {
            rowcount[c] = 1 + nfirst_before;
            if (!kGlobal) rowlen[c] = L;
        }
    }
}

// Rows of at most kRapSmall candidates in one pass, two coarse rows per
// wavefront (32 lanes and their own LDS tables each): these rows are short
// and a wavefront's time is its chain of dependent accesses, so two rows
// share it.  Same candidate order, hash, ranks and in-order sums as
// k_rap_row; each row goes to its slots tcol/tval + c * kRapSmall (moved to
// its CSR place by k_rap_compact after the scan) and its count and candidate
// length to rowcount/rowlen (-1 and the length when it is longer).
__global__ __launch_bounds__(kWave) void k_rap_pair(int nrows, const int *__restrict__ rptr,
                                                    const int *__restrict__ rmem, const int *__restrict__ aia,
                                                    const int *__restrict__ aja, const double *__restrict__ aval,
                                                    const int *__restrict__ agg, int *__restrict__ rowcount,
                                                    int *__restrict__ rowlen, int *__restrict__ tcol,
                                                    double *__restrict__ tval) {
    constexpr int kH = kWave / 2, kCap = kRapSmall, kHash = 2 * kRapSmall;
    __shared__ int s_col[2][kCap];
    __shared__ double s_val[2][kCap];
    __shared__ int s_keys[2][kHash], s_minp[2][kHash], s_maxp[2][kHash];
    __shared__ int s_diag[2];
    const int half = threadIdx.x / kH, hl = threadIdx.x % kH;
    const int c = blockIdx.x * 2 + half;
    const bool valid = c < nrows;
    int *col = s_col[half];
    double *val = s_val[half];
    int *keys = s_keys[half], *minp = s_minp[half], *maxp = s_maxp[half];
    // 1) enumerate candidates (members ascending x stored row order)
    int L = 0;
    if (valid) {
        const int q1 = rptr[c + 1];
        for (int q = rptr[c]; q < q1; ++q) {
            const int i1 = rmem[q];
            const int b = aia[i1], len = aia[i1 + 1] - b;
            for (int k = hl; k < len; k += kH) {
                const int p = L + k;
                if (p < kCap) {
                    col[p] = agg[aja[b + k]];
                    val[p] = aval[b + k];
                }
            }
            L += len;
        }
    }
    const bool over = L > kCap;   // handled by the larger variants
    if (valid && over && hl == 0) {
        rowcount[c] = -1;
        rowlen[c] = L;
    }
    const bool act = valid && !over;
    const int La = act ? L : 0;
    const long long obase = (long long)c * kCap;
    // 2) hash of distinct coarse columns (the diagonal included) -> first and last position
    for (int t = hl; t < kHash; t += kH) {
        keys[t] = -1;
        minp[t] = INT_MAX;
        maxp[t] = -1;
    }
    if (hl == 0) s_diag[half] = 0;
    __syncthreads();
    const unsigned mask = kHash - 1u;
    for (int p = hl; p < La; p += kH) {
        const int key = col[p];
        if (key <= HZ_UNPT) continue;
        unsigned slot = rap_hash(key, mask);
        while (true) {
            const int prev = atomicCAS(&keys[slot], -1, key);
            if (prev == -1 || prev == key) break;
            slot = (slot + 1) & mask;
        }
        atomicMin(&minp[slot], p);
        atomicMax(&maxp[slot], p);
    }
    __syncthreads();
    // 3) first appearances in rounds of kH (trip count uniform over the wavefront)
    int lw = La;
    lw = max(lw, __shfl_xor(lw, kH, kWave));
    int nfirst_before = 0;
    const unsigned long long hmask = 0xffffffffull << (half * kH);
    for (int base = 0; base < lw; base += kH) {
        const int p = base + hl;
        bool first = false, dfirst = false;
        int cp = -1, last = -1;
        if (p < La) {
            cp = col[p];
            if (cp > HZ_UNPT) {
                const unsigned slot = rap_find(keys, cp, mask);
                const bool f = minp[slot] == p;
                last = maxp[slot];
                first = f && cp != c;
                dfirst = f && cp == c;
            }
        }
        const unsigned long long m = __ballot(first) & hmask;
        if (first || dfirst) {
            double v = dfirst ? 0.0 + val[p] : val[p];
            for (int q = p + 1; q <= last; ++q)
                if (col[q] == cp) v = v + val[q];
            const long long pos = dfirst ? obase
                                         : obase + 1 + nfirst_before +
                                               __popcll(m & ((1ull << (half * kH + hl)) - 1ull) & hmask);
            tcol[pos] = cp;
            tval[pos] = v;
            if (dfirst) s_diag[half] = 1;
        }
        nfirst_before += __popcll(m);
    }
    __syncthreads();
    if (act && hl == 0) {
        if (!s_diag[half]) {   // no candidate in column c
            tcol[obase] = c;
            tval[obase] = 0.0;
        }
        rowcount[c] = 1 + nfirst_before;
        rowlen[c] = L;
    }
}

// Rows built by the one-pass kernel (rowlen <= kCap) to their CSR places,
// kGroup lanes per row.
template <int kCap, int kGroup>
__global__ void k_rap_compact(int nc, const int *__restrict__ rowlen, const int *__restrict__ cia,
                              const int *__restrict__ tcol, const double *__restrict__ tval,
                              int *__restrict__ cja, double *__restrict__ cval) {
    const long gt = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = (int)(gt / kGroup), j = (int)(gt % kGroup);
    if (c >= nc || rowlen[c] > kCap) return;
    const int b = cia[c], cnt = cia[c + 1] - b;
    const long long t = (long long)c * kCap;
    for (int k = j; k < cnt; k += kGroup) {
        cja[b + k] = tcol[t + k];
        cval[b + k] = tval[t + k];
    }
}

__global__ void k_collect_overflow(int nc, const int *__restrict__ rowcount, int *__restrict__ list,
                                   int *__restrict__ nlist) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x)
        if (rowcount[c] < 0) list[atomicAdd(nlist, 1)] = c;
}

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 4096) g = 4096;
    if (g < 1) g = 1;
    return (int)g;
}

}  // namespace

void exclusive_scan_int_device(const int *in, int *out, int n) {
    hipStream_t st = stream();
    size_t tmp_bytes = 0;
    // out has n+1 entries; scan n+1 inputs where the last input is treated as 0
    int *in_ext = arena().alloc_n<int>((size_t)n + 1);
    if (n > 0) HZ_CHECK(hipMemcpyAsync(in_ext, in, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
    HZ_CHECK(hipMemsetAsync(in_ext + n, 0, sizeof(int), st));
    HZ_CHECK(rocprim::exclusive_scan(nullptr, tmp_bytes, in_ext, out, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), st));
    void *tmp = arena().alloc(tmp_bytes);
    HZ_CHECK(rocprim::exclusive_scan(tmp, tmp_bytes, in_ext, out, 0, (size_t)n + 1,
                                     rocprim::plus<int>(), st));
}

long long exclusive_scan_int(const int *in, int *out, int n) {
    hipStream_t st = stream();
    exclusive_scan_int_device(in, out, n);
    int total = 0;
    HZ_CHECK(hipMemcpyAsync(&total, out + n, sizeof(int), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    return total;
}

namespace {
void strength_impl(const DevCSR &A, double th2, const double *th2_row, DevCSR &N, int *counts_ws,
                   const double *dfirst) {
    hipStream_t st = stream();
    const int n = A.n;
    const double *d = dfirst;
    if (!d) {
        double *dd = arena().alloc_n<double>((size_t)n + 1);
        k_first_diag<<<grid_for(n), kBlock, 0, st>>>(n, A.ia, A.ja, A.val, dd);
        HZ_LAUNCH_CHECK();
        d = dd;
    }
    const int gg = div_up((long)n * kStrLanes, kBlock);
    k_strength_grp<false><<<gg, kBlock, 0, st>>>(n, A.ia, A.ja, A.val, d, th2, counts_ws, nullptr, nullptr, nullptr,
                                                 th2_row);
    HZ_LAUNCH_CHECK();
    N.n = n;
    N.ncol = A.ncol;
    N.ia = arena().alloc_n<int>((size_t)n + 1);
    N.nnz = (int)exclusive_scan_int(counts_ws, N.ia, n);
    N.ja = arena().alloc_n<int>((size_t)N.nnz + 1);
    N.val = arena().alloc_n<double>((size_t)N.nnz + 1);
    k_strength_grp<true><<<gg, kBlock, 0, st>>>(n, A.ia, A.ja, A.val, d, th2, nullptr, N.ia, N.ja, N.val, th2_row);
    HZ_LAUNCH_CHECK();
}
}  // namespace

void strength_graph(const DevCSR &A, double theta, DevCSR &N, int *counts_ws, const double *dfirst) {
    strength_impl(A, theta * theta, nullptr, N, counts_ws, dfirst);
}

void strength_graph_rows(const DevCSR &A, const double *th2_row, DevCSR &N, int *counts_ws) {
    strength_impl(A, 0.0, th2_row, N, counts_ws, nullptr);
}

int hec_aggregate(const DevCSR &N, const int *perm, int *agg) {
    hipStream_t st = stream();
    const int n = N.n;
    int *target = arena().alloc_n<int>((size_t)n + 1);
    int *tpos = arena().alloc_n<int>((size_t)n + 1);
    int *first = arena().alloc_n<int>((size_t)n + 1);
    int *state = arena().alloc_n<int>((size_t)n + 1);
    int *cflag = arena().alloc_n<int>((size_t)n + 1);
    int *cid = arena().alloc_n<int>((size_t)n + 1);
    int *err = arena().alloc_n<int>(1);
    HZ_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));
    const int g = grid_for(n);
    k_hec_target<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, N.val, target, err);
    k_turn_pos<<<g, kBlock, 0, st>>>(n, perm, tpos);
    k_fill_int<<<g, kBlock, 0, st>>>(first, INT_MAX, n);
    k_hec_first<<<g, kBlock, 0, st>>>(n, target, tpos, first);
    k_hec_classify<<<g, kBlock, 0, st>>>(n, target, tpos, first, state, cflag);
    HZ_LAUNCH_CHECK();
    int herr = 0;
    HZ_CHECK(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
    const int nagg = (int)exclusive_scan_int(cflag, cid, n);   // synchronizes
    if (herr) return -1016;
    k_hec_resolve<<<g, kBlock, 0, st>>>(n, target, tpos, first, state, cid, agg);
    HZ_LAUNCH_CHECK();
    return nagg;
}

void build_restriction(int n, const int *agg, int nagg, int *rptr, int *rmem) {
    hipStream_t st = stream();
    int *cnt = arena().alloc_n<int>((size_t)nagg + 1);
    HZ_CHECK(hipMemsetAsync(cnt, 0, sizeof(int) * ((size_t)nagg + 1), st));
    k_agg_count<<<grid_for(n), kBlock, 0, st>>>(n, agg, cnt);
    HZ_LAUNCH_CHECK();
    exclusive_scan_int_device(cnt, rptr, nagg);
    // stable radix sort of (aggregate, vertex) pairs: members come out ascending
    unsigned int *keys = arena().alloc_n<unsigned int>((size_t)n + 1);
    unsigned int *keys_out = arena().alloc_n<unsigned int>((size_t)n + 1);
    int *vals = arena().alloc_n<int>((size_t)n + 1);
    int *vals_out = arena().alloc_n<int>((size_t)n + 1);
    k_agg_keys<<<grid_for(n), kBlock, 0, st>>>(n, agg, nagg, keys, vals);
    HZ_LAUNCH_CHECK();
    int bits = 1;
    while ((1u << bits) <= (unsigned)nagg && bits < 32) ++bits;
    size_t tmp_bytes = 0;
    HZ_CHECK(rocprim::radix_sort_pairs(nullptr, tmp_bytes, keys, keys_out, vals, vals_out, (size_t)n,
                                       0, bits, st));
    void *tmp = arena().alloc(tmp_bytes);
    HZ_CHECK(rocprim::radix_sort_pairs(tmp, tmp_bytes, keys, keys_out, vals, vals_out, (size_t)n, 0,
                                       bits, st));
    HZ_CHECK(hipMemcpyAsync(rmem, vals_out, sizeof(int) * (size_t)n, hipMemcpyDeviceToDevice, st));
}

void rap_agg(const DevCSR &A, const int *agg, int nagg, const int *rptr, const int *rmem,
             DevCSR &Ac) {
    hipStream_t st = stream();
    int *rowcount = arena().alloc_n<int>((size_t)nagg + 1);
    int *rowlen = arena().alloc_n<int>((size_t)nagg + 1);
    int *olist = arena().alloc_n<int>((size_t)nagg + 1);
    int *nol = arena().alloc_n<int>(1);
    // rows whose candidate list exceeds the capacity of the pass just run
    auto overflow = [&]() {
        HZ_CHECK(hipMemsetAsync(nol, 0, sizeof(int), st));
        k_collect_overflow<<<grid_for(nagg), kBlock, 0, st>>>(nagg, rowcount, olist, nol);
        HZ_LAUNCH_CHECK();
        int h = 0;
        HZ_CHECK(hipMemcpyAsync(&h, nol, sizeof(int), hipMemcpyDeviceToHost, st));
        HZ_CHECK(hipStreamSynchronize(st));
        return h;
    };
    // the small-LDS kernel over all rows (one pass), the large-LDS one over
    // its overflow (count, then fill); what still overflows goes to global
    // scratch
    int n1 = 0, n2 = 0;
    int *list1 = nullptr, *list2 = nullptr;
    // rows up to kRapSmall candidates are built in one pass into slots of
    // kRapSmall entries and moved to their CSR places after the scan
    int *tcol = arena().alloc_n<int>((size_t)nagg * kRapSmall + 1);
    double *tval = arena().alloc_n<double>((size_t)nagg * kRapSmall + 1);
    Ac.n = nagg;
    Ac.ncol = nagg;
    Ac.ia = arena().alloc_n<int>((size_t)nagg + 1);
    long long total = 0;
    if (nagg > 0) {
        k_rap_pair<<<div_up(nagg, 2), kWave, 0, st>>>(nagg, rptr, rmem, A.ia, A.ja, A.val, agg, rowcount, rowlen,
                                                      tcol, tval);
        HZ_LAUNCH_CHECK();
        HZ_CHECK(hipMemsetAsync(nol, 0, sizeof(int), st));
        k_collect_overflow<<<grid_for(nagg), kBlock, 0, st>>>(nagg, rowcount, olist, nol);
        HZ_LAUNCH_CHECK();
        // the row pointers are final unless a row overflowed: one host wait for both
        exclusive_scan_int_device(rowcount, Ac.ia, nagg);
        int h[2] = {0, 0};
        HZ_CHECK(hipMemcpyAsync(&h[0], nol, sizeof(int), hipMemcpyDeviceToHost, st));
        HZ_CHECK(hipMemcpyAsync(&h[1], Ac.ia + nagg, sizeof(int), hipMemcpyDeviceToHost, st));
        HZ_CHECK(hipStreamSynchronize(st));
        n1 = h[0];
        total = h[1];
    } else {
        HZ_CHECK(hipMemsetAsync(Ac.ia, 0, sizeof(int), st));
    }
    if (n1 > 0) {
        list1 = arena().alloc_n<int>((size_t)n1 + 1);
        HZ_CHECK(hipMemcpyAsync(list1, olist, sizeof(int) * n1, hipMemcpyDeviceToDevice, st));
        k_rap_row<false, false, kRapCap><<<n1, kWave, 0, st>>>(n1, list1, rptr, rmem, A.ia, A.ja, A.val, agg,
                                                                rowcount, rowlen, nullptr, nullptr, nullptr,
                                                                nullptr, nullptr, nullptr, nullptr, nullptr,
                                                                nullptr);
        HZ_LAUNCH_CHECK();
        n2 = overflow();
    }
    int *gcol = nullptr;
    double *gval = nullptr;
    long long *goff = nullptr;
    int *gkeys = nullptr, *gminp = nullptr;
    long long *hoffd = nullptr;
    if (n2 > 0) {
        list2 = arena().alloc_n<int>((size_t)n2 + 1);
        HZ_CHECK(hipMemcpyAsync(list2, olist, sizeof(int) * n2, hipMemcpyDeviceToDevice, st));
        // each overflow row is independent: the order of the list does not matter
        std::vector<int> hl(n2);
        HZ_CHECK(memcpy_sync(hl.data(), list2, sizeof(int) * n2, hipMemcpyDeviceToHost));
        std::vector<long long> hoff(n2 + 1, 0), hh(n2 + 1, 0);
        for (int i = 0; i < n2; ++i) {
            int L = 0;
            HZ_CHECK(memcpy_sync(&L, rowlen + hl[i], sizeof(int), hipMemcpyDeviceToHost));
            hoff[i + 1] = hoff[i] + L;
            long long H = 1;
            while (H < 2LL * L) H <<= 1;
            hh[i + 1] = hh[i] + H;
        }
        hoffd = arena().alloc_n<long long>((size_t)n2 + 1);
        HZ_CHECK(memcpy_sync(hoffd, hh.data(), sizeof(long long) * (n2 + 1), hipMemcpyHostToDevice));
        gkeys = arena().alloc_n<int>((size_t)hh[n2] + 1);
        gminp = arena().alloc_n<int>((size_t)hh[n2] + 1);
        goff = arena().alloc_n<long long>((size_t)n2 + 1);
        HZ_CHECK(memcpy_sync(goff, hoff.data(), sizeof(long long) * (n2 + 1), hipMemcpyHostToDevice));
        gcol = arena().alloc_n<int>((size_t)hoff[n2] + 1);
        gval = arena().alloc_n<double>((size_t)hoff[n2] + 1);
        k_rap_row<false, true, 1><<<n2, kWave, 0, st>>>(n2, list2, rptr, rmem, A.ia, A.ja, A.val, agg, rowcount,
                                                         rowlen, nullptr, nullptr, nullptr, gcol, gval, goff, gkeys,
                                                         gminp, hoffd);
        HZ_LAUNCH_CHECK();
    }
    if (n1 > 0) total = exclusive_scan_int(rowcount, Ac.ia, nagg);
    Ac.nnz = (int)total;
    Ac.ja = arena().alloc_n<int>((size_t)Ac.nnz + 1);
    Ac.val = arena().alloc_n<double>((size_t)Ac.nnz + 1);
    if (nagg > 0) {
        constexpr int kGroup = 16;
        k_rap_compact<kRapSmall, kGroup><<<div_up((long)nagg * kGroup, kBlock), kBlock, 0, st>>>(
            nagg, rowlen, Ac.ia, tcol, tval, Ac.ja, Ac.val);
        HZ_LAUNCH_CHECK();
    }
    if (n1 > 0) {
        k_rap_row<true, false, kRapCap><<<n1, kWave, 0, st>>>(n1, list1, rptr, rmem, A.ia, A.ja, A.val, agg,
                                                               nullptr, nullptr, Ac.ia, Ac.ja, Ac.val, nullptr,
                                                               nullptr, nullptr, nullptr, nullptr, nullptr);
        HZ_LAUNCH_CHECK();
    }
    if (n2 > 0) {
        k_rap_row<true, true, 1><<<n2, kWave, 0, st>>>(n2, list2, rptr, rmem, A.ia, A.ja, A.val, agg, nullptr,
                                                        nullptr, Ac.ia, Ac.ja, Ac.val, gcol, gval, goff, gkeys,
                                                        gminp, hoffd);
        HZ_LAUNCH_CHECK();
    }
}

void csr_transpose(const DevCSR &A, DevCSR &AT) {
    // The coarsest matrix is small: transpose on the host with the counting
    // sort of dcsr_trans (sparse.c:770-830) and upload.
    const int n = A.n, m = A.ncol, nnz = A.nnz;
    std::vector<int> ia(n + 1), ja(nnz);
    std::vector<double> va(nnz);
    HZ_CHECK(memcpy_sync(ia.data(), A.ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
    if (nnz) {
        HZ_CHECK(memcpy_sync(ja.data(), A.ja, sizeof(int) * nnz, hipMemcpyDeviceToHost));
        HZ_CHECK(memcpy_sync(va.data(), A.val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    }
    std::vector<int> tia(m + 1, 0), tja(nnz);
    std::vector<double> tva(nnz);
    for (int k = 0; k < nnz; ++k) tia[ja[k] + 1]++;
    for (int i = 0; i < m; ++i) tia[i + 1] += tia[i];
    std::vector<int> nxt(tia.begin(), tia.end());
    for (int i = 0; i < n; ++i)
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            const int pos = nxt[ja[k]]++;
            tja[pos] = i;
            tva[pos] = va[k];
        }
    AT.n = m;
    AT.ncol = n;
    AT.nnz = nnz;
    AT.ia = arena().alloc_n<int>((size_t)m + 1);
    AT.ja = arena().alloc_n<int>((size_t)nnz + 1);
    AT.val = arena().alloc_n<double>((size_t)nnz + 1);
    HZ_CHECK(memcpy_sync(AT.ia, tia.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice));
    if (nnz) {
        HZ_CHECK(memcpy_sync(AT.ja, tja.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
        HZ_CHECK(memcpy_sync(AT.val, tva.data(), sizeof(double) * nnz, hipMemcpyHostToDevice));
    }
}

}  // namespace hz
