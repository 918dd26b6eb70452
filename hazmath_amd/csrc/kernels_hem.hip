// kernels_hem.hip -- HEM aggregation (aggregation_hem, amg_setup_ua.c:913-1051)
// resolved exactly on the GPU.
//
// Pass 1 of the reference walks k = 0..n-1; a still-unmatched k pairs with its
// heaviest unmatched neighbour (ties: the last one in row order, heavy_edg
// scanned backwards, amg_setup_ua.c:864-891,983).  With a symmetric strength
// pattern every lower neighbour of an unmatched k is already matched at time k,
// so k's candidates are its higher neighbours, and
//     taken(j, t) = exists i < t adjacent to j with init(i) and pick(i) = j
//     free(k)     = not isolated(k) and not taken(k, k)
//     pick(k)     = best j > k adjacent to k with not taken(j, k)   (or none)
//     init(k)     = free(k) and pick(k) != none
// These are evaluated as a monotone fixpoint over {unknown, true, false}: a
// round decides every vertex whose inputs are decided, so the result equals
// the sequential sweep.  Pass 2 (leftovers join the smallest neighbouring
// aggregate, first in row order on ties, sizes growing as they join) runs in
// rounds in which a leftover decides only when it is the lowest undecided
// leftover touching each of its candidate aggregates.
#include <climits>
#include <cstring>

#include "amg.hpp"

namespace hz {

namespace {

constexpr int kBlock = 256;
constexpr int kUnk = -3;      // undetermined pick / state
constexpr int kNone = -1;     // no candidate
constexpr int kFreeUnk = 0, kFreeT = 1, kFreeF = 2;

inline int grid_for(long n) {
    long g = (n + kBlock - 1) / kBlock;
    if (g > 8192) g = 8192;
    if (g < 1) g = 1;
    return (int)g;
}

// Pattern symmetry: every off-diagonal (i, j) of N has (j, i).
__global__ void k_sym_check(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            int *__restrict__ asym) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            const int j = ja[k];
            if (j == i) continue;
            bool found = false;
            for (int q = ia[j]; q < ia[j + 1]; ++q)
                if (ja[q] == i) { found = true; break; }
            if (!found) { atomicExch(asym, 1); return; }
        }
    }
}

// Preference list of each row: higher neighbours sorted by |a| descending,
// then by row position descending (heavy_edg's tie rule).
__global__ void k_prefs(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                        const double *__restrict__ a, int *__restrict__ pref, int *__restrict__ npref,
                        int *__restrict__ fstate) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int b = ia[k], e = ia[k + 1];
        fstate[k] = (e - b > 1) ? kFreeUnk : kFreeF;   // isolated: never processed
        int m = 0;
        for (int q = b; q < e; ++q) {
            const int j = ja[q];
            if (j <= k) continue;
            // insertion: keep (|a| desc, position desc)
            const double w = fabs(a[q]);
            int pos = m;
            while (pos > 0) {
                const int qq = pref[b + pos - 1];
                const double wq = fabs(a[qq]);
                if (wq > w || (wq == w && qq > q)) break;
                pref[b + pos] = qq;
                --pos;
            }
            pref[b + pos] = q;
            ++m;
        }
        npref[k] = m;
    }
}

// taken(j, t): 1 true, 0 false, -1 unknown
__device__ __forceinline__ int taken(int j, int t, const int *__restrict__ ia, const int *__restrict__ ja,
                                     const volatile int *fstate, const volatile int *pick) {
    int res = 0;
    for (int q = ia[j]; q < ia[j + 1]; ++q) {
        const int i = ja[q];
        if (i == j || i >= t) continue;
        const int fi = fstate[i];
        if (fi == kFreeF) continue;
        const int pi = pick[i];
        if (pi == kUnk) { res = -1; continue; }
        if (pi != j) continue;          // picks someone else (or none)
        if (fi == kFreeT) return 1;
        res = -1;
    }
    return res;
}

// free(k) evaluation.  Returns 1 taken, 0 free, -1 unknown; when unknown and
// exactly one lower neighbour i can still take k (pick(i) == k decided,
// free(i) unknown), *single = i: then free(k) = not free(i) exactly.
__device__ __forceinline__ int free_eval(int k, const int *__restrict__ ia, const int *__restrict__ ja,
                                         const volatile int *fstate, const volatile int *pick,
                                         int *single) {
    int res = 0, cand = -1, ncand = 0;
    bool ambiguous = false;
    for (int q = ia[k]; q < ia[k + 1]; ++q) {
        const int i = ja[q];
        if (i >= k) continue;
        const int fi = fstate[i];
        if (fi == kFreeF) continue;
        const int pi = pick[i];
        if (pi == kUnk) { res = -1; ambiguous = true; continue; }
        if (pi != k) continue;
        if (fi == kFreeT) return 1;
        res = -1;
        cand = i;
        ++ncand;
    }
    *single = (!ambiguous && ncand == 1) ? cand : -1;
    return res;
}

__global__ void k_hem_round(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const int *__restrict__ pref, const int *__restrict__ npref,
                            int *fstate, int *pick, int *linkpar, int *__restrict__ undecided) {
    int local = 0;
    const int blk = xcd_block_id();
    for (int k = blk * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        int f = ((volatile int *)fstate)[k];
        if (f == kFreeF) continue;
        int p = ((volatile int *)pick)[k];
        if (f == kFreeT && p != kUnk) continue;
        if (f == kFreeUnk) {
            int single = -1;
            const int t = free_eval(k, ia, ja, fstate, pick, &single);
            if (t == 1) { f = kFreeF; fstate[k] = f; }
            else if (t == 0) { f = kFreeT; fstate[k] = f; }
            else if (single >= 0 && ((volatile int *)linkpar)[k] < 0) linkpar[k] = 2 * single + 1;
        }
        if (f != kFreeF && p == kUnk) {
            const int b = ia[k], m = npref[k];
            int np = kNone;
            for (int r = 0; r < m; ++r) {
                const int j = ja[pref[b + r]];
                const int t = taken(j, k, ia, ja, fstate, pick);
                if (t == 1) continue;
                np = (t == 0) ? j : kUnk;
                break;
            }
            if (np != kUnk) { p = np; pick[k] = p; }
        }
        if (f == kFreeUnk || (f == kFreeT && p == kUnk)) ++local;
    }
    if (local) atomicAdd(undecided, local);
}

// Pointer jumping along free(k) = free(a) XOR parity links (packed as
// 2*a + parity so a link is read in one load).
__global__ void k_hem_jump(int n, int *fstate, int *linkpar) {
    const int blk = xcd_block_id();
    for (int k = blk * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int lp = ((volatile int *)linkpar)[k];
        if (lp < 0) continue;
        if (((volatile int *)fstate)[k] != kFreeUnk) { linkpar[k] = -1; continue; }
        const int a = lp >> 1, par = lp & 1;
        const int fa = ((volatile int *)fstate)[a];
        if (fa != kFreeUnk) {
            const bool freek = (fa == kFreeT) != (par == 1);
            fstate[k] = freek ? kFreeT : kFreeF;
            linkpar[k] = -1;
            continue;
        }
        const int la = ((volatile int *)linkpar)[a];
        if (la >= 0) linkpar[k] = ((la >> 1) << 1) | (par ^ (la & 1));
    }
}

__global__ void k_init_flags(int n, const int *__restrict__ fstate, const int *__restrict__ pick,
                             int *__restrict__ flag, int *__restrict__ leftover) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const bool free_k = fstate[k] == kFreeT;
        flag[k] = (free_k && pick[k] >= 0) ? 1 : 0;
        leftover[k] = (free_k && pick[k] == kNone) ? 1 : 0;
    }
}

__global__ void k_assign_pairs(int n, const int *__restrict__ flag, const int *__restrict__ id,
                               const int *__restrict__ pick, int *__restrict__ agg,
                               int *__restrict__ size) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (flag[k]) {
            const int c = id[k];
            agg[k] = c;
            agg[pick[k]] = c;
            size[c] = 2;
        }
    }
}

__global__ void k_fill_i(int *__restrict__ x, int v, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = v;
}

// Pass 2 round: lowest-undecided-leftover claims per aggregate, then decisions.
__global__ void k_left_claim(int nl, const int *__restrict__ list, const int *__restrict__ done,
                             const int *__restrict__ ia, const int *__restrict__ ja,
                             const int *__restrict__ agg, int *__restrict__ owner) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) {
        if (done[t]) continue;
        const int k = list[t];
        for (int q = ia[k]; q < ia[k + 1]; ++q) {
            const int c = agg[ja[q]];
            if (c >= 0) atomicMin(&owner[c], k);
        }
    }
}

// A leftover decides when (a) it is the lowest undecided leftover touching each
// of its candidate aggregates (their sizes are final for it) and (b) no lower
// undecided leftover is adjacent to it (that one's choice would add a
// candidate).  Two leftovers deciding in one round then never interact.
__global__ void k_left_decide(int nl, const int *__restrict__ list, int *__restrict__ done,
                              const int *__restrict__ ia, const int *__restrict__ ja,
                              int *__restrict__ agg, int *__restrict__ size, int *__restrict__ owner,
                              const int *__restrict__ lstate, int *__restrict__ err,
                              int *__restrict__ remaining) {
    int local = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) {
        if (done[t]) continue;
        const int k = list[t];
        bool ready = true;
        int best = -1;
        double bw = 0.0;
        int iz = 0;
        for (int q = ia[k]; q < ia[k + 1]; ++q) {
            const int j = ja[q];
            if (j < k && lstate[j]) ready = false;
            const int c = agg[j];
            if (c < 0) continue;
            if (owner[c] != k) ready = false;
            const double w = (double)(-size[c]);
            if (iz == 0 || w > bw) { bw = w; best = c; }
            ++iz;
        }
        if (!ready) { ++local; continue; }
        if (iz == 0) { atomicExch(err, 1); done[t] = 1; continue; }
        // decided this round; agg[k] is written in k_left_commit so that other
        // leftovers of this round still see k as unassigned
        done[t] = 2 + best;
    }
    if (local) atomicAdd(remaining, local);
}

__global__ void k_left_commit(int nl, const int *__restrict__ list, int *__restrict__ done,
                              int *__restrict__ agg, int *__restrict__ size, int *__restrict__ lstate) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) {
        const int d = done[t];
        if (d >= 2) {
            const int c = d - 2;
            agg[list[t]] = c;
            size[c] += 1;      // c is claimed by this leftover alone in this round
            lstate[list[t]] = 0;
            done[t] = 1;
        }
    }
}

__global__ void k_reset_owner(int nl, const int *__restrict__ list, const int *__restrict__ ia,
                              const int *__restrict__ ja, const int *__restrict__ agg,
                              int *__restrict__ owner) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) {
        const int k = list[t];
        for (int q = ia[k]; q < ia[k + 1]; ++q) {
            const int c = agg[ja[q]];
            if (c >= 0) owner[c] = INT_MAX;
        }
    }
}

__global__ void k_compact_leftovers(int n, const int *__restrict__ leftover, const int *__restrict__ pos,
                                    int *__restrict__ list) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        if (leftover[k]) list[pos[k]] = k;
}

}  // namespace

// Returns SUCCESS, -1017 (reference exit(17)), or 1 when the strength pattern
// is not symmetric (caller falls back to the host sweep).
int hem_aggregate_gpu(const DevCSR &N, int *agg, int *nagg_out) {
    hipStream_t st = stream();
    const int n = N.n;
    const int g = grid_for(n);
    int *flagbuf = arena().alloc_n<int>(4);
    HZ_CHECK(hipMemsetAsync(flagbuf, 0, 4 * sizeof(int), st));
    k_sym_check<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, flagbuf);
    HZ_LAUNCH_CHECK();
    int asym = 0;
    HZ_CHECK(memcpy_sync(&asym, flagbuf, sizeof(int), hipMemcpyDeviceToHost));
    if (asym) return 1;

    int *pref = arena().alloc_n<int>((size_t)N.nnz + 1);
    int *npref = arena().alloc_n<int>((size_t)n + 1);
    int *fstate = arena().alloc_n<int>((size_t)n + 1);
    int *pick = arena().alloc_n<int>((size_t)n + 1);
    k_prefs<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, N.val, pref, npref, fstate);
    k_fill_i<<<g, kBlock, 0, st>>>(pick, kUnk, n);
    HZ_LAUNCH_CHECK();
    int *undecided = flagbuf + 1;
    int *linkpar = arena().alloc_n<int>((size_t)n + 1);
    k_fill_i<<<g, kBlock, 0, st>>>(linkpar, -1, n);
    const int kRoundsPerCheck = 3, kJumps = 4;
    int rounds1 = 0, rounds2 = 0;
    for (int round = 0;; round += kRoundsPerCheck) {
        for (int r = 0; r < kRoundsPerCheck; ++r) {
            HZ_CHECK(hipMemsetAsync(undecided, 0, sizeof(int), st));
            k_hem_round<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, pref, npref, fstate, pick, linkpar,
                                              undecided);
            for (int j = 0; j < kJumps; ++j) k_hem_jump<<<g, kBlock, 0, st>>>(n, fstate, linkpar);
        }
        HZ_LAUNCH_CHECK();
        rounds1 = round + kRoundsPerCheck;
        int und = 0;
        HZ_CHECK(memcpy_sync(&und, undecided, sizeof(int), hipMemcpyDeviceToHost));
        if (und == 0) break;
        if (round > n + 8) return 1;   // cannot happen: the lowest undecided vertex decides every round
    }
    // pass-1 ids: initiators ranked in index order
    int *flag = arena().alloc_n<int>((size_t)n + 1);
    int *leftover = arena().alloc_n<int>((size_t)n + 1);
    int *id = arena().alloc_n<int>((size_t)n + 1);
    k_init_flags<<<g, kBlock, 0, st>>>(n, fstate, pick, flag, leftover);
    HZ_LAUNCH_CHECK();
    const int nc = (int)exclusive_scan_int(flag, id, n);
    int *size = arena().alloc_n<int>((size_t)nc + 1);
    k_fill_i<<<g, kBlock, 0, st>>>(agg, -1, n);
    k_assign_pairs<<<g, kBlock, 0, st>>>(n, flag, id, pick, agg, size);
    HZ_LAUNCH_CHECK();
    // pass 2
    int *lpos = arena().alloc_n<int>((size_t)n + 1);
    const int nl = (int)exclusive_scan_int(leftover, lpos, n);
    if (nl > 0) {
        int *list = arena().alloc_n<int>((size_t)nl + 1);
        int *done = arena().alloc_n<int>((size_t)nl + 1);
        int *owner = arena().alloc_n<int>((size_t)nc + 1);
        k_compact_leftovers<<<g, kBlock, 0, st>>>(n, leftover, lpos, list);
        HZ_CHECK(hipMemsetAsync(done, 0, sizeof(int) * nl, st));
        k_fill_i<<<grid_for(nc), kBlock, 0, st>>>(owner, INT_MAX, nc);
        int *err = flagbuf + 2, *remaining = flagbuf + 3;
        HZ_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));
        // lstate[v] = 1 while v is an undecided leftover (the leftover flags themselves)
        int *lstate = leftover;
        const int gl = grid_for(nl);
        for (int round = 0;; ++round) {
            HZ_CHECK(hipMemsetAsync(remaining, 0, sizeof(int), st));
            k_left_claim<<<gl, kBlock, 0, st>>>(nl, list, done, N.ia, N.ja, agg, owner);
            k_left_decide<<<gl, kBlock, 0, st>>>(nl, list, done, N.ia, N.ja, agg, size, owner, lstate,
                                                 err, remaining);
            k_reset_owner<<<gl, kBlock, 0, st>>>(nl, list, N.ia, N.ja, agg, owner);
            k_left_commit<<<gl, kBlock, 0, st>>>(nl, list, done, agg, size, lstate);
            HZ_LAUNCH_CHECK();
            int hr[2] = {0, 0};
            HZ_CHECK(memcpy_sync(hr, err, 2 * sizeof(int), hipMemcpyDeviceToHost));
            rounds2 = round + 1;
            if (hr[0]) return -1017;
            if (hr[1] == 0) break;
            if (round > nl + 8) return 1;
        }
    }
    *nagg_out = nc;
    if (trace_enabled())
        std::fprintf(stderr, "[hz]   HEM n=%d: pass-1 rounds %d, leftovers %d, pass-2 rounds %d\n", n,
                     rounds1, nl, rounds2);
    return HZ_SUCCESS;
}

}  // namespace hz
