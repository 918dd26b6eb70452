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
#include <algorithm>
#include <climits>
#include <cstring>

#include "amg.hpp"

namespace hz {

namespace {

constexpr int kBlock = 256;
constexpr int kUnk = -3;      // undetermined pick / state
constexpr int kNone = -1;     // no candidate
constexpr int kFreeUnk = 0, kFreeT = 1, kFreeF = 2;

// State words (fstate, pick, linkpar) are read and written with relaxed
// agent-scope atomics: they are coherent across the XCDs' L2 caches, which
// lets workgroups see each other's decisions while they run.  Each word only
// moves from unknown to its final value (links only move downward), so no
// ordering between words is needed.
__device__ __forceinline__ int ld(const int *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st(int *p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Inserts position q (weight |a[q]|) into the m sorted positions at out[0..m):
// |a| descending, then position descending (heavy_edg's tie rule).
__device__ __forceinline__ void pref_insert(int *out, int m, int q, const double *__restrict__ a) {
    const double w = fabs(a[q]);
    int pos = m;
    while (pos > 0) {
        const int qq = out[pos - 1];
        const double wq = fabs(a[qq]);
        if (wq > w || (wq == w && qq > q)) break;
        out[pos] = qq;
        --pos;
    }
    out[pos] = q;
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
            if (ja[q] <= k) continue;
            pref_insert(pref + b, m, q, a);
            ++m;
        }
        npref[k] = m;
    }
}

// Compact preference lists for the host sweep.  cnt[k]: number of higher
// neighbours (0 for an isolated row, fewer than two entries); info[k]: the
// same count, or -1 for an isolated row.  A one-entry row whose entry is not
// its diagonal would be an isolated vertex that others can still pick, which
// the higher-neighbour lists cannot express: it raises *unsupported.
__global__ void k_pref_count(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             int *__restrict__ cnt, int *__restrict__ info, int *__restrict__ unsupported) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int b = ia[k], e = ia[k + 1];
        if (e - b <= 1) {
            if (e - b == 1 && ja[b] != k) atomicExch(unsupported, 1);
            cnt[k] = 0;
            info[k] = -1;
            continue;
        }
        int m = 0;
        for (int q = b; q < e; ++q) m += ja[q] > k;
        cnt[k] = m;
        info[k] = m;
    }
}

// Row k's higher neighbours as column indices at out[pp[k] ..), in preference order.
__global__ void k_pref_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                            const double *__restrict__ a, const int *__restrict__ pp, int *__restrict__ out) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        const int b = ia[k], e = ia[k + 1];
        if (e - b <= 1) continue;
        int *o = out + pp[k];
        int m = 0;
        for (int q = b; q < e; ++q) {
            if (ja[q] <= k) continue;
            pref_insert(o, m, q, a);
            ++m;
        }
        for (int r = 0; r < m; ++r) o[r] = ja[o[r]];
    }
}

__global__ void k_pairs_to_agg(int nc, const int *__restrict__ pairs, int *__restrict__ agg) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < nc; c += gridDim.x * blockDim.x) {
        agg[pairs[2 * c]] = c;
        agg[pairs[2 * c + 1]] = c;
    }
}

__global__ void k_scatter_one(int nl, const int *__restrict__ list, int *__restrict__ flag) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nl; t += gridDim.x * blockDim.x) flag[list[t]] = 1;
}

// taken(j, t): 1 true, 0 false, -1 unknown
__device__ __forceinline__ int taken(int j, int t, const int *__restrict__ ia, const int *__restrict__ ja,
                                     const int *fstate, const int *pick) {
    int res = 0;
    for (int q = ia[j]; q < ia[j + 1]; ++q) {
        const int i = ja[q];
        if (i == j || i >= t) continue;
        const int fi = ld(fstate + i);
        if (fi == kFreeF) continue;
        const int pi = ld(pick + i);
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
                                         const int *fstate, const int *pick, int *single) {
    int res = 0, cand = -1, ncand = 0;
    bool ambiguous = false;
    for (int q = ia[k]; q < ia[k + 1]; ++q) {
        const int i = ja[q];
        if (i >= k) continue;
        const int fi = ld(fstate + i);
        if (fi == kFreeF) continue;
        const int pi = ld(pick + i);
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

// One update of vertex k (free test, link creation and one pointer jump along
// free(k) = free(a) XOR parity, pick).  Links only point to lower vertices and
// every update is monotone (unknown -> decided, links strictly downward), so
// updates may be repeated and interleaved in any order.  Returns bit 0: state
// changed, bit 1: k still undecided.
__device__ __forceinline__ int hem_vertex(int k, const int *__restrict__ ia, const int *__restrict__ ja,
                                          const int *__restrict__ pref, const int *__restrict__ npref,
                                          int *fstate, int *pick, int *linkpar) {
    int f = ld(fstate + k);
    if (f == kFreeF) return 0;
    int p = ld(pick + k);
    if (f == kFreeT && p != kUnk) return 0;
    int ch = 0;
    if (f == kFreeUnk) {
        int single = -1;
        const int t = free_eval(k, ia, ja, fstate, pick, &single);
        if (t == 1) {
            st(fstate + k, kFreeF);
            return 1;
        }
        if (t == 0) {
            f = kFreeT;
            st(fstate + k, f);
            ch = 1;
        } else {
            int lp = ld(linkpar + k);
            if (lp < 0 && single >= 0) {
                lp = 2 * single + 1;
                st(linkpar + k, lp);
                ch = 1;
            }
            if (lp >= 0) {
                const int a = lp >> 1, par = lp & 1;
                const int fa = ld(fstate + a);
                if (fa != kFreeUnk) {
                    f = ((fa == kFreeT) != (par == 1)) ? kFreeT : kFreeF;
                    st(fstate + k, f);
                    ch = 1;
                    if (f == kFreeF) return 1;
                } else {
                    const int la = ld(linkpar + a);
                    if (la >= 0) {
                        st(linkpar + k, ((la >> 1) << 1) | (par ^ (la & 1)));
                        ch = 1;
                    }
                }
            }
        }
    }
    if (p == kUnk) {
        const int b = ia[k], m = npref[k];
        int np = kNone;
        for (int r = 0; r < m; ++r) {
            const int j = ja[pref[b + r]];
            const int t = taken(j, k, ia, ja, fstate, pick);
            if (t == 1) continue;
            np = (t == 0) ? j : kUnk;
            break;
        }
        if (np != kUnk) {
            p = np;
            st(pick + k, p);
            ch = 1;
        }
    }
    const int und = (f == kFreeUnk || p == kUnk) ? 2 : 0;
    return ch | und;
}

// One pass-1 round.  Each workgroup owns contiguous chunks of kHemChunk
// vertices and repeats its updates until a whole sweep over the chunk changes
// nothing (at most kHemLocalIters sweeps), so dependency chains that stay
// inside a chunk resolve within one launch.  Every workgroup finishes after a
// bounded number of sweeps; results of other workgroups are read as they
// appear but never waited for.
constexpr int kHemPer = 4;
constexpr int kHemChunk = kBlock * kHemPer;
constexpr int kHemLocalIters = 64;

__global__ __launch_bounds__(kBlock) void k_hem_round(int n, const int *__restrict__ ia,
                                                      const int *__restrict__ ja,
                                                      const int *__restrict__ pref,
                                                      const int *__restrict__ npref, int *fstate, int *pick,
                                                      int *linkpar, int *__restrict__ undecided) {
    __shared__ int changed[2];
    const int nchunks = (n + kHemChunk - 1) / kHemChunk;
    int local = 0;
    for (int c = xcd_block_id(); c < nchunks; c += gridDim.x) {
        const int base = c * kHemChunk;
        int und = 0;
        for (int it = 0; it < kHemLocalIters; ++it) {
            if (threadIdx.x == 0) changed[it & 1] = 0;
            __syncthreads();
            int mych = 0;
            und = 0;
#pragma unroll
            for (int s = 0; s < kHemPer; ++s) {
                const int k = base + s * kBlock + threadIdx.x;
                if (k < n) {
                    const int r = hem_vertex(k, ia, ja, pref, npref, fstate, pick, linkpar);
                    mych |= r & 1;
                    und += r >> 1;
                }
            }
            if (mych) changed[it & 1] = 1;
            __threadfence_block();
            __syncthreads();
            if (!changed[it & 1]) break;
        }
        __syncthreads();   // every thread has read its last flag before the next chunk resets one
        local += und;
    }
    if (local) atomicAdd(undecided, local);
}

// Pass 1 in one launch.  Workgroups take chunks of kHemChunk vertices by
// ticket in index order and sweep a chunk until every vertex in it is
// decided.  A vertex depends only on lower vertices, so the lowest unfinished
// chunk is always held by a running workgroup and decides at least its lowest
// undecided vertex in every sweep.  A chunk that needs more than
// kPipeMaxSweeps sweeps raises *status and every workgroup stops (the host
// then finishes with k_hem_round); the kernel therefore always terminates.
constexpr int kPipeMaxSweeps = 8192;
// The cap of the default form, which hands a level with long dependency chains
// to the host sweep instead.  Measured on 128^3 Kuhn levels (3-D): the Poisson
// level 0 needs fewer than 16 sweeps a chunk, the anisotropic one (eps = 1e-3)
// 513-1024, where the whole pipeline takes 17 ms against about 7 ms for the
// host sweep over preference lists.
constexpr int kPipeProbeSweeps = 128;

__global__ __launch_bounds__(kBlock) void k_hem_pipeline(int n, const int *__restrict__ ia,
                                                         const int *__restrict__ ja,
                                                         const int *__restrict__ pref,
                                                         const int *__restrict__ npref, int *fstate, int *pick,
                                                         int *linkpar, int *__restrict__ ticket,
                                                         int *__restrict__ status, int max_sweeps) {
    __shared__ int chunk;
    const int nchunks = (n + kHemChunk - 1) / kHemChunk;
    for (;;) {
        if (threadIdx.x == 0) chunk = atomicAdd(ticket, 1);
        __syncthreads();
        const int c = chunk;
        __syncthreads();
        if (c >= nchunks) return;
        const int base = c * kHemChunk;
        for (int it = 0;; ++it) {
            int und = 0;
#pragma unroll
            for (int s = 0; s < kHemPer; ++s) {
                const int k = base + s * kBlock + threadIdx.x;
                if (k < n) und |= hem_vertex(k, ia, ja, pref, npref, fstate, pick, linkpar) >> 1;
            }
            if (!__syncthreads_or(und)) break;
            if ((it & 15) == 15) {
                int stop = 0;
                if (threadIdx.x == 0) {
                    if (it >= max_sweeps) st(status, 1);
                    stop = ld(status);
                }
                if (__syncthreads_or(stop)) return;
            }
        }
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

int g_hem_pipe_max_sweeps = 0;
int g_hem_pipe_giveups = 0;

// Pass 2 of HEM (amg_setup_ua.c:1018-1046) on the GPU: the nl leftovers in
// list (ascending) join the smallest neighbouring aggregate, first in row
// order on ties, with sizes growing as they join.  agg holds the pass-1
// aggregates (-1 elsewhere), size their sizes, lstate[v] = 1 exactly for the
// leftovers (cleared as they decide).  flagbuf: 2 device ints of scratch.
// Returns SUCCESS, -1017 (a leftover without aggregated neighbours) or 1.
static int hem_pass2(const DevCSR &N, int *agg, int *size, int nc, const int *list, int nl, int *lstate,
                     int *flagbuf, int *rounds_out) {
    hipStream_t st = stream();
    int *done = arena().alloc_n<int>((size_t)nl + 1);
    int *owner = arena().alloc_n<int>((size_t)nc + 1);
    HZ_CHECK(hipMemsetAsync(done, 0, sizeof(int) * nl, st));
    k_fill_i<<<grid_for(nc), kBlock, 0, st>>>(owner, INT_MAX, nc);
    int *err = flagbuf, *remaining = flagbuf + 1;
    HZ_CHECK(hipMemsetAsync(err, 0, sizeof(int), st));
    const int gl = grid_for(nl);
    // one round, then a few per host check: a round after the last decision
    // changes nothing, and `remaining` counts the undecided leftovers of the
    // last one (most levels need a single round)
    constexpr int kRoundsPerCheck = 4;
    for (int round = 0, batch = 1;; round += batch, batch = kRoundsPerCheck) {
        for (int q = 0; q < batch; ++q) {
            HZ_CHECK(hipMemsetAsync(remaining, 0, sizeof(int), st));
            k_left_claim<<<gl, kBlock, 0, st>>>(nl, list, done, N.ia, N.ja, agg, owner);
            k_left_decide<<<gl, kBlock, 0, st>>>(nl, list, done, N.ia, N.ja, agg, size, owner, lstate, err,
                                                 remaining);
            k_reset_owner<<<gl, kBlock, 0, st>>>(nl, list, N.ia, N.ja, agg, owner);
            k_left_commit<<<gl, kBlock, 0, st>>>(nl, list, done, agg, size, lstate);
            HZ_LAUNCH_CHECK();
        }
        int hr[2] = {0, 0};
        HZ_CHECK(memcpy_sync(hr, err, 2 * sizeof(int), hipMemcpyDeviceToHost));
        *rounds_out = round + batch;
        if (hr[0]) return -1017;
        if (hr[1] == 0) break;
        if (round > nl + 8) return 1;
    }
    return HZ_SUCCESS;
}

// Returns SUCCESS, -1017 (reference exit(17)), or 1 when the strength pattern
// is not symmetric (caller falls back to the host sweep).
int hem_aggregate_gpu(const DevCSR &N, int *agg, int *nagg_out, bool host_on_giveup) {
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
    int *linkpar = arena().alloc_n<int>((size_t)n + 1);
    k_fill_i<<<g, kBlock, 0, st>>>(linkpar, -1, n);
    // one undecided counter per round, cleared once
    const int kCounterSlots = 4096;
    int *undecided = arena().alloc_n<int>(kCounterSlots);
    HZ_CHECK(hipMemsetAsync(undecided, 0, sizeof(int) * kCounterSlots, st));
    const int nchunks = (n + kHemChunk - 1) / kHemChunk;
    const int gr = std::max(1, std::min(8192, nchunks));
    int *pipe = arena().alloc_n<int>(2);   // ticket, status
    HZ_CHECK(hipMemsetAsync(pipe, 0, 2 * sizeof(int), st));
    // hz_options.hem_pipeline_max_sweeps: > 0 a cap finished by the round
    // kernel, < 0 a cap that hands the level to the host sweep, 0 the default
    bool probe = host_on_giveup;
    int max_sweeps = probe ? kPipeProbeSweeps : kPipeMaxSweeps;
    if (g_hem_pipe_max_sweeps != 0) {
        probe = g_hem_pipe_max_sweeps < 0;
        max_sweeps = probe ? -g_hem_pipe_max_sweeps : g_hem_pipe_max_sweeps;
    }
    k_hem_pipeline<<<std::min(nchunks, 2048), kBlock, 0, st>>>(n, N.ia, N.ja, pref, npref, fstate, pick,
                                                              linkpar, pipe, pipe + 1, max_sweeps);
    HZ_LAUNCH_CHECK();
    int pipe_status = 0;
    HZ_CHECK(memcpy_sync(&pipe_status, pipe + 1, sizeof(int), hipMemcpyDeviceToHost));
    const int kRoundsPerCheck = 2;
    int rounds1 = 0, rounds2 = 0;
    if (pipe_status != 0) ++g_hem_pipe_giveups;
    if (pipe_status != 0 && probe) {
        if (trace_enabled())
            std::fprintf(stderr, "[hz]   HEM n=%d: pipeline reached %d sweeps in a chunk, host sweep instead\n", n,
                         max_sweeps);
        return 2;
    }
    for (int round = 0; pipe_status != 0; round += kRoundsPerCheck) {
        if (round + kRoundsPerCheck > kCounterSlots) {
            HZ_CHECK(hipMemsetAsync(undecided, 0, sizeof(int) * kCounterSlots, st));
            round = 0;   // only the slot index restarts; progress is kept in the state arrays
        }
        for (int r = 0; r < kRoundsPerCheck; ++r)
            k_hem_round<<<gr, kBlock, 0, st>>>(n, N.ia, N.ja, pref, npref, fstate, pick, linkpar,
                                               undecided + round + r);
        HZ_LAUNCH_CHECK();
        rounds1 += kRoundsPerCheck;
        int und = 0;
        HZ_CHECK(memcpy_sync(&und, undecided + round + kRoundsPerCheck - 1, sizeof(int),
                             hipMemcpyDeviceToHost));
        if (und == 0) break;
        if (rounds1 > n + 8) return 1;   // cannot happen: the lowest undecided vertex decides every round
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
        k_compact_leftovers<<<g, kBlock, 0, st>>>(n, leftover, lpos, list);
        HZ_LAUNCH_CHECK();
        // the leftover flags double as lstate (1 while undecided)
        const int r = hem_pass2(N, agg, size, nc, list, nl, leftover, flagbuf + 2, &rounds2);
        if (r != HZ_SUCCESS) return r;
    }
    *nagg_out = nc;
    if (trace_enabled())
        std::fprintf(stderr, "[hz]   HEM n=%d: pass-1 %s, fallback rounds %d, leftovers %d, pass-2 rounds %d\n",
                     n, pipe_status ? "pipeline gave up" : "pipeline", rounds1, nl, rounds2);
    return HZ_SUCCESS;
}

// Preference lists of HEM's pass 1 for the host sweep, queued on the stream
// without a host wait: info[k] (count of higher neighbours, -1 isolated),
// their column indices in preference order concatenated in row order at
// cols[pp[k] ..) (pp[n]: the total; cols holds up to nnz entries), and
// flags[0..1] nonzero when the lists cannot express the pattern (not
// symmetric, or a one-entry row without its diagonal).  Device arena memory.
void hem_pref_lists(const DevCSR &N, HemPrefLists &out) {
    hipStream_t st = stream();
    const int n = N.n;
    const int g = grid_for(n);
    out.flags = arena().alloc_n<int>(2);
    HZ_CHECK(hipMemsetAsync(out.flags, 0, 2 * sizeof(int), st));
    k_sym_check<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, out.flags);
    int *cnt = arena().alloc_n<int>((size_t)n + 1);
    out.info = arena().alloc_n<int>((size_t)n + 1);
    out.pp = arena().alloc_n<int>((size_t)n + 1);
    k_pref_count<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, cnt, out.info, out.flags + 1);
    HZ_LAUNCH_CHECK();
    exclusive_scan_int_device(cnt, out.pp, n);
    out.cols = arena().alloc_n<int>((size_t)N.nnz + 1);
    k_pref_fill<<<g, kBlock, 0, st>>>(n, N.ia, N.ja, N.val, out.pp, out.cols);
    HZ_LAUNCH_CHECK();
}

// Aggregate map of a host pass 1 from its nc pairs (host array, pair c =
// aggregate c): -1 elsewhere.  Queued on the stream; pairs_host must stay
// untouched until the stream has passed this point.
void hem_agg_from_pairs(const int *pairs_host, int nc, int n, int *agg) {
    hipStream_t st = stream();
    int *pd = arena().alloc_n<int>(2 * (size_t)nc + 2);
    if (nc > 0) HZ_CHECK(hipMemcpyAsync(pd, pairs_host, sizeof(int) * 2 * (size_t)nc, hipMemcpyHostToDevice, st));
    k_fill_i<<<grid_for(n), kBlock, 0, st>>>(agg, -1, n);
    if (nc > 0) k_pairs_to_agg<<<grid_for(nc), kBlock, 0, st>>>(nc, pd, agg);
    HZ_LAUNCH_CHECK();
}

// Pass 2 after a host pass 1: agg_dev holds the pass-1 map (-1 for leftovers
// and isolated vertices), every pass-1 aggregate has two members, and
// list_host the nl leftovers in ascending order.
int hem_pass2_from_host(const DevCSR &N, int *agg_dev, int nc, const int *list_host, int nl, int *rounds) {
    hipStream_t st = stream();
    const int n = N.n;
    int *list = arena().alloc_n<int>((size_t)nl + 1);
    HZ_CHECK(hipMemcpyAsync(list, list_host, sizeof(int) * nl, hipMemcpyHostToDevice, st));
    int *size = arena().alloc_n<int>((size_t)nc + 1);
    k_fill_i<<<grid_for(nc), kBlock, 0, st>>>(size, 2, nc);
    int *lstate = arena().alloc_n<int>((size_t)n + 1);
    k_fill_i<<<grid_for(n), kBlock, 0, st>>>(lstate, 0, n);
    k_scatter_one<<<grid_for(nl), kBlock, 0, st>>>(nl, list, lstate);
    HZ_LAUNCH_CHECK();
    int *flagbuf = arena().alloc_n<int>(2);
    return hem_pass2(N, agg_dev, size, nc, list, nl, lstate, flagbuf, rounds);
}

}  // namespace hz
