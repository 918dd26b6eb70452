// aggregate_seq.cpp -- VMB and HEM aggregation, host-staged.
//
// Both are greedy sweeps whose every decision depends on all earlier ones in
// index order (VMB: aggregation_vmb, amg_setup_ua.c:596-727; HEM:
// aggregation_hem, amg_setup_ua.c:913-1051).  The strength graph is built on
// the GPU, copied to the host, swept once, and the aggregate map copied back.
// HEC, the reference default, is resolved fully on the GPU instead
// (kernels_setup.hip).
#include <algorithm>
#include <cmath>
#include <vector>

#include "aggregate_sweeps.h"
#include "amg.hpp"

namespace hz {


namespace {
// Grow-only pinned staging buffers, kept for the process: copies through
// pinned memory run at full link speed (pageable copies are staged twice).
struct PinnedStage {
    void *p = nullptr;
    size_t cap = 0;
    void *get(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (bytes > cap) {
            if (p) (void)hipHostFree(p);
            p = nullptr;
            HZ_CHECK(hipHostMalloc(&p, bytes, hipHostMallocDefault));
            cap = bytes;
        }
        return p;
    }
};
PinnedStage g_ia, g_ja, g_a, g_v, g_small;

// VMB's second pass on the GPU, from pass 1's map in agg (-2: not yet
// aggregated): every such vertex takes the first neighbour that is in a pass-1
// aggregate.  The sequential pass makes exactly these choices when no
// aggregate passes the cap with all of them (each then has room when its turn
// comes); flag bit 1: some aggregate would, bit 2: some vertex has no such
// neighbour (pass 3 would run) -- the host pass runs instead.
constexpr int kVmbGpuPass2Rows = 65536;
constexpr int kVmbBlock = 256;

__global__ void k_vmb2_first(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const int *__restrict__ agg, int *__restrict__ cand, int *__restrict__ count,
                             int *__restrict__ flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        int c = HZ_UNPT - 2;   // aggregated in pass 1
        if (agg[i] == HZ_UNPT - 1) {
            c = HZ_UNPT;
            for (int k = ia[i]; k < ia[i + 1]; ++k) {
                const int a = agg[ja[k]];
                if (a > HZ_UNPT) {
                    c = a;
                    break;
                }
            }
            if (c > HZ_UNPT) atomicAdd(&count[c], 1);
            else atomicOr(flag, 2);
        }
        cand[i] = c;
    }
}

__global__ void k_vmb2_caps(int na, const int *__restrict__ size, const int *__restrict__ count, int maxagg,
                            int *__restrict__ flag) {
    for (int a = blockIdx.x * blockDim.x + threadIdx.x; a < na; a += gridDim.x * blockDim.x)
        if (size[a] + count[a] > maxagg) atomicOr(flag, 1);
}

__global__ void k_vmb2_join(int n, const int *__restrict__ cand, int *__restrict__ agg) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        if (cand[i] > HZ_UNPT) agg[i] = cand[i];
}

// pass 2 of VMB on the GPU from pass 1's map v (pinned); agg_dev ends with the
// final map when it returns true
bool vmb_pass2_gpu(const DevCSR &N, const int *v, const std::vector<int> &size, int maxagg, int *agg_dev) {
    const int n = N.n, na = (int)size.size();
    hipStream_t st = stream();
    int *dsize = arena().alloc_n<int>((size_t)na + 1);
    int *count = arena().alloc_n<int>((size_t)na + 1);
    int *cand = arena().alloc_n<int>((size_t)n + 1);
    int *flag = arena().alloc_n<int>(1);
    HZ_CHECK(hipMemcpyAsync(agg_dev, v, sizeof(int) * (size_t)n, hipMemcpyHostToDevice, st));
    HZ_CHECK(hipMemcpyAsync(dsize, size.data(), sizeof(int) * (size_t)na, hipMemcpyHostToDevice, st));
    HZ_CHECK(hipMemsetAsync(count, 0, sizeof(int) * (size_t)na, st));
    HZ_CHECK(hipMemsetAsync(flag, 0, sizeof(int), st));
    const int g = (int)std::min<long>(8192, ((long)n + kVmbBlock - 1) / kVmbBlock);
    const int gc = (int)std::min<long>(8192, ((long)na + kVmbBlock - 1) / kVmbBlock);
    k_vmb2_first<<<g, kVmbBlock, 0, st>>>(n, N.ia, N.ja, agg_dev, cand, count, flag);
    k_vmb2_caps<<<gc, kVmbBlock, 0, st>>>(na, dsize, count, maxagg, flag);
    k_vmb2_join<<<g, kVmbBlock, 0, st>>>(n, cand, agg_dev);
    HZ_LAUNCH_CHECK();
    int *hf = static_cast<int *>(g_small.get(4 * sizeof(int)));
    HZ_CHECK(hipMemcpyAsync(hf, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    return hf[0] == 0;
}
}  // namespace

int aggregate_host(const DevCSR &N, int type, int max_aggregation, int *agg_dev, int *nagg) {
    const int n = N.n;
    int *ia = static_cast<int *>(g_ia.get(sizeof(int) * ((size_t)n + 1)));
    int *ja = static_cast<int *>(g_ja.get(sizeof(int) * (size_t)N.nnz));
    double *a = static_cast<double *>(g_a.get(sizeof(double) * (size_t)N.nnz));
    hipStream_t st = stream();
    HZ_CHECK(hipMemcpyAsync(ia, N.ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
    static std::vector<int> v;   // kept across calls: no fresh pages per call
    int *pv = static_cast<int *>(g_v.get(sizeof(int) * ((size_t)n + 1)));
    int st_code;
    if (type == HZ_VMB) {
        // VMB's first pass walks the rows downwards: the column indices come
        // in blocks of about equal size from the last rows down, and the pass
        // starts on the first block while the others are on the link
        constexpr int kBlocks = 8;
        static hipEvent_t ev[kBlocks];
        static bool ev_ready = false;
        if (!ev_ready) {
            for (auto &e : ev) HZ_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            ev_ready = true;
        }
        HZ_CHECK(hipStreamSynchronize(st));
        const long long total = ia[n] - ia[0];
        int lo[kBlocks];
        int nb = 0;
        for (int b = 0, hi = n; b < kBlocks && hi > 0; ++b) {
            int r = hi;
            if (b == kBlocks - 1) r = 0;
            else {
                const long long target = ia[hi] - total / kBlocks;   // rows [r, hi) hold about total / kBlocks
                while (r > 0 && ia[r - 1] >= target) --r;
                if (r == hi) --r;
            }
            const long long e0 = ia[r], e1 = ia[hi];
            if (e1 > e0)
                HZ_CHECK(hipMemcpyAsync(ja + e0, N.ja + e0, sizeof(int) * (size_t)(e1 - e0), hipMemcpyDeviceToHost,
                                        st));
            HZ_CHECK(hipEventRecord(ev[nb], st));
            lo[nb++] = r;
            hi = r;
        }
        RowBlocksDown arrival;
        arrival.nblocks = nb;
        arrival.lo = lo;
        static double waited = 0.0;
        waited = 0.0;
        const bool tr = trace_enabled();
        arrival.wait = [tr](int b) {
            const double w0 = tr ? wall_ms() : 0.0;
            HZ_CHECK(hipEventSynchronize(ev[b]));
            if (tr) waited += wall_ms() - w0;
        };
        double pass_ms[3] = {0, 0, 0};
        const double s0 = tr ? wall_ms() : 0.0;
        // pass 1 on the host, into the pinned map
        static std::vector<int> size;
        int left = 0;
        const int na = vmb_pass1(n, ia, ja, max_aggregation, pv, size, left, &arrival);
        HZ_CHECK(hipStreamSynchronize(st));
        const double s1 = tr ? wall_ms() : 0.0;
        bool on_gpu = false;
        if (na < 1) {
            *nagg = 0;
            st_code = HZ_ERROR_AMG_COARSEING;
        } else {
            st_code = HZ_SUCCESS;
            on_gpu = n >= kVmbGpuPass2Rows && vmb_pass2_gpu(N, pv, size, max_aggregation, agg_dev);
            *nagg = on_gpu ? na : vmb_pass23(n, ia, ja, max_aggregation, pv, size, na, left,
                                              tr ? pass_ms : nullptr);
        }
        if (tr)
            std::fprintf(stderr, "[hz]   VMB n=%d nnz=%lld: sweep %.2f ms (pass 1 %.2f, of which waiting for the graph "
                         "%.2f; pass 2 %s %.2f; pass 3 %.2f)\n", n, total, wall_ms() - s0, s1 - s0, waited,
                         on_gpu ? "on the GPU" : "on the host", on_gpu ? wall_ms() - s1 : pass_ms[1], pass_ms[2]);
        if (on_gpu) return st_code;
        HZ_CHECK(hipMemcpyAsync(agg_dev, pv, sizeof(int) * n, hipMemcpyHostToDevice, st));
        HZ_CHECK(hipStreamSynchronize(st));
        return st_code;
    } else {
        if (N.nnz) {
            HZ_CHECK(hipMemcpyAsync(ja, N.ja, sizeof(int) * N.nnz, hipMemcpyDeviceToHost, st));
            HZ_CHECK(hipMemcpyAsync(a, N.val, sizeof(double) * N.nnz, hipMemcpyDeviceToHost, st));
        }
        HZ_CHECK(hipStreamSynchronize(st));
        st_code = hem_sweep(n, ia, ja, a, v, *nagg);
    }
    if (st_code == -1017) return st_code;
    std::copy(v.begin(), v.end(), pv);
    HZ_CHECK(hipMemcpyAsync(agg_dev, pv, sizeof(int) * n, hipMemcpyHostToDevice, st));
    HZ_CHECK(hipStreamSynchronize(st));
    return st_code;
}

// HEM with the sequential pass 1 on the host and only the data it reads on
// the link.  With a symmetric strength pattern every lower neighbour of a
// still-unmatched k is matched by time k, so pass 1 at k takes the first
// unmatched entry of k's higher neighbours in preference order (|a|
// descending, then row position descending: heavy_edg's scan).  The GPU sorts
// those lists (hem_pref_lists); their copy to the host runs in row chunks
// that the sweep consumes as they land; pass 2 (the leftovers) runs on the
// GPU.  Returns 1 when the lists cannot express the pattern (the caller then
// runs the full host sweep).
int aggregate_hem_staged(const DevCSR &N, int *agg_dev, int *nagg) {
    constexpr int kChunks = 8;
    static hipEvent_t ev[kChunks];
    static bool ev_ready = false;
    if (!ev_ready) {
        for (auto &e : ev) HZ_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        ev_ready = true;
    }
    const int n = N.n;
    hipStream_t st = stream();
    HemPrefLists pl;
    hem_pref_lists(N, pl);
    int *small = static_cast<int *>(g_small.get(4 * sizeof(int)));
    int *info = static_cast<int *>(g_ia.get(sizeof(int) * ((size_t)n + 1)));
    HZ_CHECK(hipMemcpyAsync(small, pl.flags, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipMemcpyAsync(small + 2, pl.pp + n, sizeof(int), hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipMemcpyAsync(info, pl.info, sizeof(int) * n, hipMemcpyDeviceToHost, st));
    HZ_CHECK(hipStreamSynchronize(st));
    if (small[0] || small[1]) return 1;
    const long long total = small[2];
    int *cols = static_cast<int *>(g_ja.get(sizeof(int) * ((size_t)total + 1)));
    // row chunks of about equal entry counts; the sweep waits for each in turn
    int row_end[kChunks];
    {
        long long off = 0;
        int c = 0;
        for (int k = 0; k < n && c < kChunks - 1; ++k) {
            off += info[k] > 0 ? info[k] : 0;
            if (off * kChunks >= total * (c + 1)) row_end[c++] = k + 1;
        }
        while (c < kChunks) row_end[c++] = n;
        long long start = 0, acc = 0;
        int k = 0;
        for (int q = 0; q < kChunks; ++q) {
            for (; k < row_end[q]; ++k) acc += info[k] > 0 ? info[k] : 0;
            if (acc > start)
                HZ_CHECK(hipMemcpyAsync(cols + start, pl.cols + start, sizeof(int) * (acc - start),
                                        hipMemcpyDeviceToHost, st));
            HZ_CHECK(hipEventRecord(ev[q], st));
            start = acc;
        }
    }
    // the sweep appends its pairs (initiator, partner) in order, the GPU
    // turns them into the aggregate map; its working set is one byte per
    // vertex (matched), cache resident
    int *pairs = static_cast<int *>(g_v.get(sizeof(int) * ((size_t)n + 2)));
    static std::vector<unsigned char> matched;
    matched.assign(n, 0);
    unsigned char *mt = matched.data();
    std::vector<int> left;
    long long off = 0;
    int nc = 0, k = 0;
    for (int q = 0; q < kChunks; ++q) {
        HZ_CHECK(hipEventSynchronize(ev[q]));
        for (; k < row_end[q]; ++k) {
            const int c = info[k];
            if (c < 0) continue;   // isolated (fewer than two entries): no aggregate
            const int *pk = cols + off;
            off += c;
            if (mt[k]) continue;   // matched by a lower vertex
            int pick = -1;
            for (int r = 0; r < c; ++r)
                if (!mt[pk[r]]) {
                    pick = pk[r];
                    break;
                }
            if (pick >= 0) {
                mt[k] = 1;
                mt[pick] = 1;
                pairs[2 * nc] = k;
                pairs[2 * nc + 1] = pick;
                ++nc;
            } else {
                left.push_back(k);
            }
        }
    }
    hem_agg_from_pairs(pairs, nc, n, agg_dev);
    int rounds = 0;
    const int nl = (int)left.size();
    if (nl > 0) {
        const int p2 = hem_pass2_from_host(N, agg_dev, nc, left.data(), nl, &rounds);
        if (p2 != HZ_SUCCESS) return p2;
    } else {
        HZ_CHECK(hipStreamSynchronize(st));   // the pairs buffer is reused by the next level
    }
    *nagg = nc;
    if (trace_enabled())
        std::fprintf(stderr, "[hz]   HEM n=%d: host pass 1 over %lld preference entries, leftovers %d, "
                     "pass-2 rounds %d\n", n, total, nl, rounds);
    return HZ_SUCCESS;
}

}  // namespace hz
