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

#include "amg.hpp"

namespace hz {

namespace {

int vmb(int n, const int *ia, const int *ja, int maxagg, std::vector<int> &v, int &nagg) {
    int left = n, na = 0;
    v.assign(n, -2);
    for (int i = n - 1; i >= 0; --i) {
        if (ia[i + 1] - ia[i] == 1) {
            v[i] = HZ_UNPT;
            --left;
            continue;
        }
        bool untouched = true;
        for (int k = ia[i]; k < ia[i + 1] && untouched; ++k)
            if (v[ja[k]] >= HZ_UNPT) untouched = false;
        if (!untouched) continue;
        int cnt = 1;
        v[i] = na;
        --left;
        for (int k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] != i && cnt < maxagg) {
                v[ja[k]] = na;
                --left;
                ++cnt;
            }
        ++na;
    }
    nagg = na;
    if (na < 1) return HZ_ERROR_AMG_COARSEING;
    std::vector<int> c1(v), size(na, 0);
    for (int i = 0; i < n; ++i)
        if (v[i] >= 0) size[v[i]]++;
    for (int i = n - 1; i >= 0; --i) {
        if (v[i] >= HZ_UNPT) continue;
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            const int a = c1[ja[k]];
            if (a > HZ_UNPT && size[a] < maxagg) {
                v[i] = a;
                --left;
                size[a]++;
                break;
            }
        }
    }
    while (left > 0) {
        for (int i = 0; i < n; ++i) {
            if (v[i] >= HZ_UNPT) continue;
            int cnt = 1;
            v[i] = na;
            --left;
            for (int k = ia[i]; k < ia[i + 1]; ++k)
                if (ja[k] != i && v[ja[k]] < HZ_UNPT && cnt < maxagg) {
                    v[ja[k]] = na;
                    --left;
                    ++cnt;
                }
            ++na;
        }
    }
    nagg = na;
    return HZ_SUCCESS;
}

// heavy_edg (amg_setup_ua.c:864-891): first element wins unless a later one
// (in scan direction) is strictly heavier.
int heaviest(const std::vector<double> &w, const std::vector<int> &id, int n0, int n1) {
    if (n0 < 0 || n1 < 0) return -1;
    const int step = (n0 > n1) ? -1 : 1;
    int best = id[n0];
    double bw = w[n0];
    for (int j = n0; j != n1 + step; j += step)
        if (w[j] > bw && id[j] >= 0) {
            bw = w[j];
            best = id[j];
        }
    return best;
}

int hem(int n, const int *ia, const int *ja, const double *a, std::vector<int> &v, int &nagg) {
    v.assign(n, -1);
    int maxdeg = 0;
    for (int k = 0; k < n; ++k) {
        const int l = ia[k + 1] - ia[k];
        maxdeg = std::max(maxdeg, l);
        if (l <= 1) v[k] = -2;
    }
    std::vector<double> w(maxdeg + 1);
    std::vector<int> id(maxdeg + 1), size(n, 0);
    int nc = 0;
    for (int k = 0; k < n; ++k) {
        if (v[k] != -1) continue;
        int iz = 0;
        for (int jk = ia[k]; jk < ia[k + 1]; ++jk) {
            const int j = ja[jk];
            if (v[j] < 0 && j != k) {
                w[iz] = std::fabs(a[jk]);
                id[iz] = j;
                ++iz;
            }
        }
        if (iz) {
            const int pick = heaviest(w, id, iz - 1, 0);
            v[k] = nc;
            v[pick] = nc;
            size[nc] += 2;
            ++nc;
        }
    }
    for (int k = 0; k < n; ++k) {
        if (v[k] != -1) continue;
        int iz = 0;
        for (int jk = ia[k]; jk < ia[k + 1]; ++jk) {
            const int c = v[ja[jk]];
            if (c >= 0) {
                w[iz] = (double)(-size[c]);
                id[iz] = c;
                ++iz;
            }
        }
        if (!iz) {
            std::fprintf(stderr, "%%%%ERROR in %s: isolated point (=%d) on the second matching pass",
                         "aggregation_hem", k);
            return -1017;
        }
        const int c = heaviest(w, id, 0, iz - 1);
        v[k] = c;
        size[c]++;
    }
    for (int k = 0; k < n; ++k)
        if (v[k] < 0) v[k] = HZ_UNPT;
    nagg = nc;
    return HZ_SUCCESS;
}

}  // namespace

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
}  // namespace

int aggregate_host(const DevCSR &N, int type, int max_aggregation, int *agg_dev, int *nagg) {
    const int n = N.n;
    int *ia = static_cast<int *>(g_ia.get(sizeof(int) * ((size_t)n + 1)));
    int *ja = static_cast<int *>(g_ja.get(sizeof(int) * (size_t)N.nnz));
    double *a = static_cast<double *>(g_a.get(sizeof(double) * (size_t)N.nnz));
    hipStream_t st = stream();
    HZ_CHECK(hipMemcpyAsync(ia, N.ia, sizeof(int) * (n + 1), hipMemcpyDeviceToHost, st));
    if (N.nnz) {
        HZ_CHECK(hipMemcpyAsync(ja, N.ja, sizeof(int) * N.nnz, hipMemcpyDeviceToHost, st));
        if (type != HZ_VMB)
            HZ_CHECK(hipMemcpyAsync(a, N.val, sizeof(double) * N.nnz, hipMemcpyDeviceToHost, st));
    }
    HZ_CHECK(hipStreamSynchronize(st));
    std::vector<int> v;
    int st_code = (type == HZ_VMB) ? vmb(n, ia, ja, max_aggregation, v, *nagg) : hem(n, ia, ja, a, v, *nagg);
    if (st_code == -1017) return st_code;
    int *pv = static_cast<int *>(g_v.get(sizeof(int) * (size_t)n));
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
