// aggregate_sweeps.cpp -- the sequential VMB and HEM sweeps on host arrays
// (no HIP dependency: also built alone, with sanitizers, as a CPU check).
//
// VMB: aggregation_vmb, amg_setup_ua.c:596-727; HEM: aggregation_hem,
// amg_setup_ua.c:913-1051.  Every decision depends on all earlier ones in
// index order.
#include "aggregate_sweeps.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "hazmath_mi355x.h"

namespace hz {

int vmb_pass1(int n, const int *ia, const int *ja, int maxagg, int *v, std::vector<int> &size, int &left,
              const RowBlocksDown *arrival) {
    left = n;
    int na = 0;
    std::fill(v, v + n, -2);
    size.clear();   // aggregate sizes as the pass forms them (the vertices that take the id)
    int blk = 0, here = n;   // rows >= here have arrived
    for (int i = n - 1; i >= 0; --i) {
        while (arrival && i < here && blk < arrival->nblocks) {
            arrival->wait(blk);
            here = arrival->lo[blk++];
        }
        if (ia[i + 1] - ia[i] == 1) {
            v[i] = HZ_UNPT;
            --left;
            continue;
        }
        // "every neighbour still free": scanned from the row's end, where the
        // higher rows, swept (and aggregated) first, sit -- the same answer
        bool untouched = true;
        for (int k = ia[i + 1] - 1; k >= ia[i] && untouched; --k)
            if (v[ja[k]] >= HZ_UNPT) untouched = false;
        if (!untouched) continue;
        int cnt = 1, members = 1;
        v[i] = na;
        --left;
        for (int k = ia[i]; k < ia[i + 1]; ++k)
            if (ja[k] != i && cnt < maxagg) {
                members += v[ja[k]] != na;
                v[ja[k]] = na;
                --left;
                ++cnt;
            }
        size.push_back(members);
        ++na;
    }
    return na;
}

int vmb_pass23(int n, const int *ia, const int *ja, int maxagg, int *v, std::vector<int> &size, int na, int left,
               double *ms) {
    const auto now = [] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    const double t1 = ms ? now() : 0.0;
    // pass 2 reads only pass 1's map: its own assignments are collected and
    // written after it (each vertex is visited once, so none is read back)
    static std::vector<std::pair<int, int>> joins;
    joins.clear();
    for (int i = n - 1; i >= 0; --i) {
        if (v[i] >= HZ_UNPT) continue;
        for (int k = ia[i]; k < ia[i + 1]; ++k) {
            const int a = v[ja[k]];
            if (a > HZ_UNPT && size[a] < maxagg) {
                joins.emplace_back(i, a);
                --left;
                size[a]++;
                break;
            }
        }
    }
    for (const auto &j : joins) v[j.first] = j.second;
    const double t2 = ms ? now() : 0.0;
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
    if (ms) {
        ms[1] = t2 - t1;
        ms[2] = now() - t2;
    }
    return na;
}

int vmb_sweep(int n, const int *ia, const int *ja, int maxagg, std::vector<int> &v, int &nagg,
              const RowBlocksDown *arrival, double *ms) {
    const auto now = [] {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
    };
    const double t0 = ms ? now() : 0.0;
    static std::vector<int> size;
    int left = 0;
    v.resize(n);
    const int na = vmb_pass1(n, ia, ja, maxagg, v.data(), size, left, arrival);
    nagg = na;
    if (ms) ms[0] = now() - t0;
    if (na < 1) return HZ_ERROR_AMG_COARSEING;
    nagg = vmb_pass23(n, ia, ja, maxagg, v.data(), size, na, left, ms);
    return HZ_SUCCESS;
}

namespace {
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

}  // namespace

int hem_sweep(int n, const int *ia, const int *ja, const double *a, std::vector<int> &v, int &nagg) {
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

}  // namespace hz
