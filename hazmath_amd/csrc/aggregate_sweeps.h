// aggregate_sweeps.h -- the sequential VMB and HEM sweeps on host arrays.
#pragma once

#include <functional>
#include <vector>

namespace hz {

// Column indices that arrive in row blocks from the last rows down: block b
// holds rows [lo[b], lo[b - 1]) (lo[-1] = n), and wait(b) returns once it is
// there.  VMB's first pass walks the rows downwards and reads only the row it
// is at, so it can start before the whole graph has arrived.
struct RowBlocksDown {
    int nblocks = 0;
    const int *lo = nullptr;
    std::function<void(int)> wait;
};

// aggregation_vmb (amg_setup_ua.c:596-727) on the strength graph (ia, ja) of n
// rows: v[i] = aggregate of i or -1 (UNPT); returns SUCCESS.  arrival: ja
// arrives in row blocks (null: all there).
// ms (optional): wall milliseconds of the three passes.
int vmb_sweep(int n, const int *ia, const int *ja, int maxagg, std::vector<int> &v, int &nagg,
              const RowBlocksDown *arrival = nullptr, double *ms = nullptr);
// The same in two parts, into v[0..n): pass 1 (roots with all their neighbours
// free, in reverse order) leaves v = aggregate, UNPT or -2 (not yet aggregated),
// size[a] = the members of aggregate a and left = the reference's count of
// vertices still to aggregate, and returns the number of aggregates; passes 2
// (join the first neighbouring pass-1 aggregate below the cap, reverse order)
// and 3 (new aggregates from what is left, forward) then return the final count.
int vmb_pass1(int n, const int *ia, const int *ja, int maxagg, int *v, std::vector<int> &size, int &left,
              const RowBlocksDown *arrival = nullptr);
int vmb_pass23(int n, const int *ia, const int *ja, int maxagg, int *v, std::vector<int> &size, int na, int left,
               double *ms = nullptr);
// aggregation_hem (amg_setup_ua.c:913-1051) on the strength graph (ia, ja, a):
// SUCCESS, or -1017 (the reference's exit(17), its message printed) when a
// leftover has no aggregated neighbour.
int hem_sweep(int n, const int *ia, const int *ja, const double *a, std::vector<int> &v, int &nagg);

}  // namespace hz
