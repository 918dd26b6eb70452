// aggregate_sweeps.h -- the sequential VMB and HEM sweeps on host arrays.
#pragma once

#include <vector>

namespace hz {

// aggregation_vmb (amg_setup_ua.c:596-727) on the strength graph (ia, ja) of n
// rows: v[i] = aggregate of i or -1 (UNPT); returns SUCCESS.
int vmb_sweep(int n, const int *ia, const int *ja, int maxagg, std::vector<int> &v, int &nagg);
// aggregation_hem (amg_setup_ua.c:913-1051) on the strength graph (ia, ja, a):
// SUCCESS, or -1017 (the reference's exit(17), its message printed) when a
// leftover has no aggregated neighbour.
int hem_sweep(int n, const int *ia, const int *ja, const double *a, std::vector<int> &v, int &nagg);

}  // namespace hz
