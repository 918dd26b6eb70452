// Harness of the sanitizer build (TEST INFRASTRUCTURE): gives input.cpp the
// library entry points it calls, served by the oracle (same defaults, the
// serial solver), and exports the host sweeps of aggregate_sweeps.cpp with C
// linkage so the test can compare them with the oracle's aggregates.
#include <algorithm>
#include <vector>

#include "aggregate_sweeps.h"
#include "hz_oracle.h"
#include "tune_cache.h"

namespace hz {
int fatal_status(int code) { return code; }
}  // namespace hz

extern "C" {

void param_amg_init(AMG_param *p) { orc_param_amg_init(p); }
void param_linear_solver_init(linear_itsolver_param *p) { orc_param_linear_solver_init(p); }
INT linear_solver_dcsr_krylov_amg(dCSRmat *A, dvector *b, dvector *x, linear_itsolver_param *it, AMG_param *amg) {
    return orc_krylov_amg(A, b, x, it, amg, nullptr);
}

int san_vmb(int n, const int *ia, const int *ja, int maxagg, int *agg, int *nagg) {
    std::vector<int> v;
    const int st = hz::vmb_sweep(n, ia, ja, maxagg, v, *nagg);
    std::copy(v.begin(), v.end(), agg);
    return st;
}

int san_hem(int n, const int *ia, const int *ja, const double *a, int *agg, int *nagg) {
    std::vector<int> v;
    const int st = hz::hem_sweep(n, ia, ja, a, v, *nagg);
    if (st == 0) std::copy(v.begin(), v.end(), agg);
    return st;
}

int san_tune_reset(const char *path) { return hz::TuneCache::get().reset(path, false); }
int san_tune_lookup(const char *key) { return hz::TuneCache::get().lookup(key); }
void san_tune_store(const char *key, int shape) { hz::TuneCache::get().store(key, shape); }

}  // extern "C"
