// rotation.hpp -- CR_1 / CR_2 rotation search (internal).
#pragma once
#include <vector>

#include "dtc_internal.hpp"

namespace dtc {

struct RotJob {
    int dom, criterion, active, _pad;
    int64_t s0[3];   // first grid index of the 21^3 refinement block
    double div;      // angle per grid index (radians)
};

struct RotResult {
    double alpha = 0.0, n[3] = {0, 0, 0}, aa[9] = {0}, EE_opt[36] = {0};
    int64_t jj = 0;
    double crit = 0.0;
};

// EE_host: n_dom column-major 6x6 matrices.  out[2*i + c]: criterion c+1 of domain i.
int rotation_search_batch(const double* EE_host, int n_dom, double step_deg, double num_zero, hipStream_t s,
                          std::vector<RotResult>& out);
// The search without allocations: one domain, buffers in a fixed per-GPU work
// space of ROT_SCRATCH_BYTES (an initial step of 1 degree needs about 24 KB).
constexpr size_t ROT_SCRATCH_BYTES = 256 * 1024;
int rotation_search_noalloc(const double EE[36], double step_deg, double num_zero, RotResult out[2]);
int64_t rotation_noalloc_max_points();
void rotation_finish(const double EE[36], const int64_t ang[3], int64_t jj, double step_deg, RotResult& out);

}  // namespace dtc
