// io.hpp -- voxel block extraction and boundary tables (internal).
#pragma once
#include "dtc_internal.hpp"

namespace dtc {

int extract_block_device(const void* vol_dev, int kind, const int64_t vdim[3], const int64_t xa_n[3],
                         const int64_t xe_n[3], void* out_dev, hipStream_t s);
// ids_dev / vals_dev == null: count only; order 1: HEX08 macro (24 load cases), 2: HEX20 (60)
int boundaries_device(const double* nodes_dev, int64_t n, const double mn[3], const double mx[3], const double dim[3],
                      int order, int64_t* n_b, int64_t* ids_dev, double* vals_dev, DevMem& flag, DevMem& pos,
                      DevMem& tmp, hipStream_t s);

}  // namespace dtc
