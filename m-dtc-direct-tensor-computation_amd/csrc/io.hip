// io.hip -- voxel block extraction and boundary-condition tables on gfx950.
//
//  * k_extract: the domain block Phi(xa_n:xe_n)^3 of a whole volume, x
//    fastest (f-src/mod_struct_preprocess.f90:230-244).
//  * k_bc_flag / k_bc_write: generate_boundaries (f-src/mod_struct_preprocess.f90:
//    490-738): boundary nodes in node order, tested against the domain box
//    faces in the reference's order, and their prescribed displacements of
//    the 24 (HEX08 macro) or 60 (HEX20 macro) load cases
//    (f-src/mod_write_deck.f90:22-279), bit-identical to the reference's FP64
//    operations.
#include "bc.hpp"
#include "dtc_internal.hpp"
#include "io.hpp"

namespace dtc {

namespace {

constexpr int NT = 256;

__global__ void k_extract(const char* __restrict__ vol, int es, int64_t vx, int64_t vy, int64_t x0, int64_t y0,
                          int64_t z0, int64_t nx, int64_t ny, int64_t nz, char* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nx * ny * nz) return;
    int64_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
    int64_t src = (x0 + x) + vx * ((y0 + y) + vy * (z0 + z));
    for (int b = 0; b < es; ++b) out[i * es + b] = vol[src * es + b];
}

__global__ void k_bc_flag(const double* __restrict__ nodes, int64_t n, double mn0, double mn1, double mn2, double mx0,
                          double mx1, double mx2, double d0, double d1, double d2, int* __restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double mn[3] = {mn0, mn1, mn2}, mx[3] = {mx0, mx1, mx2}, dm[3] = {d0, d1, d2};
    flag[i] = on_box_face(nodes + 3 * i, mn, mx, dm) ? 1 : 0;
}

__global__ void k_bc_write(const double* __restrict__ nodes, int64_t n, const int* __restrict__ flag,
                           const int* __restrict__ pos, double mn0, double mn1, double mn2, double mx0, double mx1,
                           double mx2, int order, int64_t* __restrict__ ids, double* __restrict__ vals) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const double mn[3] = {mn0, mn1, mn2}, mx[3] = {mx0, mx1, mx2};
    const int64_t b = pos[i];
    const int RM = macro_lcs(order);
    ids[b] = i + 1;
    double N[20];
    macro_shape(order, nodes + 3 * i, mn, mx, N);
    for (int d = 0; d < 3; ++d)
        for (int lc = 0; lc < RM; ++lc) vals[lc + (int64_t)RM * (3 * b + d)] = bc_value_m(order, N, d, lc);
}

}  // namespace

int extract_block_device(const void* vol_dev, int kind, const int64_t vdim[3], const int64_t xa_n[3],
                         const int64_t xe_n[3], void* out_dev, hipStream_t s) {
    const int es = kind == DTC_IK1 ? 1 : kind == DTC_IK2 ? 2 : 4;
    int64_t n[3];
    for (int d = 0; d < 3; ++d) {
        n[d] = xe_n[d] - xa_n[d] + 1;
        if (n[d] <= 0 || xa_n[d] < 1 || xe_n[d] > vdim[d])
            return set_error(DTC_ERR_BAD_ARGUMENT, "block [%lld,%lld] outside 1..%lld on axis %d", (long long)xa_n[d],
                             (long long)xe_n[d], (long long)vdim[d], d);
    }
    const int64_t cnt = n[0] * n[1] * n[2];
    k_extract<<<(unsigned)((cnt + NT - 1) / NT), NT, 0, s>>>((const char*)vol_dev, es, vdim[0], vdim[1], xa_n[0] - 1,
                                                             xa_n[1] - 1, xa_n[2] - 1, n[0], n[1], n[2], (char*)out_dev);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

int boundaries_device(const double* nodes_dev, int64_t n, const double mn[3], const double mx[3], const double dim[3],
                      int order, int64_t* n_b, int64_t* ids_dev, double* vals_dev, DevMem& flag, DevMem& pos,
                      DevMem& tmp, hipStream_t s) {
    if (n <= 0) { *n_b = 0; return DTC_OK; }
    DTC_TRY(flag.alloc((size_t)n * 4));
    DTC_TRY(pos.alloc((size_t)n * 4));
    const unsigned g = (unsigned)((n + NT - 1) / NT);
    k_bc_flag<<<g, NT, 0, s>>>(nodes_dev, n, mn[0], mn[1], mn[2], mx[0], mx[1], mx[2], dim[0], dim[1], dim[2], flag.as<int>());
    DTC_HIP(hipGetLastError());
    DTC_TRY(scan_exclusive_i32(flag.as<int>(), pos.as<int>(), n, n_b, tmp, s));
    if (ids_dev && vals_dev) {
        k_bc_write<<<g, NT, 0, s>>>(nodes_dev, n, flag.as<int>(), pos.as<int>(), mn[0], mn[1], mn[2], mx[0], mx[1], mx[2],
                                    order, ids_dev, vals_dev);
        DTC_HIP(hipGetLastError());
    }
    return DTC_OK;
}

}  // namespace dtc
