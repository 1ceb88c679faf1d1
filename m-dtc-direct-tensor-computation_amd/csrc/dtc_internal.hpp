// dtc_internal.hpp -- shared declarations of the gfx950 library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/dtc_capi.h"

namespace dtc {

// ---------------------------------------------------------------- errors
int set_error(int code, const char* fmt, ...);

#define DTC_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            if (e_ == hipErrorOutOfMemory)                                               \
                return ::dtc::set_error(DTC_ERR_OOM, "%s: %s", #call, hipGetErrorString(e_)); \
            return ::dtc::set_error(DTC_ERR_DEVICE, "%s: %s (%s:%d)", #call,              \
                                    hipGetErrorString(e_), __FILE__, __LINE__);          \
        }                                                                                \
    } while (0)

#define DTC_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != DTC_OK) return rc_; \
    } while (0)

// ---------------------------------------------------------------- device buffers
// Owning device allocation; allocation failure is reported as DTC_ERR_OOM.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    DevMem(DevMem&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevMem& operator=(DevMem&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p; bytes = o.bytes;
            o.p = nullptr; o.bytes = 0;
        }
        return *this;
    }
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int alloc(size_t n) {
        release();
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return set_error(DTC_ERR_OOM, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return DTC_OK;
    }
    // keep the allocation when it is large enough (repeated runs reuse buffers)
    int ensure(size_t n) {
        if (p && bytes >= n) return DTC_OK;
        return alloc(n);
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// ---------------------------------------------------------------- constants
constexpr int R24 = 24;          // load cases of the HEX08 macro element
constexpr int BX = 8, BY = 4, BZ = 4;   // owned nodes of one solver brick
constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;   // halo node tile
constexpr int EX = BX + 1, EY = BY + 1, EZ = BZ + 1;   // element tile

// HEX08 local node offsets in gen_quadmesh order (f-src/mod_gen_quadmesh.f90:298-306)
__host__ __device__ constexpr int h08_off(int l, int d) {
    // l: 0..7, d: 0..2
    return (d == 0) ? ((l == 1 || l == 2 || l == 5 || l == 6) ? 1 : 0)
         : (d == 1) ? ((l == 2 || l == 3 || l == 6 || l == 7) ? 1 : 0)
                    : (l >= 4 ? 1 : 0);
}

// ---------------------------------------------------------------- mesher
struct MeshDev {
    int nx = 0, ny = 0, nz = 0, nen = 8, mul = 1;
    int Lx = 0, Ly = 0, Lz = 0;       // lattice extents
    int64_t n_elems0 = 0, n_nodes0 = 0;   // before connectivity filtering
    int64_t n_elems = 0, n_nodes = 0;     // kept
    DevMem solid;     // uint8 per voxel
    DevMem eidx;      // int32 per voxel: element index before filtering
    DevMem ftmask;    // uint32 per voxel: first-touched local nodes
    DevMem nbase;     // int32 per voxel: first node number - 1 of the voxel's first touches
    DevMem numLat;    // int32 per lattice node: node number (1-based), 0 = unused
    DevMem latOfNum;  // int32 per node number: lattice index
    DevMem parent;    // int32 per voxel: union-find
    DevMem touch;     // uint8 per voxel (component roots)
    DevMem keepE;     // int32 per voxel
    DevMem newE;      // int32 per voxel: kept element index
    DevMem keepN;     // int32 per node number
    DevMem cref;      // int32 per node number: kept node number - 1
    DevMem bbox;      // int32[6]
    DevMem scratch;   // scan temp storage
};

// Builds the mesh of one voxel block already resident on the device.
// Returns DTC_ERR_EMPTY_MESH when no voxel reaches llimit.
int mesh_build_device(const void* phi_dev, int kind, const int64_t n[3], int64_t llimit, int nen,
                      MeshDev& m, hipStream_t s);
// Writes nodes (3 x N, column-major) and elems (nen x N_e, 1-based) to device buffers.
int mesh_write_outputs(const MeshDev& m, const double delta[3], const int64_t xa_n[3],
                       double* nodes_dev, int64_t* elems_dev, hipStream_t s);

// exclusive prefix sum of int32 (n items), total returned through *total (host)
int scan_exclusive_i32(const int* in, int* out, int64_t n, int64_t* total, DevMem& tmp, hipStream_t s);

// ---------------------------------------------------------------- solver topology
struct DomInfo {
    int nx, ny, nz;           // voxels per edge
    int macro_order;          // 1: HEX08 macro element (24 load cases), 2: HEX20 (60)
    int lc_base;              // first macro load case of this group of 24
    int bnx, bny, bnz;        // brick grid
    int brick_begin, brick_end;
    int64_t lat_off;          // offset into the lattice map
    int64_t vox_off;          // offset into the voxel modulus array
    int64_t node_begin, node_end;
    double min_c[3], max_c[3];   // domain box (f-src/mod_struct_preprocess.f90:516-517)
    double xa1[3];               // (xa_n - 1)
    double delta[3];
};

struct BrickInfo {
    int dom;
    int x0, y0, z0;           // first owned lattice node
    int node_begin, node_end; // owned solver nodes (contiguous)
    int e_off, e_cnt;         // solid elements of the element tile in the brick element tables
};

// New Chebyshev-Jacobi direction of one unknown, c1 d + c2 D^-1 (b - t) with t = (K x)
// of the unknown: one expression for every kernel that forms it, so that a step fused
// into an operator's epilogue rounds like the separate vector kernel
__device__ __forceinline__ float cheb_dir(float c1, float d, float c2, float di, float b, float t) {
    return c1 * d + c2 * di * (b - t);
}

// brick halo table entries: solver node | BC_BIT for a Dirichlet node, -1 = no node
constexpr int BC_BIT = 1 << 30;
constexpr int NODE_MASK = BC_BIT - 1;

// 45 coefficients of the HEX08 element stiffness in the signed Walsh basis
// (E = 1); see walsh_blocks() in solver.hip.
struct WalshK {
    double g[45];
};

}  // namespace dtc
