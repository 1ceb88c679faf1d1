// mesh.hip -- voxel -> hex meshing on gfx950, bit-exact to gen_quadmesh_from_phi
// (f-src/mod_gen_quadmesh.f90:22-679).
//
// The reference works sequentially: elements in voxel scan order (x fastest),
// first-touch node numbering, max-label colouring until no change, a keep
// rule per colour and a stable compaction.  Here every step is a data-parallel
// integer pass over voxels, lattice nodes or node numbers:
//   * element index     = exclusive scan of the solid flags;
//   * node numbering    = the first toucher of a lattice node is its touching
//                         solid voxel with the smallest voxel index (element
//                         order follows voxel order), so each voxel counts the
//                         local nodes it touches first, a scan gives the base
//                         number and the local order gives the rest;
//   * colouring         = lock-free union-find over 26-neighbouring solid
//                         voxels (the colour partition is the connected-
//                         component partition; the label values are not output);
//   * keep rule         = component of the first element, or any voxel of the
//                         component on the node bounding box (the reference's
//                         0.1*delta coordinate tests are exact lattice equalities);
//   * compaction        = scans over the keep flags, with the last element and
//                         the last node always kept (:597-651).

#include "dtc_internal.hpp"

namespace dtc {

namespace {

constexpr int NT = 256;

__host__ __device__ inline int h20_off(int l, int d) {
    // HEX20 local node offsets on the doubled lattice (f-src/mod_gen_quadmesh.f90:244-265)
    constexpr signed char T[20][3] = {
        {0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {2, 1, 0}, {2, 2, 0}, {1, 2, 0}, {0, 2, 0}, {0, 1, 0},
        {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
        {0, 0, 2}, {1, 0, 2}, {2, 0, 2}, {2, 1, 2}, {2, 2, 2}, {1, 2, 2}, {0, 2, 2}, {0, 1, 2}};
    return T[l][d];
}

__device__ inline int local_of_offset(int nen, int ox, int oy, int oz) {
    if (nen == 8) {
        constexpr signed char L8[8] = {0, 1, 3, 2, 4, 5, 7, 6};   // index ox + 2oy + 4oz
        return L8[ox + 2 * oy + 4 * oz];
    }
    for (int l = 0; l < 20; ++l)
        if (h20_off(l, 0) == ox && h20_off(l, 1) == oy && h20_off(l, 2) == oz) return l;
    return -1;
}

__device__ inline int node_off(int nen, int l, int d) { return nen == 8 ? h08_off(l, d) : h20_off(l, d); }

// Candidate voxels touching lattice coordinate a along one axis, in increasing
// voxel order, with the node's offset inside that voxel.
__device__ inline int axis_candidates(int a, int mul, int n, int vox[2], int off[2]) {
    int c = 0;
    if (mul == 1) {
        if (a - 1 >= 0 && a - 1 < n) { vox[c] = a - 1; off[c] = 1; ++c; }
        if (a >= 0 && a < n) { vox[c] = a; off[c] = 0; ++c; }
    } else if ((a & 1) == 0) {
        int x = a >> 1;
        if (x - 1 >= 0 && x - 1 < n) { vox[c] = x - 1; off[c] = 2; ++c; }
        if (x >= 0 && x < n) { vox[c] = x; off[c] = 0; ++c; }
    } else {
        int x = a >> 1;
        if (x >= 0 && x < n) { vox[c] = x; off[c] = 1; ++c; }
    }
    return c;
}

// First toucher of lattice node (a,b,c): the solid voxel with the smallest
// voxel index among the voxels containing it; -1 if none.
__device__ inline int first_toucher(const uint8_t* __restrict__ solid, int nx, int ny, int nz, int mul,
                                    int a, int b, int c, int* ox, int* oy, int* oz) {
    int vx[2], fx[2], vy[2], fy[2], vz[2], fz[2];
    int cx = axis_candidates(a, mul, nx, vx, fx);
    int cy = axis_candidates(b, mul, ny, vy, fy);
    int cz = axis_candidates(c, mul, nz, vz, fz);
    for (int k = 0; k < cz; ++k)
        for (int j = 0; j < cy; ++j)
            for (int i = 0; i < cx; ++i) {
                int v = vx[i] + nx * (vy[j] + ny * vz[k]);
                if (solid[v]) { *ox = fx[i]; *oy = fy[j]; *oz = fz[k]; return v; }
            }
    return -1;
}

template <typename P>
__global__ void k_solid(const P* __restrict__ phi, int nvox, int64_t llimit, uint8_t* __restrict__ solid,
                        int* __restrict__ flag) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    int s = (int64_t)phi[v] >= llimit;   // f-src/mod_gen_quadmesh.f90:242,296
    solid[v] = (uint8_t)s;
    flag[v] = s;
}

__global__ void k_fill_i32(int* __restrict__ a, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_fill_u8(uint8_t* __restrict__ a, int64_t n, uint8_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_ftcount(const uint8_t* __restrict__ solid, int nx, int ny, int nz, int nen, int mul,
                          uint32_t* __restrict__ ftmask, int* __restrict__ cnt) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int nvox = nx * ny * nz;
    if (v >= nvox) return;
    uint32_t mask = 0;
    if (solid[v]) {
        int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
        for (int l = 0; l < nen; ++l) {
            int ox, oy, oz;
            int ft = first_toucher(solid, nx, ny, nz, mul, x * mul + node_off(nen, l, 0),
                                   y * mul + node_off(nen, l, 1), z * mul + node_off(nen, l, 2), &ox, &oy, &oz);
            if (ft == v) mask |= 1u << l;
        }
    }
    ftmask[v] = mask;
    cnt[v] = __popc(mask);
}

__global__ void k_number(const uint8_t* __restrict__ solid, const uint32_t* __restrict__ ftmask,
                         const int* __restrict__ nbase, int nx, int ny, int nz, int nen, int mul, int Lx, int Ly,
                         int* __restrict__ numLat, int* __restrict__ latOfNum) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nx * ny * nz || !solid[v]) return;
    uint32_t mask = ftmask[v];
    int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
    for (int l = 0; l < nen; ++l) {
        if (!(mask & (1u << l))) continue;
        int num = nbase[v] + __popc(mask & ((1u << l) - 1u)) + 1;
        int p = (x * mul + node_off(nen, l, 0)) + Lx * ((y * mul + node_off(nen, l, 1)) + Ly * (z * mul + node_off(nen, l, 2)));
        numLat[p] = num;
        latOfNum[num - 1] = p;
    }
}

__device__ inline int uf_find(int* parent, int x) {
    volatile int* pv = parent;
    for (;;) {
        int p = pv[x];
        if (p == x) return x;
        int gp = pv[p];
        if (gp != p) pv[x] = gp;   // path halving; any stale value is still an ancestor
        x = p;
    }
}

__device__ inline void uf_unite(int* parent, int a, int b) {
    for (;;) {
        a = uf_find(parent, a);
        b = uf_find(parent, b);
        if (a == b) return;
        if (a < b) { int t = a; a = b; b = t; }
        // hook the larger root under the smaller one: the final root of a
        // component is its smallest voxel index
        if (atomicCAS(&parent[a], a, b) == a) return;
    }
}

__global__ void k_uf_init(int* __restrict__ parent, int nvox) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvox) parent[v] = v;
}

// two solid voxels are connected when they share a node: 26-neighbourhood
__global__ void k_uf_union(const uint8_t* __restrict__ solid, int nx, int ny, int nz, int* parent) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nx * ny * nz || !solid[v]) return;
    int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
    for (int dz = -1; dz <= 0; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                if (dz == 0 && (dy > 0 || (dy == 0 && dx >= 0))) continue;   // 13 backward neighbours
                int xx = x + dx, yy = y + dy, zz = z + dz;
                if (xx < 0 || yy < 0 || zz < 0 || xx >= nx || yy >= ny) continue;
                int u = xx + nx * (yy + ny * zz);
                if (solid[u]) uf_unite(parent, v, u);
            }
}

__global__ void k_uf_compress(const uint8_t* __restrict__ solid, int nvox, int* parent) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nvox && solid[v]) parent[v] = uf_find(parent, v);
}

__global__ void k_bbox(const uint8_t* __restrict__ solid, int nx, int ny, int nz, int* __restrict__ bbox) {
    __shared__ int s[6][NT];
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    int t = threadIdx.x;
    int mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, mx[3] = {-1, -1, -1};
    if (v < nx * ny * nz && solid[v]) {
        int c[3] = {v % nx, (v / nx) % ny, v / (nx * ny)};
        for (int d = 0; d < 3; ++d) { mn[d] = c[d]; mx[d] = c[d]; }
    }
    for (int d = 0; d < 3; ++d) { s[2 * d][t] = mn[d]; s[2 * d + 1][t] = mx[d]; }
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w)
            for (int d = 0; d < 3; ++d) {
                s[2 * d][t] = min(s[2 * d][t], s[2 * d][t + w]);
                s[2 * d + 1][t] = max(s[2 * d + 1][t], s[2 * d + 1][t + w]);
            }
        __syncthreads();
    }
    if (t == 0)
        for (int d = 0; d < 3; ++d) {
            if (s[2 * d][0] != INT32_MAX) atomicMin(&bbox[2 * d], s[2 * d][0]);
            if (s[2 * d + 1][0] >= 0) atomicMax(&bbox[2 * d + 1], s[2 * d + 1][0]);
        }
}

// keep rule, f-src/mod_gen_quadmesh.f90:482-553: the first colour (component
// of node 1 = component of the first element) always; any other component
// with a node on the bounding box of all nodes
__global__ void k_touch(const uint8_t* __restrict__ solid, const int* __restrict__ eidx,
                        const int* __restrict__ parent, const int* __restrict__ bbox, int nx, int ny, int nz,
                        uint8_t* __restrict__ touch) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nx * ny * nz || !solid[v]) return;
    int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
    bool hit = (eidx[v] == 0) || x == bbox[0] || x == bbox[1] || y == bbox[2] || y == bbox[3] ||
               z == bbox[4] || z == bbox[5];
    if (hit) touch[parent[v]] = 1;
}

__global__ void k_keepE(const uint8_t* __restrict__ solid, const int* __restrict__ eidx,
                        const int* __restrict__ parent, const uint8_t* __restrict__ touch, int nvox, int ne0,
                        int* __restrict__ keepE) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    int k = 0;
    if (solid[v]) k = (touch[parent[v]] || eidx[v] == ne0 - 1) ? 1 : 0;   // last element kept (:597-614)
    keepE[v] = k;
}

__global__ void k_keepN(const uint8_t* __restrict__ solid, const int* __restrict__ latOfNum,
                        const int* __restrict__ parent, const uint8_t* __restrict__ touch, int nx, int ny, int nz,
                        int mul, int Lx, int Ly, int nn0, int* __restrict__ keepN) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn0) return;
    int p = latOfNum[k];
    int a = p % Lx, b = (p / Lx) % Ly, c = p / (Lx * Ly);
    int ox, oy, oz;
    int ft = first_toucher(solid, nx, ny, nz, mul, a, b, c, &ox, &oy, &oz);
    keepN[k] = (touch[parent[ft]] || k == nn0 - 1) ? 1 : 0;   // last node kept (:619-651)
}

__global__ void k_out_elems(const int* __restrict__ keepE, const int* __restrict__ newE,
                            const int* __restrict__ numLat, const int* __restrict__ keepN,
                            const int* __restrict__ cref, int nx, int ny, int nz, int nen, int mul, int Lx,
                            int Ly, int64_t* __restrict__ elems) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nx * ny * nz || !keepE[v]) return;
    int x = v % nx, y = (v / nx) % ny, z = v / (nx * ny);
    int64_t e = newE[v];
    for (int l = 0; l < nen; ++l) {
        int p = (x * mul + node_off(nen, l, 0)) + Lx * ((y * mul + node_off(nen, l, 1)) + Ly * (z * mul + node_off(nen, l, 2)));
        int num = numLat[p];
        elems[e * nen + l] = keepN[num - 1] ? (int64_t)cref[num - 1] + 1 : 0;   // :656-660
    }
}

// coordinates, f-src/mod_gen_quadmesh.f90:346-384 (same IEEE operations)
__global__ void k_out_nodes(const int* __restrict__ keepN, const int* __restrict__ cref,
                            const int* __restrict__ latOfNum, int nn0, int mul, int Lx, int Ly, double dx,
                            double dy, double dz, int64_t xa0, int64_t xa1, int64_t xa2, double* __restrict__ nodes) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn0 || !keepN[k]) return;
    int p = latOfNum[k];
    int64_t ii = p % Lx, jj = (p / Lx) % Ly, kk = p / (Lx * Ly);
    int64_t o = 3 * (int64_t)cref[k];
    if (mul == 2) {
        nodes[o + 2] = ((double)kk / 2.0 + (double)(xa2 - 1)) * dz;
        nodes[o + 1] = ((double)jj / 2.0 + (double)(xa1 - 1)) * dy;
        nodes[o + 0] = ((double)ii / 2.0 + (double)(xa0 - 1)) * dx;
    } else {
        nodes[o + 2] = (double)(kk + xa2 - 1) * dz;
        nodes[o + 1] = (double)(jj + xa1 - 1) * dy;
        nodes[o + 0] = (double)(ii + xa0 - 1) * dx;
    }
}

inline unsigned grid_of(int64_t n) { return (unsigned)((n + NT - 1) / NT); }

}  // namespace

namespace {

// Exclusive prefix sums of int32 values: tiles of SCAN_TILE values (256 threads x
// 4), the tile totals scanned by one workgroup, then every tile's values
// rescanned with its offset.  Integer sums: exact and independent of the order.
constexpr int SCAN_NT = 256, SCAN_ITEMS = 4, SCAN_TILE = SCAN_NT * SCAN_ITEMS;

// inclusive scan of one value per thread over the workgroup; returns the
// workgroup total in *total
__device__ inline int block_incl_scan(int v, int* s_w, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o);
        if (lane >= o) v += u;
    }
    if (lane == 63) s_w[wave] = v;
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < SCAN_NT / 64; ++w) {
        const int c = s_w[w];
        if (w < wave) base += c;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return v + base;
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_tiles(const int* __restrict__ in, int64_t n, int* __restrict__ tile_sum) {
    __shared__ int s_w[SCAN_NT / 64];
    const int64_t t0 = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (t0 + k < n) v += in[t0 + k];
    int tot;
    (void)block_incl_scan(v, s_w, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// exclusive scan of the tile totals in place (one workgroup, SCAN_NT per step)
__global__ __launch_bounds__(SCAN_NT) void k_scan_totals(int* __restrict__ tile_sum, int64_t n_tiles) {
    __shared__ int s_w[SCAN_NT / 64];
    int carry = 0;
    for (int64_t b0 = 0; b0 < n_tiles; b0 += SCAN_NT) {
        const int64_t i = b0 + threadIdx.x;
        const int v = i < n_tiles ? tile_sum[i] : 0;
        int tot;
        const int inc = block_incl_scan(v, s_w, &tot);
        if (i < n_tiles) tile_sum[i] = carry + inc - v;
        carry += tot;
    }
}

__global__ __launch_bounds__(SCAN_NT) void k_scan_apply(const int* __restrict__ in, int64_t n,
                                                       const int* __restrict__ tile_off, int* __restrict__ out) {
    __shared__ int s_w[SCAN_NT / 64];
    const int64_t t0 = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int v[SCAN_ITEMS], sum = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = t0 + k < n ? in[t0 + k] : 0;
        sum += v[k];
    }
    int tot;
    int run = block_incl_scan(sum, s_w, &tot) - sum + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (t0 + k < n) out[t0 + k] = run;
        run += v[k];
    }
}

}  // namespace

int scan_exclusive_i32(const int* in, int* out, int64_t n, int64_t* total, DevMem& tmp, hipStream_t s) {
    if (n <= 0) { *total = 0; return DTC_OK; }
    const int64_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    DTC_TRY(tmp.ensure((size_t)tiles * sizeof(int)));
    k_scan_tiles<<<(unsigned)tiles, SCAN_NT, 0, s>>>(in, n, tmp.as<int>());
    k_scan_totals<<<1, SCAN_NT, 0, s>>>(tmp.as<int>(), tiles);
    k_scan_apply<<<(unsigned)tiles, SCAN_NT, 0, s>>>(in, n, tmp.as<int>(), out);
    DTC_HIP(hipGetLastError());
    int last_in = 0, last_out = 0;
    DTC_HIP(hipMemcpyAsync(&last_in, in + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    DTC_HIP(hipMemcpyAsync(&last_out, out + n - 1, sizeof(int), hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    *total = (int64_t)last_in + last_out;
    return DTC_OK;
}

int mesh_build_device(const void* phi_dev, int kind, const int64_t n[3], int64_t llimit, int nen,
                      MeshDev& m, hipStream_t s) {
    if (nen != 8 && nen != 20) return set_error(DTC_ERR_BAD_ARGUMENT, "micro element nodes must be 8 or 20, got %d", nen);
    if (kind != DTC_IK1 && kind != DTC_IK2 && kind != DTC_IK4)
        return set_error(DTC_ERR_BAD_ARGUMENT, "voxel kind must be 1, 2 or 4, got %d", kind);
    for (int d = 0; d < 3; ++d)
        if (n[d] <= 0) return set_error(DTC_ERR_BAD_ARGUMENT, "domain extent n[%d] = %lld", d, (long long)n[d]);
    m.nx = (int)n[0]; m.ny = (int)n[1]; m.nz = (int)n[2];
    m.nen = nen; m.mul = nen == 8 ? 1 : 2;
    const int64_t Lx = n[0] * m.mul + 1, Ly = n[1] * m.mul + 1, Lz = n[2] * m.mul + 1;
    const int64_t nvox = n[0] * n[1] * n[2];
    if (Lx * Ly * Lz >= (int64_t)INT32_MAX / 2)
        return set_error(DTC_ERR_BAD_ARGUMENT, "domain of %lld^3 voxels exceeds the 32-bit lattice index", (long long)n[0]);
    m.Lx = (int)Lx; m.Ly = (int)Ly; m.Lz = (int)Lz;
    const int64_t nlat = Lx * Ly * Lz;
    const int nv = (int)nvox;

    DTC_TRY(m.solid.ensure(nvox));
    DTC_TRY(m.eidx.ensure(nvox * 4));
    DTC_TRY(m.ftmask.ensure(nvox * 4));
    DTC_TRY(m.nbase.ensure(nvox * 4));
    DTC_TRY(m.parent.ensure(nvox * 4));
    DTC_TRY(m.touch.ensure(nvox));
    DTC_TRY(m.keepE.ensure(nvox * 4));
    DTC_TRY(m.newE.ensure(nvox * 4));
    DTC_TRY(m.numLat.ensure(nlat * 4));
    DTC_TRY(m.bbox.ensure(6 * 4));
    DevMem flag;
    DTC_TRY(flag.ensure(nvox * 4));

    switch (kind) {
    case DTC_IK1: k_solid<int8_t><<<grid_of(nvox), NT, 0, s>>>((const int8_t*)phi_dev, nv, llimit, m.solid.as<uint8_t>(), flag.as<int>()); break;
    case DTC_IK2: k_solid<int16_t><<<grid_of(nvox), NT, 0, s>>>((const int16_t*)phi_dev, nv, llimit, m.solid.as<uint8_t>(), flag.as<int>()); break;
    default: k_solid<int32_t><<<grid_of(nvox), NT, 0, s>>>((const int32_t*)phi_dev, nv, llimit, m.solid.as<uint8_t>(), flag.as<int>()); break;
    }
    DTC_HIP(hipGetLastError());
    DTC_TRY(scan_exclusive_i32(flag.as<int>(), m.eidx.as<int>(), nvox, &m.n_elems0, m.scratch, s));
    if (m.n_elems0 == 0) {
        m.n_elems = m.n_nodes = m.n_nodes0 = 0;
        return set_error(DTC_ERR_EMPTY_MESH, "no voxel reaches the iso value limit %lld", (long long)llimit);
    }

    // first-touch numbering
    k_ftcount<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.nx, m.ny, m.nz, nen, m.mul, m.ftmask.as<uint32_t>(), flag.as<int>());
    DTC_HIP(hipGetLastError());
    DTC_TRY(scan_exclusive_i32(flag.as<int>(), m.nbase.as<int>(), nvox, &m.n_nodes0, m.scratch, s));
    DTC_TRY(m.latOfNum.ensure(m.n_nodes0 * 4));
    DTC_TRY(m.keepN.ensure(m.n_nodes0 * 4));
    DTC_TRY(m.cref.ensure(m.n_nodes0 * 4));
    k_fill_i32<<<grid_of(nlat), NT, 0, s>>>(m.numLat.as<int>(), nlat, 0);
    k_number<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.ftmask.as<uint32_t>(), m.nbase.as<int>(), m.nx, m.ny, m.nz,
                                          nen, m.mul, m.Lx, m.Ly, m.numLat.as<int>(), m.latOfNum.as<int>());
    DTC_HIP(hipGetLastError());

    // connected components
    k_uf_init<<<grid_of(nvox), NT, 0, s>>>(m.parent.as<int>(), nv);
    k_uf_union<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.nx, m.ny, m.nz, m.parent.as<int>());
    k_uf_compress<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), nv, m.parent.as<int>());
    DTC_HIP(hipGetLastError());

    // bounding box of all nodes (in voxel terms) and keep rule
    int bb0[6] = {INT32_MAX, -1, INT32_MAX, -1, INT32_MAX, -1};
    DTC_HIP(hipMemcpyAsync(m.bbox.p, bb0, sizeof(bb0), hipMemcpyHostToDevice, s));
    k_bbox<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.nx, m.ny, m.nz, m.bbox.as<int>());
    k_fill_u8<<<grid_of(nvox), NT, 0, s>>>(m.touch.as<uint8_t>(), nvox, 0);
    k_touch<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.eidx.as<int>(), m.parent.as<int>(), m.bbox.as<int>(),
                                         m.nx, m.ny, m.nz, m.touch.as<uint8_t>());
    k_keepE<<<grid_of(nvox), NT, 0, s>>>(m.solid.as<uint8_t>(), m.eidx.as<int>(), m.parent.as<int>(), m.touch.as<uint8_t>(),
                                         nv, (int)m.n_elems0, m.keepE.as<int>());
    DTC_HIP(hipGetLastError());
    DTC_TRY(scan_exclusive_i32(m.keepE.as<int>(), m.newE.as<int>(), nvox, &m.n_elems, m.scratch, s));
    k_keepN<<<grid_of(m.n_nodes0), NT, 0, s>>>(m.solid.as<uint8_t>(), m.latOfNum.as<int>(), m.parent.as<int>(),
                                               m.touch.as<uint8_t>(), m.nx, m.ny, m.nz, m.mul, m.Lx, m.Ly,
                                               (int)m.n_nodes0, m.keepN.as<int>());
    DTC_HIP(hipGetLastError());
    DTC_TRY(scan_exclusive_i32(m.keepN.as<int>(), m.cref.as<int>(), m.n_nodes0, &m.n_nodes, m.scratch, s));
    return DTC_OK;
}

int mesh_write_outputs(const MeshDev& m, const double delta[3], const int64_t xa_n[3], double* nodes_dev,
                       int64_t* elems_dev, hipStream_t s) {
    const int64_t nvox = (int64_t)m.nx * m.ny * m.nz;
    if (elems_dev) {
        k_out_elems<<<grid_of(nvox), NT, 0, s>>>(m.keepE.as<int>(), m.newE.as<int>(), m.numLat.as<int>(), m.keepN.as<int>(),
                                                 m.cref.as<int>(), m.nx, m.ny, m.nz, m.nen, m.mul, m.Lx, m.Ly, elems_dev);
        DTC_HIP(hipGetLastError());
    }
    if (nodes_dev) {
        k_out_nodes<<<grid_of(m.n_nodes0), NT, 0, s>>>(m.keepN.as<int>(), m.cref.as<int>(), m.latOfNum.as<int>(),
                                                       (int)m.n_nodes0, m.mul, m.Lx, m.Ly, delta[0], delta[1], delta[2],
                                                       xa_n[0], xa_n[1], xa_n[2], nodes_dev);
        DTC_HIP(hipGetLastError());
    }
    return DTC_OK;
}

}  // namespace dtc
