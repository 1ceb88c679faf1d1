// solver20.hip -- the matrix-free operator of HEX20 micro-element domains.
//
// A HEX20 micro mesh (gen_quadmesh_from_phi with 20-node elements,
// f-src/mod_gen_quadmesh.f90:244-265) is solved with the same batched PCG,
// reductions and FP64 refinement as the HEX08 voxel lattice; only the operator
// and the node-indexed auxiliaries differ:
//  * solver nodes are the mesh nodes in mesh order (node_begin + mesh id - 1);
//    consecutive runs of 128 nodes form the pseudo-bricks that the vector
//    kernels and the per-brick partial sums work on;
//  * y = K x is a gather per node: for every (element, local node a) that
//    touches the node, the three rows 3a..3a+2 of the 60x60 element stiffness
//    (E = 1, in LDS) times the element's 60 gathered values, scaled by the
//    element modulus.  Adjacency lists are sorted, so sums are deterministic;
//  * Dirichlet nodes are the mesh nodes on the domain box faces
//    (f-src/mod_struct_preprocess.f90:564-592), Jacobi uses diag(K) per DOF.
#include <algorithm>
#include <cmath>

#include "bc.hpp"
#include "solver.hpp"

namespace dtc {

namespace {

constexpr int R = R24;
constexpr int NT20 = 192;        // 8 nodes x 24 load cases per pass
constexpr int PB = 128;          // nodes per pseudo-brick
constexpr int FFD = 60;          // macro DOF slots of the force partials

// element -> voxel modulus (kept elements only)
// (and element -> voxel index, the sensitivity's output position)
__global__ void k20_emod(const int* __restrict__ keepE, const int* __restrict__ newE, const double* __restrict__ modulus,
                         double E, double escale, int64_t nvox, double* __restrict__ emod, int* __restrict__ evox) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox || !keepE[v]) return;
    emod[newE[v]] = (modulus ? modulus[v] : E) * escale;
    evox[newE[v]] = (int)v;
}

// 1-based int64 connectivity -> 0-based int32 local node ids (-1: node dropped)
__global__ void k20_elems(const int64_t* __restrict__ e64, int64_t n, int* __restrict__ e32) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) e32[i] = (int)e64[i] - 1;
}

__global__ void k20_bc(const double* __restrict__ xyz, int64_t nn, DomInfo D, const double* __restrict__ dim,
                       unsigned char* __restrict__ bcf, int* __restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int f = on_box_face(xyz + 3 * i, D.min_c, D.max_c, dim) ? 1 : 0;
    bcf[i] = (unsigned char)f;
    if (f) atomicAdd(count, 1);
}

__global__ void k20_adj_count(const int* __restrict__ elems, int64_t ne, int* __restrict__ cnt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * 20) return;
    const int q = elems[i];
    if (q >= 0) atomicAdd(&cnt[q], 1);
}

__global__ void k20_adj_fill(const int* __restrict__ elems, int64_t ne, const int* __restrict__ off,
                             int* __restrict__ cursor, int* __restrict__ adj) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne * 20) return;
    const int q = elems[i];
    if (q < 0) return;
    const int e = (int)(i / 20), a = (int)(i % 20);
    adj[off[q] + atomicAdd(&cursor[q], 1)] = (e << 5) | a;
}

// half-voxel lattice point -> local kept node id (-1: no kept node)
__global__ void k20_latmap(const int* __restrict__ numLat, const int* __restrict__ keepN, const int* __restrict__ cref,
                           int64_t nl, int* __restrict__ lat20) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nl) return;
    const int num = numLat[i];
    lat20[i] = (num > 0 && keepN[num - 1]) ? cref[num - 1] : -1;
}

// sort each node's (element, local node) list: fixed summation order
__global__ void k20_adj_sort(const int* __restrict__ off, int64_t nn, int* __restrict__ adj) {
    const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nn) return;
    const int b = off[q], e = off[q + 1];
    for (int i = b + 1; i < e; ++i) {
        const int v = adj[i];
        int j = i - 1;
        while (j >= b && adj[j] > v) { adj[j + 1] = adj[j]; --j; }
        adj[j + 1] = v;
    }
}

// Jacobi: diag(K) per DOF from the node's elements; 0 on Dirichlet DOFs, 1 for a zero diagonal
__global__ void k20_dinv(Dom20Dev G, const double* __restrict__ K, int64_t node0, double* __restrict__ d64,
                         float* __restrict__ d32) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= G.n_nodes) return;
    for (int d = 0; d < 3; ++d) {
        double s = 0.0;
        for (int k = G.adj_off[q]; k < G.adj_off[q + 1]; ++k) {
            const int e = G.adj[k] >> 5, a = G.adj[k] & 31;
            s += G.emod[e] * K[(3 * a + d) * 60 + 3 * a + d];
        }
        double inv = s != 0.0 ? 1.0 / s : 1.0;
        if (G.bcf[q]) inv = 0.0;
        d64[3 * (node0 + q) + d] = inv;
        d32[3 * (node0 + q) + d] = (float)inv;
    }
}

// boundary value of node ql (local) for load case lc of the domain's group
__device__ inline double bc20(const DomInfo& D, const Dom20Dev& G, int ql, int d, int lc) {
    double N[20];
    macro_shape(D.macro_order, G.coords + 3 * ql, D.min_c, D.max_c, N);
    return bc_value_m(D.macro_order, N, d, D.lc_base + lc);
}

struct Mv20 {
    const BrickInfo* bricks;
    const DomInfo* doms;
    const Dom20Dev* d20;
    const int* done;
    double* partial;
};

// y = K x on the pseudo-bricks; modes as the HEX08 kernel (IN_V / IN_U input,
// MASKED_DOT / RESID / PLAIN / MASKED output)
// VT: the vectors' type (T, or float with T = double: FP32 vectors, FP64 arithmetic)
template <typename T, int IN, int OUT, typename VT = T>
__global__ __launch_bounds__(NT20) void k_matvec20(Mv20 A, const T* __restrict__ Kg, const VT* __restrict__ xin,
                                                   VT* __restrict__ yout) {
    __shared__ T sK[3600];
    __shared__ double s_red[NT20];
    const BrickInfo B = A.bricks[blockIdx.x];
    if (A.done && A.done[B.dom]) return;
    for (int i = threadIdx.x; i < 3600; i += NT20) sK[i] = Kg[i];
    __syncthreads();
    const DomInfo D = A.doms[B.dom];
    const Dom20Dev G = A.d20[B.dom];
    const int tid = threadIdx.x, r = tid % R, g = tid / R;
    double acc = 0.0;
    for (int q = B.node_begin + g; q < B.node_end; q += NT20 / R) {
        const int ql = (int)(q - G.node_begin);
        T y[3] = {T(0), T(0), T(0)};
        for (int k = G.adj_off[ql]; k < G.adj_off[ql + 1]; ++k) {
            const int e = G.adj[k] >> 5, a = G.adj[k] & 31;
            const int* en = G.elems + (int64_t)e * 20;
            const T* Ka = sK + 3 * a * 60;
            T s[3] = {T(0), T(0), T(0)};
            for (int bb = 0; bb < 20; ++bb) {
                const int qb = en[bb];
                if (qb < 0) continue;
                T xv[3];
#pragma unroll
                for (int d2 = 0; d2 < 3; ++d2) {
                    if (IN == 1 && G.bcf[qb]) xv[d2] = (T)bc20(D, G, qb, d2, r);
                    else if (IN == 1 && !xin) xv[d2] = T(0);
                    else xv[d2] = (T)xin[((int64_t)(G.node_begin + qb) * 3 + d2) * R + r];
                }
#pragma unroll
                for (int dd = 0; dd < 3; ++dd)
#pragma unroll
                    for (int d2 = 0; d2 < 3; ++d2) s[dd] += Ka[dd * 60 + 3 * bb + d2] * xv[d2];
            }
            const T E = (T)G.emod[e];
#pragma unroll
            for (int dd = 0; dd < 3; ++dd) y[dd] += E * s[dd];
        }
        const bool bc = G.bcf[ql] != 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int64_t o = ((int64_t)q * 3 + d) * R + r;
            T v = y[d];
            if (OUT == 0) {          // masked, x . y
                if (bc) v = T(0);
                acc += (double)xin[o] * (double)v;
            } else if (OUT == 1) {   // -y on free rows, y . y
                v = bc ? T(0) : -v;
                acc += (double)v * (double)v;
            } else if (OUT == 3) {
                if (bc) v = T(0);
            }
            yout[o] = (VT)v;
        }
    }
    if (OUT == 0 || OUT == 1) {
        s_red[tid] = acc;
        __syncthreads();
        if (tid < R) {
            double sum = 0.0;
            for (int gg = 0; gg < NT20 / R; ++gg) sum += s_red[gg * R + tid];
            A.partial[(int64_t)blockIdx.x * R + tid] = sum;
        }
    }
}

// force projection over the domain's mesh nodes (HEX20 micro), same partial
// layout as the lattice kernel: [brick][24 lc][FFD]
template <int NEN>
__global__ __launch_bounds__(NT20) void k_project20(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                                                    const Dom20Dev* __restrict__ d20, const double* __restrict__ rf,
                                                    double* __restrict__ pff) {
    constexpr int G8 = NT20 / R, FD = 3 * NEN, PASS = 12;
    __shared__ double s[G8][R][PASS];
    const BrickInfo B = bricks[blockIdx.x];
    const DomInfo D = doms[B.dom];
    if (macro_nodes(D.macro_order) != NEN) return;
    const Dom20Dev G = d20[B.dom];
    const int tid = threadIdx.x, lc = tid % R, g = tid / R;
    double acc[FD];
#pragma unroll
    for (int k = 0; k < FD; ++k) acc[k] = 0.0;
    for (int q = B.node_begin + g; q < B.node_end; q += G8) {
        const int ql = (int)(q - G.node_begin);
        double N[20];
        macro_shape(D.macro_order, G.coords + 3 * ql, D.min_c, D.max_c, N);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double rv = rf[((int64_t)q * 3 + d) * R + lc];
#pragma unroll
            for (int l = 0; l < NEN; ++l) acc[d * NEN + l] += rv * N[l];
        }
    }
#pragma unroll
    for (int k0 = 0; k0 < FD; k0 += PASS) {
#pragma unroll
        for (int k = 0; k < PASS; ++k) s[g][lc][k] = k0 + k < FD ? acc[k0 + k] : 0.0;
        __syncthreads();
        for (int i = tid; i < R * PASS; i += NT20) {
            const int l2 = i / PASS, k = i % PASS;
            if (k0 + k >= FD) continue;
            double v = 0.0;
            for (int gg = 0; gg < G8; ++gg) v += s[gg][l2][k];
            pff[((int64_t)blockIdx.x * R + l2) * FFD + k0 + k] = v;
        }
        __syncthreads();
    }
}

// u = x + x_bc and r of domain dom in mesh numbering, column-major (3N, ncol)
__global__ __launch_bounds__(NT20) void k_fields20(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                                                   const Dom20Dev* __restrict__ d20, const double* __restrict__ x,
                                                   const double* __restrict__ rf, int dom, int ncol,
                                                   double* __restrict__ uo, double* __restrict__ ro) {
    const BrickInfo B = bricks[blockIdx.x];
    if (B.dom != dom) return;
    const DomInfo D = doms[B.dom];
    const Dom20Dev G = d20[B.dom];
    const int64_t nn = G.n_nodes;
    for (int i = threadIdx.x; i < (B.node_end - B.node_begin) * 3 * R; i += NT20) {
        const int o = i / (3 * R), rem = i % (3 * R), d = rem / R, lc = rem % R;
        if (lc >= ncol) continue;
        const int64_t q = B.node_begin + o;
        const int ql = (int)(q - G.node_begin);
        const int64_t out = (3 * (int64_t)ql + d) + 3 * nn * lc;
        const int64_t idx = (q * 3 + d) * R + lc;
        if (uo) uo[out] = G.bcf[ql] ? bc20(D, G, ql, d, lc) : x[idx];
        if (ro) ro[out] = rf[idx];
    }
}

// ---------------------------------------------------------------- element-wise FP32 operator
// y = K x in two passes (FP32, IN_V): k20_elem forms every element's forces
// y_e = E_e Khat x_e for the 24 load cases into a per-element buffer (60 rows x
// 24 load cases), k20_gather sums them per node over the node's sorted
// (element, local node) list.  The products and sums run in the node-gather
// kernel's order (k = 3 b + d per row, elements in list order).
constexpr int E20_EPB = 4;                 // elements per workgroup
constexpr int E20_NT = 240;                // 24 load cases x 10 row groups of 6 rows

__global__ __launch_bounds__(E20_NT) void k20_elem(Dom20Dev G, int dom, const int* __restrict__ done,
                                                  const float* __restrict__ Kg, const float* __restrict__ x,
                                                  float* __restrict__ ye, int64_t ne) {
    __shared__ float sK[3600];
    __shared__ float sx[E20_EPB][60][R];
    if (done && done[dom]) return;
    const int tid = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * E20_EPB;
    for (int i = tid; i < 3600; i += E20_NT) sK[i] = Kg[i];
    // x of the block's elements: 20 nodes x 72 contiguous values each
    for (int i = tid; i < E20_EPB * 20 * 72; i += E20_NT) {
        const int el = i / (20 * 72), rem = i % (20 * 72), a = rem / 72, v = rem % 72;
        const int64_t e = e0 + el;
        float val = 0.f;
        if (e < ne) {
            const int q = G.elems[e * 20 + a];
            if (q >= 0) val = x[(G.node_begin + q) * 72 + v];
        }
        sx[el][3 * a + v / R][v % R] = val;
    }
    __syncthreads();
    const int r = tid % R, g = tid / R;
    float acc[E20_EPB][6];
#pragma unroll
    for (int el = 0; el < E20_EPB; ++el)
#pragma unroll
        for (int j = 0; j < 6; ++j) acc[el][j] = 0.f;
    for (int k = 0; k < 60; ++k) {
        float xk[E20_EPB], kk[6];
#pragma unroll
        for (int el = 0; el < E20_EPB; ++el) xk[el] = sx[el][k][r];
#pragma unroll
        for (int j = 0; j < 6; ++j) kk[j] = sK[(g * 6 + j) * 60 + k];
#pragma unroll
        for (int el = 0; el < E20_EPB; ++el)
#pragma unroll
            for (int j = 0; j < 6; ++j) acc[el][j] += kk[j] * xk[el];
    }
#pragma unroll
    for (int el = 0; el < E20_EPB; ++el) {
        const int64_t e = e0 + el;
        if (e >= ne) break;
        const float E = (float)G.emod[e];
#pragma unroll
        for (int j = 0; j < 6; ++j) ye[(e * 60 + g * 6 + j) * R + r] = E * acc[el][j];
    }
}

// The same element forces on the matrix cores, FP32 (v_mfma_f32_16x16x4_f32) or
// FP64 (v_mfma_f64_16x16x4_f64) products and sums: for a group of EM elements,
// Y (64 x 24 EM) = Khat (60 x 60, rows padded to 64) X (60 x 24 EM), column
// c = 24 e + load case.  Wave w owns rows 16 w .. 16 w + 15 and keeps its Khat
// fragments in registers for the whole block; X (FP32 vectors, widened to T) is
// staged in LDS with a padded row stride.  Lane layout of the 16x16x4 products
// (l = lane, measured on gfx950): A[i][k] in lane i + 16 k, B[k][j] in lane
// j + 16 k; D[i][j] in lane j + 16 (i / 4), component i % 4 (f32), and in lane
// j + 16 (i % 4), component i / 4 (f64).
typedef float em_float4 __attribute__((ext_vector_type(4)));
typedef double em_double4 __attribute__((ext_vector_type(4)));
template <typename T> struct EmCfg;
template <> struct EmCfg<float> {
    static constexpr int EM = 8;                 // elements per workgroup: 192 columns, 12 tiles
    static constexpr int XS = 24 * EM + 16;      // LDS row stride (the 32-lane halves of a read hit distinct banks)
    using V4 = em_float4;
    __device__ static V4 mfma(float a, float b, V4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }
    __device__ static int drow(int k4, int c) { return 4 * k4 + c; }   // row of D component c in lane group k4
};
template <> struct EmCfg<double> {
    static constexpr int EM = 4;                 // 96 columns, 6 tiles
    static constexpr int XS = 24 * EM + 8;
    using V4 = em_double4;
    __device__ static V4 mfma(double a, double b, V4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
    __device__ static int drow(int k4, int c) { return k4 + 4 * c; }
};

template <typename T>
__global__ __launch_bounds__(256) void k20_elem_mf(Dom20Dev G, int dom, const int* __restrict__ done,
                                                   const T* __restrict__ Kg, const float* __restrict__ x,
                                                   T* __restrict__ ye, int64_t ne) {
    constexpr int EM = EmCfg<T>::EM, XS = EmCfg<T>::XS, NCT = 24 * EM / 16;
    using V4 = typename EmCfg<T>::V4;
    __shared__ T sX[60 * XS];
    __shared__ T sE[EM];
    if (done && done[dom]) return;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t e0 = (int64_t)blockIdx.x * EM;
    // X: 20 nodes x 72 contiguous floats per element, read as float4 (18 per node)
    for (int i = tid; i < EM * 20 * 18; i += 256) {
        const int el = i / 360, rem = i % 360, a = rem / 18, v = 4 * (rem % 18), d = v / R, r = v % R;
        const int64_t e = e0 + el;
        float4 val = make_float4(0.f, 0.f, 0.f, 0.f);
        if (e < ne) {
            const int q = G.elems[e * 20 + a];
            if (q >= 0) val = *reinterpret_cast<const float4*>(x + (G.node_begin + q) * 72 + v);
        }
        T* dst = sX + (3 * a + d) * XS + el * R + r;
        dst[0] = (T)val.x;
        dst[1] = (T)val.y;
        dst[2] = (T)val.z;
        dst[3] = (T)val.w;
    }
    if (tid < EM) sE[tid] = e0 + tid < ne ? (T)G.emod[e0 + tid] : T(0);
    // Khat fragments of this wave's row tile: A[i][k] = Khat[16 w + i][4 ks + k]
    const int i16 = lane & 15, k4 = lane >> 4;
    T a[15];
    const int row = 16 * w + i16;
#pragma unroll
    for (int ks = 0; ks < 15; ++ks) a[ks] = row < 60 ? Kg[row * 60 + 4 * ks + k4] : T(0);
    __syncthreads();
    V4 acc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) acc[ct] = V4{T(0), T(0), T(0), T(0)};
#pragma unroll
    for (int ks = 0; ks < 15; ++ks) {
        const T* xb = sX + (4 * ks + k4) * XS + i16;
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) acc[ct] = EmCfg<T>::mfma(a[ks], xb[16 * ct], acc[ct]);
    }
    // D: row 16 w + drow(lane / 16, c), column 16 ct + lane % 16
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) {
        const int col = 16 * ct + i16, el = col / R, r = col % R;
        const int64_t e = e0 + el;
        if (e >= ne) continue;
        const T E = sE[el];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int rr = 16 * w + EmCfg<T>::drow(k4, c);
            if (rr < 60) ye[(e * 60 + rr) * R + r] = E * acc[ct][c];
        }
    }
}

// Operands of the Chebyshev epilogue of k20_gather (OUT 4)
struct Cheb20 {
    const float* b;      // right-hand side
    const float* dinv;   // Jacobi scaling per (node, direction)
    float* d;            // direction, read and written
    float c1, c2;
};

// y on the pseudo-bricks: the sum of the adjacent elements' forces; MASKED zeroes
// Dirichlet rows, MASKED_DOT also writes per-brick sums of x . y; OUT 4 (FP32) applies
// one Chebyshev-Jacobi step to x in place instead of writing y: with t the masked sum,
// d = c1 d + c2 dinv (b - t), x = x + d (yout is x)
// (T: the type of the element forces and of the sums; x is an FP32 vector, y one
// of type TO)
template <typename T, int OUT, typename TO = float>
__global__ __launch_bounds__(NT20) void k20_gather(const BrickInfo* __restrict__ bricks,
                                                  const Dom20Dev* __restrict__ d20,
                                                  const int64_t* __restrict__ ebase, const int* __restrict__ done,
                                                  const T* __restrict__ ye, const float* __restrict__ xin,
                                                  TO* __restrict__ yout, double* __restrict__ partial,
                                                  Cheb20 ch) {
    static_assert(OUT != 4 || (sizeof(T) == 4 && sizeof(TO) == 4), "Chebyshev epilogue: FP32");
    __shared__ double s_red[NT20];
    const BrickInfo B = bricks[blockIdx.x];
    if (done && done[B.dom]) return;
    const Dom20Dev G = d20[B.dom];
    const int64_t eb = ebase[B.dom];
    const int tid = threadIdx.x, r = tid % R, g = tid / R;
    double acc = 0.0;
    for (int q = B.node_begin + g; q < B.node_end; q += NT20 / R) {
        const int ql = (int)(q - G.node_begin);
        // the node's (element, local node) entries in groups of 8, all loads of a
        // group issued before its sums (list order kept; absent entries add 0)
        T y[3] = {T(0), T(0), T(0)};
        const int k1 = G.adj_off[ql + 1];
        for (int kb = G.adj_off[ql]; kb < k1; kb += 8) {
            int ent[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) ent[u] = kb + u < k1 ? G.adj[kb + u] : -1;
            T v[8][3];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (ent[u] >= 0) {
                    const T* src = ye + ((eb + (ent[u] >> 5)) * 60 + 3 * (ent[u] & 31)) * R + r;
#pragma unroll
                    for (int d = 0; d < 3; ++d) v[u][d] = src[d * R];
                } else {
#pragma unroll
                    for (int d = 0; d < 3; ++d) v[u][d] = T(0);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
#pragma unroll
                for (int d = 0; d < 3; ++d) y[d] += v[u][d];
        }
        const bool bc = G.bcf[ql] != 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int64_t o = ((int64_t)q * 3 + d) * R + r;
            const TO v = bc ? TO(0) : (TO)y[d];
            if constexpr (OUT == 4) {
                const float dv = cheb_dir(ch.c1, ch.d[o], ch.c2, ch.dinv[o / R], ch.b[o], (float)v);
                ch.d[o] = dv;
                yout[o] = (TO)(yout[o] + dv);
            } else {
                if (OUT == 0) acc += (double)xin[o] * (double)v;
                yout[o] = v;
            }
        }
    }
    if (OUT == 0) {
        s_red[tid] = acc;
        __syncthreads();
        if (tid < R) {
            double sum = 0.0;
            for (int gg = 0; gg < NT20 / R; ++gg) sum += s_red[gg * R + tid];
            partial[(int64_t)blockIdx.x * R + tid] = sum;
        }
    }
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

template <typename T, int IN, int OUT>
int mv20_launch(SolverBatch& b, const void* xin, void* yout, double* part, const int* done) {
    if (b.n_bricks == 0) return DTC_OK;
    Mv20 a{b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(), b.d20dev.as<Dom20Dev>(), done, part};
    const T* K = sizeof(T) == 8 ? (const T*)b.k20_64.p : (const T*)b.k20_32.p;
    k_matvec20<T, IN, OUT><<<b.n_bricks, NT20, 0, b.stream>>>(a, K, (const T*)xin, (T*)yout);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

}  // namespace

// FP32 element forces of every domain group into ye20 (on the matrix cores;
// DTC_HEX20_MFMA=0: the VALU form)
int elem_pass32(SolverBatch& b, const float* xin, const int* done) {
    const char* me = std::getenv("DTC_HEX20_MFMA");
    const bool use_mf = !(me && me[0] == '0');
    hipStream_t s = b.stream;
    for (int g = 0; g < b.n_dom; ++g) {
        const Dom20Dev& G = b.h_d20dev[g];
        const int64_t ne = b.h_elems[g];   // the count ye20 was sized for (0 for a domain that failed set-up)
        if (ne <= 0 || !G.elems) continue;
        float* ye = b.ye20.as<float>() + b.h_ebase20[g] * 60 * R;
        if (use_mf)
            k20_elem_mf<float><<<nb(ne, EmCfg<float>::EM), 256, 0, s>>>(G, g, done, b.k20_32.as<float>(), xin, ye, ne);
        else
            k20_elem<<<nb(ne, E20_EPB), E20_NT, 0, s>>>(G, g, done, b.k20_32.as<float>(), xin, ye, ne);
    }
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

int mv20_cheb(SolverBatch& b, float* x, const int* done, const float* bv, const float* dinv, float* d, float c1,
              float c2) {
    if (hex20_gather_form()) return set_error(DTC_ERR_UNSUPPORTED, "Chebyshev epilogue of the node-gather form");
    if (b.n_bricks == 0) return DTC_OK;
    DTC_TRY(elem_pass32(b, x, done));
    k20_gather<float, 4><<<b.n_bricks, NT20, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.d20dev.as<Dom20Dev>(),
                                                         b.ebase20.as<int64_t>(), done, b.ye20.as<float>(), nullptr, x,
                                                         nullptr, Cheb20{bv, dinv, d, c1, c2});
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

int mv20(SolverBatch& b, bool fp64, int in, int out, const void* xin, void* yout, double* part, const int* done) {
    if (fp64) {
        if (in == 0 && out == 0) return mv20_launch<double, 0, 0>(b, xin, yout, part, done);
        if (in == 1 && out == 1) return mv20_launch<double, 1, 1>(b, xin, yout, part, done);
        if (in == 1 && out == 2) return mv20_launch<double, 1, 2>(b, xin, yout, part, done);
    } else if (in == 0 && (out == 0 || out == 3)) {
        if (hex20_gather_form()) {   // the node-gather kernel (reference form of the operator)
            if (out == 0) return mv20_launch<float, 0, 0>(b, xin, yout, part, done);
            return mv20_launch<float, 0, 3>(b, xin, yout, part, done);
        }
        if (b.n_bricks == 0) return DTC_OK;
        hipStream_t s = b.stream;
        DTC_TRY(elem_pass32(b, (const float*)xin, done));
        if (out == 0)
            k20_gather<float, 0><<<b.n_bricks, NT20, 0, s>>>(b.brick.as<BrickInfo>(), b.d20dev.as<Dom20Dev>(),
                                                      b.ebase20.as<int64_t>(), done, b.ye20.as<float>(),
                                                      (const float*)xin, (float*)yout, part, Cheb20{});
        else
            k20_gather<float, 3><<<b.n_bricks, NT20, 0, s>>>(b.brick.as<BrickInfo>(), b.d20dev.as<Dom20Dev>(),
                                                      b.ebase20.as<int64_t>(), done, b.ye20.as<float>(),
                                                      (const float*)xin, (float*)yout, part, Cheb20{});
        DTC_HIP(hipGetLastError());
        return DTC_OK;
    }
    return set_error(DTC_ERR_UNSUPPORTED, "HEX20 micro operator: mode (%d, %d)", in, out);
}

int mv20_q64(SolverBatch& b, const float* p, float* q, const int* done, double* q64) {
    if (b.n_bricks == 0) return DTC_OK;
    hipStream_t s = b.stream;
    if (hex20_gather_form()) {   // the node-gather form
        Mv20 a{b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(), b.d20dev.as<Dom20Dev>(), done, nullptr};
        if (q64) return set_error(DTC_ERR_UNSUPPORTED, "HEX20 node-gather operator with an FP64 output");
        k_matvec20<double, 0, 3, float><<<b.n_bricks, NT20, 0, s>>>(a, b.k20_64.as<double>(), p, q);
        DTC_HIP(hipGetLastError());
        return DTC_OK;
    }
    for (int g = 0; g < b.n_dom; ++g) {
        const Dom20Dev& G = b.h_d20dev[g];
        const int64_t ne = b.h_elems[g];
        if (ne <= 0 || !G.elems) continue;
        k20_elem_mf<double><<<nb(ne, EmCfg<double>::EM), 256, 0, s>>>(G, g, done, b.k20_64.as<double>(), p,
                                                                      b.ye20.as<double>() + b.h_ebase20[g] * 60 * R, ne);
    }
    if (q64)
        k20_gather<double, 3, double><<<b.n_bricks, NT20, 0, s>>>(b.brick.as<BrickInfo>(), b.d20dev.as<Dom20Dev>(),
                                                                  b.ebase20.as<int64_t>(), done, b.ye20.as<double>(),
                                                                  p, q64, nullptr, Cheb20{});
    else
        k20_gather<double, 3><<<b.n_bricks, NT20, 0, s>>>(b.brick.as<BrickInfo>(), b.d20dev.as<Dom20Dev>(),
                                                          b.ebase20.as<int64_t>(), done, b.ye20.as<double>(), p, q,
                                                          nullptr, Cheb20{});
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

int batch_add_domain20(SolverBatch& b, const MeshDev& m, const dtc_domain_desc& d, const double* modulus_dev,
                       double escale, int lc_base, hipStream_t s) {
    if (m.nen != 20) return set_error(DTC_ERR_BAD_ARGUMENT, "batch_add_domain20 needs a HEX20 mesh");
    DomInfo D{};
    D.nx = m.nx; D.ny = m.ny; D.nz = m.nz;
    D.macro_order = d.macro_order;
    D.lc_base = lc_base;
    for (int k = 0; k < 3; ++k) {
        D.delta[k] = d.delta[k];
        D.xa1[k] = (double)(d.xa_n[k] - 1);
        D.min_c[k] = (double)(d.xa_n[k] - 1) * d.delta[k];
        D.max_c[k] = (double)(d.xa_n[k] + d.n[k] - 1) * d.delta[k];
    }
    const double dim[3] = {d.n[0] * d.delta[0], d.n[1] * d.delta[1], d.n[2] * d.delta[2]};
    const int64_t nn = m.n_nodes, ne = m.n_elems;
    if (b.n_nodes + nn >= INT32_MAX / 72 || ne >= (1 << 26))
        return set_error(DTC_ERR_BAD_ARGUMENT, "HEX20 batch too large (%lld nodes, %lld elements)", (long long)nn,
                         (long long)ne);
    Dom20 G;
    DevMem e64, dimd, cnt, cursor;
    DTC_TRY(G.coords.alloc((size_t)nn * 3 * 8));
    DTC_TRY(e64.alloc((size_t)ne * 20 * 8));
    DTC_TRY(mesh_write_outputs(m, d.delta, d.xa_n, G.coords.as<double>(), e64.as<int64_t>(), s));
    DTC_TRY(G.elems.alloc((size_t)ne * 20 * 4));
    k20_elems<<<nb(ne * 20, 256), 256, 0, s>>>(e64.as<int64_t>(), ne * 20, G.elems.as<int>());
    DTC_TRY(G.emod.alloc((size_t)ne * 8));
    DTC_TRY(G.evox.alloc((size_t)ne * 4));
    const int64_t nvox = (int64_t)m.nx * m.ny * m.nz;
    G.nvox = nvox;
    k20_emod<<<nb(nvox, 256), 256, 0, s>>>(m.keepE.as<int>(), m.newE.as<int>(), modulus_dev, d.E, escale, nvox,
                                          G.emod.as<double>(), G.evox.as<int>());
    // boundary nodes (fewer than 6 constrained DOFs fails the domain, f-src/mod_struct_preprocess.f90:471-473)
    DTC_TRY(G.bcf.alloc((size_t)nn));
    DTC_TRY(dimd.alloc(3 * 8));
    DTC_TRY(cnt.alloc((size_t)(nn + 1) * 4));
    DTC_TRY(cursor.alloc((size_t)(nn + 1) * 4));
    DTC_HIP(hipMemcpyAsync(dimd.p, dim, 3 * 8, hipMemcpyHostToDevice, s));
    DTC_HIP(hipMemsetAsync(cnt.p, 0, (size_t)(nn + 1) * 4, s));
    DTC_HIP(hipMemsetAsync(cursor.p, 0, (size_t)(nn + 1) * 4, s));
    k20_bc<<<nb(nn, 256), 256, 0, s>>>(G.coords.as<double>(), nn, D, dimd.as<double>(), G.bcf.as<unsigned char>(),
                                       cnt.as<int>() + nn);
    int hb = 0;
    DTC_HIP(hipMemcpyAsync(&hb, cnt.as<int>() + nn, 4, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    if (3 * hb < 6) {
        DTC_TRY(batch_add_empty(b, d, DTC_ERR_FEW_CONSTRAINTS, lc_base));
        b.h_elems.back() = ne;
        b.h_nodes.back() = nn;
        b.h_bnodes.back() = hb;
        return DTC_OK;
    }
    // node -> (element, local node) adjacency, sorted per node
    DTC_HIP(hipMemsetAsync(cnt.p, 0, (size_t)(nn + 1) * 4, s));
    k20_adj_count<<<nb(ne * 20, 256), 256, 0, s>>>(G.elems.as<int>(), ne, cnt.as<int>());
    DTC_TRY(G.adj_off.alloc((size_t)(nn + 1) * 4));
    int64_t total = 0;
    DTC_TRY(scan_exclusive_i32(cnt.as<int>(), G.adj_off.as<int>(), nn + 1, &total, G.scan_tmp, s));
    DTC_TRY(G.adj.alloc((size_t)std::max<int64_t>(total, 1) * 4));
    k20_adj_fill<<<nb(ne * 20, 256), 256, 0, s>>>(G.elems.as<int>(), ne, G.adj_off.as<int>(), cursor.as<int>(),
                                                  G.adj.as<int>());
    k20_adj_sort<<<nb(nn, 256), 256, 0, s>>>(G.adj_off.as<int>(), nn, G.adj.as<int>());
    // lattice map for the p-multigrid transfers
    G.Lx = m.Lx; G.Ly = m.Ly; G.Lz = m.Lz;
    const int64_t nl = (int64_t)m.Lx * m.Ly * m.Lz;
    DTC_TRY(G.lat20.alloc((size_t)nl * 4));
    k20_latmap<<<nb(nl, 256), 256, 0, s>>>(m.numLat.as<int>(), m.keepN.as<int>(), m.cref.as<int>(), nl,
                                           G.lat20.as<int>());
    DTC_HIP(hipGetLastError());
    // pseudo-bricks of PB consecutive nodes
    D.brick_begin = (int)b.h_brick.size();
    D.node_begin = b.n_nodes;
    for (int64_t k = 0; k < nn; k += PB) {
        BrickInfo B{};
        B.dom = b.n_dom;
        B.node_begin = (int)(b.n_nodes + k);
        B.node_end = (int)(b.n_nodes + std::min<int64_t>(nn, k + PB));
        b.h_brick.push_back(B);
    }
    D.brick_end = (int)b.h_brick.size();
    D.node_end = b.n_nodes + nn;
    D.lat_off = b.n_lat;
    D.vox_off = b.n_vox;
    G.node_begin = b.n_nodes;
    G.n_nodes = nn;
    DTC_HIP(hipStreamSynchronize(s));   // temporaries go out of scope
    b.d20.push_back(std::move(G));
    b.h_dom.push_back(D);
    b.n_nodes += nn;
    b.n_lat += (int64_t)(m.nx + 1) * (m.ny + 1) * (m.nz + 1);
    b.n_vox += nvox;
    b.n_dom += 1;
    b.h_elems.push_back(ne);
    b.h_nodes.push_back(nn);
    b.h_bnodes.push_back(hb);
    b.h_status.push_back(DTC_OK);
    return DTC_OK;
}

int finalize20(SolverBatch& b) {
    hipStream_t s = b.stream;
    if (b.khat20.size() != 3600) return set_error(DTC_ERR_BAD_ARGUMENT, "HEX20 element stiffness missing");
    std::vector<Dom20Dev> h(b.n_dom);
    for (int i = 0; i < b.n_dom; ++i) {
        const Dom20& G = b.d20[i];
        h[i] = Dom20Dev{G.coords.as<double>(), G.elems.as<int>(), G.emod.as<double>(), G.adj_off.as<int>(),
                        G.adj.as<int>(), G.bcf.as<unsigned char>(), G.lat20.as<int>(), G.node_begin, (int)G.n_nodes,
                        G.Lx, G.Ly, G.Lz};
    }
    b.h_d20dev = h;
    // per-element force buffer of the FP32 operator: element offsets per domain group
    b.h_ebase20.assign(b.n_dom, 0);
    int64_t ne_all = 0;
    for (int i = 0; i < b.n_dom; ++i) {
        b.h_ebase20[i] = ne_all;
        ne_all += b.h_elems[i];
    }
    DTC_TRY(b.ebase20.ensure(std::max<size_t>(b.h_ebase20.size(), 1) * 8));
    if (!b.h_ebase20.empty())
        DTC_HIP(hipMemcpyAsync(b.ebase20.p, b.h_ebase20.data(), b.h_ebase20.size() * 8, hipMemcpyHostToDevice, s));
    // FP64 element forces of the block's Q = K P, FP32 ones of the smoothing steps (one buffer)
    DTC_TRY(b.ye20.ensure((size_t)std::max<int64_t>(ne_all, 1) * 60 * R * 8));
    DTC_TRY(b.d20dev.ensure(std::max<size_t>(h.size(), 1) * sizeof(Dom20Dev)));
    if (!h.empty()) DTC_HIP(hipMemcpyAsync(b.d20dev.p, h.data(), h.size() * sizeof(Dom20Dev), hipMemcpyHostToDevice, s));
    std::vector<float> k32(3600);
    for (int i = 0; i < 3600; ++i) k32[i] = (float)b.khat20[i];
    DTC_TRY(b.k20_64.ensure(3600 * 8));
    DTC_TRY(b.k20_32.ensure(3600 * 4));
    DTC_HIP(hipMemcpyAsync(b.k20_64.p, b.khat20.data(), 3600 * 8, hipMemcpyHostToDevice, s));
    DTC_HIP(hipMemcpyAsync(b.k20_32.p, k32.data(), 3600 * 4, hipMemcpyHostToDevice, s));
    for (int i = 0; i < b.n_dom; ++i)
        if (h[i].n_nodes > 0)
            k20_dinv<<<nb(h[i].n_nodes, 256), 256, 0, s>>>(h[i], b.k20_64.as<double>(), h[i].node_begin,
                                                           b.dinv64.as<double>(), b.dinv32.as<float>());
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipStreamSynchronize(s));
    return DTC_OK;
}

int project20(SolverBatch& b) {
    if (b.n_bricks == 0) return DTC_OK;
    k_project20<8><<<b.n_bricks, NT20, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                      b.d20dev.as<Dom20Dev>(), b.r64.as<double>(), b.part_ff.as<double>());
    k_project20<20><<<b.n_bricks, NT20, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                       b.d20dev.as<Dom20Dev>(), b.r64.as<double>(),
                                                       b.part_ff.as<double>());
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

// ---------------------------------------------------------------- sensitivity (HEX20 micro, HEX08 macro)
// h[q][j][d] (+)= sum_lc M[lc][j] u_lc(q, d), u = x + x_bc, over the 24 load cases of
// one domain group (the HEX08 form in k_grad_h); q is the node's local number, the
// same in every group of a domain; add = 0 for the first group
__global__ __launch_bounds__(256) void k20_grad_h(Dom20Dev G, DomInfo D, const double* __restrict__ x,
                                                  const double* __restrict__ Mj, int add, double* __restrict__ h) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)G.n_nodes * 18) return;
    const int q = (int)(i / 18), j = (int)(i % 18) / 3, d = (int)(i % 3);
    const bool bc = G.bcf[q] != 0;
    double s = 0.0;
    for (int lc = 0; lc < R; ++lc) {
        const double u = bc ? bc20(D, G, q, d, lc) : x[((G.node_begin + q) * 3 + d) * R + lc];
        s += Mj[lc * 6 + j] * u;
    }
    h[i] = add ? h[i] + s : s;
}

// dL/dE_v = sum_j h_j,e^T Khat u_j,e per element (one wave per element: lane r < 60
// forms row r of Khat u_j, the wave sums h_j . (Khat u_j) in a fixed order), written
// at the element's voxel
constexpr int G20_EPB = 4;   // elements (waves) per workgroup
__global__ __launch_bounds__(64 * G20_EPB) void k20_grad_elem(Dom20Dev G, DomInfo D, const double* __restrict__ Kg,
                                                              const double* __restrict__ x,
                                                              const double* __restrict__ h,
                                                              const int* __restrict__ evox, int64_t ne,
                                                              double* __restrict__ g) {
    __shared__ double sK[3600];
    __shared__ double su[G20_EPB][6][60];
    __shared__ double sh[G20_EPB][6][60];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 3600; i += 64 * G20_EPB) sK[i] = Kg[i];
    const int64_t e = (int64_t)blockIdx.x * G20_EPB + w;
    if (e < ne) {
        for (int i = lane; i < 6 * 60; i += 64) {
            const int j = i / 60, k = i % 60, a = k / 3, d = k % 3;
            const int q = G.elems[e * 20 + a];
            double u = 0.0, hv = 0.0;
            if (q >= 0) {
                u = G.bcf[q] ? bc20(D, G, q, d, j) : x[((G.node_begin + q) * 3 + d) * R + j];
                hv = h[((int64_t)q * 6 + j) * 3 + d];
            }
            su[w][j][k] = u;
            sh[w][j][k] = hv;
        }
    }
    __syncthreads();
    if (e >= ne) return;
    double part = 0.0;
    if (lane < 60) {
        for (int j = 0; j < 6; ++j) {
            double ku = 0.0;
            for (int c = 0; c < 60; ++c) ku += sK[lane * 60 + c] * su[w][j][c];
            part += sh[w][j][lane] * ku;
        }
    }
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (lane == 0) g[evox[e]] = part;
}

int grad20(SolverBatch& b, int i, int ng, const double* M, double* dL_dE_host) {
    if (i < 0 || ng < 1 || i + ng > b.n_dom) return set_error(DTC_ERR_BAD_ARGUMENT, "domain groups %d..%d", i, i + ng - 1);
    const Dom20& G = b.d20[i];
    const Dom20Dev& Gd = b.h_d20dev[i];
    const DomInfo& D = b.h_dom[i];
    hipStream_t s = b.stream;
    const int64_t nvox = G.nvox, ne = b.h_elems[i];
    DevMem h, Md, g;
    DTC_TRY(Md.alloc((size_t)ng * R * 6 * 8));
    DTC_TRY(g.alloc((size_t)std::max<int64_t>(nvox, 1) * 8));
    DTC_HIP(hipMemcpyAsync(Md.p, M, (size_t)ng * R * 6 * 8, hipMemcpyHostToDevice, s));
    DTC_HIP(hipMemsetAsync(g.p, 0, (size_t)std::max<int64_t>(nvox, 1) * 8, s));   // voxels without an element: 0
    if (ne > 0 && Gd.n_nodes > 0) {
        DTC_TRY(h.alloc((size_t)Gd.n_nodes * 18 * 8));
        for (int k = 0; k < ng; ++k) {
            const Dom20Dev& Gk = b.h_d20dev[i + k];
            if (Gk.n_nodes != Gd.n_nodes) return set_error(DTC_ERR_DEVICE, "load-case groups of one domain differ");
            k20_grad_h<<<nb((int64_t)Gd.n_nodes * 18, 256), 256, 0, s>>>(Gk, b.h_dom[i + k], b.x64.as<double>(),
                                                                         Md.as<double>() + (size_t)k * R * 6, k > 0,
                                                                         h.as<double>());
        }
        k20_grad_elem<<<nb(ne, G20_EPB), 64 * G20_EPB, 0, s>>>(Gd, D, b.k20_64.as<double>(), b.x64.as<double>(),
                                                               h.as<double>(), G.evox.as<int>(), ne, g.as<double>());
    }
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipMemcpyAsync(dL_dE_host, g.p, (size_t)nvox * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    return DTC_OK;
}

int fields20(SolverBatch& b, int i, int ncol, double* u_dev, double* r_dev) {
    k_fields20<<<b.n_bricks, NT20, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(), b.d20dev.as<Dom20Dev>(),
                                                  b.x64.as<double>(), b.r64.as<double>(), i, ncol, u_dev, r_dev);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

}  // namespace dtc
