// solver.hip -- batched matrix-free Jacobi-PCG for the 24 load cases of HEX08
// voxel domains on gfx950.
//
// Replaces the PETSc path of f-src/mod_dtc_main_subroutines.f90:426-1238:
// MatSetValues assembly of K_e = Hexe08(bone) per element, Dirichlet lifting
// f = -A x_bc with f[bc] = 0, MatZeroRowsColumns(diag 0.0), KSPSolve with
// CG + Jacobi and x0 = 0 for each load case, and the reactions r = A_org u.
//
// Design
//  * Operator: never assembled.  Every element is a voxel with modulus E_v,
//    so K_e = E_v * Khat.  In the signed Walsh basis of the 8 element nodes
//    Khat is block diagonal (3x3, 2x2, 1x1 blocks, 45 coefficients), so one
//    element product costs two 8-point Hadamard butterflies per direction
//    plus 45 multiplies instead of a dense 24x24 product.
//  * Tiling: the lattice of kept nodes is cut into bricks of 8x4x4 owned
//    nodes.  A workgroup stages the halo (10x6x6 nodes) of its brick in LDS,
//    computes every element touching an owned node (9x5x5 voxels) and
//    accumulates into LDS in 8 parity classes (elements of one class share
//    no node), then writes its owned nodes once.  No atomics, deterministic.
//  * Vectors: node-major [node][3][24], nodes numbered brick by brick so a
//    brick's owned nodes are contiguous.  All 24 load cases advance in
//    lockstep with their own alpha/beta; a converged load case freezes.
//  * Reductions: per-brick partial sums in FP64, then one workgroup per
//    domain sums them in brick order (deterministic).
//  * Precision: DTC_FP64 runs PCG in FP64; DTC_FP32_REFINE runs the inner
//    PCG in FP32 on the residual equation and keeps x, the true residual
//    r = -(K u)_free and the reactions in FP64.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include "bc.hpp"
#include "solver.hpp"

namespace dtc {

namespace {

constexpr int R = R24;
constexpr int NT_UPD = 192;    // multiple of 24: a thread keeps one load case

// IN_JS (FP32 Chebyshev step only): x = cin D^-1 v, the first smoothing step from
// x = 0 formed while staging v, so that it is never stored as a vector
enum InMode { IN_V = 0, IN_U = 1, IN_JS = 2 };
enum OutMode { OUT_MASKED_DOT = 0, OUT_RESID = 1, OUT_PLAIN = 2, OUT_MASKED = 3, OUT_CHEB = 4, OUT_BRES = 5 };

template <typename T> struct WK { T g[45]; };

template <typename T> WK<T> cast_wk(const WalshK& w) {
    WK<T> o;
    for (int i = 0; i < 45; ++i) o.g[i] = (T)w.g[i];
    return o;
}

__device__ inline bool is_bc(const DomInfo& D, int gx, int gy, int gz) {
    return gx == 0 || gy == 0 || gz == 0 || gx == D.nx || gy == D.ny || gz == D.nz;
}

__device__ inline void node_coord(const DomInfo& D, int gx, int gy, int gz, double c[3]) {
    c[0] = ((double)gx + D.xa1[0]) * D.delta[0];
    c[1] = ((double)gy + D.xa1[1]) * D.delta[1];
    c[2] = ((double)gz + D.xa1[2]) * D.delta[2];
}

// ---------------------------------------------------------------- element operator
// modes after the forward butterflies: index bit0 <-> x, bit1 <-> y, bit2 <-> z
template <typename T>
__device__ inline void fht_fwd(T v[8]) {
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!(b & s)) {
                T a = v[b], c = v[b + s];
                v[b] = a + c;
                v[b + s] = c - a;
            }
}
template <typename T>
__device__ inline void fht_inv(T v[8]) {
#pragma unroll
    for (int s = 1; s < 8; s <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!(b & s)) {
                T a = v[b], c = v[b + s];
                v[b] = a - c;
                v[b + s] = a + c;
            }
}
// fht_inv after its first butterfly stage (s = 1)
template <typename T>
__device__ inline void fht_inv_tail(T v[8]) {
#pragma unroll
    for (int s = 2; s < 8; s <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
            if (!(b & s)) {
                T a = v[b], c = v[b + s];
                v[b] = a - c;
                v[b + s] = a + c;
            }
}

// y = Kw w on the (mode, direction) groups; w, y indexed [mode][dir]
template <typename T>
__device__ inline void walsh_apply(const WK<T>& k, const T w[8][3], T y[8][3]) {
#define M3(o, i0, d0, i1, d1, i2, d2)                                                         \
    {                                                                                         \
        T a0 = w[i0][d0], a1 = w[i1][d1], a2 = w[i2][d2];                                     \
        y[i0][d0] = k.g[o + 0] * a0 + k.g[o + 1] * a1 + k.g[o + 2] * a2;                      \
        y[i1][d1] = k.g[o + 3] * a0 + k.g[o + 4] * a1 + k.g[o + 5] * a2;                      \
        y[i2][d2] = k.g[o + 6] * a0 + k.g[o + 7] * a1 + k.g[o + 8] * a2;                      \
    }
#define M2(o, i0, d0, i1, d1)                                                                 \
    {                                                                                         \
        T a0 = w[i0][d0], a1 = w[i1][d1];                                                     \
        y[i0][d0] = k.g[o + 0] * a0 + k.g[o + 1] * a1;                                        \
        y[i1][d1] = k.g[o + 2] * a0 + k.g[o + 3] * a1;                                        \
    }
    y[0][0] = y[0][1] = y[0][2] = T(0);
    M3(0, 1, 0, 2, 1, 4, 2)          // x.ux, y.uy, z.uz
    M2(9, 1, 1, 2, 0)                // x.uy, y.ux
    M2(13, 1, 2, 4, 0)               // x.uz, z.ux
    M2(17, 2, 2, 4, 1)               // y.uz, z.uy
    M2(21, 3, 0, 6, 2)               // xy.ux, yz.uz
    M2(25, 3, 1, 5, 2)               // xy.uy, xz.uz
    M3(29, 3, 2, 5, 1, 6, 0)         // xy.uz, xz.uy, yz.ux
    M2(38, 5, 0, 6, 1)               // xz.ux, yz.uy
    y[7][0] = k.g[42] * w[7][0];     // xyz
    y[7][1] = k.g[43] * w[7][1];
    y[7][2] = k.g[44] * w[7][2];
#undef M3
#undef M2
}

// ---------------------------------------------------------------- wave sums
// DPP lane move of an FP64 value (two 32-bit moves); lanes whose source lies
// outside their 16-lane row receive 0.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
// RC partial sums per lane (RC = 4 or 2) -> per 16-lane row, the row sum of each
// of the RC values, in fixed order: log2(RC) halving exchanges between lane pairs
// (each lane keeps half of the values and adds its partner's copies), then
// row_shr by RC, 2 RC, .. 8 accumulates each residue class of the row into its
// last lanes.  Returns true on the RC lanes 16-RC..15 of each row, which then
// hold the sum of value `*lc` in `*v`.
template <int RC>
__device__ inline bool row_sums(const double (&a)[RC], int lane, double* v, int* lc) {
    const bool b0 = lane & 1, b1 = lane & 2;
    double k;
    int l;
    if constexpr (RC == 4) {
        double k0 = b0 ? a[2] : a[0], k1 = b0 ? a[3] : a[1];
        const double s0 = b0 ? a[0] : a[2], s1 = b0 ? a[1] : a[3];
        k0 += dpp_f64<0xB1>(s0);   // quad_perm [1,0,3,2]
        k1 += dpp_f64<0xB1>(s1);
        k = b1 ? k1 : k0;
        const double s = b1 ? k0 : k1;
        k += dpp_f64<0x4E>(s);     // quad_perm [2,3,0,1]
        l = 2 * (int)b0 + (int)b1;
        k += dpp_f64<0x114>(k);    // row_shr:4
        k += dpp_f64<0x118>(k);    // row_shr:8
    } else {
        static_assert(RC == 2, "two or four load cases per chunk");
        k = b0 ? a[1] : a[0];
        k += dpp_f64<0xB1>(b0 ? a[0] : a[1]);
        l = (int)b0;
        k += dpp_f64<0x112>(k);    // row_shr:2
        k += dpp_f64<0x114>(k);
        k += dpp_f64<0x118>(k);
    }
    *v = k;
    *lc = l;
    return (lane & 15) >= 16 - RC;
}

// ---------------------------------------------------------------- matvec
// Load cases are processed in chunks of RC: one (node, direction) segment of a
// chunk is one 16-byte vector V (float4 / double2).
template <typename T> struct MvCfg;
template <> struct MvCfg<float> { static constexpr int RC = 4; using V = float4; };
template <> struct MvCfg<double> { static constexpr int RC = 2; using V = double2; };
constexpr int MV_NT = 512;
constexpr int MV_NT_SP = 256;     // sparse-brick variant
constexpr int MV_SLOTS = 384;     // (owned node, element corner) pairs of a sparse brick, at most
constexpr int NH = HX * HY * HZ, NET = EX * EY * EZ, NO = BX * BY * BZ;
constexpr int OB_BIAS = 1 + BX + BX * BY;   // owned-layout base of tile element (0,0,0) is -OB_BIAS

// offsets of element corner b (bit 0 x, bit 1 y, bit 2 z) in the halo and owned node layouts
__host__ __device__ constexpr int hcorner(int b) { return (b & 1) + HX * ((b >> 1) & 1) + HX * HY * (b >> 2); }
__host__ __device__ constexpr int ocorner(int b) { return (b & 1) + BX * ((b >> 1) & 1) + BX * BY * (b >> 2); }
// halo index of owned node o
__host__ __device__ constexpr int ohalo(int o) {
    return (o % BX + 1) + HX * ((o / BX) % BY + 1) + HX * HY * (o / (BX * BY) + 1);
}

__device__ inline float vcomp(const float4& v, int w) { return w == 0 ? v.x : w == 1 ? v.y : w == 2 ? v.z : v.w; }
__device__ inline double vcomp(const double2& v, int w) { return w == 0 ? v.x : v.y; }
__device__ inline void vzero(float4& v) { v = make_float4(0.f, 0.f, 0.f, 0.f); }
__device__ inline void vzero(double2& v) { v = make_double2(0.0, 0.0); }
__device__ inline float4 vneg(const float4& v) { return make_float4(-v.x, -v.y, -v.z, -v.w); }
__device__ inline double2 vneg(const double2& v) { return make_double2(-v.x, -v.y); }
__device__ inline void vadd(float4& a, const float4& b) { a.x += b.x; a.y += b.y; a.z += b.z; a.w += b.w; }
__device__ inline void vadd(double2& a, const double2& b) { a.x += b.x; a.y += b.y; }

struct MvArgs {
    const BrickInfo* bricks;
    const int* blist;           // bricks of this launch (blist[boff + blockIdx.x]), or null: brick = blockIdx.x
    int boff;
    const DomInfo* doms;
    const int* halo;            // [brick][NH]
    const unsigned char* nmask; // [brick][NO]: solid tile elements around each owned node
    const int* einfo;           // tile elements, packed
    const void* eE;             // tile element moduli (T)
    const int* done;            // may be null
    double* partial;            // [brick][24]
    // OUT_CHEB / OUT_BRES only
    const void* bvec;           // right-hand side b
    const void* dinv;           // Jacobi scaling
    void* dvec;                 // Chebyshev direction
    float c1, c2;               // Chebyshev coefficients of this step
    int cflags;                 // CH_* : where the previous direction comes from, whether d is stored
    float cin = 0.f;            // IN_JS: the zero-start step's coefficient
};
// OUT_CHEB direction handling: the previous direction d is read from dvec
// (CH_READ_D), or equals the current iterate (CH_D_FROM_X: the step after a
// zero-start step, where x_1 = d_0), or is absent (c1 = 0); the new direction
// is stored only when a later step reads it (CH_WRITE_D).
constexpr int CH_READ_D = 1, CH_D_FROM_X = 2, CH_WRITE_D = 4;

// y = K x for every brick of the batch.  One workgroup per (brick, chunk
// range): blockIdx.y splits the NCH chunks into gridDim.y equal ranges, so that
// small grids (coarse multigrid levels, single domains) still fill the chip.
// Chunk loop over the load cases:
//   stage   the halo of x (10x6x6 nodes) goes to LDS (16-byte loads, issued
//           one chunk ahead into registers);
//   product every solid element of the 9x5x5 tile is one task per load case
//           (Walsh-basis product); the output of corner b lands in the corner
//           buffer b at the owned node it belongs to.  A (corner, node) slot
//           has exactly one element, so plain stores suffice: no atomics, no
//           colouring barriers;
//   gather  each owned (node, direction) sums its corner buffers in corner
//           order (deterministic) and is written once.
// OUT_MASKED_DOT zeroes Dirichlet rows and writes per-brick sums of x.y;
// the zero-start smoothing step fused into the next step (DTC_CHEB_FUSE=0: stored by k_cheb)
inline bool cheb_fuse() {
    const char* e = std::getenv("DTC_CHEB_FUSE");
    return !(e && e[0] == '0');
}

// whole-segment output stores of the FP32 matvec (DTC_MV_DEFER=0: per-chunk stores)
inline bool mv_defer() {
    const char* e = std::getenv("DTC_MV_DEFER");
    return !(e && e[0] == '0');
}

// OUT_RESID writes -y on free rows and sums y.y; OUT_PLAIN writes y;
// OUT_MASKED zeroes Dirichlet rows; OUT_BRES writes b - y on free rows and 0 on
// Dirichlet rows; OUT_CHEB applies one Chebyshev-Jacobi step to the owned nodes,
//   d = c1 d + c2 D^-1 (b - y),  yout = x + d   (yout must not alias x),
// and, when partial is set, writes per-brick sums of b . yout.
// launch bounds: at least 4 waves per SIMD (2 workgroups of 8 waves per CU), which
// caps the kernel at 128 VGPRs
// IO32 (T = double, IN_V only): 1 = x and y are FP32 vectors, 2 = x is an FP32
// vector and y an FP64 one; the arithmetic is FP64
// SPARSE: the variant for bricks with at most MV_SLOTS (owned node, element corner)
// pairs: 256 threads, and the corner outputs go to compact slots (the pairs of
// each owned node, in corner order, after those of the nodes before it) instead
// of 8 full corner buffers, so that four workgroups fit one CU.  The sums are
// those of the dense form without its zero terms: the results are identical.
// PART (OUT_CHEB): per-brick sums of b . x are written (A.partial set).
template <typename T, int IN, int OUT, int IO32 = 0, bool SPARSE = false, bool PART = false, bool DEFER_Y = true>
__global__ __launch_bounds__(SPARSE ? MV_NT_SP : MV_NT, 4) void k_matvec(MvArgs A, WK<T> wk, const T* __restrict__ xin,
                                                                      T* __restrict__ yout) {
    constexpr int RC = MvCfg<T>::RC, SX = 3 * RC, NCH = R / RC;
    using V = typename MvCfg<T>::V;
    constexpr int NT = SPARSE ? MV_NT_SP : MV_NT;
    constexpr int NLD = (NH * 3 + NT - 1) / NT;
    constexpr int NW = NT / 64;
    constexpr bool SUMS = OUT == OUT_MASKED_DOT || OUT == OUT_RESID || (OUT == OUT_CHEB && PART);
    constexpr int NSLOT = SPARSE ? MV_SLOTS * 3 : (OB_BIAS + 8 * NO) * 3 + 64 + 2;
    __shared__ int s_node[NH];
    __shared__ int s_info[NET];
    __shared__ T s_E[NET];
    __shared__ unsigned char s_nm[NO];
    __shared__ double s_wred[SUMS ? NCH : 1][NW * 4][RC];   // per chunk, wave row and load case
    __shared__ V s_x[NH * 3];
    // corner buffers (the bias keeps element bases non-negative), then a trash
    // area for the outputs of corners outside the owned nodes (lane l writes
    // vectors l .. l + 2; the contents are never read)
    __shared__ V s_c[NSLOT];
    __shared__ short s_hn[NH];                   // halo positions that hold a node, compact
    __shared__ unsigned char s_ol[NO];           // owned position of the brick's k-th solver node
    __shared__ int s_wcnt[((NH + NT - 1) / NT) * NW];
    __shared__ short s_sb[SPARSE ? NO : 1];      // SPARSE: first slot of each owned position
    static_assert(!IO32 || (sizeof(T) == 8 && IN == IN_V), "FP32 vectors with FP64 arithmetic");
    static_assert(IN != IN_JS || (sizeof(T) == 4 && OUT == OUT_CHEB && !SPARSE), "zero-start input: FP32 Chebyshev");

    const int bid = A.blist ? A.blist[A.boff + blockIdx.x] : (int)blockIdx.x;
    const int tid = threadIdx.x;
    const int* __restrict__ halo = A.halo + (int64_t)bid * NH;
    // Set-up loads in as few dependent steps as possible: the halo row and the node
    // masks need only the brick index and are issued before the brick record
    // arrives; the element list needs the record and is issued before the done
    // flag is tested (a finished domain's brick wastes these loads)
    constexpr int NHP = (NH + NT - 1) / NT;
    int hq[NHP];
#pragma unroll
    for (int kk = 0; kk < NHP; ++kk) {
        const int h = tid + kk * NT;
        hq[kk] = h < NH ? halo[h] : -1;
    }
    const unsigned char nmv = tid < NO ? A.nmask[(int64_t)bid * NO + tid] : (unsigned char)0;
    // the brick record's fields are all loaded here, in one round trip (the compiler
    // otherwise loads each field where it is first used, one round trip each)
    BrickInfo B = A.bricks[bid];
    asm volatile("" : "+s"(B.dom), "+s"(B.x0), "+s"(B.y0), "+s"(B.z0), "+s"(B.node_begin), "+s"(B.node_end),
                 "+s"(B.e_off), "+s"(B.e_cnt));
    static_assert(NET <= (SPARSE ? MV_NT_SP : MV_NT), "one tile element per thread");
    int eiv = 0;
    T eEv = T(0);
    if (tid < B.e_cnt) {
        eiv = A.einfo[B.e_off + tid];
        eEv = static_cast<const T*>(A.eE)[B.e_off + tid];
    }
    if (A.done && A.done[B.dom]) return;
    const V* xv = reinterpret_cast<const V*>(xin);
    V* yv = reinterpret_cast<V*>(yout);
    const int c0 = (int)blockIdx.y * NCH / (int)gridDim.y, c1 = ((int)blockIdx.y + 1) * NCH / (int)gridDim.y;

    // Only halo positions and owned positions that hold a node carry work: on
    // sparse (trabecular) domains a brick holds about a third of its 128
    // positions.  s_ol = owned position of solver node node_begin + k (owned
    // nodes are numbered in position order).  Staging, measured per precision:
    //  * FP32 (EARLY_X false): compact list s_hn of halo positions with a node
    //    (wave ballots, prefix over the waves); x loads issued after the barrier
    //    for the existing (node, direction) items only;
    //  * FP64 (EARLY_X true): x loads for every halo slot issued first, straight
    //    from the brick table, before any barrier.
    constexpr bool EARLY_X = sizeof(T) == 8;
    const int n_own = B.node_end - B.node_begin;
    int nh = NH;
    int pf_off[NLD], pf_xi[NLD];   // 16-byte offset in x and destination in s_x of this thread's items
    V reg[NLD];                    // x segments of the next chunk to stage, loaded one chunk ahead
    T pf_sc[IN == IN_JS ? NLD : 1]; // IN_JS: cin times the item's Jacobi scaling
    auto prefetch = [&](int chunk) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            V v;
            vzero(v);
            if (pf_off[k] >= 0) {
                if constexpr (IO32) {
                    const float2 f = reinterpret_cast<const float2*>(xin)[pf_off[k] + chunk];
                    v = make_double2((double)f.x, (double)f.y);
                } else {
                    v = xv[pf_off[k] + chunk];
                }
            }
            reg[k] = v;
        }
    };
    if constexpr (EARLY_X) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            const int i = tid + k * NT;
            int o = -1;
            if (IN == IN_V && i < NH * 3) {
                const int q = halo[i / 3];
                if (q >= 0) o = ((q & NODE_MASK) * 3 + i % 3) * NCH;
            }
            pf_off[k] = o;
            pf_xi[k] = i;
        }
        if (IN != IN_U) prefetch(c0);
    }
    {
        // halo positions h = tid + k NT (one pass with 512 threads, two with 256)
        bool exk[NHP];
#pragma unroll
        for (int kk = 0; kk < NHP; ++kk) {
            const int h = tid + kk * NT;
            const int q = hq[kk];
            if (h < NH) s_node[h] = q;
            const bool ex = q >= 0;
            exk[kk] = ex;
            const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
            if (ex && hx >= 1 && hx <= BX && hy >= 1 && hy <= BY && hz >= 1 && hz <= BZ)
                s_ol[(q & NODE_MASK) - B.node_begin] = (unsigned char)((hx - 1) + BX * ((hy - 1) + BY * (hz - 1)));
        }
        if (tid < NO) s_nm[tid] = nmv;
        if constexpr (!EARLY_X) {
            // compact list of the halo positions that hold a node, in position order
            const int lane = tid & 63, wv = tid >> 6;
            unsigned long long bal[NHP];
#pragma unroll
            for (int kk = 0; kk < NHP; ++kk) {
                bal[kk] = __ballot(exk[kk]);
                if (lane == 0) s_wcnt[kk * NW + wv] = __popcll(bal[kk]);
            }
            __syncthreads();
            nh = 0;
#pragma unroll
            for (int kk = 0; kk < NHP; ++kk) {
                int base = nh;
#pragma unroll
                for (int w = 0; w < NW; ++w) {
                    const int c = s_wcnt[kk * NW + w];
                    if (w < wv) base += c;
                    nh += c;
                }
                if (exk[kk]) s_hn[base + __popcll(bal[kk] & ((1ull << lane) - 1))] = (short)(tid + kk * NT);
            }
        }
    }
    const int ne = B.e_cnt;
    if (tid < ne) {
        s_info[tid] = eiv;
        s_E[tid] = eEv;
    }
    __syncthreads();
    if constexpr (SPARSE) {
        // slot bases: exclusive prefix of the element counts of the owned positions
        // (wave 0, two positions per lane); visible after the barrier below
        if (tid < 64) {
            const int c0 = __popc((unsigned)s_nm[2 * tid]), c1 = __popc((unsigned)s_nm[2 * tid + 1]);
            int v = c0 + c1;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int u = __shfl_up(v, o);
                if (tid >= o) v += u;
            }
            const int ex = v - c0 - c1;
            s_sb[2 * tid] = (short)ex;
            s_sb[2 * tid + 1] = (short)(ex + c0);
        }
    }

    if constexpr (!EARLY_X) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            const int i = tid + k * NT;
            int o = -1, xi = 0;
            if (i < 3 * nh) {
                const int h = s_hn[i / 3], d = i % 3;
                o = ((s_node[h] & NODE_MASK) * 3 + d) * NCH;
                xi = h * 3 + d;
            }
            pf_off[k] = o;
            pf_xi[k] = xi;
            if constexpr (IN == IN_JS)   // the product order of k_cheb's zero start: (c2 dinv) v
                pf_sc[k] = o >= 0 ? (T)A.cin * reinterpret_cast<const T*>(A.dinv)[o / NCH] : T(0);
        }
        if (IN != IN_U) prefetch(c0);
    }

    // s_x <- x of `chunk` (IN_V: from the prefetch registers; IN_U: x + BC values)
    auto stage = [&](int chunk) {
        if constexpr (IN == IN_JS) {
#pragma unroll
            for (int k = 0; k < NLD; ++k)
                if (pf_off[k] >= 0) {
                    V v = reg[k];
                    T* vp = reinterpret_cast<T*>(&v);
#pragma unroll
                    for (int w = 0; w < RC; ++w) vp[w] = pf_sc[k] * vp[w];
                    s_x[pf_xi[k]] = v;
                }
        } else if (IN == IN_V) {
#pragma unroll
            for (int k = 0; k < NLD; ++k)
                if (pf_off[k] >= 0) s_x[pf_xi[k]] = reg[k];
        } else {
            const DomInfo D = A.doms[B.dom];
            T* sx = reinterpret_cast<T*>(s_x);
            for (int i = tid; i < nh * SX; i += NT) {
                const int h = EARLY_X ? i / SX : (int)s_hn[i / SX], rem = i % SX, d = rem / RC, r = rem % RC;
                const int lc = chunk * RC + r;
                const int q = s_node[h];
                T v = T(0);
                if (q >= 0) {
                    if (q & BC_BIT) {
                        const int hx = h % HX, hy = (h / HX) % HY, hz = h / (HX * HY);
                        double c[3], N[20];
                        node_coord(D, B.x0 - 1 + hx, B.y0 - 1 + hy, B.z0 - 1 + hz, c);
                        macro_shape(D.macro_order, c, D.min_c, D.max_c, N);
                        v = (T)bc_value_m(D.macro_order, N, d, D.lc_base + lc);
                    } else if (xin) {
                        v = xin[((int64_t)(q & NODE_MASK) * 3 + d) * R + lc];
                    }
                }
                sx[h * SX + rem] = v;
            }
        }
    };

    // chunk-invariant gather items of this thread: (owned solver node, direction)
    // i = tid + j * NT < 3 n_own
    constexpr int NG = (NO * 3 + NT - 1) / NT;
    int g_off[NG], g_xi[NG];
    unsigned g_m[NG];
    T e_di[NG];   // OUT_CHEB: Jacobi scaling of the item
#pragma unroll
    for (int j = 0; j < NG; ++j) {
        const int i = tid + j * NT;
        int off = -1, xi = 0;
        unsigned m = 0;
        if (i < 3 * n_own) {
            const int o = s_ol[i / 3], d = i % 3, h = ohalo(o);
            const int q = s_node[h];
            off = ((q & NODE_MASK) * 3 + d) * NCH;
            xi = h * 3 + d;
            m = (unsigned)s_nm[o] | ((q & BC_BIT) ? 256u : 0u) | ((unsigned)o << 9) | ((unsigned)d << 16);
        }
        g_off[j] = off;
        g_xi[j] = xi;
        g_m[j] = m;
        e_di[j] = T(0);
        if (OUT == OUT_CHEB && off >= 0 && !(m & 256u)) e_di[j] = reinterpret_cast<const T*>(A.dinv)[off / NCH];
    }

    // corner slots that no element writes (element absent) are zero for the whole
    // launch, so the gather adds all 8 corner buffers without tests
    if constexpr (!SPARSE)
        for (int i = tid; i < n_own * 24; i += NT) {
            const int o = s_ol[i / 24], b = (i / 3) % 8, d = i % 3;
            if (!((s_nm[o] >> b) & 1)) vzero(s_c[(OB_BIAS + b * NO + o) * 3 + d]);
        }
    stage(c0);
    if (IN != IN_U && c0 + 1 < c1) prefetch(c0 + 1);
    __syncthreads();

    // FP32 dense Chebyshev step (one gather item per thread): the outputs of all
    // chunks stay in registers and each item's 24 load cases are stored together
    // after the last chunk, so every 96-byte node segment is written whole
    // (per-chunk 16-byte stores left partly written lines to be written back more
    // than once: 1.05 GB of writes per launch on the bench batch, 0.54 GB with
    // whole segments, 5 % faster; the Q = K P step measured 5 % slower with it).
    // The host launcher picks DEFER_Y = false with DTC_MV_DEFER=0 (A/B measurements).
    constexpr bool DEFER = DEFER_Y && OUT == OUT_CHEB && !SPARSE && IO32 == 0 && sizeof(T) == 4 && NG == 1;
    V ybuf[DEFER ? NCH : 1];
#pragma unroll
    for (int k = 0; k < (DEFER ? NCH : 1); ++k) vzero(ybuf[k]);

    // Per chunk: element products -> barrier -> stage the next chunk and issue
    // the loads of the one after -> gather -> barrier.  The next chunk is staged
    // before the gather, so the wait for its x loads (and for any earlier
    // stores) sits behind a whole element phase instead of right behind the
    // gather's stores.
    for (int chunk = c0; chunk < c1; ++chunk) {
        // gather operands of this chunk: b and d (OUT_BRES / OUT_CHEB) from memory,
        // x from LDS before the next chunk overwrites it
        V e_b[NG], e_d[NG], e_x[NG];
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            vzero(e_b[j]);
            vzero(e_d[j]);
            vzero(e_x[j]);
            if (g_off[j] >= 0) {
                const int vi = g_off[j] + chunk;
                if (OUT == OUT_BRES || OUT == OUT_CHEB) e_b[j] = reinterpret_cast<const V*>(A.bvec)[vi];
                if (OUT == OUT_CHEB && (A.cflags & CH_READ_D)) e_d[j] = reinterpret_cast<const V*>(A.dvec)[vi];
                if (OUT == OUT_MASKED_DOT || OUT == OUT_CHEB) e_x[j] = s_x[g_xi[j]];
            }
        }

        // ---- element products: task t = (tile element t / RC, load case t % RC)
        {
            const T* sx = reinterpret_cast<const T*>(s_x);
            T* sc = reinterpret_cast<T*>(s_c);
            const int ntask = ne * RC;
#pragma unroll 1
            for (int t = tid; t < ntask; t += NT) {
                const int r = t % RC, k = t / RC;
                const int info = s_info[k];
                const T E = s_E[k];
                const unsigned mask = (unsigned)info >> 24;
                const T* xb = sx + (info & 511) * SX + r;
                T* cb = sc + ((info >> 9) & 255) * SX + r;
                T w[8][3], y[8][3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    T v[8];
#pragma unroll
                    for (int b = 0; b < 8; ++b) v[b] = xb[hcorner(b) * SX + d * RC];
                    fht_fwd(v);
#pragma unroll
                    for (int b = 0; b < 8; ++b) w[b][d] = v[b];
                }
                walsh_apply(wk, w, y);
                T o[3][8];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    // mode 0 (rigid translation) carries no force: the first
                    // inverse butterfly of the pair (0, 1) reduces to (-y1, y1)
                    const T y1 = y[1][d] * E;
                    o[d][0] = -y1;
                    o[d][1] = y1;
#pragma unroll
                    for (int b = 2; b < 8; b += 2) {
                        const T a = y[b][d] * E, c = y[b + 1][d] * E;
                        o[d][b] = a - c;
                        o[d][b + 1] = a + c;
                    }
                    fht_inv_tail(o[d]);
                }
                if constexpr (SPARSE) {
                    // owned corners only: slot of node ob = first slot + rank of corner b among its elements
                    const int ob0 = ((info >> 9) & 255) - OB_BIAS;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        if (!(mask & (1u << b))) continue;
                        const int ob = ob0 + ocorner(b);
                        const int slot = s_sb[ob] + __popc((unsigned)s_nm[ob] & ((1u << b) - 1u));
                        T* dst = sc + slot * SX + r;
#pragma unroll
                        for (int d = 0; d < 3; ++d) dst[d * RC] = o[d][b];
                    }
                } else {
                    // corners outside the owned nodes write to this lane's trash slot
                    T* trash = sc + ((OB_BIAS + 8 * NO) * 3 + (tid & 63)) * RC;
#pragma unroll
                    for (int b = 0; b < 8; ++b) {
                        T* dst = (mask & (1u << b)) ? cb + (b * NO + ocorner(b)) * SX : trash;
#pragma unroll
                        for (int d = 0; d < 3; ++d) dst[d * RC] = o[d][b];
                    }
                }
            }
        }
        __syncthreads();
        if (chunk + 1 < c1) {
            stage(chunk + 1);
            if (IN != IN_U && chunk + 2 < c1) prefetch(chunk + 2);
        }

        // ---- gather: thread per owned (node, direction), all RC load cases in one vector
        double acc[RC];
#pragma unroll
        for (int w = 0; w < RC; ++w) acc[w] = 0.0;
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            if (g_off[j] < 0) continue;
            const unsigned m = g_m[j];
            const int od = (int)(m >> 9);          // o | d << 7
            const int o = od & 127, d = od >> 7;
            V y;
            vzero(y);
            if constexpr (SPARSE) {
                const int sb = s_sb[o], cnt = __popc((unsigned)s_nm[o]);
#pragma unroll
                for (int b = 0; b < 8; ++b)
                    if (b < cnt) vadd(y, s_c[(sb + b) * 3 + d]);
            } else {
#pragma unroll
                for (int b = 0; b < 8; ++b) vadd(y, s_c[(OB_BIAS + b * NO + o) * 3 + d]);
            }
            const bool bc = (m & 256u) != 0;
            const int vi = g_off[j] + chunk;
            if (OUT == OUT_MASKED_DOT) {
                if (bc) vzero(y);
#pragma unroll
                for (int w = 0; w < RC; ++w) acc[w] += (double)vcomp(e_x[j], w) * (double)vcomp(y, w);
            } else if (OUT == OUT_RESID) {
                if (bc) vzero(y); else y = vneg(y);
#pragma unroll
                for (int w = 0; w < RC; ++w) acc[w] += (double)vcomp(y, w) * (double)vcomp(y, w);
            } else if (OUT == OUT_MASKED) {
                if (bc) vzero(y);
            } else if (OUT == OUT_BRES || OUT == OUT_CHEB) {
                const V bv = e_b[j];
                if (OUT == OUT_BRES) {
                    if (bc) vzero(y);
                    else {
                        V t = bv;
                        vadd(t, vneg(y));
                        y = t;
                    }
                } else {
                    V* dp = reinterpret_cast<V*>(A.dvec) + vi;
                    V dn, xn;
                    vzero(dn);
                    vzero(xn);
                    if (!bc) {
                        const T di = e_di[j];
                        const V xv = e_x[j];
                        const V dv = (A.cflags & CH_D_FROM_X) ? xv : e_d[j];
                        T* dnp = reinterpret_cast<T*>(&dn);
                        T* xnp = reinterpret_cast<T*>(&xn);
#pragma unroll
                        for (int w = 0; w < RC; ++w) {
                            const T r = vcomp(bv, w) - vcomp(y, w);
                            dnp[w] = (T)A.c1 * vcomp(dv, w) + (T)A.c2 * di * r;
                            xnp[w] = vcomp(xv, w) + dnp[w];
                        }
                    }
                    if (A.cflags & CH_WRITE_D) *dp = dn;
                    y = xn;
                    if (PART)
#pragma unroll
                        for (int w = 0; w < RC; ++w) acc[w] += (double)vcomp(bv, w) * (double)vcomp(y, w);
                }
            }
            if constexpr (IO32 == 1) {
                reinterpret_cast<float2*>(yout)[vi] = make_float2((float)y.x, (float)y.y);
            } else if constexpr (DEFER) {
#pragma unroll
                for (int k = 0; k < NCH; ++k)
                    if (k == chunk - c0) ybuf[k] = y;
            } else {
                yv[vi] = y;
            }
        }
        if constexpr (SUMS) {
            // every lane holds the RC load cases of the chunk: fixed-order row sums
            // (waves past the last gather item hold zeros)
            const int lane = tid & 63, row = (tid >> 6) * 4 + (lane >> 4);
            double v = 0.0;
            int lc = lane & (RC - 1);
            bool out = (lane & 15) >= 16 - RC;
            if ((tid & ~63) < 3 * n_own) out = row_sums<RC>(acc, lane, &v, &lc);
            if (out) s_wred[chunk][row][lc] = v;
        }
        __syncthreads();
    }
    if constexpr (DEFER) {
        if (g_off[0] >= 0) {
#pragma unroll
            for (int k = 0; k < NCH; ++k)
                if (c0 + k < c1) yv[g_off[0] + c0 + k] = ybuf[k];
        }
    }
    if (SUMS && tid < R &&
        tid / RC >= c0 && tid / RC < c1) {
        const int ch = tid / RC, r = tid % RC;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW * 4; ++w) s += s_wred[ch][w][r];
        A.partial[(int64_t)bid * R + tid] = s;
    }
}

#include "matvec_mf.hpp"

// ---------------------------------------------------------------- vector kernels
// per brick, thread t keeps load case t % 24
template <typename T>
__global__ __launch_bounds__(NT_UPD) void k_cg_init(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                                    const int* __restrict__ active, T* __restrict__ x,
                                                    const T* __restrict__ r, T* __restrict__ p,
                                                    const T* __restrict__ dinv, double* __restrict__ prz,
                                                    double* __restrict__ prr, int zero_x, int set_p,
                                                    const int* __restrict__ keep_p, const T* __restrict__ zin) {
    __shared__ double s1[NT_UPD], s2[NT_UPD];
    const BrickInfo B = bricks[blockIdx.x];
    const int tid = threadIdx.x, lc = tid % R;
    double a1 = 0.0, a2 = 0.0;
    if (!(done && done[B.dom])) {
        const int act = active[B.dom * R + lc];
        // keep the search direction of a load case still iterating when its
        // residual was replaced; restart it otherwise
        const int sp = set_p || !(keep_p && keep_p[B.dom * R + lc]);
        const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
        for (int64_t i = tid; i < cnt; i += NT_UPD) {
            int64_t idx = base + i;
            T rv = r[idx];
            T z = act ? (zin ? zin[idx] : rv * dinv[idx / R]) : T(0);
            if (act && sp) p[idx] = z;
            if (zero_x) x[idx] = T(0);
            a1 += (double)rv * (double)z;
            a2 += (double)rv * (double)rv;
        }
    }
    s1[tid] = a1; s2[tid] = a2;
    __syncthreads();
    if (tid < R) {
        double u = 0.0, v = 0.0;
        for (int k = tid; k < NT_UPD; k += R) { u += s1[k]; v += s2[k]; }
        prz[(int64_t)blockIdx.x * R + tid] = u;
        prr[(int64_t)blockIdx.x * R + tid] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(NT_UPD) void k_cg_update(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                                      const double* __restrict__ alpha, T* __restrict__ x,
                                                      T* __restrict__ r, const T* __restrict__ p,
                                                      const T* __restrict__ q, const T* __restrict__ dinv,
                                                      double* __restrict__ prz, double* __restrict__ prr) {
    __shared__ double s1[NT_UPD], s2[NT_UPD];
    const BrickInfo B = bricks[blockIdx.x];
    const int tid = threadIdx.x, lc = tid % R;
    if (done && done[B.dom]) return;
    const T a = (T)alpha[B.dom * R + lc];
    double a1 = 0.0, a2 = 0.0;
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = tid; i < cnt; i += NT_UPD) {
        int64_t idx = base + i;
        T rv = r[idx];
        if (a != T(0)) {      // frozen load cases keep x and r
            rv -= a * q[idx];
            x[idx] += a * p[idx];
            r[idx] = rv;
        }
        if (prz) a1 += (double)rv * (double)(rv * dinv[idx / R]);
        a2 += (double)rv * (double)rv;
    }
    s1[tid] = a1; s2[tid] = a2;
    __syncthreads();
    if (tid < R) {
        double u = 0.0, v = 0.0;
        for (int k = tid; k < NT_UPD; k += R) { u += s1[k]; v += s2[k]; }
        if (prz) prz[(int64_t)blockIdx.x * R + tid] = u;
        prr[(int64_t)blockIdx.x * R + tid] = v;
    }
}

template <typename T>
__global__ __launch_bounds__(NT_UPD) void k_cg_dir(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                                   const double* __restrict__ beta, const T* __restrict__ r,
                                                   T* __restrict__ p, const T* __restrict__ dinv,
                                                   const T* __restrict__ zin) {
    const BrickInfo B = bricks[blockIdx.x];
    if (done && done[B.dom]) return;
    const int tid = threadIdx.x, lc = tid % R;
    const T b = (T)beta[B.dom * R + lc];
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = tid; i < cnt; i += NT_UPD) {
        int64_t idx = base + i;
        p[idx] = (zin ? zin[idx] : r[idx] * dinv[idx / R]) + b * p[idx];
    }
}

// ---------------------------------------------------------------- per-domain reductions (one block per domain)
constexpr int RED_NT = 256;

// sum over the bricks of a domain of part[b][lc] for the 24 load cases:
// strided per thread, then a fixed shuffle/LDS tree (deterministic)
__device__ inline void dom_sum24(const double* __restrict__ part, int b0, int b1, double* s_out /* [4][24] */,
                                 double out[R]) {
    const int tid = threadIdx.x;
    double acc[R];
#pragma unroll
    for (int lc = 0; lc < R; ++lc) acc[lc] = 0.0;
    for (int b = b0 + tid; b < b1; b += RED_NT)
#pragma unroll
        for (int lc = 0; lc < R; ++lc) acc[lc] += part[(int64_t)b * R + lc];
#pragma unroll
    for (int lc = 0; lc < R; ++lc) {
        double v = acc[lc];
        for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m);
        acc[lc] = v;
    }
    const int wave = tid / 64, lane = tid % 64;
    if (lane == 0)
        for (int lc = 0; lc < R; ++lc) s_out[wave * R + lc] = acc[lc];
    __syncthreads();
    for (int lc = 0; lc < R; ++lc) out[lc] = ((s_out[lc] + s_out[R + lc]) + s_out[2 * R + lc]) + s_out[3 * R + lc];
    __syncthreads();
}

__global__ __launch_bounds__(RED_NT) void k_red_init(const DomInfo* __restrict__ doms, const int* __restrict__ skip,
                                                     const double* __restrict__ prz, const double* __restrict__ prr,
                                                     const double* __restrict__ thr2, int* __restrict__ active,
                                                     double* __restrict__ rz, int* __restrict__ done) {
    __shared__ double s_out[4 * R];
    const int d = blockIdx.x, tid = threadIdx.x;
    const DomInfo D = doms[d];
    double z[R], rr[R];
    dom_sum24(prz, D.brick_begin, D.brick_end, s_out, z);
    dom_sum24(prr, D.brick_begin, D.brick_end, s_out, rr);
    if (tid == 0) {
        int any = 0;
        for (int lc = 0; lc < R; ++lc) {
            int a = (skip && skip[d]) ? 0 : active[d * R + lc];
            if (a && !(rr[lc] > thr2[d * R + lc])) a = 0;
            active[d * R + lc] = a;
            rz[d * R + lc] = z[lc];
            any |= a;
        }
        done[d] = any ? 0 : 1;
    }
}

__global__ __launch_bounds__(RED_NT) void k_red_alpha(const DomInfo* __restrict__ doms, const int* __restrict__ done,
                                                      const double* __restrict__ ppq, int* __restrict__ active,
                                                      const double* __restrict__ rz, double* __restrict__ alpha) {
    __shared__ double s_out[4 * R];
    const int d = blockIdx.x, tid = threadIdx.x;
    if (done[d]) return;
    const DomInfo D = doms[d];
    double pq[R];
    dom_sum24(ppq, D.brick_begin, D.brick_end, s_out, pq);
    if (tid < R) {
        const int lc = tid;
        int a = active[d * R + lc];
        double al = 0.0;
        if (a) {
            if (pq[lc] > 0.0) al = rz[d * R + lc] / pq[lc];
            else active[d * R + lc] = 0;      // breakdown: freeze this load case
        }
        alpha[d * R + lc] = al;
    }
}

__global__ __launch_bounds__(RED_NT) void k_red_beta(const DomInfo* __restrict__ doms, int* __restrict__ done,
                                                     const double* __restrict__ prz, const double* __restrict__ prr,
                                                     const double* __restrict__ thr2, int* __restrict__ active,
                                                     double* __restrict__ rz, double* __restrict__ beta,
                                                     int64_t* __restrict__ iters, int64_t max_iters) {
    __shared__ double s_out[4 * R];
    const int d = blockIdx.x, tid = threadIdx.x;
    if (done[d]) return;
    const DomInfo D = doms[d];
    double z[R], rr[R];
    dom_sum24(prz, D.brick_begin, D.brick_end, s_out, z);
    dom_sum24(prr, D.brick_begin, D.brick_end, s_out, rr);
    if (tid == 0) {
        int any = 0;
        for (int lc = 0; lc < R; ++lc) {
            int a = active[d * R + lc];
            double bt = 0.0;
            if (a) {
                if (!(rr[lc] > thr2[d * R + lc])) a = 0;
                else {
                    const double old = rz[d * R + lc];
                    bt = old != 0.0 ? z[lc] / old : 0.0;
                }
            }
            active[d * R + lc] = a;
            rz[d * R + lc] = z[lc];
            beta[d * R + lc] = bt;
            any |= a;
        }
        const int64_t it = iters[d] + 1;
        iters[d] = it;
        if (!any || it >= max_iters) done[d] = 1;
    }
}

// outer (FP64) residual bookkeeping after a RESID pass: rr = ||f - K u||^2 per load case;
// load cases above tol^2 ||f||^2 stay active.  The recursive PCG residual target
// is the larger of (tol/2)^2 ||f||^2 (so the true residual check passes after
// drift) and itol2 * rr (an inner solve reduces the residual it started from by
// at most sqrt(itol2), itol2 per domain; null: no relative target)
__global__ __launch_bounds__(RED_NT) void k_red_outer(const DomInfo* __restrict__ doms, const double* __restrict__ prr,
                                                      int first, double tol2, const double* __restrict__ itol2v,
                                                      double* __restrict__ rr0,
                                                      double* __restrict__ rrcur, int* __restrict__ active64,
                                                      int* __restrict__ active, double* __restrict__ thr2,
                                                      int* __restrict__ done64, int* __restrict__ done) {
    __shared__ double s_out[4 * R];
    const int d = blockIdx.x, tid = threadIdx.x;
    const DomInfo D = doms[d];
    double rr[R];
    dom_sum24(prr, D.brick_begin, D.brick_end, s_out, rr);
    if (tid == 0) {
        int any = 0;
        const double itol2 = itol2v ? itol2v[d] : 0.0;
        for (int lc = 0; lc < R; ++lc) {
            if (first) rr0[d * R + lc] = rr[lc];
            const double r0 = rr0[d * R + lc];
            rrcur[d * R + lc] = rr[lc];
            const int a = (r0 > 0.0) && (rr[lc] > tol2 * r0);
            active64[d * R + lc] = a;
            active[d * R + lc] = a;
            thr2[d * R + lc] = fmax(0.25 * tol2 * r0, itol2 * rr[lc]);
            any |= a;
        }
        done64[d] = any ? 0 : 1;
        done[d] = any ? 0 : 1;
    }
}

// column scale of the FP32 correction problem: 1 / ||r64|| of the load case
// (rr = ||r64||^2 per load case), 1 without scaling or for a zero residual
__device__ inline double col_scale(const double* __restrict__ rr, int dom, int lc) {
    if (!rr) return 1.0;
    const double v = rr[dom * R + lc];
    return v > 0.0 ? 1.0 / sqrt(v) : 1.0;
}

// r32 = r64 (times its column scale when rr is given), x32 = 0 when given (residual replacement)
__global__ __launch_bounds__(NT_UPD) void k_to32(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                                 const double* __restrict__ r64, float* __restrict__ r32,
                                                 float* __restrict__ x32, const double* __restrict__ rr) {
    const BrickInfo B = bricks[blockIdx.x];
    if (done[B.dom]) return;
    const double sc = col_scale(rr, B.dom, threadIdx.x % R);      // NT_UPD is a multiple of 24
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = threadIdx.x; i < cnt; i += NT_UPD) {
        r32[base + i] = (float)(r64[base + i] * sc);
        if (x32) x32[base + i] = 0.f;
    }
}

// x64 += x32 (divided by its column scale when rr is given), x32 = 0 (group
// update of the FP32 correction)
__global__ __launch_bounds__(NT_UPD) void k_acc64(const BrickInfo* __restrict__ bricks, const int* __restrict__ skip,
                                                  float* __restrict__ x32, double* __restrict__ x64,
                                                  const double* __restrict__ rr) {
    const BrickInfo B = bricks[blockIdx.x];
    if (skip[B.dom]) return;
    const double sc = 1.0 / col_scale(rr, B.dom, threadIdx.x % R);
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = threadIdx.x; i < cnt; i += NT_UPD) {
        x64[base + i] += (double)x32[base + i] * sc;
        x32[base + i] = 0.f;
    }
}

// ---------------------------------------------------------------- block PCG
// The 24 load cases of a domain share one Krylov block: block PCG in the
// A-orthogonal form
//   Q = A P,  W = P^T Q,  alpha = W^+ (P^T R),  X += P alpha,  R -= Q alpha,
//   Z = M R,  beta = -W^+ (Q^T Z),  P = Z + P beta.
// A vector row is one (node, direction) pair holding the 24 load cases.
constexpr int GR_NT = 256;          // 4 waves
constexpr int SLICE_NODES = 512;    // nodes per Gram workgroup (at least)
constexpr int MAX_SLICES = 1024;    // Gram workgroups per domain (at most)
constexpr int RR = R * R;           // entries of a 24 x 24 coefficient matrix
constexpr double PIVOT_DROP = 1e-9;   // of P^T A P
constexpr double ORTH_DROP = 1e-10;   // of P^T P

// Gram partial sums over the rows of one slice of a domain's nodes:
//   g[0] = A^T B, g[1] = A^T C (NGR = 2; zero for NGR = 1), FP64 sums of the
// products of FP32 (or FP64) operands.  Lane (i, j) of a wave keeps rows 3i..3i+2 x columns 3j..3j+2; the
// four waves take every fourth vector row.  The loads of the next GU rows of a
// wave are issued before the products of the current ones (a slice is a long
// sequential walk per wave); the waves' sums are added in wave order.
template <int NGR, typename TA = float, typename TB = float, int GU = 4>
__global__ __launch_bounds__(GR_NT) void k_gram(const int4* __restrict__ slices, const int* __restrict__ done,
                                               const TA* __restrict__ A, const TB* __restrict__ B,
                                               const TB* __restrict__ C, double* __restrict__ gpart) {
    __shared__ double s_g[NGR * RR];
    const int4 sl = slices[blockIdx.x];   // domain, first node, end node
    if (done[sl.x]) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ci = 3 * (lane >> 3), cj = 3 * (lane & 7);
    double g[NGR][9];
#pragma unroll
    for (int t = 0; t < NGR; ++t)
#pragma unroll
        for (int k = 0; k < 9; ++k) g[t][k] = 0.0;
    const int64_t r1 = (int64_t)sl.z * 3;
    // two register sets: the next GU rows load into one while the products of the
    // other run; full steps (all GU rows in the slice) load without guards
    TA av[2][GU][3];
    TB bv[2][NGR][GU][3];
    auto load = [&](int set, int64_t row0) {
#pragma unroll
        for (int u = 0; u < GU; ++u) {
            const int64_t row = row0 + 4 * u;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                av[set][u][i] = A[row * R + ci + i];
                bv[set][0][u][i] = B[row * R + cj + i];
                if (NGR == 2) bv[set][NGR - 1][u][i] = C[row * R + cj + i];
            }
        }
    };
    auto product = [&](const TA (&a)[3], const TB (&b)[3], double (&gg)[9]) {
        double ad[3], bd[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            ad[i] = (double)a[i];
            bd[i] = (double)b[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) gg[3 * i + j] = fma(ad[i], bd[j], gg[3 * i + j]);
    };
    auto products = [&](int set) {
#pragma unroll
        for (int u = 0; u < GU; ++u)
#pragma unroll
            for (int t = 0; t < NGR; ++t) product(av[set][u], bv[set][t][u], g[t]);
    };
    int64_t row0 = (int64_t)sl.y * 3 + wave;
    const int64_t full_end = r1 - 4 * (GU - 1);   // row0 < full_end: rows row0 + 4u < r1
    if (row0 < full_end) load(0, row0);
    while (row0 < full_end) {
        if (row0 + 4 * GU < full_end) load(1, row0 + 4 * GU);
        products(0);
        row0 += 4 * GU;
        if (row0 >= full_end) break;
        if (row0 + 4 * GU < full_end) load(0, row0 + 4 * GU);
        products(1);
        row0 += 4 * GU;
    }
    for (; row0 < r1; row0 += 4) {      // the last rows of the wave (fewer than GU)
        TA a[3];
        TB b[NGR][3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            a[i] = A[row0 * R + ci + i];
            b[0][i] = B[row0 * R + cj + i];
            if (NGR == 2) b[NGR - 1][i] = C[row0 * R + cj + i];
        }
#pragma unroll
        for (int t = 0; t < NGR; ++t) product(a, b[t], g[t]);
    }
    for (int w = 0; w < 4; ++w) {
        if (wave == w)
#pragma unroll
            for (int t = 0; t < NGR; ++t)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int e = t * RR + (ci + i) * R + cj + j;
                        s_g[e] = w == 0 ? g[t][3 * i + j] : s_g[e] + g[t][3 * i + j];
                    }
        __syncthreads();
    }
    double* out = gpart + (int64_t)blockIdx.x * 2 * RR;
    for (int e = tid; e < 2 * RR; e += GR_NT) out[e] = e < NGR * RR ? s_g[e] : 0.0;
}

// The same Gram partials with the rows staged through LDS: the workgroup loads
// blocks of GB_ROWS vector rows of A, B (and C) with coalesced 16-byte loads (the
// next block's loads issued before the current block's products), and lane (i, j)
// reads its 3 + 3 (+ 3) values of a row from LDS, where the 8 lanes sharing an
// address are served by one broadcast.  (Loading them from global memory sends
// each value through the vector memory path 8 times.)  Wave w takes rows w, w + 4,
// ... of each block in order, so every wave sums the same rows in the same order
// as k_gram and the results are identical.
constexpr int GB_ROWS = 32;
template <int NGR, typename TA = float, typename TB = float>
__global__ __launch_bounds__(GR_NT) void k_gram_lds(const int4* __restrict__ slices, const int* __restrict__ done,
                                                   const TA* __restrict__ A, const TB* __restrict__ B,
                                                   const TB* __restrict__ C, double* __restrict__ gpart) {
    constexpr int EA = 16 / sizeof(TA), EB = 16 / sizeof(TB);      // elements per 16-byte vector
    constexpr int VA = GB_ROWS * R / EA, VB = GB_ROWS * R / EB;     // 16-byte vectors per block and array
    constexpr int VT = VA + NGR * VB;
    constexpr int NPF = (VT + GR_NT - 1) / GR_NT;                   // prefetch vectors per thread
    __shared__ uint4 sAv[VA];
    __shared__ uint4 sBv[NGR][VB];
    __shared__ double s_g[NGR * RR];
    const int4 sl = slices[blockIdx.x];   // domain, first node, end node
    if (done[sl.x]) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int ci = 3 * (lane >> 3), cj = 3 * (lane & 7);
    const int64_t r0 = (int64_t)sl.y * 3, r1 = (int64_t)sl.z * 3;
    double g[NGR][9];
#pragma unroll
    for (int t = 0; t < NGR; ++t)
#pragma unroll
        for (int k = 0; k < 9; ++k) g[t][k] = 0.0;
    uint4 pf[NPF];
    // vector v of block row0: array (0 = A, 1 = B, 2 = C), rows beyond r1 read as zero
    auto load = [&](int64_t row0) {
#pragma unroll
        for (int k = 0; k < NPF; ++k) {
            const int v = tid + k * GR_NT;
            uint4 val = make_uint4(0u, 0u, 0u, 0u);
            if (v < VA) {
                const int64_t e = row0 * R + (int64_t)v * EA;
                if (e < r1 * R) val = reinterpret_cast<const uint4*>(A + e)[0];
            } else if (v < VT) {
                const int w = v - VA, t = w / VB, u = w % VB;
                const int64_t e = row0 * R + (int64_t)u * EB;
                const TB* src = t == 0 ? B : C;
                if (e < r1 * R) val = reinterpret_cast<const uint4*>(src + e)[0];
            }
            pf[k] = val;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int k = 0; k < NPF; ++k) {
            const int v = tid + k * GR_NT;
            if (v < VA) sAv[v] = pf[k];
            else if (v < VT) sBv[(v - VA) / VB][(v - VA) % VB] = pf[k];
        }
    };
    const TA* sA = reinterpret_cast<const TA*>(sAv);
    if (r0 < r1) load(r0);
    for (int64_t row0 = r0; row0 < r1; row0 += GB_ROWS) {
        __syncthreads();                              // the previous block's products are done
        store();
        __syncthreads();
        if (row0 + GB_ROWS < r1) load(row0 + GB_ROWS);
#pragma unroll 2
        for (int k = 0; k < GB_ROWS / 4; ++k) {
            const int rr = wave + 4 * k;
            if (row0 + rr >= r1) break;
            double ad[3], bd[NGR][3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                ad[i] = (double)sA[rr * R + ci + i];
#pragma unroll
                for (int t = 0; t < NGR; ++t) bd[t][i] = (double)reinterpret_cast<const TB*>(sBv[t])[rr * R + cj + i];
            }
#pragma unroll
            for (int t = 0; t < NGR; ++t)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[t][3 * i + j] = fma(ad[i], bd[t][j], g[t][3 * i + j]);
        }
    }
    __syncthreads();
    for (int w = 0; w < 4; ++w) {
        if (wave == w)
#pragma unroll
            for (int t = 0; t < NGR; ++t)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int e = t * RR + (ci + i) * R + cj + j;
                        s_g[e] = w == 0 ? g[t][3 * i + j] : s_g[e] + g[t][3 * i + j];
                    }
        __syncthreads();
    }
    double* out = gpart + (int64_t)blockIdx.x * 2 * RR;
    for (int e = tid; e < 2 * RR; e += GR_NT) out[e] = e < NGR * RR ? s_g[e] : 0.0;
}

// gsum[d] = sum over the domain's slices of the Gram partials, in a fixed order:
// workgroup (d, c) sums entries 64c..64c+63 of the 2 x 576; wave w takes the
// slices w, w + 4, ... and the four waves are added in order
__global__ __launch_bounds__(256) void k_gram_sum(const int* __restrict__ dslice, const int* __restrict__ done,
                                                 const double* __restrict__ gpart, double* __restrict__ gsum) {
    __shared__ double s_w[4][64];
    const int d = blockIdx.x;
    if (done[d]) return;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, e = (int)blockIdx.y * 64 + lane;
    const int s0 = dslice[d], s1 = dslice[d + 1];
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    int s = s0 + wave;
    for (; s + 12 < s1; s += 16)
#pragma unroll
        for (int u = 0; u < 4; ++u) a[u] += gpart[(int64_t)(s + 4 * u) * 2 * RR + e];
    for (int u = 0; s < s1; s += 4, ++u) a[u & 3] += gpart[(int64_t)s * 2 * RR + e];
    s_w[wave][lane] = ((a[0] + a[1]) + a[2]) + a[3];
    __syncthreads();
    if (wave == 0) gsum[(int64_t)d * 2 * RR + e] = ((s_w[0][lane] + s_w[1][lane]) + s_w[2][lane]) + s_w[3][lane];
}

// Per domain: takes the Gram sums of k_gram_sum, then
//   mode 0: W = sym(g[0]) = P^T A P is factored as L L^T and L kept; a pivot at
//           or below 1e-9 of its diagonal entry before elimination (or not
//           finite) marks a direction that depends on the previous ones: its
//           column of L is zero and the solves give it a zero coefficient.
//           coef = W^+ g[1];
//   mode 1: coef = -W^+ g[0] with the kept factor;
//   mode 2: G = sym(g[0]) = P^T P is factored the same way (pivots at or below
//           1e-10 dropped, L not kept) and coef = L^-T, so that P coef has
//           orthonormal columns (zero columns for the dropped directions).
// coef[k * 24 + j] multiplies vector column k in the update of column j.
constexpr int BS_NT = 256;

__global__ __launch_bounds__(BS_NT) void k_bsolve(const int* __restrict__ done, const double* __restrict__ gsum,
                                                 int mode, double* __restrict__ lfac, float* __restrict__ coef) {
    __shared__ double s_l[RR], s_b[RR], s_d[R];
    const int d = blockIdx.x, lane = threadIdx.x;
    if (done[d]) return;
    double* L = lfac + (int64_t)d * RR;
    for (int e = lane; e < RR; e += BS_NT) {
        const double w = gsum[(int64_t)d * 2 * RR + e];
        const double g = mode == 0 ? gsum[(int64_t)d * 2 * RR + RR + e] : 0.0;
        if (mode == 0) {
            s_l[e] = w;
            s_b[e] = g;
        } else if (mode == 1) {
            s_b[e] = w;
            s_l[e] = L[e];
        } else {
            s_l[e] = w;
            s_b[e] = (e / R == e % R) ? 1.0 : 0.0;
        }
    }
    __syncthreads();
    if (mode != 1) {
        const double drop = mode == 0 ? PIVOT_DROP : ORTH_DROP;
        constexpr int NV = (RR + BS_NT - 1) / BS_NT;
        double v[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) {
            const int e = lane + BS_NT * n, i = e / R, j = e % R;
            if (e < RR) v[n] = 0.5 * (s_l[i * R + j] + s_l[j * R + i]);
        }
        __syncthreads();
#pragma unroll
        for (int n = 0; n < NV; ++n)
            if (lane + BS_NT * n < RR) s_l[lane + BS_NT * n] = v[n];
        __syncthreads();
        if (lane < R) s_d[lane] = s_l[lane * (R + 1)];
        __syncthreads();
        for (int k = 0; k < R; ++k) {
            const double dk = s_l[k * R + k];
            const bool keep = dk > drop * s_d[k] && dk < INFINITY;
            const double lk = keep ? sqrt(dk) : 0.0;
            __syncthreads();
            if (lane >= k && lane < R) s_l[lane * R + k] = keep ? (lane == k ? lk : s_l[lane * R + k] / lk) : 0.0;
            __syncthreads();
            for (int e = lane; e < RR; e += BS_NT) {
                const int i = e / R, j = e % R;
                if (j > k && i >= j) s_l[e] -= s_l[i * R + k] * s_l[j * R + k];
            }
            __syncthreads();
        }
        if (mode == 0)
            for (int e = lane; e < RR; e += BS_NT) L[e] = s_l[e];
    }
    __syncthreads();
    // L L^T y = b (modes 0, 1) or L^T y = I (mode 2) for column j of b, in place
    // (lane j owns column j of s_b)
    if (lane < R) {
        const int j = lane;
        for (int k = 0; k < R && mode != 2; ++k) {
            double t = s_b[k * R + j];
            for (int m = 0; m < k; ++m) t -= s_l[k * R + m] * s_b[m * R + j];
            const double lkk = s_l[k * R + k];
            s_b[k * R + j] = lkk > 0.0 ? t / lkk : 0.0;
        }
        for (int k = R - 1; k >= 0; --k) {
            double t = s_b[k * R + j];
            for (int m = k + 1; m < R; ++m) t -= s_l[m * R + k] * s_b[m * R + j];
            const double lkk = s_l[k * R + k];
            s_b[k * R + j] = lkk > 0.0 ? t / lkk : 0.0;
        }
        for (int k = 0; k < R; ++k)
            coef[(int64_t)d * RR + k * R + j] = (float)(mode == 1 ? -s_b[k * R + j] : s_b[k * R + j]);
    }
}

// X += P alpha, R -= Q alpha on the rows of each node slice of a domain (the Gram
// slices; alpha: k_bsolve mode 0), and per-slice sums of r.r per load case.  Of the 6 threads of a row, thread q
// keeps load cases 4q..4q+3.
__global__ __launch_bounds__(NT_UPD) void k_bupdate(const int4* __restrict__ slices, const int* __restrict__ done,
                                                   const float* __restrict__ coef, float* __restrict__ x,
                                                   float* __restrict__ r, const float* __restrict__ p,
                                                   const float* __restrict__ q, double* __restrict__ prr) {
    __shared__ float4 s_c[R * 6];
    __shared__ double s_a[NT_UPD][4];
    __shared__ float4 s_pq[2][NT_UPD / 6][6];   // P and Q rows of the current block
    const int4 sl = slices[blockIdx.x];   // domain, first node, end node
    if (done[sl.x]) return;
    const int tid = threadIdx.x, qd = tid % 6, rw = tid / 6;
    const float4* c4 = reinterpret_cast<const float4*>(coef + (int64_t)sl.x * RR);
    for (int e = tid; e < R * 6; e += NT_UPD) s_c[e] = c4[e];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t r0 = (int64_t)sl.y * 3, r1 = (int64_t)sl.z * 3;
    // the block's P and Q rows are loaded once (thread t: 16-byte vector t of P and
    // of Q, coalesced) and read from LDS by the 6 threads of a row (a broadcast),
    // instead of every thread loading both whole rows; the next block's vectors are
    // in flight during the current block's products
    constexpr int RB = NT_UPD / 6;   // rows per block (one per row lane)
    float4 pf_p, pf_q;
    auto load = [&](int64_t b0) {
        const int64_t row = b0 + tid / 6;
        pf_p = pf_q = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < r1) {
            pf_p = reinterpret_cast<const float4*>(p + row * R)[qd];
            pf_q = reinterpret_cast<const float4*>(q + row * R)[qd];
        }
    };
    if (r0 < r1) load(r0);
    for (int64_t b0 = r0; b0 < r1; b0 += RB) {
        __syncthreads();                 // the previous block's reads of s_pq are done (and s_c is set)
        s_pq[0][tid / 6][qd] = pf_p;
        s_pq[1][tid / 6][qd] = pf_q;
        __syncthreads();
        if (b0 + RB < r1) load(b0 + RB);
        const int64_t row = b0 + rw;
        if (row >= r1) continue;
        float4* xr = reinterpret_cast<float4*>(x + row * R) + qd;
        float4* rr = reinterpret_cast<float4*>(r + row * R) + qd;
        float4 xv = *xr, rv = *rr;
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const float4 pv = s_pq[0][rw][c], qv = s_pq[1][rw][c];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const float4 a = s_c[(4 * c + w) * 6 + qd];
                const float pk = vcomp(pv, w), qk = vcomp(qv, w);
                xv.x += pk * a.x; xv.y += pk * a.y; xv.z += pk * a.z; xv.w += pk * a.w;
                rv.x -= qk * a.x; rv.y -= qk * a.y; rv.z -= qk * a.z; rv.w -= qk * a.w;
            }
        }
        *xr = xv;
        *rr = rv;
        acc[0] += (double)rv.x * (double)rv.x;
        acc[1] += (double)rv.y * (double)rv.y;
        acc[2] += (double)rv.z * (double)rv.z;
        acc[3] += (double)rv.w * (double)rv.w;
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) s_a[tid][w] = acc[w];
    __syncthreads();
    if (tid < R) {
        const int qq = tid / 4, w = tid % 4;
        double s = 0.0;
        for (int k = 0; k < NT_UPD / 6; ++k) s += s_a[k * 6 + qq][w];
        prr[(int64_t)blockIdx.x * R + tid] = s;
    }
}

// The FP64 form of k_bupdate (block PCG on FP64 vectors): x += P alpha,
// r -= Q alpha with x, r, Q in FP64 and the directions P in FP32, per-brick r.r.
// Of the 6 threads of a row, thread q keeps load cases 4q..4q+3 (as k_bupdate).
__global__ __launch_bounds__(NT_UPD) void k_bupdate64(const int4* __restrict__ slices, const int* __restrict__ done,
                                                     const float* __restrict__ coef, double* __restrict__ x,
                                                     double* __restrict__ r, const float* __restrict__ p,
                                                     const double* __restrict__ q, double* __restrict__ prr) {
    __shared__ float4 s_c[R * 6];
    __shared__ double s_a[NT_UPD][4];
    __shared__ float4 s_p[NT_UPD / 6][6];      // the block's P rows and
    __shared__ double2 s_q[NT_UPD / 6][12];    // Q rows (staged as in k_bupdate)
    const int4 sl = slices[blockIdx.x];   // domain, first node, end node
    if (done[sl.x]) return;
    const int tid = threadIdx.x, qd = tid % 6, rw = tid / 6;
    const float4* c4 = reinterpret_cast<const float4*>(coef + (int64_t)sl.x * RR);
    for (int e = tid; e < R * 6; e += NT_UPD) s_c[e] = c4[e];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    constexpr int RB = NT_UPD / 6;
    const int64_t r0b = (int64_t)sl.y * 3, r1 = (int64_t)sl.z * 3;
    float4 pf_p;
    double2 pf_q[2];
    auto load = [&](int64_t b0) {
        const int64_t row = b0 + tid / 6;
        pf_p = make_float4(0.f, 0.f, 0.f, 0.f);
        pf_q[0] = pf_q[1] = make_double2(0.0, 0.0);
        if (row < r1) {
            pf_p = reinterpret_cast<const float4*>(p + row * R)[qd];
            pf_q[0] = reinterpret_cast<const double2*>(q + row * R)[2 * qd];
            pf_q[1] = reinterpret_cast<const double2*>(q + row * R)[2 * qd + 1];
        }
    };
    if (r0b < r1) load(r0b);
    for (int64_t b0 = r0b; b0 < r1; b0 += RB) {
        __syncthreads();
        s_p[tid / 6][qd] = pf_p;
        s_q[tid / 6][2 * qd] = pf_q[0];
        s_q[tid / 6][2 * qd + 1] = pf_q[1];
        __syncthreads();
        if (b0 + RB < r1) load(b0 + RB);
        const int64_t row = b0 + rw;
        if (row >= r1) continue;
        const float4* pr = s_p[rw];
        const double2* qr = s_q[rw];
        double2* xr = reinterpret_cast<double2*>(x + row * R) + 2 * qd;
        double2* rr = reinterpret_cast<double2*>(r + row * R) + 2 * qd;
        double xv[4], rv[4];
        {
            const double2 x0 = xr[0], x1 = xr[1], r0 = rr[0], r1v = rr[1];
            xv[0] = x0.x; xv[1] = x0.y; xv[2] = x1.x; xv[3] = x1.y;
            rv[0] = r0.x; rv[1] = r0.y; rv[2] = r1v.x; rv[3] = r1v.y;
        }
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const float4 pv = pr[c];
            const double2 qa = qr[2 * c], qb = qr[2 * c + 1];
            const double qk[4] = {qa.x, qa.y, qb.x, qb.y};
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const float4 a = s_c[(4 * c + w) * 6 + qd];
                const double pk = (double)vcomp(pv, w);
                xv[0] += pk * (double)a.x; xv[1] += pk * (double)a.y; xv[2] += pk * (double)a.z; xv[3] += pk * (double)a.w;
                rv[0] -= qk[w] * (double)a.x; rv[1] -= qk[w] * (double)a.y;
                rv[2] -= qk[w] * (double)a.z; rv[3] -= qk[w] * (double)a.w;
            }
        }
        xr[0] = make_double2(xv[0], xv[1]);
        xr[1] = make_double2(xv[2], xv[3]);
        rr[0] = make_double2(rv[0], rv[1]);
        rr[1] = make_double2(rv[2], rv[3]);
#pragma unroll
        for (int w = 0; w < 4; ++w) acc[w] += rv[w] * rv[w];
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) s_a[tid][w] = acc[w];
    __syncthreads();
    if (tid < R) {
        const int qq = tid / 4, w = tid % 4;
        double s = 0.0;
        for (int k = 0; k < NT_UPD / 6; ++k) s += s_a[k * 6 + qq][w];
        prr[(int64_t)blockIdx.x * R + tid] = s;
    }
}

// pn = Z + P coef (z null: pn = P coef); pn aliases neither p nor z.  With beta
// (k_bsolve mode 1) the new direction block, with L^-T (mode 2) its orthonormal
// columns
// GRAM: also the per-slice Gram partials pnᵀpn of the rows written (for the
// orthonormalisation that follows): the block's new rows are staged in LDS and wave w
// sums rows w, w + 3, ... of each block, lane (i, j) its 3 x 3 entries, in FP64 (the
// layout of k_gram_lds's partials, another summation order)
template <bool GRAM>
__global__ __launch_bounds__(NT_UPD) void k_bdir(const int4* __restrict__ slices, const int* __restrict__ done,
                                                const float* __restrict__ coef, const float* __restrict__ p,
                                                const float* __restrict__ z, float* __restrict__ pn,
                                                double* __restrict__ gpart) {
    __shared__ float4 s_c[R * 6];
    __shared__ float4 s_p[NT_UPD / 6][6];   // P rows of the current block (staged as in k_bupdate)
    __shared__ float4 s_n[GRAM ? NT_UPD / 6 : 1][6];   // GRAM: the block's new rows
    __shared__ double s_g[GRAM ? RR : 1];
    const int4 sl = slices[blockIdx.x];   // domain, first node, end node
    if (done[sl.x]) return;
    const int tid = threadIdx.x, qd = tid % 6, rw = tid / 6;
    const int wave = tid >> 6, lane = tid & 63, ci = 3 * (lane >> 3), cj = 3 * (lane & 7);
    const float4* c4 = reinterpret_cast<const float4*>(coef + (int64_t)sl.x * RR);
    for (int e = tid; e < R * 6; e += NT_UPD) s_c[e] = c4[e];
    constexpr int RB = NT_UPD / 6, NWV = NT_UPD / 64;
    double g[GRAM ? 9 : 1];
#pragma unroll
    for (int k = 0; k < (GRAM ? 9 : 1); ++k) g[k] = 0.0;
    const int64_t r0 = (int64_t)sl.y * 3, r1 = (int64_t)sl.z * 3;
    float4 pf = make_float4(0.f, 0.f, 0.f, 0.f);
    auto load = [&](int64_t b0) {
        const int64_t row = b0 + tid / 6;
        pf = row < r1 ? reinterpret_cast<const float4*>(p + row * R)[qd] : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    if (r0 < r1) load(r0);
    for (int64_t b0 = r0; b0 < r1; b0 += RB) {
        __syncthreads();
        s_p[tid / 6][qd] = pf;
        __syncthreads();
        if (b0 + RB < r1) load(b0 + RB);
        const int64_t row = b0 + rw;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (row < r1) {
            if (z) v = reinterpret_cast<const float4*>(z + row * R)[qd];
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const float4 pv = s_p[rw][c];
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const float4 a = s_c[(4 * c + w) * 6 + qd];
                    const float pk = vcomp(pv, w);
                    v.x += pk * a.x; v.y += pk * a.y; v.z += pk * a.z; v.w += pk * a.w;
                }
            }
            reinterpret_cast<float4*>(pn + row * R)[qd] = v;
        }
        if constexpr (GRAM) {
            s_n[rw][qd] = v;   // rows past the slice hold zeros
            __syncthreads();
            const float* sn = reinterpret_cast<const float*>(s_n);
            for (int rr = wave; rr < RB; rr += NWV) {
                double ad[3], bd[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    ad[i] = (double)sn[rr * R + ci + i];
                    bd[i] = (double)sn[rr * R + cj + i];
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) g[3 * i + j] = fma(ad[i], bd[j], g[3 * i + j]);
            }
        }
    }
    if constexpr (GRAM) {
        __syncthreads();
        for (int w = 0; w < NWV; ++w) {
            if (wave == w)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const int e = (ci + i) * R + cj + j;
                        s_g[e] = w == 0 ? g[3 * i + j] : s_g[e] + g[3 * i + j];
                    }
            __syncthreads();
        }
        double* out = gpart + (int64_t)blockIdx.x * 2 * RR;
        for (int e = tid; e < 2 * RR; e += NT_UPD) out[e] = e < RR ? s_g[e] : 0.0;
    }
}

// per domain: the inner block iteration has finished when every load case's
// recursive residual (scaled back by its column scale, see k_to32) is at or
// below its target (a non-finite one is not), or at the iteration cap
__global__ __launch_bounds__(RED_NT) void k_bcheck(const int* __restrict__ dslice, int* __restrict__ done,
                                                  const double* __restrict__ prr, const double* __restrict__ thr2,
                                                  const double* __restrict__ rrcur, int64_t* __restrict__ iters,
                                                  int64_t max_iters) {
    __shared__ double s_out[4 * R];
    const int d = blockIdx.x, tid = threadIdx.x;
    if (done[d]) return;
    double rr[R];
    dom_sum24(prr, dslice[d], dslice[d + 1], s_out, rr);   // per-slice partials of k_bupdate
    if (tid == 0) {
        int any = 0;
        for (int lc = 0; lc < R; ++lc) {
            const double sc = col_scale(rrcur, d, lc);
            any |= !(rr[lc] <= thr2[d * R + lc] * sc * sc);
        }
        const int64_t it = iters[d] + 1;
        iters[d] = it;
        if (!any || it >= max_iters) done[d] = 1;
    }
}

// algorithmic bytes of one timed launch: over the domains not done, vpass
// vectors (FP32, or FP64 with fp64), the element moduli (4 or 8 bytes each) and,
// with jac, the FP32 Jacobi scaling; and its dense element-product flops
// 2 (3 nen)^2 R N_e; dom_bytes[d] = {V32, E32, J32, V64, F}.  out = {bytes, flops},
// each summed by one workgroup in a fixed order.
__global__ __launch_bounds__(256) void k_launch_bytes(const int* __restrict__ done, const double* __restrict__ dom_bytes,
                                                      int n_dom, int vpass, int fp64, int jac, double* __restrict__ out) {
    __shared__ double sh[2][256];
    double acc = 0.0, fl = 0.0;
    for (int d = threadIdx.x; d < n_dom; d += 256) {
        if (done && done[d]) continue;
        const double* q = dom_bytes + 5 * d;
        acc += vpass * (fp64 ? q[3] : q[0]) + (fp64 ? 2.0 : 1.0) * q[1] + (jac ? q[2] : 0.0);
        fl += q[4];
    }
    sh[0][threadIdx.x] = acc;
    sh[1][threadIdx.x] = fl;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sh[0][0];
        out[1] = sh[1][0];
    }
}

__global__ void k_fill_f64(double* __restrict__ a, int64_t n, double v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_f32(float* __restrict__ a, int64_t n, float v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_i32b(int* __restrict__ a, int64_t n, int v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}
__global__ void k_fill_i64(int64_t* __restrict__ a, int64_t n, int64_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------- topology
__global__ void k_brick_count(const int* __restrict__ numLat, const int* __restrict__ keepN, int Lx, int Ly,
                              int Lz, int bnx, int bny, int* __restrict__ cnt) {
    __shared__ int s;
    const int b = blockIdx.x, t = threadIdx.x;     // blockDim = BX*BY*BZ
    if (t == 0) s = 0;
    __syncthreads();
    const int bx = b % bnx, by = (b / bnx) % bny, bz = b / (bnx * bny);
    const int gx = bx * BX + t % BX, gy = by * BY + (t / BX) % BY, gz = bz * BZ + t / (BX * BY);
    if (gx < Lx && gy < Ly && gz < Lz) {
        int num = numLat[gx + Lx * (gy + Ly * gz)];
        if (num > 0 && keepN[num - 1]) atomicAdd(&s, 1);
    }
    __syncthreads();
    if (t == 0) cnt[b] = s;
}

__global__ void k_brick_assign(const BrickInfo* __restrict__ bl, int brick0, const int* __restrict__ numLat,
                               const int* __restrict__ keepN, const int* __restrict__ cref, int Lx, int Ly, int Lz,
                               int64_t lat_off, int* __restrict__ lat2node, int* __restrict__ sol2ref) {
    constexpr int NO = BX * BY * BZ;
    __shared__ int flag[NO];
    const BrickInfo B = bl[brick0 + blockIdx.x];
    const int t = threadIdx.x;
    const int gx = B.x0 + t % BX, gy = B.y0 + (t / BX) % BY, gz = B.z0 + t / (BX * BY);
    int kept = 0, num = 0;
    if (gx < Lx && gy < Ly && gz < Lz) {
        num = numLat[gx + Lx * (gy + Ly * gz)];
        kept = (num > 0 && keepN[num - 1]) ? 1 : 0;
    }
    flag[t] = kept;
    __syncthreads();
    if (kept) {
        int rank = 0;
        for (int k = 0; k < t; ++k) rank += flag[k];   // lexicographic order inside the brick
        int s = B.node_begin + rank;
        lat2node[lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)] = s;
        sol2ref[s] = cref[num - 1] + 1;
    }
}

__global__ void k_count_bc(const int* __restrict__ numLat, const int* __restrict__ keepN, int Lx, int Ly, int Lz,
                           int* __restrict__ out) {
    __shared__ int c;
    if (threadIdx.x == 0) c = 0;
    __syncthreads();
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < (int64_t)Lx * Ly * Lz) {
        int x = p % Lx, y = (p / Lx) % Ly, z = p / ((int64_t)Lx * Ly);
        int num = numLat[p];
        if (num > 0 && keepN[num - 1] && (x == 0 || y == 0 || z == 0 || x == Lx - 1 || y == Ly - 1 || z == Lz - 1))
            atomicAdd(&c, 1);
    }
    __syncthreads();
    if (threadIdx.x == 0 && c) atomicAdd(out, c);
}

// element modulus of every voxel; escale folds a domain's edge length a into
// the modulus (K_e is linear in a), so one Walsh table serves the batch
__global__ void k_emod(const int* __restrict__ keepE, const double* __restrict__ modulus, double E, double escale,
                       int64_t nvox, double* __restrict__ emod) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    emod[v] = keepE[v] ? (modulus ? modulus[v] : E) * escale : 0.0;
}

// Jacobi: diag(K)_node,d = kdiag * sum of the moduli of the adjacent elements
__global__ void k_dinv(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                       const int* __restrict__ lat2node, const double* __restrict__ emod, double kdiag,
                       double* __restrict__ d64, float* __restrict__ d32) {
    const BrickInfo B = bricks[blockIdx.x];
    const DomInfo D = doms[B.dom];
    const int t = threadIdx.x;
    const int gx = B.x0 + t % BX, gy = B.y0 + (t / BX) % BY, gz = B.z0 + t / (BX * BY);
    const int Lx = D.nx + 1, Ly = D.ny + 1, Lz = D.nz + 1;
    if (gx >= Lx || gy >= Ly || gz >= Lz) return;
    int q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
    if (q < 0) return;
    double s = 0.0;
    for (int dz = -1; dz <= 0; ++dz)
        for (int dy = -1; dy <= 0; ++dy)
            for (int dx = -1; dx <= 0; ++dx) {
                int vx = gx + dx, vy = gy + dy, vz = gz + dz;
                if (vx < 0 || vy < 0 || vz < 0 || vx >= D.nx || vy >= D.ny || vz >= D.nz) continue;
                s += emod[D.vox_off + vx + (int64_t)D.nx * (vy + (int64_t)D.ny * vz)];
            }
    double dg = kdiag * s;
    double inv = dg != 0.0 ? 1.0 / dg : 1.0;   // PCJACOBI uses 1 for a zero diagonal
    if (is_bc(D, gx, gy, gz)) inv = 0.0;        // constrained DOFs stay zero
    for (int d = 0; d < 3; ++d) {              // one value per DOF (isotropic for the HEX08 voxel)
        d64[3 * q + d] = inv;
        d32[3 * q + d] = (float)inv;
    }
}

// ---------------------------------------------------------------- matvec tables
// solid elements of the element tile of each brick (blockDim >= NET)
__global__ void k_brick_ecount(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                               const double* __restrict__ emod, int* __restrict__ cnt) {
    __shared__ int s;
    const BrickInfo B = bricks[blockIdx.x];
    const DomInfo D = doms[B.dom];
    const int i = threadIdx.x;
    if (i == 0) s = 0;
    __syncthreads();
    if (i < NET) {
        const int vx = B.x0 - 1 + i % EX, vy = B.y0 - 1 + (i / EX) % EY, vz = B.z0 - 1 + i / (EX * EY);
        if (vx >= 0 && vy >= 0 && vz >= 0 && vx < D.nx && vy < D.ny && vz < D.nz &&
            emod[D.vox_off + vx + (int64_t)D.nx * (vy + (int64_t)D.ny * vz)] != 0.0)
            atomicAdd(&s, 1);
    }
    __syncthreads();
    if (i == 0) cnt[blockIdx.x] = s;
}

// halo node table (solver node | BC_BIT) and the packed tile element list of
// each brick, elements in lexicographic tile order (blockDim >= NET)
__global__ void k_brick_tables(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                               const int* __restrict__ lat2node, const double* __restrict__ emod,
                               int* __restrict__ halo, unsigned char* __restrict__ nmask, int* __restrict__ einfo,
                               double* __restrict__ eE64, float* __restrict__ eE32) {
    __shared__ int s_flag[NET];
    const BrickInfo B = bricks[blockIdx.x];
    const DomInfo D = doms[B.dom];
    const int tid = threadIdx.x;
    const int Lx = D.nx + 1, Ly = D.ny + 1, Lz = D.nz + 1;
    for (int i = tid; i < NH; i += blockDim.x) {
        const int gx = B.x0 - 1 + i % HX, gy = B.y0 - 1 + (i / HX) % HY, gz = B.z0 - 1 + i / (HX * HY);
        int q = -1;
        if (gx >= 0 && gy >= 0 && gz >= 0 && gx < Lx && gy < Ly && gz < Lz) {
            q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
            if (q >= 0 && is_bc(D, gx, gy, gz)) q |= BC_BIT;
        }
        halo[(int64_t)blockIdx.x * NH + i] = q;
    }
    double e = 0.0;
    const int ex = tid % EX, ey = (tid / EX) % EY, ez = tid / (EX * EY);
    if (tid < NET) {
        const int vx = B.x0 - 1 + ex, vy = B.y0 - 1 + ey, vz = B.z0 - 1 + ez;
        if (vx >= 0 && vy >= 0 && vz >= 0 && vx < D.nx && vy < D.ny && vz < D.nz)
            e = emod[D.vox_off + vx + (int64_t)D.nx * (vy + (int64_t)D.ny * vz)];
        s_flag[tid] = e != 0.0 ? 1 : 0;
    }
    __syncthreads();
    if (tid < NO) {   // solid tile elements around owned node tid, bit b = element at corner offset -b
        const int hx = tid % BX + 1, hy = (tid / BX) % BY + 1, hz = tid / (BX * BY) + 1;
        unsigned m = 0;
        for (int b = 0; b < 8; ++b)
            if (s_flag[(hx - (b & 1)) + EX * ((hy - ((b >> 1) & 1)) + EY * (hz - (b >> 2)))]) m |= 1u << b;
        nmask[(int64_t)blockIdx.x * NO + tid] = (unsigned char)m;
    }
    if (tid < NET && e != 0.0) {
        int rank = 0;
        for (int k = 0; k < tid; ++k) rank += s_flag[k];
        unsigned mask = 0;
        for (int b = 0; b < 8; ++b) {
            const int hx = ex + (b & 1), hy = ey + ((b >> 1) & 1), hz = ez + (b >> 2);
            if (hx >= 1 && hx <= BX && hy >= 1 && hy <= BY && hz >= 1 && hz <= BZ) mask |= 1u << b;
        }
        const int hb = ex + HX * (ey + HY * ez);
        const int ob = (ex - 1) + BX * ((ey - 1) + BY * (ez - 1)) + OB_BIAS;
        const int64_t o = (int64_t)B.e_off + rank;
        einfo[o] = (int)((unsigned)hb | ((unsigned)ob << 9) | (mask << 24));
        eE64[o] = e;
        eE32[o] = (float)e;
    }
}

// ---------------------------------------------------------------- force projection
// ff(d*NEN + l, lc) = sum over owned nodes r_d(node, lc) N_l(xi(node))
// (f-src/mod_struct_calcmat.f90:254-280) with the NEN macro shape functions;
// partial per brick [brick][24 lc][FFD]; bricks of the other macro order are skipped
constexpr int FFD = 60;   // macro DOF slots of the force partials (3 x 20)

template <int NEN>
__global__ __launch_bounds__(NT_UPD) void k_project(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                                                    const int* __restrict__ lat2node, const double* __restrict__ rf,
                                                    double* __restrict__ pff) {
    constexpr int G = NT_UPD / R;    // 8 node groups
    constexpr int FD = 3 * NEN, PASS = 12;
    __shared__ double s[G][R][PASS];
    const BrickInfo B = bricks[blockIdx.x];
    const DomInfo D = doms[B.dom];
    if (macro_nodes(D.macro_order) != NEN) return;
    const int tid = threadIdx.x, lc = tid % R, g = tid / R;
    const int Lx = D.nx + 1, Ly = D.ny + 1;
    double acc[FD];
#pragma unroll
    for (int k = 0; k < FD; ++k) acc[k] = 0.0;
    for (int o = g; o < BX * BY * BZ; o += G) {
        int gx = B.x0 + o % BX, gy = B.y0 + (o / BX) % BY, gz = B.z0 + o / (BX * BY);
        if (gx > D.nx || gy > D.ny || gz > D.nz) continue;
        int q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
        if (q < 0) continue;
        double c[3], N[20];
        node_coord(D, gx, gy, gz, c);
        macro_shape(D.macro_order, c, D.min_c, D.max_c, N);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double rv = rf[((int64_t)q * 3 + d) * R + lc];
#pragma unroll
            for (int l = 0; l < NEN; ++l) acc[d * NEN + l] += rv * N[l];
        }
    }
#pragma unroll
    for (int k0 = 0; k0 < FD; k0 += PASS) {
#pragma unroll
        for (int k = 0; k < PASS; ++k) s[g][lc][k] = k0 + k < FD ? acc[k0 + k] : 0.0;
        __syncthreads();
        for (int i = tid; i < R * PASS; i += NT_UPD) {
            const int l2 = i / PASS, k = i % PASS;
            if (k0 + k >= FD) continue;
            double v = 0.0;
            for (int gg = 0; gg < G; ++gg) v += s[gg][l2][k];
            pff[((int64_t)blockIdx.x * R + l2) * FFD + k0 + k] = v;
        }
        __syncthreads();
    }
}

// ff[d] = sum over the domain's bricks of the force partials: workgroup (d, g) takes
// entries 32 g .. 32 g + 31 of the R x FFD block; its 8 thread rows sum every 8th
// brick, and the 8 row sums are added in row order (fixed order)
constexpr int RFF_E = 32, RFF_P = 8;   // entries per workgroup, brick partitions
__global__ __launch_bounds__(RFF_E * RFF_P) void k_red_ff(const DomInfo* __restrict__ doms,
                                                         const double* __restrict__ pff, double* __restrict__ ff) {
    __shared__ double s_p[RFF_P][RFF_E];
    const int d = blockIdx.x;
    const DomInfo D = doms[d];
    const int FD = 3 * macro_nodes(D.macro_order);
    const int e = threadIdx.x % RFF_E, part = threadIdx.x / RFF_E;
    const int i = (int)blockIdx.y * RFF_E + e;
    double v = 0.0;
    if (i < R * FFD && i % FFD < FD)
        for (int b = D.brick_begin + part; b < D.brick_end; b += RFF_P) v += pff[(int64_t)b * R * FFD + i];
    s_p[part][e] = v;
    __syncthreads();
    if (part == 0 && i < R * FFD) {
        double t = 0.0;
        for (int k = 0; k < RFF_P; ++k) t += s_p[k][e];
        ff[(int64_t)d * R * FFD + i] = t;
    }
}

// u (mesh numbering) = x + x_bc, column-major (3N, ncol): load cases lc_base ..
// lc_base + ncol - 1 of the domain's group
__global__ __launch_bounds__(NT_UPD) void k_fields(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                                                   const int* __restrict__ lat2node, const int* __restrict__ sol2ref,
                                                   const double* __restrict__ x, const double* __restrict__ rf,
                                                   int dom, int64_t nn, int ncol, double* __restrict__ uo,
                                                   double* __restrict__ ro) {
    const BrickInfo B = bricks[blockIdx.x];
    if (B.dom != dom) return;
    const DomInfo D = doms[B.dom];
    const int Lx = D.nx + 1, Ly = D.ny + 1;
    for (int i = threadIdx.x; i < BX * BY * BZ * 3 * R; i += NT_UPD) {
        int o = i / (3 * R), rem = i % (3 * R), d = rem / R, lc = rem % R;
        if (lc >= ncol) continue;
        int gx = B.x0 + o % BX, gy = B.y0 + (o / BX) % BY, gz = B.z0 + o / (BX * BY);
        if (gx > D.nx || gy > D.ny || gz > D.nz) continue;
        int q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
        if (q < 0) continue;
        int64_t ref = sol2ref[q] - 1;
        int64_t out = (3 * ref + d) + 3 * nn * lc;
        double v = x[((int64_t)q * 3 + d) * R + lc];
        if (is_bc(D, gx, gy, gz)) {
            double c[3], N[20];
            node_coord(D, gx, gy, gz, c);
            macro_shape(D.macro_order, c, D.min_c, D.max_c, N);
            v = bc_value_m(D.macro_order, N, d, D.lc_base + lc);
        }
        if (uo) uo[out] = v;
        if (ro) ro[out] = rf[((int64_t)q * 3 + d) * R + lc];
    }
}

// ---------------------------------------------------------------- sensitivity
// h[node][j][d] (+)= sum_lc M[lc][j] u_lc(node, d), u = x + x_bc, over the load cases
// of domain group dom (M: its 24 rows); h is indexed by the node numbers of group
// h_begin's range (the groups of one domain share one mesh and its numbering),
// add = 0 for the first group
__global__ __launch_bounds__(NT_UPD) void k_grad_h(const BrickInfo* __restrict__ bricks, const DomInfo* __restrict__ doms,
                                                   const int* __restrict__ lat2node, const double* __restrict__ x,
                                                   const double* __restrict__ Mj, int dom, int64_t h_begin, int add,
                                                   double* __restrict__ h) {
    const BrickInfo B = bricks[blockIdx.x];
    if (B.dom != dom) return;
    const DomInfo D = doms[B.dom];
    const int Lx = D.nx + 1, Ly = D.ny + 1;
    for (int i = threadIdx.x; i < BX * BY * BZ * 18; i += NT_UPD) {
        int o = i / 18, j = (i % 18) / 3, d = i % 3;
        int gx = B.x0 + o % BX, gy = B.y0 + (o / BX) % BY, gz = B.z0 + o / (BX * BY);
        if (gx > D.nx || gy > D.ny || gz > D.nz) continue;
        int q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
        if (q < 0) continue;
        bool bc = is_bc(D, gx, gy, gz);
        double N[20];
        if (bc) {
            double c[3];
            node_coord(D, gx, gy, gz, c);
            macro_shape(D.macro_order, c, D.min_c, D.max_c, N);
        }
        double s = 0.0;
        for (int lc = 0; lc < R; ++lc) {
            double u = bc ? bc_value_m(D.macro_order, N, d, D.lc_base + lc) : x[((int64_t)q * 3 + d) * R + lc];
            s += Mj[lc * 6 + j] * u;
        }
        double* hp = h + ((q - D.node_begin + h_begin) * 6 + j) * 3 + d;
        *hp = add ? *hp + s : s;
    }
}

__device__ inline double u_at(const DomInfo& D, const int* lat2node, const double* x, int gx, int gy, int gz, int d,
                              int lc) {
    const int Lx = D.nx + 1, Ly = D.ny + 1;
    int q = lat2node[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
    if (q < 0) return 0.0;
    if (is_bc(D, gx, gy, gz)) {
        double c[3], N[20];
        node_coord(D, gx, gy, gz, c);
        macro_shape(D.macro_order, c, D.min_c, D.max_c, N);
        return bc_value_m(D.macro_order, N, d, D.lc_base + lc);
    }
    return x[((int64_t)q * 3 + d) * R + lc];
}

// dL/dE_v = sum_j h_j,e^T Khat u_j,e for every voxel of domain dom
__global__ void k_grad_elem(const DomInfo* __restrict__ doms, const int* __restrict__ lat2node,
                            const double* __restrict__ emod, const double* __restrict__ x,
                            const double* __restrict__ h, WK<double> wk, int dom, double* __restrict__ g) {
    const DomInfo D = doms[dom];
    const int64_t nvox = (int64_t)D.nx * D.ny * D.nz;
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvox) return;
    double out = 0.0;
    if (emod[D.vox_off + v] != 0.0) {
        const int vx = v % D.nx, vy = (v / D.nx) % D.ny, vz = v / ((int64_t)D.nx * D.ny);
        const int Lx = D.nx + 1, Ly = D.ny + 1;
        int qb[8];
#pragma unroll
        for (int b = 0; b < 8; ++b)
            qb[b] = lat2node[D.lat_off + (vx + (b & 1)) + (int64_t)Lx * ((vy + ((b >> 1) & 1)) + (int64_t)Ly * (vz + (b >> 2)))];
        for (int j = 0; j < 6; ++j) {
            double w[8][3], y[8][3];
            for (int d = 0; d < 3; ++d) {
                double t[8];
                for (int b = 0; b < 8; ++b)
                    t[b] = u_at(D, lat2node, x, vx + (b & 1), vy + ((b >> 1) & 1), vz + (b >> 2), d, j);
                fht_fwd(t);
                for (int b = 0; b < 8; ++b) w[b][d] = t[b];
            }
            walsh_apply(wk, w, y);
            for (int d = 0; d < 3; ++d) {
                double t[8];
                for (int b = 0; b < 8; ++b) t[b] = y[b][d];
                fht_inv(t);
                for (int b = 0; b < 8; ++b)
                    if (qb[b] >= 0) out += t[b] * h[((int64_t)qb[b] * 6 + j) * 3 + d];
            }
        }
    }
    g[v] = out;
}

// ---------------------------------------------------------------- multigrid (FP32 preconditioner)
// Level l+1 has voxels of 2x2x2 voxels of level l.  Its voxel modulus is the sum
// of the 8 children / 4: their mean times 2, because the element stiffness is
// linear in the edge length and the Walsh table is the one of the fine edge.
__global__ void k_coarsen(const double* __restrict__ ef, int fx, int fy, int fz, double* __restrict__ ec, int cx,
                          int cy, int cz) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= (int64_t)cx * cy * cz) return;
    const int x = (int)(v % cx), y = (int)((v / cx) % cy), z = (int)(v / ((int64_t)cx * cy));
    double s = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int j = 0; j < 2; ++j)
            for (int i = 0; i < 2; ++i) {
                const int X = 2 * x + i, Y = 2 * y + j, Z = 2 * z + k;
                if (X < fx && Y < fy && Z < fz) s += ef[X + (int64_t)fx * (Y + (int64_t)fy * Z)];
            }
    ec[v] = 0.25 * s;
}

// lattice node flag: 1 if a voxel touching the node carries a modulus
__global__ void k_exist(const double* __restrict__ e, int nx, int ny, int nz, int* __restrict__ ex) {
    const int Lx = nx + 1, Ly = ny + 1, Lz = nz + 1;
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= (int64_t)Lx * Ly * Lz) return;
    const int x = (int)(p % Lx), y = (int)((p / Lx) % Ly), z = (int)(p / ((int64_t)Lx * Ly));
    int f = 0;
    for (int dz = -1; dz <= 0; ++dz)
        for (int dy = -1; dy <= 0; ++dy)
            for (int dx = -1; dx <= 0; ++dx) {
                const int vx = x + dx, vy = y + dy, vz = z + dz;
                if (vx < 0 || vy < 0 || vz < 0 || vx >= nx || vy >= ny || vz >= nz) continue;
                if (e[vx + (int64_t)nx * (vy + (int64_t)ny * vz)] != 0.0) f = 1;
            }
    ex[p] = f;
}

// nodes per brick from a node flag array (blockDim = BX*BY*BZ)
__global__ void k_brick_count_ex(const int* __restrict__ ex, int Lx, int Ly, int Lz, int bnx, int bny,
                                 int* __restrict__ cnt) {
    __shared__ int s;
    const int b = blockIdx.x, t = threadIdx.x;
    if (t == 0) s = 0;
    __syncthreads();
    const int bx = b % bnx, by = (b / bnx) % bny, bz = b / (bnx * bny);
    const int gx = bx * BX + t % BX, gy = by * BY + (t / BX) % BY, gz = bz * BZ + t / (BX * BY);
    if (gx < Lx && gy < Ly && gz < Lz && ex[gx + Lx * (gy + Ly * gz)]) atomicAdd(&s, 1);
    __syncthreads();
    if (t == 0) cnt[b] = s;
}

// solver node numbers of the flagged nodes, lexicographic inside each brick
__global__ void k_brick_assign_ex(const BrickInfo* __restrict__ bl, const int* __restrict__ ex, int Lx, int Ly, int Lz,
                                  int64_t lat_off, int* __restrict__ lat2node) {
    __shared__ int flag[NO];
    const BrickInfo B = bl[blockIdx.x];
    const int t = threadIdx.x;
    const int gx = B.x0 + t % BX, gy = B.y0 + (t / BX) % BY, gz = B.z0 + t / (BX * BY);
    int f = 0;
    if (gx < Lx && gy < Ly && gz < Lz) f = ex[gx + Lx * (gy + Ly * gz)];
    flag[t] = f;
    __syncthreads();
    if (f) {
        int rank = 0;
        for (int k = 0; k < t; ++k) rank += flag[k];
        lat2node[lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)] = B.node_begin + rank;
    }
}

// Transfer kernels: 18 lanes per node, each lane one float4 = 4 load cases of one
// direction of the node's [3][24] segment.
constexpr int T_LPN = 3 * R / 4;           // float4 per node
constexpr int T_NPP = 256 / T_LPN;         // nodes per pass of a 256-thread workgroup

__device__ inline void f4_axpy(float4& s, float w, const float4& a) {
    s.x += w * a.x; s.y += w * a.y; s.z += w * a.z; s.w += w * a.w;
}

// s_pos[k] = position (0..NO-1) of the brick's k-th solver node (owned nodes are
// numbered in position order), so the transfer loops visit occupied positions only
__device__ inline int brick_positions(const BrickInfo& B, const DomInfo& D, const int* __restrict__ l2n,
                                      unsigned char* s_pos) {
    const int t = threadIdx.x;
    if (t < NO) {
        const int gx = B.x0 + t % BX, gy = B.y0 + (t / BX) % BY, gz = B.z0 + t / (BX * BY);
        const int Lx = D.nx + 1, Ly = D.ny + 1;
        if (gx <= D.nx && gy <= D.ny && gz <= D.nz) {
            const int q = l2n[D.lat_off + gx + (int64_t)Lx * (gy + (int64_t)Ly * gz)];
            if (q >= 0) s_pos[q - B.node_begin] = (unsigned char)t;
        }
    }
    __syncthreads();
    return B.node_end - B.node_begin;
}

// b_c = P^T r_f: trilinear weights over the free (interior) fine nodes; one
// workgroup per (coarse brick, blockIdx.y-th share of its nodes)
__global__ __launch_bounds__(256) void k_restrict(const BrickInfo* __restrict__ cbr, const DomInfo* __restrict__ cdom,
                                                  const int* __restrict__ cl2n, const DomInfo* __restrict__ fdom,
                                                  const int* __restrict__ fl2n, const float* __restrict__ rf,
                                                  const int* __restrict__ done, float* __restrict__ bc) {
    // fine lattice nodes of the brick's restriction stencils, staged once: positions
    // 2 x0 - 1 .. 2 (x0 + BX - 1) + 1 (and likewise in y, z); -1 = no node
    constexpr int FX = 2 * BX + 1, FY = 2 * BY + 1, FZ = 2 * BZ + 1;
    __shared__ unsigned char s_pos[NO];
    __shared__ int s_f2n[FX * FY * FZ];
    const BrickInfo B = cbr[blockIdx.x];
    if (done && done[B.dom]) return;
    const int tid = threadIdx.x;
    const DomInfo C = cdom[B.dom], F = fdom[B.dom];
    const int fLx = F.nx + 1, fLy = F.ny + 1;
    const int fx0 = 2 * B.x0 - 1, fy0 = 2 * B.y0 - 1, fz0 = 2 * B.z0 - 1;
    for (int i = tid; i < FX * FY * FZ; i += 256) {
        const int X = fx0 + i % FX, Y = fy0 + (i / FX) % FY, Z = fz0 + i / (FX * FY);
        s_f2n[i] = (X >= 0 && Y >= 0 && Z >= 0 && X <= F.nx && Y <= F.ny && Z <= F.nz)
                       ? fl2n[F.lat_off + X + (int64_t)fLx * (Y + (int64_t)fLy * Z)]
                       : -1;
    }
    const int n = brick_positions(B, C, cl2n, s_pos);   // (its barrier also publishes s_f2n)
    if (tid >= T_NPP * T_LPN) return;
    const int slot = tid / T_LPN, j = tid % T_LPN;
    const float4* rf4 = reinterpret_cast<const float4*>(rf);
    float4* bc4 = reinterpret_cast<float4*>(bc);
    for (int k = slot + (int)blockIdx.y * T_NPP; k < n; k += T_NPP * (int)gridDim.y) {
        const int o = s_pos[k], qc = B.node_begin + k;
        const int gx = B.x0 + o % BX, gy = B.y0 + (o / BX) % BY, gz = B.z0 + o / (BX * BY);
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
        if (!is_bc(C, gx, gy, gz)) {
#pragma unroll
            for (int k = -1; k <= 1; ++k)
#pragma unroll
                for (int jj = -1; jj <= 1; ++jj)
#pragma unroll
                    for (int i = -1; i <= 1; ++i) {
                        const int X = 2 * gx + i, Y = 2 * gy + jj, Z = 2 * gz + k;
                        if (X > 0 && Y > 0 && Z > 0 && X < F.nx && Y < F.ny && Z < F.nz) {   // free fine node
                            const int qf = s_f2n[(X - fx0) + FX * ((Y - fy0) + FY * (Z - fz0))];
                            if (qf >= 0) {
                                constexpr float W[3] = {0.5f, 1.f, 0.5f};
                                const float w = W[i + 1] * W[jj + 1] * W[k + 1];
                                f4_axpy(sum, w, rf4[(int64_t)qf * T_LPN + j]);
                            }
                        }
                    }
        }
        bc4[(int64_t)qc * T_LPN + j] = sum;
    }
}

// x_f += P x_c on the free nodes of each fine brick (trilinear interpolation);
// blockIdx.y selects a share of the brick's nodes
__global__ __launch_bounds__(256) void k_prolong(const BrickInfo* __restrict__ fbr, const DomInfo* __restrict__ fdom,
                                                 const int* __restrict__ fl2n, const DomInfo* __restrict__ cdom,
                                                 const int* __restrict__ cl2n, const float* __restrict__ xc,
                                                 const int* __restrict__ done, float* __restrict__ xf) {
    // coarse lattice nodes the brick's fine nodes interpolate from, staged once:
    // positions x0 / 2 .. (x0 + BX - 1) / 2 + 1 (and likewise in y, z); -1 = no node
    constexpr int CXM = BX / 2 + 2, CYM = BY / 2 + 2, CZM = BZ / 2 + 2;   // room for odd origins
    __shared__ unsigned char s_pos[NO];
    __shared__ int s_c2n[CXM * CYM * CZM];
    const BrickInfo B = fbr[blockIdx.x];
    if (done && done[B.dom]) return;
    const int tid = threadIdx.x;
    const DomInfo F = fdom[B.dom], C = cdom[B.dom];
    const int cLx = C.nx + 1, cLy = C.ny + 1;
    const int cx0 = B.x0 >> 1, cy0 = B.y0 >> 1, cz0 = B.z0 >> 1;
    for (int i = tid; i < CXM * CYM * CZM; i += 256) {
        const int X = cx0 + i % CXM, Y = cy0 + (i / CXM) % CYM, Z = cz0 + i / (CXM * CYM);
        s_c2n[i] = (X <= C.nx && Y <= C.ny && Z <= C.nz)
                       ? cl2n[C.lat_off + X + (int64_t)cLx * (Y + (int64_t)cLy * Z)]
                       : -1;
    }
    const int n = brick_positions(B, F, fl2n, s_pos);   // (its barrier also publishes s_c2n)
    if (tid >= T_NPP * T_LPN) return;
    const int slot = tid / T_LPN, j = tid % T_LPN;
    const float4* xc4 = reinterpret_cast<const float4*>(xc);
    float4* xf4 = reinterpret_cast<float4*>(xf);
    for (int k = slot + (int)blockIdx.y * T_NPP; k < n; k += T_NPP * (int)gridDim.y) {
        const int o = s_pos[k], qf = B.node_begin + k;
        const int gx = B.x0 + o % BX, gy = B.y0 + (o / BX) % BY, gz = B.z0 + o / (BX * BY);
        if (!(gx > 0 && gy > 0 && gz > 0 && gx < F.nx && gy < F.ny && gz < F.nz)) continue;   // free nodes only
        const int X0 = gx >> 1, Y0 = gy >> 1, Z0 = gz >> 1;
        const int ox = gx & 1, oy = gy & 1, oz = gz & 1;
        const float w = (ox ? 0.5f : 1.f) * (oy ? 0.5f : 1.f) * (oz ? 0.5f : 1.f);
        float4 sum = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int jj = 0; jj < 2; ++jj)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    if ((i && !ox) || (jj && !oy) || (k && !oz)) continue;
                    const int qc = s_c2n[(X0 + i - cx0) + CXM * ((Y0 + jj - cy0) + CYM * (Z0 + k - cz0))];
                    if (qc >= 0) f4_axpy(sum, 1.f, xc4[(int64_t)qc * T_LPN + j]);
                }
        float4 v = xf4[(int64_t)qf * T_LPN + j];
        f4_axpy(v, w, sum);
        xf4[(int64_t)qf * T_LPN + j] = v;
    }
}

// one Chebyshev-Jacobi step on the owned nodes of each brick:
//   zero:      d = c2 D^-1 b,              x = d        (first step from x = 0; t unused)
//   otherwise: d = c1 d + c2 D^-1 (b - t), x = x + d    (t = A x)
// d == null: the new direction is not stored (zero start: it equals x).
// prz (optional): per-brick sums of b . x per load case
__global__ __launch_bounds__(NT_UPD) void k_cheb(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                                 const float* __restrict__ b, const float* __restrict__ t,
                                                 const float* __restrict__ dinv, float* __restrict__ d,
                                                 float* __restrict__ x, float c1, float c2, int zero,
                                                 double* __restrict__ prz) {
    __shared__ double s1[NT_UPD];
    const BrickInfo B = bricks[blockIdx.x];
    if (done && done[B.dom]) return;
    const int tid = threadIdx.x;
    double a1 = 0.0;
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = tid; i < cnt; i += NT_UPD) {
        const int64_t idx = base + i;
        const float bv = b[idx];
        const float di = dinv[idx / R];
        float dv, xv;
        if (zero) {
            dv = c2 * di * bv;
            xv = dv;
        } else {
            dv = cheb_dir(c1, d[idx], c2, di, bv, t[idx]);
            xv = x[idx] + dv;
        }
        if (d) d[idx] = dv;
        x[idx] = xv;
        a1 += (double)bv * (double)xv;
    }
    if (prz) {
        s1[tid] = a1;
        __syncthreads();
        if (tid < R) {
            double u = 0.0;
            for (int k = tid; k < NT_UPD; k += R) u += s1[k];
            prz[(int64_t)blockIdx.x * R + tid] = u;
        }
    }
}

// ---------------------------------------------------------------- p-multigrid transfers (HEX20 -> HEX08)
// The HEX20 nodes of a voxel mesh sit on the half-voxel lattice: corners at even
// coordinates (the HEX08 nodes of the same voxels), mid-edge nodes with one odd
// coordinate.  P (HEX08 -> HEX20) is the trilinear interpolation of the HEX08
// element: identity at corners, the mean of the edge's two ends at a mid-edge
// node.  Restriction is P^T in gather form, in a fixed order per coarse node.

// r08 = P^T res for one domain group: coarse lattice node (i, j, k) takes the fine
// node at (2i, 2j, 2k) and half of each mid-edge node at (2i +- 1, 2j, 2k),
// (2i, 2j +- 1, 2k), (2i, 2j, 2k +- 1); 0 on box-face (Dirichlet) nodes.
__global__ __launch_bounds__(256) void k_pmg_restrict(Dom20Dev G, DomInfo D08, const int* __restrict__ lat2node08,
                                                     const int* __restrict__ done, int dom,
                                                     const float* __restrict__ res, float* __restrict__ r08) {
    if (done && done[dom]) return;
    const int Lx = D08.nx + 1, Ly = D08.ny + 1, Lz = D08.nz + 1;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)Lx * Ly * Lz * R) return;
    const int lc = (int)(i % R);
    const int64_t p = i / R;
    const int cx = (int)(p % Lx), cy = (int)((p / Lx) % Ly), cz = (int)(p / ((int64_t)Lx * Ly));
    const int c = lat2node08[D08.lat_off + p];
    if (c < 0) return;
    const int fx = 2 * cx, fy = 2 * cy, fz = 2 * cz;
    auto fine = [&](int x, int y, int z) -> int {
        if (x < 0 || y < 0 || z < 0 || x >= G.Lx || y >= G.Ly || z >= G.Lz) return -1;
        return G.lat20[x + (int64_t)G.Lx * (y + (int64_t)G.Ly * z)];
    };
    const int q0 = fine(fx, fy, fz);
    const bool bc = q0 >= 0 && G.bcf[q0];
    const int nb6[6] = {fine(fx - 1, fy, fz), fine(fx + 1, fy, fz), fine(fx, fy - 1, fz),
                        fine(fx, fy + 1, fz), fine(fx, fy, fz - 1), fine(fx, fy, fz + 1)};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        float v = 0.f;
        if (!bc) {
            if (q0 >= 0) v = res[((G.node_begin + q0) * 3 + d) * R + lc];
#pragma unroll
            for (int k = 0; k < 6; ++k)
                if (nb6[k] >= 0) v += 0.5f * res[((G.node_begin + nb6[k]) * 3 + d) * R + lc];
        }
        r08[((int64_t)c * 3 + d) * R + lc] = v;
    }
}

// z += P z08 on the free HEX20 nodes of the pseudo-bricks
__global__ __launch_bounds__(NT_UPD) void k_pmg_prolong(const BrickInfo* __restrict__ bricks,
                                                       const DomInfo* __restrict__ doms,
                                                       const Dom20Dev* __restrict__ d20,
                                                       const DomInfo* __restrict__ doms08,
                                                       const int* __restrict__ lat2node08,
                                                       const int* __restrict__ done, const float* __restrict__ z08,
                                                       float* __restrict__ z) {
    const BrickInfo B = bricks[blockIdx.x];
    if (done && done[B.dom]) return;
    const DomInfo D = doms[B.dom];
    const DomInfo C = doms08[B.dom];
    const Dom20Dev G = d20[B.dom];
    const int Lx = C.nx + 1, Ly = C.ny + 1;
    for (int i = threadIdx.x; i < (B.node_end - B.node_begin) * R; i += NT_UPD) {
        const int64_t q = B.node_begin + i / R;
        const int lc = i % R;
        const int ql = (int)(q - G.node_begin);
        if (G.bcf[ql]) continue;
        int l[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) l[k] = (int)lrint(2.0 * (G.coords[3 * ql + k] / D.delta[k] - D.xa1[k]));
        int odd = -1;
#pragma unroll
        for (int k = 0; k < 3; ++k)
            if (l[k] & 1) odd = k;
        int lo[3] = {l[0] >> 1, l[1] >> 1, l[2] >> 1};
        const int c0 = lat2node08[C.lat_off + lo[0] + (int64_t)Lx * (lo[1] + (int64_t)Ly * lo[2])];
        int c1 = -1;
        if (odd >= 0) {
            lo[odd] += 1;
            c1 = lat2node08[C.lat_off + lo[0] + (int64_t)Lx * (lo[1] + (int64_t)Ly * lo[2])];
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            float v = 0.f;
            if (odd < 0) {
                if (c0 >= 0) v = z08[((int64_t)c0 * 3 + d) * R + lc];
            } else {
                if (c0 >= 0) v += 0.5f * z08[((int64_t)c0 * 3 + d) * R + lc];
                if (c1 >= 0) v += 0.5f * z08[((int64_t)c1 * 3 + d) * R + lc];
            }
            z[(q * 3 + d) * R + lc] += v;
        }
    }
}

// res = b - t on the rows of the pseudo-bricks (t = K x, masked on Dirichlet rows)
__global__ __launch_bounds__(NT_UPD) void k_sub(const BrickInfo* __restrict__ bricks, const int* __restrict__ done,
                                               const float* __restrict__ b, const float* __restrict__ t,
                                               float* __restrict__ res) {
    const BrickInfo B = bricks[blockIdx.x];
    if (done && done[B.dom]) return;
    const int64_t base = (int64_t)B.node_begin * 3 * R, cnt = (int64_t)(B.node_end - B.node_begin) * 3 * R;
    for (int64_t i = threadIdx.x; i < cnt; i += NT_UPD) res[base + i] = b[base + i] - t[base + i];
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

// Grids of fewer than FILL_BLOCKS workgroups leave CUs idle: the matvec then
// splits each brick's load-case chunks over blockIdx.y (a divisor of nch), the
// transfer kernels split each brick's nodes.
constexpr int FILL_BLOCKS = 2048;
inline int chunk_split(int n_bricks, int nch) {
    int want = n_bricks > 0 ? (FILL_BLOCKS + n_bricks - 1) / n_bricks : 1;
    if (want >= nch) return nch;
    int best = 1;
    for (int s = 1; s <= want; ++s)
        if (nch % s == 0) best = s;
    return best;
}
// Launch-shape autotuner.  The split of a brick's load-case chunks over
// gridDim.y workgroups (1, 2, 3 or 6) changes how the chip is filled, not the
// arithmetic (every chunk is computed whole by one workgroup), so it can be
// picked by timing without changing any result.  autotune_splits() times the
// candidates of each FP32 matvec kernel on each multigrid level of a batch
// before its first solve and caches the fastest per (kernel, variant, bricks);
// the launches look the choice up (chunk_split's rule when absent, or with
// DTC_TUNE=0).  DTC_TUNE_LOG=1 prints each decision.
enum TuneKind { TK_MV = 0, TK_CHEB = 1, TK_BRES = 2 };
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, int> g_tune;

int tuned_split(int kind, int variant, int n_bricks, int nch) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune.find(std::make_tuple(kind, variant, n_bricks));
    return it != g_tune.end() ? it->second : chunk_split(n_bricks, nch);
}

inline int node_split(int n_bricks) {
    const int want = n_bricks > 0 ? (FILL_BLOCKS + n_bricks - 1) / n_bricks : 1;
    return want < 1 ? 1 : want > 8 ? 8 : want;
}

}  // namespace

// ---------------------------------------------------------------- host side
int element_stiffness(int nen, double E, double nu, double a, double* K) {
    if (nen != 8) return set_error(DTC_ERR_UNSUPPORTED, "element stiffness: only HEX08 is implemented here");
    if (!(a > 0.0) || !(nu > -1.0 && nu < 0.5)) return set_error(DTC_ERR_BAD_ARGUMENT, "element stiffness: a=%g nu=%g", a, nu);
    const int nd = 24;
    double D[6][6] = {};
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D[i][j] = lam;
    for (int i = 0; i < 3; ++i) D[i][i] += 2.0 * mu;
    D[3][3] = D[4][4] = D[5][5] = mu;
    std::memset(K, 0, sizeof(double) * nd * nd);
    const double g = 1.0 / std::sqrt(3.0), jac = 2.0 / a, detJ = (a / 2.0) * (a / 2.0) * (a / 2.0);
    for (int qi = 0; qi < 8; ++qi) {
        const double q[3] = {(qi & 1) ? g : -g, (qi & 2) ? g : -g, (qi & 4) ? g : -g};
        double B[6][24] = {};
        for (int l = 0; l < 8; ++l) {
            const double sx = 2.0 * h08_off(l, 0) - 1, sy = 2.0 * h08_off(l, 1) - 1, sz = 2.0 * h08_off(l, 2) - 1;
            const double dx = sx * (1 + sy * q[1]) * (1 + sz * q[2]) / 8.0 * jac;
            const double dy = (1 + sx * q[0]) * sy * (1 + sz * q[2]) / 8.0 * jac;
            const double dz = (1 + sx * q[0]) * (1 + sy * q[1]) * sz / 8.0 * jac;
            B[0][3 * l] = dx; B[1][3 * l + 1] = dy; B[2][3 * l + 2] = dz;
            B[3][3 * l] = dy; B[3][3 * l + 1] = dx;
            B[4][3 * l] = dz; B[4][3 * l + 2] = dx;
            B[5][3 * l + 1] = dz; B[5][3 * l + 2] = dy;
        }
        for (int r = 0; r < nd; ++r)
            for (int c = 0; c < nd; ++c) {
                double s = 0.0;
                for (int m = 0; m < 6; ++m)
                    for (int p = 0; p < 6; ++p) s += B[m][r] * D[m][p] * B[p][c];
                K[r + nd * c] += detJ * s;
            }
    }
    return DTC_OK;
}

int walsh_blocks(const double* Khat, WalshK* out) {
    // node l <-> bit index b = ox + 2 oy + 4 oz; H[m][b] = prod over set bits i of m of (2 bit_i(b) - 1)
    int l_of_b[8];
    for (int l = 0; l < 8; ++l) l_of_b[h08_off(l, 0) + 2 * h08_off(l, 1) + 4 * h08_off(l, 2)] = l;
    auto H = [](int m, int b) {
        double s = 1.0;
        for (int i = 0; i < 3; ++i)
            if (m & (1 << i)) s *= ((b >> i) & 1) ? 1.0 : -1.0;
        return s;
    };
    // Kw[(m,d),(m',d')] = sum H[m][b] Khat[(l(b),d),(l(b'),d')] H[m'][b'] / 64
    double Kw[24][24];
    double kmax = 0.0;
    for (int m = 0; m < 8; ++m)
        for (int d = 0; d < 3; ++d)
            for (int m2 = 0; m2 < 8; ++m2)
                for (int d2 = 0; d2 < 3; ++d2) {
                    double s = 0.0;
                    for (int b = 0; b < 8; ++b)
                        for (int b2 = 0; b2 < 8; ++b2)
                            s += H(m, b) * Khat[(3 * l_of_b[b] + d) + 24 * (3 * l_of_b[b2] + d2)] * H(m2, b2);
                    Kw[3 * m + d][3 * m2 + d2] = s / 64.0;
                    kmax = std::fmax(kmax, std::fabs(s / 64.0));
                }
    // groups in the order walsh_apply reads them
    const int groups[11][3][2] = {
        {{1, 0}, {2, 1}, {4, 2}}, {{1, 1}, {2, 0}, {-1, 0}}, {{1, 2}, {4, 0}, {-1, 0}}, {{2, 2}, {4, 1}, {-1, 0}},
        {{3, 0}, {6, 2}, {-1, 0}}, {{3, 1}, {5, 2}, {-1, 0}}, {{3, 2}, {5, 1}, {6, 0}}, {{5, 0}, {6, 1}, {-1, 0}},
        {{7, 0}, {-1, 0}, {-1, 0}}, {{7, 1}, {-1, 0}, {-1, 0}}, {{7, 2}, {-1, 0}, {-1, 0}}};
    int grp_of[24];
    for (int i = 0; i < 24; ++i) grp_of[i] = -1;
    int o = 0;
    for (int g = 0; g < 11; ++g) {
        int n = 0;
        while (n < 3 && groups[g][n][0] >= 0) ++n;
        for (int i = 0; i < n; ++i) grp_of[3 * groups[g][i][0] + groups[g][i][1]] = g;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j)
                out->g[o++] = Kw[3 * groups[g][i][0] + groups[g][i][1]][3 * groups[g][j][0] + groups[g][j][1]];
    }
    if (o != 45) return set_error(DTC_ERR_DEVICE, "walsh_blocks: %d coefficients", o);
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j)
            if ((grp_of[i] < 0 || grp_of[i] != grp_of[j]) && std::fabs(Kw[i][j]) > 1e-12 * kmax)
                return set_error(DTC_ERR_UNSUPPORTED, "element stiffness is not block diagonal in the Walsh basis");
    return DTC_OK;
}

int batch_begin(SolverBatch& b, const dtc_domain_desc* d, int64_t n_dom) {
    int64_t nlat = 0, nvox = 0;
    for (int64_t i = 0; i < n_dom; ++i) {
        nlat += (d[i].n[0] + 1) * (d[i].n[1] + 1) * (d[i].n[2] + 1);
        nvox += d[i].n[0] * d[i].n[1] * d[i].n[2];
    }
    DTC_TRY(b.lat2node.ensure((size_t)nlat * 4));
    DTC_TRY(b.sol2ref.ensure((size_t)nlat * 4));
    DTC_TRY(b.emod.ensure((size_t)nvox * 8));
    k_fill_i32b<<<nb(nlat, 256), 256, 0, b.stream>>>(b.lat2node.as<int>(), nlat, -1);
    k_fill_f64<<<nb(nvox, 256), 256, 0, b.stream>>>(b.emod.as<double>(), nvox, 0.0);
    DTC_HIP(hipGetLastError());
    b.n_dom = 0; b.n_nodes = 0; b.n_lat = 0; b.n_vox = 0;
    b.h_dom.clear(); b.h_brick.clear(); b.h_elems.clear(); b.h_nodes.clear(); b.h_bnodes.clear(); b.h_status.clear();
    b.d20.clear();
    b.h_contrast.clear();
    b.n_slices = 0;   // the block PCG slices follow the batch's domains (block_setup)
    b.block_graph.reset();   // captured with the previous composition's buffers
    return DTC_OK;
}

// Empty or failed domains keep a DomInfo with no bricks so indices stay aligned.
int batch_add_empty(SolverBatch& b, const dtc_domain_desc& d, int status, int lc_base) {
    DomInfo D{};
    D.nx = (int)d.n[0]; D.ny = (int)d.n[1]; D.nz = (int)d.n[2];
    D.macro_order = d.macro_order;
    D.lc_base = lc_base;
    D.brick_begin = D.brick_end = (int)b.h_brick.size();
    D.node_begin = D.node_end = b.n_nodes;
    D.lat_off = b.n_lat; D.vox_off = b.n_vox;
    b.h_dom.push_back(D);
    b.n_lat += (d.n[0] + 1) * (d.n[1] + 1) * (d.n[2] + 1);
    b.n_vox += d.n[0] * d.n[1] * d.n[2];
    b.n_dom += 1;
    b.h_elems.push_back(0);
    b.h_nodes.push_back(0);
    b.h_bnodes.push_back(0);
    b.h_status.push_back(status);
    if (b.micro20) b.d20.emplace_back();
    return DTC_OK;
}

int batch_add_domain(SolverBatch& b, const MeshDev& m, const dtc_domain_desc& d, const double* modulus_dev,
                     double escale, int lc_base, hipStream_t s) {
    if (m.nen != 8) return set_error(DTC_ERR_UNSUPPORTED, "the solver implements HEX08 micro elements");
    DomInfo D{};
    D.nx = m.nx; D.ny = m.ny; D.nz = m.nz;
    D.macro_order = d.macro_order;
    D.lc_base = lc_base;
    const int Lx = m.Lx, Ly = m.Ly, Lz = m.Lz;
    D.bnx = (Lx + BX - 1) / BX; D.bny = (Ly + BY - 1) / BY; D.bnz = (Lz + BZ - 1) / BZ;
    const int nbr = D.bnx * D.bny * D.bnz;
    for (int k = 0; k < 3; ++k) {
        D.delta[k] = d.delta[k];
        D.xa1[k] = (double)(d.xa_n[k] - 1);
        D.min_c[k] = (double)(d.xa_n[k] - 1) * d.delta[k];
        D.max_c[k] = (double)(d.xa_n[k] + d.n[k] - 1) * d.delta[k];
    }
    {   // boundary nodes: fewer than 6 constrained DOFs fails the domain (f-src/mod_struct_preprocess.f90:471-473)
        DevMem nbc;
        DTC_TRY(nbc.alloc(4));
        int zero = 0, hb = 0;
        DTC_HIP(hipMemcpyAsync(nbc.p, &zero, 4, hipMemcpyHostToDevice, s));
        const int64_t nl = (int64_t)Lx * Ly * Lz;
        k_count_bc<<<nb(nl, 256), 256, 0, s>>>(m.numLat.as<int>(), m.keepN.as<int>(), Lx, Ly, Lz, nbc.as<int>());
        DTC_HIP(hipGetLastError());
        DTC_HIP(hipMemcpyAsync(&hb, nbc.p, 4, hipMemcpyDeviceToHost, s));
        DTC_HIP(hipStreamSynchronize(s));
        if (3 * hb < 6) {
            DTC_TRY(batch_add_empty(b, d, DTC_ERR_FEW_CONSTRAINTS, lc_base));
            b.h_elems.back() = m.n_elems;
            b.h_nodes.back() = m.n_nodes;
            b.h_bnodes.back() = hb;
            return DTC_OK;
        }
        b.h_bnodes.push_back(hb);
    }
    DevMem cnt;
    DTC_TRY(cnt.alloc((size_t)nbr * 4));
    k_brick_count<<<nbr, BX * BY * BZ, 0, s>>>(m.numLat.as<int>(), m.keepN.as<int>(), Lx, Ly, Lz, D.bnx, D.bny, cnt.as<int>());
    DTC_HIP(hipGetLastError());
    std::vector<int> hc(nbr);
    DTC_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)nbr * 4, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    D.brick_begin = (int)b.h_brick.size();
    D.node_begin = b.n_nodes;
    int64_t node = b.n_nodes;
    std::vector<BrickInfo> mine;
    for (int i = 0; i < nbr; ++i) {
        if (!hc[i]) continue;
        BrickInfo B{};
        B.dom = b.n_dom;
        B.x0 = (i % D.bnx) * BX; B.y0 = ((i / D.bnx) % D.bny) * BY; B.z0 = (i / (D.bnx * D.bny)) * BZ;
        B.node_begin = (int)node;
        node += hc[i];
        B.node_end = (int)node;
        mine.push_back(B);
    }
    if (node >= INT32_MAX / 72) return set_error(DTC_ERR_BAD_ARGUMENT, "batch exceeds %d solver nodes", INT32_MAX / 72);
    D.brick_end = D.brick_begin + (int)mine.size();
    D.node_end = node;
    D.lat_off = b.n_lat;
    D.vox_off = b.n_vox;
    const int64_t nvox = (int64_t)m.nx * m.ny * m.nz;
    if (!mine.empty()) {
        DevMem bl;
        DTC_TRY(bl.alloc(mine.size() * sizeof(BrickInfo)));
        DTC_HIP(hipMemcpyAsync(bl.p, mine.data(), mine.size() * sizeof(BrickInfo), hipMemcpyHostToDevice, s));
        k_brick_assign<<<(unsigned)mine.size(), BX * BY * BZ, 0, s>>>(bl.as<BrickInfo>(), 0, m.numLat.as<int>(), m.keepN.as<int>(),
                                                                   m.cref.as<int>(), Lx, Ly, Lz, D.lat_off, b.lat2node.as<int>(),
                                                                   b.sol2ref.as<int>());
        DTC_HIP(hipGetLastError());
        DTC_HIP(hipStreamSynchronize(s));   // bl goes out of scope
    }
    k_emod<<<nb(nvox, 256), 256, 0, s>>>(m.keepE.as<int>(), modulus_dev, d.E, escale, nvox, b.emod.as<double>() + D.vox_off);
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipStreamSynchronize(s));
    b.h_brick.insert(b.h_brick.end(), mine.begin(), mine.end());
    b.h_dom.push_back(D);
    b.n_nodes = node;
    b.n_lat += (int64_t)Lx * Ly * Lz;
    b.n_vox += nvox;
    b.n_dom += 1;
    b.h_elems.push_back(m.n_elems);
    b.h_nodes.push_back(m.n_nodes);
    b.h_status.push_back(DTC_OK);
    return DTC_OK;
}

namespace {

// largest eigenvalue of a symmetric positive semidefinite n x n matrix (power iteration)
double sym_lmax(const double* A, int n) {
    std::vector<double> v(n, 1.0), w(n);
    for (int i = 0; i < n; ++i) v[i] += 0.01 * i;
    double lam = 0.0;
    for (int it = 0; it < 2000; ++it) {
        double nrm = 0.0;
        for (int i = 0; i < n; ++i) {
            double t = 0.0;
            for (int j = 0; j < n; ++j) t += A[i + n * j] * v[j];
            w[i] = t;
            nrm += t * t;
        }
        nrm = std::sqrt(nrm);
        double num = 0.0, den = 0.0;
        for (int i = 0; i < n; ++i) { num += v[i] * w[i]; den += v[i] * v[i]; }
        lam = num / den;
        for (int i = 0; i < n; ++i) v[i] = w[i] / nrm;
    }
    return lam;
}

// matvec tables of a brick list: tile element counts -> offsets -> halo table,
// owned-node element masks and packed element lists (bricks re-uploaded with offsets)
// (owned node, element corner) pairs of each brick: the population count of its
// owned-node element masks
__global__ __launch_bounds__(256) void k_brick_slots(const unsigned char* __restrict__ nmask, int nbr,
                                                    int* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbr) return;
    const unsigned* w = reinterpret_cast<const unsigned*>(nmask + (int64_t)b * NO);
    int c = 0;
#pragma unroll 8
    for (int k = 0; k < NO / 4; ++k) c += __popc(w[k]);
    out[b] = c;
}

// bricks of a level for the FP32 matvecs: those with at most MV_SLOTS corner
// slots first (the sparse kernel), then the rest (the dense kernel)
int brick_lists(hipStream_t s, int nbr, const DevMem& nmask, DevMem& blist, int* n_sparse) {
    // DTC_SPARSE=1: bricks with few corner slots run the 256-thread variant (measured
    // 10% slower over the whole mixed solve on trabecular batches, so not the default);
    // otherwise no list is kept (brick = workgroup)
    *n_sparse = 0;
    const char* se = std::getenv("DTC_SPARSE");
    if (nbr == 0 || !(se && se[0] == '1')) {
        blist.release();
        return DTC_OK;
    }
    DevMem cnt;
    DTC_TRY(cnt.alloc((size_t)nbr * 4));
    k_brick_slots<<<(nbr + 255) / 256, 256, 0, s>>>(nmask.as<unsigned char>(), nbr, cnt.as<int>());
    DTC_HIP(hipGetLastError());
    std::vector<int> hc(nbr), order;
    DTC_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)nbr * 4, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    order.reserve(nbr);
    for (int i = 0; i < nbr; ++i)
        if (hc[i] <= MV_SLOTS) order.push_back(i);
    *n_sparse = (int)order.size();
    for (int i = 0; i < nbr; ++i)
        if (hc[i] > MV_SLOTS) order.push_back(i);
    DTC_TRY(blist.ensure((size_t)nbr * 4));
    DTC_HIP(hipMemcpyAsync(blist.p, order.data(), (size_t)nbr * 4, hipMemcpyHostToDevice, s));
    DTC_HIP(hipStreamSynchronize(s));
    return DTC_OK;
}

int build_tables(hipStream_t s, std::vector<BrickInfo>& hb, DevMem& brick, const DevMem& dom, const DevMem& lat2node,
                 const DevMem& emod, DevMem& halo, DevMem& nmask, DevMem& einfo, DevMem& eE64, DevMem& eE32,
                 int64_t* n_tile) {
    const int nbr = (int)hb.size();
    *n_tile = 0;
    if (nbr == 0) return DTC_OK;
    DevMem cnt;
    DTC_TRY(cnt.alloc((size_t)nbr * 4));
    k_brick_ecount<<<nbr, 256, 0, s>>>(brick.as<BrickInfo>(), dom.as<DomInfo>(), emod.as<double>(), cnt.as<int>());
    DTC_HIP(hipGetLastError());
    std::vector<int> hc(nbr);
    DTC_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)nbr * 4, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    int64_t off = 0;
    for (int i = 0; i < nbr; ++i) {
        hb[i].e_off = (int)off;
        hb[i].e_cnt = hc[i];
        off += hc[i];
    }
    if (off >= INT32_MAX) return set_error(DTC_ERR_BAD_ARGUMENT, "batch exceeds %d tile elements", INT32_MAX);
    *n_tile = off;
    DTC_HIP(hipMemcpyAsync(brick.p, hb.data(), hb.size() * sizeof(BrickInfo), hipMemcpyHostToDevice, s));
    const size_t ne = (size_t)(off > 0 ? off : 1);
    DTC_TRY(halo.ensure((size_t)nbr * NH * 4));
    DTC_TRY(nmask.ensure((size_t)nbr * NO));
    DTC_TRY(einfo.ensure(ne * 4));
    DTC_TRY(eE64.ensure(ne * 8));
    DTC_TRY(eE32.ensure(ne * 4));
    k_brick_tables<<<nbr, 256, 0, s>>>(brick.as<BrickInfo>(), dom.as<DomInfo>(), lat2node.as<int>(), emod.as<double>(),
                                       halo.as<int>(), nmask.as<unsigned char>(), einfo.as<int>(), eE64.as<double>(),
                                       eE32.as<float>());
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipStreamSynchronize(s));   // cnt goes out of scope
    return DTC_OK;
}

}  // namespace

// B operands of k_matvec_mf: B[4 b' + d'][4 b + d] = Khat(3 b + d, 3 b' + d') / kmax in
// corner-bit order (b = ox + 2 oy + 4 oz), direction 3 a zero pad, as FP16 pairs
// hi + lo; lane l of column tile t holds column j = l % 16 (n = 16 t + j) and
// rows 8 (l / 16) .. 8 (l / 16) + 7.  Also which FP32 matvecs use it: none unless
// DTC_MFMA=1 (then Q = K P with 3 FP16 terms and the smoothing steps with
// DTC_MG_TERMS = 1 or 3, default 3).
int mf_table(const double* Khat, SolverBatch& b) {
    int l_of_b[8];
    for (int l = 0; l < 8; ++l) l_of_b[h08_off(l, 0) + 2 * h08_off(l, 1) + 4 * h08_off(l, 2)] = l;
    double kmax = 0.0;
    for (int i = 0; i < 576; ++i) kmax = std::fmax(kmax, std::fabs(Khat[i]));
    if (!(kmax > 0.0)) return set_error(DTC_ERR_DEVICE, "element stiffness is zero");
    std::vector<_Float16> tab((size_t)2 * 2 * 64 * 8);
    for (int t = 0; t < 2; ++t)
        for (int q = 0; q < 2; ++q)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int k = 8 * (l / 16) + e, n = 16 * t + (l % 16);
                    const int bo = n / 4, dd = n % 4, bi = k / 4, di = k % 4;
                    const double v = (dd < 3 && di < 3)
                                         ? Khat[(3 * l_of_b[bo] + dd) + 24 * (3 * l_of_b[bi] + di)] / kmax
                                         : 0.0;
                    const _Float16 hi = (_Float16)v;
                    tab[(((size_t)t * 2 + q) * 64 + l) * 8 + e] = q == 0 ? hi : (_Float16)(v - (double)hi);
                }
    DTC_TRY(b.kmf.ensure(tab.size() * sizeof(_Float16)));
    DTC_HIP(hipMemcpyAsync(b.kmf.p, tab.data(), tab.size() * sizeof(_Float16), hipMemcpyHostToDevice, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    b.kmf_max = (float)kmax;
    const char* e = std::getenv("DTC_MFMA");
    const bool on = e && e[0] == '1';   // opt-in: slower than the Walsh kernel on the bench batch (measured)
    b.mf_cg = on ? 1 : 0;
    int terms = 3;
    if (const char* t = std::getenv("DTC_MG_TERMS")) terms = std::atoi(t) == 1 ? 1 : 3;
    b.mf_smooth = on ? terms : 0;
    return DTC_OK;
}

int batch_finalize(SolverBatch& b, int precision) {
    b.n_bricks = (int)b.h_brick.size();
    hipStream_t s = b.stream;
    DTC_TRY(b.dom.ensure(b.h_dom.size() * sizeof(DomInfo)));
    DTC_HIP(hipMemcpyAsync(b.dom.p, b.h_dom.data(), b.h_dom.size() * sizeof(DomInfo), hipMemcpyHostToDevice, s));
    DTC_TRY(b.brick.ensure((b.h_brick.size() + 1) * sizeof(BrickInfo)));
    if (!b.h_brick.empty())
        DTC_HIP(hipMemcpyAsync(b.brick.p, b.h_brick.data(), b.h_brick.size() * sizeof(BrickInfo), hipMemcpyHostToDevice, s));
    const int64_t nn = b.n_nodes > 0 ? b.n_nodes : 1;
    const size_t v64 = (size_t)nn * 3 * R * 8, v32 = (size_t)nn * 3 * R * 4;
    DTC_TRY(b.dinv64.ensure((size_t)nn * 3 * 8));
    DTC_TRY(b.dinv32.ensure((size_t)nn * 3 * 4));
    DTC_TRY(b.x64.ensure(v64));
    DTC_TRY(b.r64.ensure(v64));
    if (precision == DTC_FP64) {
        DTC_TRY(b.p64.ensure(v64));
        DTC_TRY(b.q64.ensure(v64));
    } else {
        DTC_TRY(b.x32.ensure(v32));
        DTC_TRY(b.r32.ensure(v32));
        DTC_TRY(b.p32.ensure(v32));
        DTC_TRY(b.q32.ensure(v32));
    }
    const size_t nbk = (size_t)(b.n_bricks > 0 ? b.n_bricks : 1);
    DTC_TRY(b.part_a.ensure(nbk * R * 8));
    DTC_TRY(b.part_b.ensure(nbk * R * 8));
    DTC_TRY(b.part_c.ensure(nbk * R * 8));
    DTC_TRY(b.part_ff.ensure(nbk * R * FFD * 8));
    const size_t nd = (size_t)(b.n_dom > 0 ? b.n_dom : 1);
    for (DevMem* m : {&b.alpha, &b.beta, &b.rz, &b.rr, &b.rr0, &b.thr2, &b.scale}) DTC_TRY(m->alloc(nd * R * 8));
    DTC_TRY(b.active.ensure(nd * R * 4));
    DTC_TRY(b.active64.ensure(nd * R * 4));
    DTC_TRY(b.active_prev.ensure(nd * R * 4));
    DTC_TRY(b.done.ensure(nd * 4));
    DTC_TRY(b.done64.ensure(nd * 4));
    DTC_TRY(b.iters.ensure(nd * 8));
    DTC_TRY(b.ff.ensure(nd * R * FFD * 8));
    // element stiffness (E = 1) and its Walsh blocks
    double Khat[576];
    DTC_TRY(element_stiffness(8, 1.0, b.nu, b.a, Khat));
    DTC_TRY(walsh_blocks(Khat, &b.wk64));
    b.kdiag = Khat[0];
    DTC_TRY(mf_table(Khat, b));
    // element-by-element bound lambda_max(D^-1 K) <= lambda_max(diag(Khat)^-1/2 Khat diag(Khat)^-1/2);
    // the diagonal of the cube element is constant
    b.cheb_lmax = sym_lmax(Khat, 24) / b.kdiag;
    if (b.micro20) {
        DTC_TRY(finalize20(b));
    } else if (b.n_bricks > 0) {
        k_dinv<<<b.n_bricks, BX * BY * BZ, 0, s>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(), b.lat2node.as<int>(),
                                                   b.emod.as<double>(), b.kdiag, b.dinv64.as<double>(), b.dinv32.as<float>());
        DTC_HIP(hipGetLastError());
        DTC_TRY(build_tables(s, b.h_brick, b.brick, b.dom, b.lat2node, b.emod, b.halo, b.nmask, b.einfo, b.eE64,
                             b.eE32, &b.n_tile_elems));
        DTC_TRY(brick_lists(s, b.n_bricks, b.nmask, b.blist, &b.n_sparse));
    }
    DTC_HIP(hipStreamSynchronize(s));
    b.precision = precision;
    return DTC_OK;
}

int batch_build_mg(SolverBatch& b, int max_levels) {
    if (b.micro20) return set_error(DTC_ERR_UNSUPPORTED, "the multigrid preconditioner implements HEX08 micro elements");
    b.block_graph.reset();   // the levels' buffers are rebuilt
    hipStream_t s = b.stream;
    int maxn = 0;
    for (const DomInfo& D : b.h_dom) maxn = std::max(maxn, std::max(D.nx, std::max(D.ny, D.nz)));
    int L = 0;
    for (int n = maxn; n > 8 && L < max_levels; n = (n + 1) / 2) ++L;
    b.mg.resize(L);
    const size_t v0 = (size_t)(b.n_nodes > 0 ? b.n_nodes : 1) * 3 * R * 4;
    DTC_TRY(b.z32.ensure(v0));
    DTC_TRY(b.t32.ensure(v0));
    DTC_TRY(b.d32.ensure(v0));
    for (int l = 0; l < L; ++l) {
        Level& C = b.mg[l];
        const std::vector<DomInfo>& FD = l == 0 ? b.h_dom : b.mg[l - 1].h_dom;
        const double* fe = l == 0 ? b.emod.as<double>() : b.mg[l - 1].emod.as<double>();
        C.h_dom.assign(FD.begin(), FD.end());
        int64_t nlat = 0, nvox = 0, nbt = 0;
        std::vector<int64_t> boff(C.h_dom.size());
        for (size_t d = 0; d < C.h_dom.size(); ++d) {
            DomInfo& D = C.h_dom[d];
            D.nx = (D.nx + 1) / 2; D.ny = (D.ny + 1) / 2; D.nz = (D.nz + 1) / 2;
            for (int k = 0; k < 3; ++k) D.delta[k] *= 2.0;
            D.bnx = (D.nx + BX) / BX; D.bny = (D.ny + BY) / BY; D.bnz = (D.nz + BZ) / BZ;
            D.lat_off = nlat;
            D.vox_off = nvox;
            boff[d] = nbt;
            nlat += (int64_t)(D.nx + 1) * (D.ny + 1) * (D.nz + 1);
            nvox += (int64_t)D.nx * D.ny * D.nz;
            nbt += (int64_t)D.bnx * D.bny * D.bnz;
        }
        C.n_lat = nlat;
        C.n_vox = nvox;
        DTC_TRY(C.emod.ensure((size_t)std::max<int64_t>(nvox, 1) * 8));
        DTC_TRY(C.lat2node.ensure((size_t)std::max<int64_t>(nlat, 1) * 4));
        DevMem ex, cnt;
        DTC_TRY(ex.alloc((size_t)std::max<int64_t>(nlat, 1) * 4));
        DTC_TRY(cnt.alloc((size_t)std::max<int64_t>(nbt, 1) * 4));
        k_fill_i32b<<<nb(nlat, 256), 256, 0, s>>>(C.lat2node.as<int>(), nlat, -1);
        for (size_t d = 0; d < C.h_dom.size(); ++d) {
            const DomInfo &F = FD[d], &D = C.h_dom[d];
            const int64_t cv = (int64_t)D.nx * D.ny * D.nz, cl = (int64_t)(D.nx + 1) * (D.ny + 1) * (D.nz + 1);
            k_coarsen<<<nb(cv, 256), 256, 0, s>>>(fe + F.vox_off, F.nx, F.ny, F.nz, C.emod.as<double>() + D.vox_off,
                                                 D.nx, D.ny, D.nz);
            k_exist<<<nb(cl, 256), 256, 0, s>>>(C.emod.as<double>() + D.vox_off, D.nx, D.ny, D.nz,
                                               ex.as<int>() + D.lat_off);
            k_brick_count_ex<<<D.bnx * D.bny * D.bnz, NO, 0, s>>>(ex.as<int>() + D.lat_off, D.nx + 1, D.ny + 1,
                                                                   D.nz + 1, D.bnx, D.bny, cnt.as<int>() + boff[d]);
        }
        DTC_HIP(hipGetLastError());
        std::vector<int> hc((size_t)std::max<int64_t>(nbt, 1));
        DTC_HIP(hipMemcpyAsync(hc.data(), cnt.p, (size_t)nbt * 4, hipMemcpyDeviceToHost, s));
        DTC_HIP(hipStreamSynchronize(s));
        C.h_brick.clear();
        int64_t node = 0;
        for (size_t d = 0; d < C.h_dom.size(); ++d) {
            DomInfo& D = C.h_dom[d];
            D.brick_begin = (int)C.h_brick.size();
            D.node_begin = node;
            const int nbr = D.bnx * D.bny * D.bnz;
            for (int i = 0; i < nbr; ++i) {
                const int c = hc[boff[d] + i];
                if (!c) continue;
                BrickInfo B{};
                B.dom = (int)d;
                B.x0 = (i % D.bnx) * BX; B.y0 = ((i / D.bnx) % D.bny) * BY; B.z0 = (i / (D.bnx * D.bny)) * BZ;
                B.node_begin = (int)node;
                node += c;
                B.node_end = (int)node;
                C.h_brick.push_back(B);
            }
            D.brick_end = (int)C.h_brick.size();
            D.node_end = node;
        }
        C.n_bricks = (int)C.h_brick.size();
        C.n_nodes = node;
        DTC_TRY(C.dom.ensure(C.h_dom.size() * sizeof(DomInfo)));
        DTC_HIP(hipMemcpyAsync(C.dom.p, C.h_dom.data(), C.h_dom.size() * sizeof(DomInfo), hipMemcpyHostToDevice, s));
        DTC_TRY(C.brick.ensure((C.h_brick.size() + 1) * sizeof(BrickInfo)));
        if (C.n_bricks > 0)
            DTC_HIP(hipMemcpyAsync(C.brick.p, C.h_brick.data(), C.h_brick.size() * sizeof(BrickInfo),
                                   hipMemcpyHostToDevice, s));
        for (size_t d = 0; d < C.h_dom.size(); ++d) {
            const DomInfo& D = C.h_dom[d];
            if (D.brick_end > D.brick_begin)
                k_brick_assign_ex<<<D.brick_end - D.brick_begin, NO, 0, s>>>(
                    C.brick.as<BrickInfo>() + D.brick_begin, ex.as<int>() + D.lat_off, D.nx + 1, D.ny + 1, D.nz + 1,
                    D.lat_off, C.lat2node.as<int>());
        }
        DTC_HIP(hipGetLastError());
        const size_t nn = (size_t)(C.n_nodes > 0 ? C.n_nodes : 1);
        DTC_TRY(C.dinv64.ensure(nn * 3 * 8));
        DTC_TRY(C.dinv32.ensure(nn * 3 * 4));
        for (DevMem* m : {&C.b, &C.x, &C.t, &C.d}) DTC_TRY(m->ensure(nn * 3 * R * 4));
        if (C.n_bricks > 0) {
            k_dinv<<<C.n_bricks, NO, 0, s>>>(C.brick.as<BrickInfo>(), C.dom.as<DomInfo>(), C.lat2node.as<int>(),
                                            C.emod.as<double>(), b.kdiag, C.dinv64.as<double>(), C.dinv32.as<float>());
            DTC_HIP(hipGetLastError());
            DTC_TRY(build_tables(s, C.h_brick, C.brick, C.dom, C.lat2node, C.emod, C.halo, C.nmask, C.einfo, C.eE64,
                                 C.eE32, &C.n_tile_elems));
            DTC_TRY(brick_lists(s, C.n_bricks, C.nmask, C.blist, &C.n_sparse));
        }
        DTC_HIP(hipStreamSynchronize(s));   // ex, cnt go out of scope
    }
    return DTC_OK;
}

namespace {

constexpr int MAX_SLOTS = 4096;   // timed launches between two stream synchronisations

// Start of a timed launch of kernel kind `kind`: records the first event on the
// batch stream.  *slot = -1 when timing is off (or the slot table is full).
int timed_begin(SolverBatch& b, int kind, int* slot) {
    *slot = -1;
    if (!b.timing || b.n_slots >= MAX_SLOTS) return DTC_OK;
    if (b.ev_free.size() < 2) {
        for (int k = 0; k < 2; ++k) {
            hipEvent_t e;
            DTC_HIP(hipEventCreate(&e));
            b.ev.push_back(e);
            b.ev_free.push_back(e);
        }
    }
    SolverBatch::Timed t;
    t.e1 = b.ev_free.back(); b.ev_free.pop_back();
    t.e0 = b.ev_free.back(); b.ev_free.pop_back();
    t.kind = kind;
    t.slot = b.n_slots++;
    DTC_HIP(hipEventRecord(t.e0, b.stream));
    b.ev_used.push_back(t);
    *slot = t.slot;
    return DTC_OK;
}

// End of a timed launch: second event, then the launch's algorithmic bytes over
// the domains still iterating: vpass vectors of the working precision, the
// element moduli and (jac) the Jacobi scaling
int timed_end(SolverBatch& b, int slot, int vpass, bool fp64, bool jac, const int* done) {
    if (slot < 0) return DTC_OK;
    DTC_HIP(hipEventRecord(b.ev_used.back().e1, b.stream));
    k_launch_bytes<<<1, 256, 0, b.stream>>>(done, b.dom_bytes.as<double>(), b.n_dom, vpass, fp64 ? 1 : 0, jac ? 1 : 0,
                                           b.launch_bytes.as<double>() + 2 * slot);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

// after a stream synchronisation: accumulate the timed launches
int harvest_events(SolverBatch& b) {
    if (b.ev_used.empty()) return DTC_OK;
    std::vector<double> by((size_t)b.n_slots * 2);
    DTC_HIP(hipMemcpy(by.data(), b.launch_bytes.p, by.size() * 8, hipMemcpyDeviceToHost));
    for (const SolverBatch::Timed& t : b.ev_used) {
        float ms = 0.f;
        DTC_HIP(hipEventElapsedTime(&ms, t.e0, t.e1));
        KernelStat& k = b.kstat[t.kind];
        k.ms_sum += ms;
        k.bytes_sum += by[2 * t.slot];
        k.flops_sum += by[2 * t.slot + 1];
        k.count += 1;
        b.ev_free.push_back(t.e0);
        b.ev_free.push_back(t.e1);
    }
    b.ev_used.clear();
    b.n_slots = 0;
    return DTC_OK;
}

// split > 0: that gridDim.y (the autotuner's candidates), else the tuned one
template <typename T, int IN, int OUT>
int launch_mv(SolverBatch& b, const WK<T>& wk, const T* xin, T* yout, double* part, const int* done, int split = 0) {
    if (b.n_bricks == 0) return DTC_OK;
    MvArgs a{b.brick.as<BrickInfo>(), nullptr, 0, b.dom.as<DomInfo>(), b.halo.as<int>(), b.nmask.as<unsigned char>(),
             b.einfo.as<int>(), sizeof(T) == 4 ? b.eE32.p : b.eE64.p, done, part};
    // timed: the CG matvec (OUT_MASKED: of the block PCG) and the FP64 true residual
    constexpr int kind = (IN == IN_V && (OUT == OUT_MASKED_DOT || OUT == OUT_MASKED)) ? DTC_KSTAT_MV
                       : (IN == IN_U && OUT == OUT_RESID)                              ? DTC_KSTAT_RESID
                                                                                       : -1;
    int slot = -1;
    if (kind >= 0) DTC_TRY(timed_begin(b, kind, &slot));
    if (b.micro20) {
        DTC_TRY(mv20(b, sizeof(T) == 8, IN, OUT, xin, yout, part, done));
    } else if constexpr (sizeof(T) == 4 && IN == IN_V && (OUT == OUT_MASKED || OUT == OUT_MASKED_DOT)) {
        if (b.mf_cg) {
            const dim3 grid(b.n_bricks, split > 0 ? split : tuned_split(TK_MV, 3, b.n_bricks, R / MF_RC));
            k_matvec_mf<OUT, 3><<<grid, MV_NT, 0, b.stream>>>(a, b.kmf.as<uint4>(), b.kmf_max,
                                                             reinterpret_cast<const float*>(xin),
                                                             reinterpret_cast<float*>(yout));
        } else {
            const int gy = split > 0 ? split : tuned_split(TK_MV, 0, b.n_bricks, R / MvCfg<T>::RC);
            if (b.blist.p) {
                // sparse bricks (256-thread kernel), then the dense ones
                a.blist = b.blist.as<int>();
                a.boff = 0;
                if (b.n_sparse > 0)
                    k_matvec<T, IN, OUT, false, true><<<dim3(b.n_sparse, gy), MV_NT_SP, 0, b.stream>>>(a, wk, xin, yout);
                a.boff = b.n_sparse;
                if (b.n_bricks > b.n_sparse) {
                    const dim3 gd(b.n_bricks - b.n_sparse, gy);
                    if (mv_defer()) k_matvec<T, IN, OUT, false, false><<<gd, MV_NT, 0, b.stream>>>(a, wk, xin, yout);
                    else k_matvec<T, IN, OUT, false, false, false, false><<<gd, MV_NT, 0, b.stream>>>(a, wk, xin, yout);
                }
            } else if (mv_defer()) {
                k_matvec<T, IN, OUT><<<dim3(b.n_bricks, gy), MV_NT, 0, b.stream>>>(a, wk, xin, yout);
            } else {
                k_matvec<T, IN, OUT, false, false, false, false><<<dim3(b.n_bricks, gy), MV_NT, 0, b.stream>>>(a, wk, xin,
                                                                                                         yout);
            }
        }
        DTC_HIP(hipGetLastError());
    } else {
        const dim3 grid(b.n_bricks, chunk_split(b.n_bricks, R / MvCfg<T>::RC));
        k_matvec<T, IN, OUT><<<grid, MV_NT, 0, b.stream>>>(a, wk, xin, yout);
        DTC_HIP(hipGetLastError());
    }
    // IN_U without x (x = 0) reads no vector
    DTC_TRY(timed_end(b, slot, (IN == IN_U && !xin) ? 1 : 2, sizeof(T) == 8, false, done));
    return DTC_OK;
}

// topology of one level for the FP32 matvec
struct MvTopo {
    int n_bricks;
    const BrickInfo* bricks;
    const int* blist;              // sparse bricks first (n_sparse of them), then the dense ones; null: all dense
    int n_sparse;
    const DomInfo* doms;
    const int* halo;
    const unsigned char* nmask;
    const int* einfo;
    const void* eE32;
};

// element products of the smoothing matvecs: 0 Walsh VALU kernel, else MFMA with
// `terms` FP16 terms (table and normalisation of the batch)
struct MfUse {
    int terms;
    const uint4* tab;
    float kmax;
};

// the dense FP32 Chebyshev step without a brick list: whole-segment stores unless
// DTC_MV_DEFER=0 (df false)
template <int IN, int OUT, bool PART>
void mv32_dense(bool df, dim3 grid, hipStream_t s, const MvArgs& a, const WK<float>& wk, const float* x, float* y) {
    if (df) k_matvec<float, IN, OUT, 0, false, PART, true><<<grid, MV_NT, 0, s>>>(a, wk, x, y);
    else k_matvec<float, IN, OUT, 0, false, PART, false><<<grid, MV_NT, 0, s>>>(a, wk, x, y);
}

// tb: the batch whose fine level this is (its launches are timed), else null
template <int OUT>
// cin != 0 (OUT_CHEB, dense bricks, VALU element products only): x is cin D^-1 x
// (the zero-start step formed on the fly, IN_JS)
int launch_mv32(const MvTopo& T, hipStream_t s, const WK<float>& wk, const float* x, float* y, const int* done,
                const float* bv = nullptr, const float* dinv = nullptr, float* dv = nullptr, float c1 = 0.f,
                float c2 = 0.f, double* partial = nullptr, int cflags = CH_READ_D | CH_WRITE_D,
                SolverBatch* tb = nullptr, MfUse mf = MfUse{0, nullptr, 0.f}, int split = 0, float cin = 0.f) {
    if (T.n_bricks == 0) return DTC_OK;
    if (cin != 0.f && (OUT != OUT_CHEB || mf.terms != 0 || T.n_sparse > 0))
        return set_error(DTC_ERR_BAD_ARGUMENT, "zero-start input needs the dense FP32 Chebyshev step");
    MvArgs a{T.bricks, nullptr, 0, T.doms, T.halo, T.nmask, T.einfo, T.eE32, done, partial, bv, dinv, dv, c1, c2, cflags};
    a.cin = cin;
    const int tk = OUT == OUT_CHEB ? TK_CHEB : TK_BRES;
    const dim3 grid(T.n_bricks, split > 0 ? split : tuned_split(tk, mf.terms, T.n_bricks, R / MvCfg<float>::RC));
    constexpr int kind = OUT == OUT_CHEB ? DTC_KSTAT_CHEB : OUT == OUT_BRES ? DTC_KSTAT_BRES : -1;
    int slot = -1;
    if (tb && kind >= 0) DTC_TRY(timed_begin(*tb, kind, &slot));
    if (mf.terms == 3 && partial) k_matvec_mf<OUT, 3, true><<<grid, MV_NT, 0, s>>>(a, mf.tab, mf.kmax, x, y);
    else if (mf.terms == 3) k_matvec_mf<OUT, 3><<<grid, MV_NT, 0, s>>>(a, mf.tab, mf.kmax, x, y);
    else if (mf.terms == 1 && partial) k_matvec_mf<OUT, 1, true><<<grid, MV_NT, 0, s>>>(a, mf.tab, mf.kmax, x, y);
    else if (mf.terms == 1) k_matvec_mf<OUT, 1><<<grid, MV_NT, 0, s>>>(a, mf.tab, mf.kmax, x, y);
    else if (T.blist) {
        // sparse bricks (256-thread kernel), then the dense ones
        constexpr bool CP = OUT == OUT_CHEB;
        MvArgs as = a;
        as.blist = T.blist;
        as.boff = 0;
        if (T.n_sparse > 0) {
            const dim3 gs(T.n_sparse, grid.y);
            if (CP && partial) k_matvec<float, IN_V, OUT, false, true, CP><<<gs, MV_NT_SP, 0, s>>>(as, wk, x, y);
            else k_matvec<float, IN_V, OUT, false, true, false><<<gs, MV_NT_SP, 0, s>>>(as, wk, x, y);
        }
        as.boff = T.n_sparse;
        if (T.n_bricks > T.n_sparse) {
            const dim3 gd(T.n_bricks - T.n_sparse, grid.y);
            const bool df = mv_defer();
            bool js = false;
            if constexpr (OUT == OUT_CHEB) {
                js = cin != 0.f;
                if (js && CP && partial && df) k_matvec<float, IN_JS, OUT, false, false, CP><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
                else if (js && CP && partial) k_matvec<float, IN_JS, OUT, false, false, CP, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
                else if (js && df) k_matvec<float, IN_JS, OUT, false, false, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
                else if (js) k_matvec<float, IN_JS, OUT, false, false, false, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
            }
            if (js) {
            } else if (CP && partial && df) k_matvec<float, IN_V, OUT, false, false, CP><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
            else if (CP && partial) k_matvec<float, IN_V, OUT, false, false, CP, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
            else if (df) k_matvec<float, IN_V, OUT, false, false, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
            else k_matvec<float, IN_V, OUT, false, false, false, false><<<gd, MV_NT, 0, s>>>(as, wk, x, y);
        }
    } else if constexpr (OUT == OUT_CHEB) {
        const bool df = mv_defer();
        if (cin != 0.f && partial) mv32_dense<IN_JS, OUT, true>(df, grid, s, a, wk, x, y);
        else if (cin != 0.f) mv32_dense<IN_JS, OUT, false>(df, grid, s, a, wk, x, y);
        else if (partial) mv32_dense<IN_V, OUT, true>(df, grid, s, a, wk, x, y);
        else mv32_dense<IN_V, OUT, false>(df, grid, s, a, wk, x, y);
    } else {
        k_matvec<float, IN_V, OUT><<<grid, MV_NT, 0, s>>>(a, wk, x, y);
    }
    DTC_HIP(hipGetLastError());
    if (tb) {
        // x, y, b (with the zero-start input x is b itself: 2 vectors); Chebyshev: the
        // direction read and written as the flags say
        const int vp = (cin != 0.f ? 2 : 3) + (OUT == OUT_CHEB && (cflags & CH_READ_D) ? 1 : 0) +
                       (OUT == OUT_CHEB && (cflags & CH_WRITE_D) ? 1 : 0);
        DTC_TRY(timed_end(*tb, slot, vp, false, OUT == OUT_CHEB, done));
    }
    return DTC_OK;
}

// vectors and topology of multigrid level l (0 = the fine level, whose right-hand
// side is the PCG residual); the iterate alternates between buf[0] and buf[1]
struct LvView {
    MvTopo T;
    const int* l2n;
    const float* dinv;
    float* b;
    float* buf[2];
    float* d;
};

// Chebyshev smoothing parameters: interval [lmax / ratio, lmax] of D^-1 K
struct MgParams {
    double ratio = 7.0;        // smoothing passes of the fine and intermediate levels
    int coarse_degree = 4;     // coarsest level
    double coarse_ratio = 50.0;
    double lscale = 1.0;       // multiplies the element-by-element bound of lambda_max
};

MgParams mg_params() {
    MgParams p;
    if (const char* e = std::getenv("DTC_MG_RATIO")) p.ratio = std::atof(e);
    if (const char* e = std::getenv("DTC_MG_CDEG")) p.coarse_degree = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("DTC_MG_CRATIO")) p.coarse_ratio = std::atof(e);
    if (const char* e = std::getenv("DTC_MG_LSCALE")) p.lscale = std::atof(e);
    return p;
}

// z = V(r): one V-cycle over the levels with Chebyshev-Jacobi pre- and
// post-smoothing by the same polynomial (a symmetric preconditioner), trilinear
// prolongation and its transpose as restriction.  Every smoothing step after the
// first is one fused matvec (OUT_CHEB); the residual before restriction is one
// OUT_BRES matvec.  prz (optional) receives per-brick sums of r . z.
int pmg_vcycle(SolverBatch& b, const float* r, float* z, double* prz);

int mg_vcycle(SolverBatch& b, const WK<float>& wk, const float* r, float* z, double* prz) {
    if (b.micro20) return pmg_vcycle(b, r, z, prz);
    hipStream_t s = b.stream;
    const int L = (int)b.mg.size();
    const int* done = b.done_ext ? b.done_ext : b.done.as<int>();
    static const MgParams P = mg_params();
    const int deg = b.mg_degree > 0 ? b.mg_degree : 2;
    const MfUse mfu{b.mf_smooth, b.kmf.as<uint4>(), b.kmf_max};
    auto view = [&](int l) {
        LvView v;
        if (l == 0) {
            v.T = {b.n_bricks, b.brick.as<BrickInfo>(), b.blist.as<int>(), b.n_sparse, b.dom.as<DomInfo>(), b.halo.as<int>(),
                   b.nmask.as<unsigned char>(), b.einfo.as<int>(), b.eE32.p};
            v.l2n = b.lat2node.as<int>();
            v.dinv = b.dinv32.as<float>();
            v.b = const_cast<float*>(r);
            v.buf[0] = z;
            v.buf[1] = b.t32.as<float>();
            v.d = b.d32.as<float>();
        } else {
            Level& C = b.mg[l - 1];
            v.T = {C.n_bricks, C.brick.as<BrickInfo>(), C.blist.as<int>(), C.n_sparse, C.dom.as<DomInfo>(), C.halo.as<int>(),
                   C.nmask.as<unsigned char>(), C.einfo.as<int>(), C.eE32.p};
            v.l2n = C.lat2node.as<int>();
            v.dinv = C.dinv32.as<float>();
            v.b = C.b.as<float>();
            v.buf[0] = C.x.as<float>();
            v.buf[1] = C.t.as<float>();
            v.d = C.d.as<float>();
        }
        return v;
    };
    std::vector<int> cur(L + 1, 0);
    auto smooth = [&](int l, int m, bool zero, double ratio, double* dot) -> int {
        const LvView v = view(l);
        if (v.T.n_bricks == 0) return DTC_OK;
        const double hi = b.cheb_lmax * P.lscale, lo = hi / ratio;
        const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
        double rho = 1.0 / sigma;
        float c0 = 0.f;   // the zero-start step's coefficient, when that step is fused into the next
        for (int k = 0; k < m; ++k) {
            float c1 = 0.f, c2 = (float)(1.0 / theta);
            if (k > 0) {
                const double rn = 1.0 / (2.0 * sigma - rho);
                c1 = (float)(rn * rho);
                c2 = (float)(2.0 * rn / delta);
                rho = rn;
            }
            double* dk = k == m - 1 ? dot : nullptr;
            // the zero-start step x_1 = c2 D^-1 b is formed inside the next step's
            // staging (IN_JS) instead of being stored, where that kernel form runs
            const bool fuse = zero && m >= 2 && mfu.terms == 0 && v.T.n_sparse == 0 && cheb_fuse();
            if (zero && k == 0 && fuse) {
                c0 = c2;
            } else if (zero && k == 0) {
                // x_1 = d_0: the next step takes its previous direction from x
                k_cheb<<<v.T.n_bricks, NT_UPD, 0, s>>>(v.T.bricks, done, v.b, nullptr, v.dinv, nullptr, v.buf[cur[l]],
                                                       c1, c2, 1, dk);
                DTC_HIP(hipGetLastError());
            } else {
                const int fl = (k == 0 ? 0 : (zero && k == 1) ? CH_D_FROM_X : CH_READ_D) | (k < m - 1 ? CH_WRITE_D : 0);
                const bool js = fuse && k == 1;
                DTC_TRY(launch_mv32<OUT_CHEB>(v.T, s, wk, js ? v.b : v.buf[cur[l]], v.buf[1 - cur[l]], done, v.b,
                                              v.dinv, v.d, c1, c2, dk, fl, l == 0 ? &b : nullptr, mfu, 0,
                                              js ? c0 : 0.f));
                cur[l] = 1 - cur[l];
            }
        }
        return DTC_OK;
    };
    // the fine iterate makes (deg - 1) + deg swaps (or coarse_degree - 1 without
    // coarse levels): start in the buffer that makes it end in z
    cur[0] = (L > 0 ? 2 * deg - 1 : P.coarse_degree - 1) % 2;
    for (int l = 0; l < L; ++l) {
        DTC_TRY(smooth(l, deg, true, P.ratio, nullptr));
        const LvView f = view(l), c = view(l + 1);
        cur[l + 1] = 0;
        if (c.T.n_bricks == 0 || f.T.n_bricks == 0) continue;
        float* res = f.buf[1 - cur[l]];
        DTC_TRY(launch_mv32<OUT_BRES>(f.T, s, wk, f.buf[cur[l]], res, done, f.b, nullptr, nullptr, 0.f, 0.f, nullptr,
                                      0, l == 0 ? &b : nullptr, mfu));
        k_restrict<<<dim3(c.T.n_bricks, node_split(c.T.n_bricks)), 256, 0, s>>>(c.T.bricks, c.T.doms, c.l2n, f.T.doms, f.l2n, res, done, c.b);
        DTC_HIP(hipGetLastError());
    }
    DTC_TRY(smooth(L, P.coarse_degree, true, P.coarse_ratio, L == 0 ? prz : nullptr));
    for (int l = L - 1; l >= 0; --l) {
        const LvView f = view(l), c = view(l + 1);
        if (f.T.n_bricks > 0 && c.T.n_bricks > 0) {
            k_prolong<<<dim3(f.T.n_bricks, node_split(f.T.n_bricks)), 256, 0, s>>>(f.T.bricks, f.T.doms, f.l2n, c.T.doms, c.l2n, c.buf[cur[l + 1]], done,
                                                   f.buf[cur[l]]);
            DTC_HIP(hipGetLastError());
        }
        DTC_TRY(smooth(l, deg, false, P.ratio, l == 0 ? prz : nullptr));
    }
    return DTC_OK;
}

// p-multigrid V-cycle of HEX20 micro domains (FP32): Chebyshev-Jacobi smoothing
// on the HEX20 operator (degree mg_degree, zero start), the residual restricted
// to the HEX08 discretisation of the same voxels (P^T), that batch's geometric
// V-cycle, the correction prolongated (P), and the same Chebyshev polynomial
// after it (a symmetric preconditioner).
int pmg_vcycle(SolverBatch& b, const float* r, float* z, double* prz) {
    if (!b.p08) return set_error(DTC_ERR_DEVICE, "p-multigrid without its HEX08 level");
    SolverBatch& c = *b.p08;
    hipStream_t s = b.stream;
    const int* done = b.done.as<int>();
    static const MgParams P = mg_params();
    const int deg = b.mg_degree > 0 ? b.mg_degree : 2;
    const BrickInfo* bl = b.brick.as<BrickInfo>();
    float* t = b.t32.as<float>();
    float* d = b.d32.as<float>();
    const double hi = b.cheb_lmax * P.lscale, lo = hi / P.ratio;
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    auto smooth = [&](bool zero, double* dot) -> int {
        double rho = 1.0 / sigma;
        for (int k = 0; k < deg; ++k) {
            float c1 = 0.f, c2 = (float)(1.0 / theta);
            if (k > 0) {
                const double rn = 1.0 / (2.0 * sigma - rho);
                c1 = (float)(rn * rho);
                c2 = (float)(2.0 * rn / delta);
                rho = rn;
            }
            double* dk = k == deg - 1 ? dot : nullptr;
            if (zero && k == 0) {
                k_cheb<<<b.n_bricks, NT_UPD, 0, s>>>(bl, done, r, nullptr, b.dinv32.as<float>(), d, z, c1, c2, 1, dk);
            } else if (!dk && cheb_fuse() && !hex20_gather_form()) {
                // the step in the epilogue of the operator's node gather (no t vector)
                DTC_TRY(mv20_cheb(b, z, done, r, b.dinv32.as<float>(), d, c1, c2));
            } else {
                DTC_TRY(mv20(b, false, IN_V, OUT_MASKED, z, t, nullptr, done));
                k_cheb<<<b.n_bricks, NT_UPD, 0, s>>>(bl, done, r, t, b.dinv32.as<float>(), d, z, c1, c2, 0, dk);
            }
            DTC_HIP(hipGetLastError());
        }
        return DTC_OK;
    };
    DTC_TRY(smooth(true, nullptr));
    // residual, restricted to the HEX08 nodes of each domain group
    DTC_TRY(mv20(b, false, IN_V, OUT_MASKED, z, t, nullptr, done));
    k_sub<<<b.n_bricks, NT_UPD, 0, s>>>(bl, done, r, t, t);
    DTC_HIP(hipGetLastError());
    float* r08 = c.r32.as<float>();
    float* z08 = c.z32.as<float>();
    for (int g = 0; g < b.n_dom && g < c.n_dom; ++g) {
        const DomInfo& D08 = c.h_dom[g];
        if (b.h_status[g] != DTC_OK || c.h_status[g] != DTC_OK) continue;
        const int64_t n = (int64_t)(D08.nx + 1) * (D08.ny + 1) * (D08.nz + 1) * R;
        k_pmg_restrict<<<nb(n, 256), 256, 0, s>>>(b.h_d20dev[g], D08, c.lat2node.as<int>(), done, g, t, r08);
    }
    DTC_HIP(hipGetLastError());
    c.done_ext = done;
    DTC_TRY(mg_vcycle(c, cast_wk<float>(c.wk64), r08, z08, nullptr));
    k_pmg_prolong<<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.dom.as<DomInfo>(), b.d20dev.as<Dom20Dev>(), c.dom.as<DomInfo>(),
                                                c.lat2node.as<int>(), done, z08, z);
    DTC_HIP(hipGetLastError());
    return smooth(false, prz);
}

// p-multigrid set-up of a HEX20 batch over its HEX08 twin c (same domain groups,
// finalised): the twin's geometric levels, the fine Chebyshev bound and vectors
int build_pmg(SolverBatch& b, SolverBatch& c) {
    if (!b.micro20 || c.micro20) return set_error(DTC_ERR_BAD_ARGUMENT, "p-multigrid needs a HEX20 batch and its HEX08 twin");
    if (c.n_dom != b.n_dom) return set_error(DTC_ERR_DEVICE, "p-multigrid: %d HEX08 groups for %d HEX20 groups", c.n_dom,
                                             b.n_dom);
    b.block_graph.reset();
    c.stream = b.stream;
    c.mg_degree = b.mg_degree;
    DTC_TRY(batch_build_mg(c, 8));
    // element-by-element bound of lambda_max(D^-1 K) for HEX20:
    // lambda_max(diag(Khat)^-1/2 Khat diag(Khat)^-1/2)
    std::vector<double> S(3600);
    for (int i = 0; i < 60; ++i)
        for (int j = 0; j < 60; ++j)
            S[i + 60 * j] = b.khat20[i + 60 * j] / std::sqrt(b.khat20[i + 60 * i] * b.khat20[j + 60 * j]);
    b.cheb_lmax = sym_lmax(S.data(), 60);
    const size_t v0 = (size_t)(b.n_nodes > 0 ? b.n_nodes : 1) * 3 * R * 4;
    DTC_TRY(b.z32.ensure(v0));
    DTC_TRY(b.t32.ensure(v0));
    DTC_TRY(b.d32.ensure(v0));
    b.p08 = &c;
    return DTC_OK;
}

// one PCG iteration on (x, r, p, q) in precision T; zmg != null: multigrid
// preconditioner (FP32 only), z = V(r) stored in zmg
template <typename T>
int pcg_iteration(SolverBatch& b, const WK<T>& wk, T* x, T* r, T* p, T* q, const T* dinv, int64_t max_iters, T* zmg) {
    const int D = b.n_dom, NBk = b.n_bricks;
    int* done = b.done.as<int>();
    DTC_TRY((launch_mv<T, IN_V, OUT_MASKED_DOT>(b, wk, p, q, b.part_a.as<double>(), done)));
    k_red_alpha<<<D, RED_NT, 0, b.stream>>>(b.dom.as<DomInfo>(), done, b.part_a.as<double>(), b.active.as<int>(),
                                        b.rz.as<double>(), b.alpha.as<double>());
    k_cg_update<T><<<NBk, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), done, b.alpha.as<double>(), x, r, p, q, dinv,
                                                 zmg ? nullptr : b.part_a.as<double>(), b.part_b.as<double>());
    if constexpr (std::is_same<T, float>::value) {
        if (zmg) DTC_TRY(mg_vcycle(b, wk, r, zmg, b.part_a.as<double>()));
    }
    k_red_beta<<<D, RED_NT, 0, b.stream>>>(b.dom.as<DomInfo>(), done, b.part_a.as<double>(), b.part_b.as<double>(),
                                       b.thr2.as<double>(), b.active.as<int>(), b.rz.as<double>(), b.beta.as<double>(),
                                       b.iters.as<int64_t>(), max_iters);
    k_cg_dir<T><<<NBk, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), done, b.beta.as<double>(), r, p, dinv, zmg);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

int all_done(SolverBatch& b, const DevMem& flags, bool* out) {
    std::vector<int> h(b.n_dom);
    DTC_HIP(hipMemcpyAsync(h.data(), flags.p, (size_t)b.n_dom * 4, hipMemcpyDeviceToHost, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    DTC_TRY(harvest_events(b));
    bool all = true;
    for (int v : h) all = all && (v != 0);
    *out = all;
    return DTC_OK;
}

template <typename T>
int pcg_run(SolverBatch& b, const WK<T>& wk, T* x, T* r, T* p, T* q, const T* dinv, int64_t max_iters,
            int64_t max_steps, T* zmg) {
    const int CHECK = zmg ? 4 : 20;
    for (int64_t steps = 0; steps < max_steps; steps += CHECK) {
        bool fin = false;
        DTC_TRY(all_done(b, b.done, &fin));
        if (fin) return DTC_OK;
        for (int k = 0; k < CHECK; ++k) DTC_TRY(pcg_iteration<T>(b, wk, x, r, p, q, dinv, max_iters, zmg));
    }
    return DTC_OK;
}

// Q = A P of the block PCG: FP32 vectors, FP64 arithmetic.  The Gram matrix
// W = P^T Q resolves the directions of the block only as far as Q is accurate:
// for smooth directions FP32 sums lose most digits of A P to cancellation, and
// W's small eigenvalues (the ones that steer the block) drown in that error.
// q64 set: Q is written as an FP64 vector (the block PCG on FP64 vectors)
int launch_mv_q64(SolverBatch& b, const float* p, float* q, const int* done, double* q64 = nullptr) {
    MvArgs a{b.brick.as<BrickInfo>(), nullptr, 0, b.dom.as<DomInfo>(), b.halo.as<int>(), b.nmask.as<unsigned char>(),
             b.einfo.as<int>(), b.eE64.p, done, nullptr};
    int slot = -1;
    DTC_TRY(timed_begin(b, DTC_KSTAT_MV, &slot));
    const dim3 grid(b.n_bricks, chunk_split(b.n_bricks, R / MvCfg<double>::RC));
    if (q64)
        k_matvec<double, IN_V, OUT_MASKED, 2><<<grid, MV_NT, 0, b.stream>>>(
            a, cast_wk<double>(b.wk64), reinterpret_cast<const double*>(p), q64);
    else
        k_matvec<double, IN_V, OUT_MASKED, 1><<<grid, MV_NT, 0, b.stream>>>(
            a, cast_wk<double>(b.wk64), reinterpret_cast<const double*>(p), reinterpret_cast<double*>(q));
    DTC_HIP(hipGetLastError());
    DTC_TRY(timed_end(b, slot, 2, false, false, done));
    return DTC_OK;
}

// block PCG set-up: slices of SLICE_NODES nodes (more for domains of over
// MAX_SLICES * SLICE_NODES nodes) within each domain for the Gram sums, and the
// per-domain factor and coefficient storage
int block_setup(SolverBatch& b) {
    if (b.n_slices > 0) return DTC_OK;
    std::vector<int4> sl;
    std::vector<int> ds((size_t)b.n_dom + 1);
    for (int d = 0; d < b.n_dom; ++d) {
        ds[d] = (int)sl.size();
        const DomInfo& D = b.h_dom[d];
        const int64_t nn = D.node_end - D.node_begin;
        const int64_t step = std::max<int64_t>(SLICE_NODES, (nn + MAX_SLICES - 1) / MAX_SLICES);
        for (int64_t n0 = D.node_begin; n0 < D.node_end; n0 += step)
            sl.push_back(make_int4(d, (int)n0, (int)std::min<int64_t>(n0 + step, D.node_end), 0));
    }
    ds[b.n_dom] = (int)sl.size();
    DTC_TRY(b.slices.ensure(std::max<size_t>(sl.size(), 1) * sizeof(int4)));
    DTC_TRY(b.dslice.ensure(ds.size() * 4));
    DTC_TRY(b.gpart.ensure(std::max<size_t>(sl.size(), 1) * 2 * RR * 8));
    DTC_TRY(b.gsum.ensure((size_t)std::max(b.n_dom, 1) * 2 * RR * 8));
    DTC_TRY(b.lfac.ensure((size_t)std::max(b.n_dom, 1) * RR * 8));
    DTC_TRY(b.bcoef.ensure((size_t)std::max(b.n_dom, 1) * RR * 4));
    // per-slice r.r partials of the block update (fewer slices than bricks, but no assumption)
    DTC_TRY(b.part_b.ensure(std::max<size_t>({sl.size(), (size_t)b.n_bricks, 1}) * R * 8));
    if (!sl.empty())
        DTC_HIP(hipMemcpyAsync(b.slices.p, sl.data(), sl.size() * sizeof(int4), hipMemcpyHostToDevice, b.stream));
    DTC_HIP(hipMemcpyAsync(b.dslice.p, ds.data(), ds.size() * 4, hipMemcpyHostToDevice, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    b.n_slices = (int)sl.size();
    return DTC_OK;
}

// per-domain Grams a^T b and a^T c (c may be null) into gsum
template <typename T> struct NoDeduce { using type = T; };
// the Gram partials of the orthonormalisation formed in the direction kernel that
// writes the new block (DTC_ORTH_FUSE=0: a separate Gram kernel reads it back)
inline bool orth_fuse() {
    const char* e = std::getenv("DTC_ORTH_FUSE");
    return !(e && e[0] == '0');
}

// Gram form: rows staged through LDS (default) or read by every lane from global
// memory (DTC_GRAM_LDS=0; the same sums, for measurements)
inline bool gram_lds() {
    const char* e = std::getenv("DTC_GRAM_LDS");
    return !(e && e[0] == '0');
}
template <typename TA = float, typename TB = float>
int block_gram(SolverBatch& b, const TA* a, const TB* bb, const typename NoDeduce<TB>::type* c) {
    hipStream_t s = b.stream;
    int* done = b.done.as<int>();
    if (b.n_slices > 0) {
        const int4* sl = b.slices.as<int4>();
        double* gp = b.gpart.as<double>();
        if (gram_lds()) {
            if (c) k_gram_lds<2, TA, TB><<<b.n_slices, GR_NT, 0, s>>>(sl, done, a, bb, c, gp);
            else k_gram_lds<1, TA, TB><<<b.n_slices, GR_NT, 0, s>>>(sl, done, a, bb, nullptr, gp);
        } else {
            if (c) k_gram<2, TA, TB><<<b.n_slices, GR_NT, 0, s>>>(sl, done, a, bb, c, gp);
            else k_gram<1, TA, TB><<<b.n_slices, GR_NT, 0, s>>>(sl, done, a, bb, nullptr, gp);
        }
    }
    k_gram_sum<<<dim3(b.n_dom, 2 * RR / 64), 256, 0, s>>>(b.dslice.as<int>(), done, b.gpart.as<double>(),
                                                         b.gsum.as<double>());
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

// dst = src with orthonormal columns per domain (dropped directions zero).  In
// FP32, combinations of nearly dependent columns with large coefficients that
// cancel lose the directions they should keep: the residual columns of a
// domain line up with its slowest modes after the first residual replacement.
int block_orth(SolverBatch& b, const float* src, float* dst, bool partials_done = false) {
    hipStream_t s = b.stream;
    int* done = b.done.as<int>();
    if (partials_done) {   // the direction kernel wrote src's Gram partials
        k_gram_sum<<<dim3(b.n_dom, 2 * RR / 64), 256, 0, s>>>(b.dslice.as<int>(), done, b.gpart.as<double>(),
                                                             b.gsum.as<double>());
    } else {
        DTC_TRY(block_gram(b, src, src, nullptr));
    }
    k_bsolve<<<b.n_dom, BS_NT, 0, s>>>(done, b.gsum.as<double>(), 2, b.lfac.as<double>(), b.bcoef.as<float>());
    k_bdir<false><<<b.n_slices, NT_UPD, 0, s>>>(b.slices.as<int4>(), done, b.bcoef.as<float>(), src, nullptr, dst,
                                                nullptr);
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

// the new direction block Z + P beta and, with orth_fuse(), its Gram partials in the
// same kernel; then the orthonormalisation into p
int block_dir_orth(SolverBatch& b, float* p, const float* z, float* q) {
    hipStream_t s = b.stream;
    int* done = b.done.as<int>();
    const int4* sl = b.slices.as<int4>();
    const bool fuse = orth_fuse() && b.n_slices > 0;
    if (fuse)
        k_bdir<true><<<b.n_slices, NT_UPD, 0, s>>>(sl, done, b.bcoef.as<float>(), p, z, q, b.gpart.as<double>());
    else
        k_bdir<false><<<b.n_slices, NT_UPD, 0, s>>>(sl, done, b.bcoef.as<float>(), p, z, q, nullptr);
    DTC_HIP(hipGetLastError());
    return block_orth(b, q, p, fuse);
}

// one block PCG iteration on FP64 vectors (x64, r64) with the FP32 multigrid as
// preconditioner: Q = K P in FP64 (FP32 directions P, FP64 arithmetic), the
// Grams in FP64, x += P alpha and r -= Q alpha in FP64; r is rounded to FP32
// only as the V-cycle's input.  Used for high-contrast (greyscale) batches, whose
// FP32 recursive residual loses the soft regions' part to the rounding of the
// stiff regions' terms.
int block_iteration64(SolverBatch& b, const WK<float>& wk, float* p, float* q, float* z, int64_t max_iters) {
    hipStream_t s = b.stream;
    const int D = b.n_dom, NBk = b.n_bricks, NSl = b.n_slices;
    const int4* sl = b.slices.as<int4>();
    int* done = b.done.as<int>();
    const BrickInfo* bl = b.brick.as<BrickInfo>();
    double* x = b.x64.as<double>();
    double* r = b.r64.as<double>();
    double* q64 = b.q64.as<double>();
    if (b.micro20) DTC_TRY(mv20_q64(b, p, nullptr, done, q64));
    else DTC_TRY(launch_mv_q64(b, p, nullptr, done, q64));
    DTC_TRY((block_gram<float, double>(b, p, q64, r)));
    k_bsolve<<<D, BS_NT, 0, s>>>(done, b.gsum.as<double>(), 0, b.lfac.as<double>(), b.bcoef.as<float>());
    k_bupdate64<<<NSl, NT_UPD, 0, s>>>(sl, done, b.bcoef.as<float>(), x, r, p, q64, b.part_b.as<double>());
    k_bcheck<<<D, RED_NT, 0, s>>>(b.dslice.as<int>(), done, b.part_b.as<double>(), b.thr2.as<double>(), nullptr,
                                  b.iters.as<int64_t>(), max_iters);
    k_to32<<<NBk, NT_UPD, 0, s>>>(bl, done, r, b.r32.as<float>(), nullptr, nullptr);
    DTC_HIP(hipGetLastError());
    DTC_TRY(mg_vcycle(b, wk, b.r32.as<float>(), z, nullptr));
    DTC_TRY((block_gram<double, float>(b, q64, z, nullptr)));
    k_bsolve<<<D, BS_NT, 0, s>>>(done, b.gsum.as<double>(), 1, b.lfac.as<double>(), b.bcoef.as<float>());
    return block_dir_orth(b, p, z, q);
}

// one block PCG iteration on (x, r); p is the direction block (orthonormal
// columns), q and z are work vectors
int block_iteration(SolverBatch& b, const WK<float>& wk, float* x, float* r, float* p, float* q, float* z,
                    int64_t max_iters) {
    if (b.block64) return block_iteration64(b, wk, p, q, z, max_iters);
    hipStream_t s = b.stream;
    const int D = b.n_dom, NSl = b.n_slices;
    const int4* sl = b.slices.as<int4>();
    int* done = b.done.as<int>();
    if (b.block_q64 && b.micro20) DTC_TRY(mv20_q64(b, p, q, done));
    else if (b.block_q64) DTC_TRY(launch_mv_q64(b, p, q, done));
    else DTC_TRY((launch_mv<float, IN_V, OUT_MASKED>(b, wk, p, q, nullptr, done)));
    DTC_TRY(block_gram(b, p, q, r));
    k_bsolve<<<D, BS_NT, 0, s>>>(done, b.gsum.as<double>(), 0, b.lfac.as<double>(), b.bcoef.as<float>());
    k_bupdate<<<NSl, NT_UPD, 0, s>>>(sl, done, b.bcoef.as<float>(), x, r, p, q, b.part_b.as<double>());
    k_bcheck<<<D, RED_NT, 0, s>>>(b.dslice.as<int>(), done, b.part_b.as<double>(), b.thr2.as<double>(),
                                  b.rr.as<double>(), b.iters.as<int64_t>(), max_iters);
    DTC_HIP(hipGetLastError());
    DTC_TRY(mg_vcycle(b, wk, r, z, nullptr));
    DTC_TRY(block_gram(b, q, z, nullptr));
    k_bsolve<<<D, BS_NT, 0, s>>>(done, b.gsum.as<double>(), 1, b.lfac.as<double>(), b.bcoef.as<float>());
    return block_dir_orth(b, p, z, q);
}

// *out = 1 when every domain's flag is set (one workgroup, fixed order)
__global__ __launch_bounds__(256) void k_all_set(const int* __restrict__ flags, int n, int* __restrict__ out) {
    __shared__ int s_any[256];
    int miss = 0;
    for (int d = threadIdx.x; d < n; d += 256) miss |= flags[d] == 0;
    s_any[threadIdx.x] = miss;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_any[threadIdx.x] |= s_any[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = s_any[0] ? 0 : 1;
}

// block PCG iterations until every domain is done or max_steps iterations ran.
//
// Without kernel timing (the default), GROUP iterations are captured once per
// solve into a HIP graph together with a kernel that writes "every domain done"
// to a device word; the graph is replayed, and after each replay that word is
// copied into one of two page-locked host words.  The host decides to stop
// from the previous replay's word while the current replay already runs, so the
// GPU never waits for the host; a replay issued after every domain finished is
// a sequence of early exits (every kernel skips done domains), which leaves the
// results unchanged.  DTC_GRAPH=0 (or kernel timing) launches the iterations one
// by one with a synchronous check every GROUP iterations.
inline bool trace_on() {
    static const bool on = std::getenv("DTC_TRACE") && std::getenv("DTC_TRACE")[0] == '1';
    return on;
}

int block_run(SolverBatch& b, const WK<float>& wk, float* x, float* r, float* p, float* q, float* z,
              int64_t max_iters, int64_t max_steps) {
    constexpr int GROUP = 4;
    const char* ge = std::getenv("DTC_GRAPH");
    const bool use_graph = !(ge && ge[0] == '0') && !b.timing;
    bool fin = false;
    DTC_TRY(all_done(b, b.done, &fin));
    if (fin) return DTC_OK;
    if (!use_graph) {
        for (int64_t steps = 0; steps < max_steps; steps += GROUP) {
            if (steps > 0) {
                DTC_TRY(all_done(b, b.done, &fin));
                if (fin) return DTC_OK;
            }
            for (int k = 0; k < GROUP; ++k) DTC_TRY(block_iteration(b, wk, x, r, p, q, z, max_iters));
        }
        return DTC_OK;
    }
    hipStream_t s = b.stream;
    const void* key[6] = {x, r, p, q, z, b.done.p};
    bool same = b.block_graph.e && b.graph_iters == max_iters && b.graph_bricks == b.n_bricks &&
                b.graph_mode == (int)b.block64;
    for (int k = 0; k < 6; ++k) same = same && b.graph_key[k] == key[k];
    if (!same) {
        b.block_graph.reset();
        DTC_TRY(b.all_done_dev.ensure(4));
        if (!b.all_done_host.p) DTC_HIP(hipHostMalloc((void**)&b.all_done_host.p, 2 * sizeof(int), hipHostMallocDefault));
        hipGraph_t g = nullptr;
        DTC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        int rc = DTC_OK;
        for (int k = 0; k < GROUP && rc == DTC_OK; ++k) rc = block_iteration(b, wk, x, r, p, q, z, max_iters);
        if (rc == DTC_OK) k_all_set<<<1, 256, 0, s>>>(b.done.as<int>(), b.n_dom, b.all_done_dev.as<int>());
        const hipError_t ec = hipStreamEndCapture(s, &g);
        if (rc != DTC_OK && g) (void)hipGraphDestroy(g);
        DTC_TRY(rc);
        if (ec != hipSuccess) return set_error(DTC_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        const hipError_t ei = hipGraphInstantiate(&b.block_graph.e, g, nullptr, nullptr, 0);
        b.block_graph.g = g;
        if (ei != hipSuccess) {
            b.block_graph.e = nullptr;
            b.block_graph.reset();
            return set_error(DTC_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
        }
        for (int k = 0; k < 6; ++k) b.graph_key[k] = key[k];
        b.graph_iters = max_iters;
        b.graph_bricks = b.n_bricks;
        b.graph_mode = (int)b.block64;
    }
    // launch n copies the summary to host word n % 2 and records event n % 2
    hipEvent_t ev[2];
    for (hipEvent_t& e : ev) DTC_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    int rc = DTC_OK;
    int64_t launched = 0;
    auto launch = [&](int64_t n) -> int {
        DTC_HIP(hipGraphLaunch(b.block_graph.e, s));
        DTC_HIP(hipMemcpyAsync(b.all_done_host.p + (n & 1), b.all_done_dev.p, sizeof(int), hipMemcpyDeviceToHost, s));
        DTC_HIP(hipEventRecord(ev[n & 1], s));
        ++b.graph_launches;
        return DTC_OK;
    };
    rc = launch(0);
    launched = 1;
    for (int64_t steps = GROUP; rc == DTC_OK && steps < max_steps; steps += GROUP) {
        rc = launch(launched);
        if (rc != DTC_OK) break;
        ++launched;
        // the replay before this one has finished: its summary decides
        const hipError_t e = hipEventSynchronize(ev[(launched - 2) & 1]);
        if (e != hipSuccess) {
            rc = set_error(DTC_ERR_DEVICE, "hipEventSynchronize: %s", hipGetErrorString(e));
            break;
        }
        if (b.all_done_host.p[(launched - 2) & 1]) break;
        // DTC_TRACE=1: a progress line every 1,000 block iterations of a long phase
        if (trace_on() && (steps + GROUP) % 1000 < GROUP)
            std::fprintf(stderr, "[dtc] block phase: %lld iterations\n", (long long)(steps + GROUP));
    }
    const hipError_t es = hipStreamSynchronize(s);
    for (hipEvent_t& e : ev) (void)hipEventDestroy(e);
    DTC_TRY(rc);
    if (es != hipSuccess) return set_error(DTC_ERR_DEVICE, "hipStreamSynchronize: %s", hipGetErrorString(es));
    return DTC_OK;
}

// FP64 true residual r64 = -(K u)_free, u = x64 + x_bc (x64 == null: u = x_bc),
// then per load case activity and targets (k_red_outer)
int residual64(SolverBatch& b, const WK<double>& wk, const double* x64, int first, const dtc_solve_opts& o) {
    DTC_TRY((launch_mv<double, IN_U, OUT_RESID>(b, wk, x64, b.r64.as<double>(), b.part_c.as<double>(), nullptr)));
    const double* itol2 = b.precision == DTC_FP32_REFINE && !b.block64 ? b.itol2.as<double>() : nullptr;
    k_red_outer<<<b.n_dom, RED_NT, 0, b.stream>>>(b.dom.as<DomInfo>(), b.part_c.as<double>(), first,
                                                  o.rel_tol * o.rel_tol, itol2, b.rr0.as<double>(), b.rr.as<double>(),
                                                  b.active64.as<int>(), b.active.as<int>(), b.thr2.as<double>(),
                                                  b.done64.as<int>(), b.done.as<int>());
    DTC_HIP(hipGetLastError());
    return DTC_OK;
}

// Progress of the FP64 true residual over the residual replacements of the mixed
// solve: a domain whose worst load case has not halved its relative residual in
// STALL replacements, or whose residual is not finite, is given up (the caller
// repeats it in FP64).
struct Progress {
    static constexpr int STALL = 10;
    std::vector<double> best;
    std::vector<int64_t> at;
};

// inner iteration cap reached (or no progress, see Progress) for some domains:
// mark them finished
int apply_cap(SolverBatch& b, int64_t cap, std::vector<int64_t>& nouter, bool* any_left, Progress* pg = nullptr) {
    const int D = b.n_dom;
    std::vector<int> h(D);
    std::vector<int64_t> hi(D);
    std::vector<double> hrr, hr0;
    DTC_HIP(hipMemcpyAsync(h.data(), b.done64.p, (size_t)D * 4, hipMemcpyDeviceToHost, b.stream));
    DTC_HIP(hipMemcpyAsync(hi.data(), b.iters.p, (size_t)D * 8, hipMemcpyDeviceToHost, b.stream));
    if (pg) {
        hrr.resize((size_t)D * R);
        hr0.resize((size_t)D * R);
        DTC_HIP(hipMemcpyAsync(hrr.data(), b.rr.p, hrr.size() * 8, hipMemcpyDeviceToHost, b.stream));
        DTC_HIP(hipMemcpyAsync(hr0.data(), b.rr0.p, hr0.size() * 8, hipMemcpyDeviceToHost, b.stream));
        if ((int)pg->best.size() != D) {
            pg->best.assign(D, -1.0);
            pg->at.assign(D, 0);
        }
    }
    DTC_HIP(hipStreamSynchronize(b.stream));
    // DTC_TRACE=1: one stderr line per residual replacement of the mixed solve
    // (domain: PCG iterations so far / worst relative FP64 residual)
    static const bool trace = std::getenv("DTC_TRACE") && std::getenv("DTC_TRACE")[0] == '1';
    if (trace && pg) {
        std::fprintf(stderr, "[dtc] replacement:");
        for (int d = 0; d < D; ++d) {
            double worst = 0.0;
            for (int lc = 0; lc < R; ++lc) {
                const double r0 = hr0[(size_t)d * R + lc];
                if (r0 > 0.0) worst = std::fmax(worst, std::sqrt(hrr[(size_t)d * R + lc] / r0));
            }
            std::fprintf(stderr, " %d:%lld/%.1e", d, (long long)hi[d], worst);
        }
        std::fprintf(stderr, "\n");
    }
    bool any = false, changed = false;
    for (int d = 0; d < D; ++d) {
        bool stop = !h[d] && hi[d] >= cap;
        if (pg && !h[d] && !stop) {
            double worst = 0.0;
            for (int lc = 0; lc < R; ++lc) {
                const double r0 = hr0[(size_t)d * R + lc], rr = hrr[(size_t)d * R + lc];
                const double v = r0 > 0.0 ? std::sqrt(rr / r0) : 0.0;
                worst = std::isfinite(v) && std::isfinite(r0) ? std::fmax(worst, v) : INFINITY;
            }
            if (!std::isfinite(worst)) stop = true;
            else if (pg->best[d] < 0.0 || worst < 0.5 * pg->best[d]) {
                pg->best[d] = worst;
                pg->at[d] = nouter[d];
            } else if (nouter[d] - pg->at[d] >= Progress::STALL) {
                stop = true;
            }
        }
        if (stop) { h[d] = 1; changed = true; }
        if (!h[d]) { nouter[d] += 1; any = true; }
    }
    if (changed) {
        DTC_HIP(hipMemcpyAsync(b.done64.p, h.data(), (size_t)D * 4, hipMemcpyHostToDevice, b.stream));
        DTC_HIP(hipMemcpyAsync(b.done.p, h.data(), (size_t)D * 4, hipMemcpyHostToDevice, b.stream));
    }
    *any_left = any;
    return DTC_OK;
}

}  // namespace

// Times every candidate split of the FP32 matvec kernels the mixed multigrid solve
// launches (Q = K P on the fine level; the Chebyshev step and the restriction
// residual on every level) on this batch's own tables, and caches the fastest
// (see tuned_split).  The timed launches read the level's work vectors and write
// its scratch vector, before the solve starts, with no done flags.
int autotune_splits(SolverBatch& b, const WK<float>& wk) {
    const char* te = std::getenv("DTC_TUNE");
    if ((te && te[0] == '0') || b.n_bricks == 0 || b.micro20) return DTC_OK;
    const char* le = std::getenv("DTC_TUNE_LOG");
    const bool log = le && le[0] == '1';
    hipStream_t s = b.stream;
    hipEvent_t e0, e1;
    DTC_HIP(hipEventCreate(&e0));
    DTC_HIP(hipEventCreate(&e1));
    int rc = DTC_OK;
    auto measure = [&](int kind, int variant, int nbr, const char* what, int level,
                       const std::function<int(int)>& launch) {
        if (rc != DTC_OK || nbr == 0) return;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            if (g_tune.count(std::make_tuple(kind, variant, nbr))) return;
        }
        const int cand[4] = {1, 2, 3, 6};
        int best = chunk_split(nbr, R / MF_RC);
        float best_ms = 1e30f;
        for (int c : cand) {
            if ((rc = launch(c)) != DTC_OK) return;   // warm-up
            if (hipEventRecord(e0, s) != hipSuccess) { rc = set_error(DTC_ERR_DEVICE, "hipEventRecord"); return; }
            for (int k = 0; k < 3 && rc == DTC_OK; ++k) rc = launch(c);
            if (rc != DTC_OK) return;
            float ms = 0.f;
            if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = set_error(DTC_ERR_DEVICE, "autotune timing");
                return;
            }
            if (log)
                std::fprintf(stderr, "dtc autotune:   %s level %d candidate gridDim.y %d: %.1f us\n", what, level, c,
                             1e3f * ms / 3.f);
            if (ms < best_ms) { best_ms = ms; best = c; }
        }
        if (log)
            std::fprintf(stderr, "dtc autotune: %s level %d, %d bricks: gridDim.y %d (%.1f us)\n", what, level, nbr,
                         best, 1e3f * best_ms / 3.f);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[std::make_tuple(kind, variant, nbr)] = best;
    };
    float* p = b.p32.as<float>();
    float* q = b.q32.as<float>();
    measure(TK_MV, b.mf_cg ? 3 : 0, b.n_bricks, "Q = K P", 0, [&](int c) {
        return launch_mv<float, IN_V, OUT_MASKED>(b, wk, p, q, nullptr, nullptr, c);
    });
    const MfUse mfu{b.mf_smooth, b.kmf.as<uint4>(), b.kmf_max};
    for (int l = 0; l <= (int)b.mg.size() && rc == DTC_OK; ++l) {
        MvTopo T;
        float *x, *bb, *y, *d;
        const float* dinv;
        if (l == 0) {
            T = {b.n_bricks, b.brick.as<BrickInfo>(), b.blist.as<int>(), b.n_sparse, b.dom.as<DomInfo>(), b.halo.as<int>(),
                 b.nmask.as<unsigned char>(),
                 b.einfo.as<int>(), b.eE32.p};
            x = b.z32.as<float>(); bb = b.r32.as<float>(); y = b.t32.as<float>(); d = b.d32.as<float>();
            dinv = b.dinv32.as<float>();
        } else {
            Level& C = b.mg[l - 1];
            T = {C.n_bricks, C.brick.as<BrickInfo>(), C.blist.as<int>(), C.n_sparse, C.dom.as<DomInfo>(), C.halo.as<int>(),
                 C.nmask.as<unsigned char>(),
                 C.einfo.as<int>(), C.eE32.p};
            x = C.x.as<float>(); bb = C.b.as<float>(); y = C.t.as<float>(); d = C.d.as<float>();
            dinv = C.dinv32.as<float>();
        }
        if (!x || !bb || !y) continue;
        measure(TK_CHEB, mfu.terms, T.n_bricks, "Chebyshev step", l, [&](int c) {
            return launch_mv32<OUT_CHEB>(T, s, wk, x, y, nullptr, bb, dinv, d, 0.f, 1.f, nullptr, 0, nullptr, mfu, c);
        });
        measure(TK_BRES, mfu.terms, T.n_bricks, "restriction residual", l, [&](int c) {
            return launch_mv32<OUT_BRES>(T, s, wk, x, y, nullptr, bb, nullptr, nullptr, 0.f, 0.f, nullptr, 0, nullptr,
                                         mfu, c);
        });
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int batch_build_pmg(SolverBatch& b, SolverBatch& c) { return build_pmg(b, c); }

// the launch-shape tuning of a mixed multigrid solve, ahead of it (after the
// multigrid hierarchy is built): sub-batches solved at the same time are tuned one
// after another before their solves start, so that no tuning launch overlaps
// another sub-batch's kernels (batch_solve then finds the choices cached)
int batch_autotune(SolverBatch& b) {
    if (b.precision != DTC_FP32_REFINE || b.precond != DTC_PC_MG || (b.micro20 && !b.p08)) return DTC_OK;
    return autotune_splits(b, cast_wk<float>(b.wk64));
}

int batch_solve(SolverBatch& b, const dtc_solve_opts& opts, std::vector<int64_t>& iters,
                std::vector<int64_t>& outer, std::vector<double>& relres) {
    const int D = b.n_dom;
    // with the multigrid, the 24 load cases of a domain iterate as one block
    // (DTC_BLOCK_PCG=0: one PCG per load case)
    const char* be = std::getenv("DTC_BLOCK_PCG");
    const bool block = b.precision == DTC_FP32_REFINE && b.precond == DTC_PC_MG && (!b.micro20 || b.p08) &&
                       !(be && be[0] == '0');
    // inner_tol outside (0, 1): automatic, per domain.  The block restarts at every
    // FP64 residual replacement, so it takes its FP32 phases as deep as FP32
    // allows: 1e-6 (two phases to 1e-8 instead of three) where the moduli of the
    // domain's solid voxels differ by at most 100x; high-contrast (greyscale)
    // domains stall in FP32 before that and use 1e-4, as does one PCG per load
    // case (it keeps its direction over a replacement)
    const dtc_solve_opts& o = opts;
    if (D > 0) {
        std::vector<double> it2(D);
        for (int d = 0; d < D; ++d) {
            double t = o.inner_tol;
            if (!(t > 0.0 && t < 1.0)) {
                const double c = d < (int)b.h_contrast.size() ? b.h_contrast[d] : 1.0;
                t = block && c <= 100.0 ? 1e-6 : 1e-4;
            }
            it2[d] = t * t;
        }
        DTC_TRY(b.itol2.ensure((size_t)D * 8));
        DTC_HIP(hipMemcpyAsync(b.itol2.p, it2.data(), (size_t)D * 8, hipMemcpyHostToDevice, b.stream));
        DTC_HIP(hipStreamSynchronize(b.stream));
    }
    iters.assign(D, 0);
    outer.assign(D, 0);
    relres.assign(D, 0.0);
    if (D == 0 || b.n_bricks == 0) return DTC_OK;
    hipStream_t s = b.stream;
    const int64_t nv = b.n_nodes * 3 * R;
    const WK<double> wk64 = cast_wk<double>(b.wk64);
    const WK<float> wk32 = cast_wk<float>(b.wk64);
    const int64_t cap = o.max_iter > 0 ? o.max_iter : 100000;
    const BrickInfo* bl = b.brick.as<BrickInfo>();
    const DomInfo* dl = b.dom.as<DomInfo>();
    k_fill_i64<<<nb(D, 64), 64, 0, s>>>(b.iters.as<int64_t>(), D, 0);
    k_fill_f64<<<nb(nv, 256), 256, 0, s>>>(b.x64.as<double>(), nv, 0.0);
    if (b.precision == DTC_FP64) {
        k_fill_f64<<<nb(nv, 256), 256, 0, s>>>(b.p64.as<double>(), nv, 0.0);
        k_fill_f64<<<nb(nv, 256), 256, 0, s>>>(b.q64.as<double>(), nv, 0.0);
    } else {
        k_fill_f32<<<nb(nv, 256), 256, 0, s>>>(b.p32.as<float>(), nv, 0.f);
        k_fill_f32<<<nb(nv, 256), 256, 0, s>>>(b.q32.as<float>(), nv, 0.f);
    }
    DTC_HIP(hipGetLastError());
    for (KernelStat& k : b.kstat) k = KernelStat{};
    b.reactions_valid = false;
    b.graph_launches = 0;
    if (b.timing) {
        // per domain group: vector bytes (FP32), moduli (FP32), Jacobi scaling (FP32), vector
        // bytes (FP64), dense element-product flops 2 (3 nen)^2 R N_e
        std::vector<double> db((size_t)D * 5, 0.0);
        const double nd = 3.0 * (b.micro20 ? 20 : 8);
        for (int d = 0; d < D; ++d) {
            const double nn = (double)(b.h_dom[d].node_end - b.h_dom[d].node_begin);
            const double ne = b.h_status[d] == DTC_OK ? (double)b.h_elems[d] : 0.0;
            db[5 * d] = 3.0 * nn * R * 4.0;
            db[5 * d + 1] = 4.0 * ne;
            db[5 * d + 2] = 3.0 * nn * 4.0;
            db[5 * d + 3] = 3.0 * nn * R * 8.0;
            db[5 * d + 4] = 2.0 * nd * nd * R * ne;
        }
        DTC_TRY(b.dom_bytes.ensure(db.size() * 8));
        DTC_TRY(b.launch_bytes.ensure((size_t)MAX_SLOTS * 2 * 8));
        DTC_HIP(hipMemcpyAsync(b.dom_bytes.p, db.data(), db.size() * 8, hipMemcpyHostToDevice, s));
        DTC_HIP(hipStreamSynchronize(s));
        b.n_slots = 0;
    }
    std::vector<int64_t> nouter(D, 0);
    // high-contrast batches (solid moduli differing by more than 100x): the block
    // iterates on FP64 vectors with the FP32 V-cycle (DTC_BLOCK64=1 / 0 overrides)
    {
        double cmax = 1.0;
        for (double c : b.h_contrast) cmax = std::fmax(cmax, c);
        const char* e64 = std::getenv("DTC_BLOCK64");
        b.block64 = block && (e64 ? e64[0] == '1' : cmax > 100.0);
        if (b.block64) DTC_TRY(b.q64.ensure((size_t)nv * 8));
    }
    DTC_TRY(residual64(b, wk64, nullptr, 1, o));          // r64 = f (lifted right-hand side)

    if (b.precision == DTC_FP64) {
        // PCG in FP64 on (x64, r64); a true-residual check restarts it if the
        // recursive residual drifted
        for (int it = 0; it < 16; ++it) {
            bool left = false;
            DTC_TRY(apply_cap(b, cap, nouter, &left));
            if (!left) break;
            double* x = b.x64.as<double>();
            double* r = b.r64.as<double>();
            k_cg_init<double><<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), b.active.as<int>(), x, r,
                                                            b.p64.as<double>(), b.dinv64.as<double>(),
                                                            b.part_a.as<double>(), b.part_b.as<double>(), 0, 1, nullptr,
                                                            nullptr);
            k_red_init<<<D, RED_NT, 0, s>>>(dl, b.done64.as<int>(), b.part_a.as<double>(), b.part_b.as<double>(),
                                            b.thr2.as<double>(), b.active.as<int>(), b.rz.as<double>(), b.done.as<int>());
            DTC_HIP(hipGetLastError());
            DTC_TRY(pcg_run<double>(b, wk64, x, r, b.p64.as<double>(), b.q64.as<double>(), b.dinv64.as<double>(), cap,
                                    INT64_MAX, nullptr));
            DTC_TRY(residual64(b, wk64, b.x64.as<double>(), 0, o));
        }
    } else {
        // FP32 PCG with FP64 residual replacement: the FP32 correction x32 is
        // folded into x64 when the recursive residual has dropped by inner_tol
        // (or meets the final target, or after REPLACE iterations), r32 is
        // replaced by the true FP64 residual and the search direction p32 is
        // kept.
        const bool mg = b.precond == DTC_PC_MG && (!b.micro20 || b.p08);
        Progress progress;   // domains that stop converging are given up (the caller repeats them in FP64)
        const int64_t REPLACE = mg ? 100 : 1000;
        float* x32 = b.x32.as<float>();
        float* r32 = b.r32.as<float>();
        float* p32 = b.p32.as<float>();
        float* z32 = mg ? b.z32.as<float>() : nullptr;
        // the block's A P in FP64 arithmetic (FP32 vectors) for HEX20 micro meshes,
        // in FP32 for HEX08 (DTC_BLOCK_Q64=1 / 0 overrides).  HEX20 rows sum up to
        // 480 products of mixed sign; in FP32 that sum loses the small eigenvalues
        // of W = P^T A P, and on meshes with corner- or edge-hinged voxel clusters
        // (near-mechanisms) the block then drifts along them and diverges.
        const char* bq64 = std::getenv("DTC_BLOCK_Q64");
        b.block_q64 = bq64 ? bq64[0] == '1' : b.micro20;
        const char* brs = std::getenv("DTC_BLOCK_RESTART");
        const bool block_restart = !(brs && brs[0] == '0');
        // a block phase ends at its targets; the cap only bounds FP32 drift
        const char* brp = std::getenv("DTC_BLOCK_REPLACE");
        const int64_t block_replace = brp ? std::max(1, std::atoi(brp)) : 400;
        // FP64 vectors: the recursion drifts far less, and a restart discards the
        // block's Krylov space, so a phase runs to its target (the true residual
        // is checked then)
        const int64_t block_replace64 = brp ? block_replace : 20000;
        float* bp = p32;
        float* bq = b.q32.as<float>();
        float* bz = z32;
        if (block) DTC_TRY(block_setup(b));
        if (mg) DTC_TRY(autotune_splits(b, wk32));
        bool first = true;
        for (int it = 0; it < 100000; ++it) {
            bool left = false;
            DTC_TRY(apply_cap(b, cap, nouter, &left, &progress));
            if (!left) break;
            // the block solves the correction problem with unit-norm columns: FP32
            // directions mix all 24 load cases, whose residuals drift apart in size
            k_to32<<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), b.r64.as<double>(), r32, x32,
                                                 block ? b.rr.as<double>() : nullptr);
            if (block && b.block64) {
                // FP64 vectors: the block restarts from M r at every true-residual check
                k_to32<<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), b.r64.as<double>(), r32, nullptr, nullptr);
                DTC_HIP(hipGetLastError());
                DTC_TRY(mg_vcycle(b, wk32, r32, bz, nullptr));
                DTC_TRY(block_orth(b, bz, bp));
                first = false;
                DTC_TRY(block_run(b, wk32, x32, r32, bp, bq, bz, cap, block_replace64));
                DTC_TRY(residual64(b, wk64, b.x64.as<double>(), 0, o));
                continue;
            }
            if (block) {
                // the direction block restarts from M r at every residual replacement
                // (DTC_BLOCK_RESTART=0: it is kept, spanning the same space after the
                // column scaling)
                if (first || block_restart) {
                    DTC_TRY(mg_vcycle(b, wk32, r32, bz, nullptr));
                    DTC_TRY(block_orth(b, bz, bp));
                }
                first = false;
                DTC_TRY(block_run(b, wk32, x32, r32, bp, bq, bz, cap, block_replace));
                k_acc64<<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), x32, b.x64.as<double>(), b.rr.as<double>());
                DTC_HIP(hipGetLastError());
                DTC_TRY(residual64(b, wk64, b.x64.as<double>(), 0, o));
                continue;
            }
            if (mg) DTC_TRY(mg_vcycle(b, wk32, r32, z32, nullptr));
            k_cg_init<float><<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), b.active.as<int>(), x32, r32, p32,
                                                           b.dinv32.as<float>(), b.part_a.as<double>(),
                                                           b.part_b.as<double>(), 0, first ? 1 : 0,
                                                           b.active_prev.as<int>(), z32);
            k_red_init<<<D, RED_NT, 0, s>>>(dl, b.done64.as<int>(), b.part_a.as<double>(), b.part_b.as<double>(),
                                            b.thr2.as<double>(), b.active.as<int>(), b.rz.as<double>(), b.done.as<int>());
            DTC_HIP(hipGetLastError());
            first = false;
            DTC_TRY(pcg_run<float>(b, wk32, x32, r32, p32, b.q32.as<float>(), b.dinv32.as<float>(), cap, REPLACE,
                                   z32));
            DTC_HIP(hipMemcpyAsync(b.active_prev.p, b.active.p, (size_t)D * R * 4, hipMemcpyDeviceToDevice, s));
            k_acc64<<<b.n_bricks, NT_UPD, 0, s>>>(bl, b.done64.as<int>(), x32, b.x64.as<double>(), nullptr);
            DTC_HIP(hipGetLastError());
            DTC_TRY(residual64(b, wk64, b.x64.as<double>(), 0, o));
        }
    }
    // results: iterations, true relative residual, convergence status
    std::vector<int64_t> hi(D);
    std::vector<double> hrr(D * R), hr0(D * R);
    DTC_HIP(hipMemcpyAsync(hi.data(), b.iters.p, (size_t)D * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipMemcpyAsync(hrr.data(), b.rr.p, (size_t)D * R * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipMemcpyAsync(hr0.data(), b.rr0.p, (size_t)D * R * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    DTC_TRY(harvest_events(b));
    for (int d = 0; d < D; ++d) {
        iters[d] = hi[d];
        outer[d] = nouter[d];
        double mx = 0.0;
        bool finite = true;
        for (int lc = 0; lc < R; ++lc) {
            const double r0 = hr0[d * R + lc], rr = hrr[d * R + lc];
            if (!std::isfinite(r0) || !std::isfinite(rr)) finite = false;
            else if (r0 > 0.0) mx = std::fmax(mx, std::sqrt(rr / r0));
        }
        relres[d] = finite ? mx : NAN;
        if (b.h_status[d] == DTC_OK && !(finite && mx <= o.rel_tol)) b.h_status[d] = DTC_ERR_NO_CONVERGENCE;
    }
    return DTC_OK;
}

int batch_forces(SolverBatch& b, std::vector<double>& ff) {
    const int D = b.n_dom;
    ff.assign((size_t)D * R * FFD, 0.0);
    if (D == 0 || b.n_bricks == 0) return DTC_OK;
    const WK<double> wk64 = cast_wk<double>(b.wk64);
    // reactions on the unmodified operator: rf = K (x + x_bc), stored in r64
    DTC_TRY((launch_mv<double, IN_U, OUT_PLAIN>(b, wk64, b.x64.as<double>(), b.r64.as<double>(), nullptr, nullptr)));
    if (b.micro20) {
        DTC_TRY(project20(b));
    } else {
        k_project<8><<<b.n_bricks, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                          b.lat2node.as<int>(), b.r64.as<double>(),
                                                          b.part_ff.as<double>());
        k_project<20><<<b.n_bricks, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                           b.lat2node.as<int>(), b.r64.as<double>(),
                                                           b.part_ff.as<double>());
    }
    k_red_ff<<<dim3(D, (R * FFD + RFF_E - 1) / RFF_E), RFF_E * RFF_P, 0, b.stream>>>(b.dom.as<DomInfo>(),
                                                                                b.part_ff.as<double>(), b.ff.as<double>());
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipMemcpyAsync(ff.data(), b.ff.p, ff.size() * 8, hipMemcpyDeviceToHost, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    b.reactions_valid = true;
    return DTC_OK;
}

int batch_fields(SolverBatch& b, int i, int ncol, double* u_out, double* r_out) {
    if (i < 0 || i >= b.n_dom) return set_error(DTC_ERR_BAD_ARGUMENT, "domain index %d", i);
    const int64_t nn = b.h_nodes[i];
    if (nn == 0) return DTC_OK;
    if (r_out && !b.reactions_valid) return set_error(DTC_ERR_BAD_ARGUMENT, "reactions requested before the force pass");
    DevMem du, dr;
    if (ncol < 1 || ncol > R) return set_error(DTC_ERR_BAD_ARGUMENT, "field columns %d", ncol);
    const size_t bytes = (size_t)nn * 3 * ncol * 8;
    if (u_out) DTC_TRY(du.alloc(bytes));
    if (r_out) DTC_TRY(dr.alloc(bytes));
    if (b.micro20) {
        DTC_TRY(fields20(b, i, ncol, u_out ? du.as<double>() : nullptr, r_out ? dr.as<double>() : nullptr));
    } else {
        k_fields<<<b.n_bricks, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                      b.lat2node.as<int>(), b.sol2ref.as<int>(), b.x64.as<double>(),
                                                      b.r64.as<double>(), i, nn, ncol, u_out ? du.as<double>() : nullptr,
                                                      r_out ? dr.as<double>() : nullptr);
        DTC_HIP(hipGetLastError());
    }
    if (u_out) DTC_HIP(hipMemcpyAsync(u_out, du.p, bytes, hipMemcpyDeviceToHost, b.stream));
    if (r_out) DTC_HIP(hipMemcpyAsync(r_out, dr.p, bytes, hipMemcpyDeviceToHost, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    return DTC_OK;
}

int batch_grad(SolverBatch& b, int i, int ng, const double* M, double* dL_dE_host) {
    if (i < 0 || ng < 1 || i + ng > b.n_dom) return set_error(DTC_ERR_BAD_ARGUMENT, "domain groups %d..%d", i, i + ng - 1);
    if (b.micro20) return grad20(b, i, ng, M, dL_dE_host);
    const DomInfo& D = b.h_dom[i];
    const int64_t nvox = (int64_t)D.nx * D.ny * D.nz;
    const int64_t nn = b.n_nodes > 0 ? b.n_nodes : 1;
    DevMem h, Md, g;
    DTC_TRY(h.alloc((size_t)nn * 18 * 8));
    DTC_TRY(Md.alloc((size_t)ng * R * 6 * 8));
    DTC_TRY(g.alloc((size_t)nvox * 8));
    DTC_HIP(hipMemcpyAsync(Md.p, M, (size_t)ng * R * 6 * 8, hipMemcpyHostToDevice, b.stream));
    for (int k = 0; k < ng; ++k)
        if (b.h_dom[i + k].brick_end > b.h_dom[i + k].brick_begin)
            k_grad_h<<<b.n_bricks, NT_UPD, 0, b.stream>>>(b.brick.as<BrickInfo>(), b.dom.as<DomInfo>(),
                                                          b.lat2node.as<int>(), b.x64.as<double>(),
                                                          Md.as<double>() + (size_t)k * R * 6, i + k, D.node_begin,
                                                          k > 0, h.as<double>());
    k_grad_elem<<<nb(nvox, 128), 128, 0, b.stream>>>(b.dom.as<DomInfo>(), b.lat2node.as<int>(), b.emod.as<double>(),
                                                     b.x64.as<double>(), h.as<double>(), cast_wk<double>(b.wk64), i,
                                                     g.as<double>());
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipMemcpyAsync(dL_dE_host, g.p, (size_t)nvox * 8, hipMemcpyDeviceToHost, b.stream));
    DTC_HIP(hipStreamSynchronize(b.stream));
    return DTC_OK;
}

}  // namespace dtc
