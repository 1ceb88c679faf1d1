// matvec_mf.hpp -- the FP32 EbE matvec with the element products on the matrix
// cores (v_mfma_f32_16x16x32_f16).  Included by solver.hip inside its
// anonymous namespace, after k_matvec, whose tables, LDS corner buffers,
// gather and output modes it shares.
//
// Element products as small GEMMs: for a group of 4 tile elements and the RC = 4
// load cases of a chunk, D (16 x 32) = A (16 x 32) B (32 x 32), with
//   A[(e, r)][4 b + d] = x of corner b, direction d, load case r of element e
//                        (d = 3 is a zero pad), and
//   B[4 b' + d'][4 b + d] = Khat(3 b + d, 3 b' + d') / kmax (the E = 1 element
//                        stiffness in corner-bit order, normalised),
// so that D[(e, r)][4 b + d] * E_e * kmax is the force of corner b, direction d.
// Two 16 x 32 x 16 column tiles, each one MFMA per operand term.
//
// Precision: x is split per (brick, chunk, load case) into a scaled FP16 pair
// x s = hi + lo (s a power of two putting max |x s| in [2^13, 2^14), so hi
// and, for all but values 2^-18 below the brick's largest, lo are normal FP16
// numbers), and Khat / kmax likewise.  TERMS = 3 sums hi.hi + lo.hi + hi.lo in
// FP32 (error about 2^-21 relative to the brick's largest |x| times |Khat|: the
// FP32 operator's own rounding level); TERMS = 1 keeps hi.hi (2^-11, for the
// multigrid smoother).  The products are exact in the MFMA and accumulate in
// FP32, so the operator is deterministic.
//
// MFMA lane layout (16x16x32 f16, 64 lanes, l = lane): A[i][k] in lane
// i + 16 (k / 8), half k % 8; B[k][j] in lane j + 16 (k / 8), half k % 8;
// D[i][j] in lane j + 16 (i / 4), float i % 4.

typedef _Float16 mf_half8 __attribute__((ext_vector_type(8)));
typedef float mf_float4 __attribute__((ext_vector_type(4)));

constexpr int MF_RC = 4;   // load cases per chunk (one float4 per (node, direction))

// B operands of one lane: [tile][term] 8 halves as a uint4
struct MfB {
    uint4 w[2][2];
};

__device__ inline mf_half8 mf_as_half8(uint4 v) {
    mf_half8 h;
    __builtin_memcpy(&h, &v, 16);
    return h;
}

__device__ inline mf_half8 mf_pack(uint2 a, uint2 b) {
    uint4 v = make_uint4(a.x, a.y, b.x, b.y);
    return mf_as_half8(v);
}

// y = K x (FP32 vectors, IN_V) for every brick: k_matvec's work split, tables,
// corner buffers, gather and output modes, with the element products on MFMA.
// PART (OUT_CHEB only): per-brick sums of b . x are written (A.partial set); the
// dot-product LDS exists only then, so that the plain step fits two workgroups per CU.
template <int OUT, int TERMS, bool PART = false>
__global__ __launch_bounds__(MV_NT, 4) void k_matvec_mf(MvArgs A, const uint4* __restrict__ kmf, float kmax,
                                                     const float* __restrict__ xin, float* __restrict__ yout) {
    constexpr int RC = MF_RC, NCH = R / RC;
    using V = float4;
    constexpr int NLD = (NH * 3 + MV_NT - 1) / MV_NT;
    constexpr int NW = MV_NT / 64;
    constexpr bool DOT = OUT == OUT_MASKED_DOT || OUT == OUT_RESID;
    static_assert(OUT == OUT_MASKED || OUT == OUT_MASKED_DOT || OUT == OUT_CHEB || OUT == OUT_BRES, "mode");
    constexpr bool SUMS = DOT || (OUT == OUT_CHEB && PART);
    constexpr int NRED = SUMS ? NCH : 1;   // per-brick dot products (LDS only where used)
    __shared__ int s_node[NH];
    __shared__ int s_info[NET];
    __shared__ float s_E[NET];
    __shared__ unsigned char s_nm[NO];
    __shared__ double s_wred[NRED][NW * 4][RC];
    // x of the chunk as scaled FP16 pairs, [halo position][load case] x 4 directions (pad 0)
    __shared__ uint2 s_xh[NH * RC];
    __shared__ uint2 s_xl[TERMS == 3 ? NH * RC : 1];
    __shared__ V s_c[(OB_BIAS + 8 * NO) * 3];
    __shared__ short s_hn[NH];
    __shared__ unsigned char s_ol[NO];
    __shared__ int s_wcnt[NW];
    __shared__ float s_wmax[NW][RC];

    const BrickInfo B = A.bricks[blockIdx.x];
    if (A.done && A.done[B.dom]) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int* __restrict__ halo = A.halo + (int64_t)blockIdx.x * NH;
    const V* xv = reinterpret_cast<const V*>(xin);
    V* yv = reinterpret_cast<V*>(yout);
    const int c0 = (int)blockIdx.y * NCH / (int)gridDim.y, c1 = ((int)blockIdx.y + 1) * NCH / (int)gridDim.y;
    const int n_own = B.node_end - B.node_begin;

    // B operands of this lane (the same for every element)
    MfB kb;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int q = 0; q < 2; ++q) kb.w[t][q] = kmf[(t * 2 + q) * 64 + lane];

    // halo staging list (compact over the positions that hold a node), owned positions
    int nh;
    {
        const int q = tid < NH ? halo[tid] : -1;
        if (tid < NH) s_node[tid] = q;
        const bool ex = q >= 0;
        const int hx = tid % HX, hy = (tid / HX) % HY, hz = tid / (HX * HY);
        if (ex && hx >= 1 && hx <= BX && hy >= 1 && hy <= BY && hz >= 1 && hz <= BZ)
            s_ol[(q & NODE_MASK) - B.node_begin] = (unsigned char)((hx - 1) + BX * ((hy - 1) + BY * (hz - 1)));
        if (tid < NO) s_nm[tid] = A.nmask[(int64_t)blockIdx.x * NO + tid];
        const unsigned long long bal = __ballot(ex);
        if (lane == 0) s_wcnt[wave] = __popcll(bal);
        // FP16 planes start at zero: the direction pad and the positions without a node
        for (int i = tid; i < NH * RC; i += MV_NT) {
            s_xh[i] = make_uint2(0u, 0u);
            if (TERMS == 3) s_xl[i] = make_uint2(0u, 0u);
        }
        __syncthreads();
        int base = 0;
        nh = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const int c = s_wcnt[w];
            if (w < wave) base += c;
            nh += c;
        }
        if (ex) s_hn[base + __popcll(bal & ((1ull << lane) - 1))] = (short)tid;
    }
    const int ne = B.e_cnt;
    {
        const float* __restrict__ eE = static_cast<const float*>(A.eE) + B.e_off;
        const int* __restrict__ ei = A.einfo + B.e_off;
        for (int i = tid; i < ne; i += MV_NT) {
            s_info[i] = ei[i];
            s_E[i] = eE[i];
        }
    }
    __syncthreads();

    // staging items of this thread: (halo position, direction), x offset in float4s
    int pf_off[NLD], pf_h[NLD], pf_d[NLD];
    V reg[NLD];
#pragma unroll
    for (int k = 0; k < NLD; ++k) {
        const int i = tid + k * MV_NT;
        int o = -1, h = 0, d = 0;
        if (i < 3 * nh) {
            h = s_hn[i / 3];
            d = i % 3;
            o = ((s_node[h] & NODE_MASK) * 3 + d) * NCH;
        }
        pf_off[k] = o;
        pf_h[k] = h;
        pf_d[k] = d;
    }
    auto prefetch = [&](int chunk) {
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            V v = make_float4(0.f, 0.f, 0.f, 0.f);
            if (pf_off[k] >= 0) v = xv[pf_off[k] + chunk];
            reg[k] = v;
        }
    };
    // per-wave maxima of |x| of the prefetched chunk, per load case
    auto wave_max = [&]() {
        float m[RC] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            m[0] = fmaxf(m[0], fabsf(reg[k].x));
            m[1] = fmaxf(m[1], fabsf(reg[k].y));
            m[2] = fmaxf(m[2], fabsf(reg[k].z));
            m[3] = fmaxf(m[3], fabsf(reg[k].w));
        }
#pragma unroll
        for (int r = 0; r < RC; ++r) {
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m[r] = fmaxf(m[r], __shfl_xor(m[r], o));
        }
        if (lane == 0)
#pragma unroll
            for (int r = 0; r < RC; ++r) s_wmax[wave][r] = m[r];
    };
    // the chunk's scales from the wave maxima (after a barrier), then the FP16 planes
    float osc[RC];   // 1 / scale of each load case of the staged chunk
    auto stage = [&]() {
        float sc[RC];
#pragma unroll
        for (int r = 0; r < RC; ++r) {
            float m = 0.f;
#pragma unroll
            for (int w = 0; w < NW; ++w) m = fmaxf(m, s_wmax[w][r]);
            int e = 0;
            (void)frexpf(m, &e);                    // m < 2^e
            sc[r] = m > 0.f ? ldexpf(1.f, 14 - e) : 1.f;
            osc[r] = m > 0.f ? ldexpf(1.f, e - 14) : 1.f;
        }
        _Float16* xh = reinterpret_cast<_Float16*>(s_xh);
        _Float16* xl = reinterpret_cast<_Float16*>(s_xl);
#pragma unroll
        for (int k = 0; k < NLD; ++k) {
            if (pf_off[k] < 0) continue;
            const float v[RC] = {reg[k].x, reg[k].y, reg[k].z, reg[k].w};
#pragma unroll
            for (int r = 0; r < RC; ++r) {
                const float xs = v[r] * sc[r];
                const _Float16 hi = (_Float16)xs;
                const int at = (pf_h[k] * RC + r) * 4 + pf_d[k];
                xh[at] = hi;
                if (TERMS == 3) xl[at] = (_Float16)(xs - (float)hi);
            }
        }
    };

    // chunk-invariant gather items of this thread: (owned solver node, direction)
    constexpr int NG = (NO * 3 + MV_NT - 1) / MV_NT;
    int g_off[NG];
    unsigned g_m[NG];
    float e_di[NG];
#pragma unroll
    for (int j = 0; j < NG; ++j) {
        const int i = tid + j * MV_NT;
        int off = -1;
        unsigned m = 0;
        if (i < 3 * n_own) {
            const int o = s_ol[i / 3], d = i % 3, h = ohalo(o);
            const int q = s_node[h];
            off = ((q & NODE_MASK) * 3 + d) * NCH;
            m = (unsigned)s_nm[o] | ((q & BC_BIT) ? 256u : 0u) | ((unsigned)o << 9) | ((unsigned)d << 16);
        }
        g_off[j] = off;
        g_m[j] = m;
        e_di[j] = 0.f;
        if (OUT == OUT_CHEB && off >= 0 && !(m & 256u)) e_di[j] = reinterpret_cast<const float*>(A.dinv)[off / NCH];
    }
    // corner slots that no element writes stay zero for the whole launch
    for (int i = tid; i < n_own * 24; i += MV_NT) {
        const int o = s_ol[i / 24], b = (i / 3) % 8, d = i % 3;
        if (!((s_nm[o] >> b) & 1)) s_c[(OB_BIAS + b * NO + o) * 3 + d] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    prefetch(c0);
    wave_max();
    __syncthreads();
    stage();
    if (c0 + 1 < c1) prefetch(c0 + 1);
    __syncthreads();

    // element-product lanes: A row m = (element m / 4 of the group, load case m % 4),
    // k slice kq = corners 2 kq and 2 kq + 1; D: column m = (corner, direction) of a
    // tile, rows 4 kq .. 4 kq + 3 = the load cases of element kq of the group
    const int m16 = lane & 15, kq = lane >> 4, er = m16 >> 2, rr = m16 & 3;
    const int hoff = ((kq & 1) ? HX : 0) + ((kq & 2) ? HX * HY : 0);   // halo offset of corner 2 kq
    const int ng = (ne + 3) >> 2;
    const int dcol = m16 & 3, bcol = m16 >> 2;                        // tile 0: corner bcol, tile 1: 4 + bcol

    for (int chunk = c0; chunk < c1; ++chunk) {
        // gather operands of this chunk from memory (x, b, d)
        V e_b[NG], e_d[NG], e_x[NG];
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            e_b[j] = e_d[j] = e_x[j] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (g_off[j] >= 0) {
                const int vi = g_off[j] + chunk;
                if (OUT == OUT_BRES || OUT == OUT_CHEB) e_b[j] = reinterpret_cast<const V*>(A.bvec)[vi];
                if (OUT == OUT_CHEB && (A.cflags & CH_READ_D)) e_d[j] = reinterpret_cast<const V*>(A.dvec)[vi];
                if (OUT == OUT_MASKED_DOT || OUT == OUT_CHEB) e_x[j] = xv[vi];
            }
        }
        // maxima of the next chunk (its loads were issued one phase ago)
        if (chunk + 1 < c1) wave_max();

        // ---- element products, one group of 4 elements per wave step
        float isc[RC];
#pragma unroll
        for (int r = 0; r < RC; ++r) isc[r] = osc[r] * kmax;
#pragma unroll 1
        for (int g = wave; g < ng; g += NW) {
            int e = g * 4 + er;
            e = e < ne ? e : ne - 1;
            const int hb = (s_info[e] & 511) + hoff;
            const uint2 h0 = s_xh[hb * RC + rr], h1 = s_xh[(hb + 1) * RC + rr];
            const mf_half8 ahi = mf_pack(h0, h1);
            mf_float4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
            d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, mf_as_half8(kb.w[0][0]), d0, 0, 0, 0);
            d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, mf_as_half8(kb.w[1][0]), d1, 0, 0, 0);
            if (TERMS == 3) {
                const uint2 l0 = s_xl[hb * RC + rr], l1 = s_xl[(hb + 1) * RC + rr];
                const mf_half8 alo = mf_pack(l0, l1);
                d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, mf_as_half8(kb.w[0][0]), d0, 0, 0, 0);
                d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(alo, mf_as_half8(kb.w[1][0]), d1, 0, 0, 0);
                d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, mf_as_half8(kb.w[0][1]), d0, 0, 0, 0);
                d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(ahi, mf_as_half8(kb.w[1][1]), d1, 0, 0, 0);
            }
            const int eo = g * 4 + kq;
            if (eo < ne && dcol < 3) {
                const int info = s_info[eo];
                const float Ee = s_E[eo];
                const int ob = (info >> 9) & 255;
                const unsigned mask = (unsigned)info >> 24;
                const V f0 = make_float4(d0[0] * Ee * isc[0], d0[1] * Ee * isc[1], d0[2] * Ee * isc[2],
                                         d0[3] * Ee * isc[3]);
                const V f1 = make_float4(d1[0] * Ee * isc[0], d1[1] * Ee * isc[1], d1[2] * Ee * isc[2],
                                         d1[3] * Ee * isc[3]);
                const int b0 = bcol, b1 = 4 + bcol;
                if (mask & (1u << b0)) s_c[(ob + b0 * NO + ocorner(b0)) * 3 + dcol] = f0;
                if (mask & (1u << b1)) s_c[(ob + b1 * NO + ocorner(b1)) * 3 + dcol] = f1;
            }
        }
        __syncthreads();
        if (chunk + 1 < c1) {
            stage();
            if (chunk + 2 < c1) prefetch(chunk + 2);
        }

        // ---- gather: thread per owned (node, direction), the RC load cases in one vector
        double acc[RC];
#pragma unroll
        for (int w = 0; w < RC; ++w) acc[w] = 0.0;
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            if (g_off[j] < 0) continue;
            const unsigned m = g_m[j];
            const int od = (int)(m >> 9);
            const int o = od & 127, d = od >> 7;
            V y = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
            for (int b = 0; b < 8; ++b) vadd(y, s_c[(OB_BIAS + b * NO + o) * 3 + d]);
            const bool bc = (m & 256u) != 0;
            const int vi = g_off[j] + chunk;
            if (OUT == OUT_MASKED_DOT) {
                if (bc) y = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
                for (int w = 0; w < RC; ++w) acc[w] += (double)vcomp(e_x[j], w) * (double)vcomp(y, w);
            } else if (OUT == OUT_MASKED) {
                if (bc) y = make_float4(0.f, 0.f, 0.f, 0.f);
            } else if (OUT == OUT_BRES) {
                if (bc) y = make_float4(0.f, 0.f, 0.f, 0.f);
                else y = make_float4(e_b[j].x - y.x, e_b[j].y - y.y, e_b[j].z - y.z, e_b[j].w - y.w);
            } else {   // OUT_CHEB
                V dn = make_float4(0.f, 0.f, 0.f, 0.f), xn = make_float4(0.f, 0.f, 0.f, 0.f);
                if (!bc) {
                    const float di = e_di[j];
                    const V xq = e_x[j];
                    const V dv = (A.cflags & CH_D_FROM_X) ? xq : e_d[j];
                    float* dnp = reinterpret_cast<float*>(&dn);
                    float* xnp = reinterpret_cast<float*>(&xn);
#pragma unroll
                    for (int w = 0; w < RC; ++w) {
                        const float rres = vcomp(e_b[j], w) - vcomp(y, w);
                        dnp[w] = A.c1 * vcomp(dv, w) + A.c2 * di * rres;
                        xnp[w] = vcomp(xq, w) + dnp[w];
                    }
                }
                if (A.cflags & CH_WRITE_D) reinterpret_cast<V*>(A.dvec)[vi] = dn;
                y = xn;
                if (PART)
#pragma unroll
                    for (int w = 0; w < RC; ++w) acc[w] += (double)vcomp(e_b[j], w) * (double)vcomp(y, w);
            }
            yv[vi] = y;
        }
        if (SUMS) {
            const int row = wave * 4 + (lane >> 4);
            double v = 0.0;
            int lc = lane & (RC - 1);
            bool out = (lane & 15) >= 16 - RC;
            if ((tid & ~63) < 3 * n_own) out = row_sums<RC>(acc, lane, &v, &lc);
            if (out) s_wred[chunk][row][lc] = v;
        }
        __syncthreads();
    }
    if (SUMS && tid < R && tid / RC >= c0 && tid / RC < c1) {
        const int ch = tid / RC, r = tid % RC;
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < NW * 4; ++w) s += s_wred[ch][w][r];
        A.partial[(int64_t)blockIdx.x * R + tid] = s;
    }
}
