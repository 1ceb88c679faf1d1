// capi.cpp -- the extern "C" entry points of libdtc_amd.so (include/dtc_capi.h).
//
// Host-side pieces here are the O(24^3) fit of calc_effective_material_parameters
// (f-src/mod_struct_calcmat.f90:216-473): the load-case matrix vv and its
// inverse, K_M = ff vv^-1, the face-averaged forces fv, C = fv meps and
// EE = (C + C^T)/2.  Everything proportional to the mesh size runs on the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "bc.hpp"
#include "dtc_internal.hpp"
#include "io.hpp"
#include "rotation.hpp"
#include "solver.hpp"

namespace dtc {

static thread_local std::string g_err;

int set_error(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

constexpr int R = R24;

struct Stream {
    hipStream_t s = nullptr;
    ~Stream() { if (s) (void)hipStreamDestroy(s); }
    int create() {
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) return set_error(DTC_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
        return DTC_OK;
    }
};

int kind_size(int kind) { return kind == DTC_IK1 ? 1 : kind == DTC_IK2 ? 2 : kind == DTC_IK4 ? 4 : 0; }
// signed voxel value v of a host block of kind ik1/ik2/ik4
int64_t voxel_value(const void* phi, int kind, size_t v) {
    return kind == DTC_IK1 ? (int64_t) static_cast<const int8_t*>(phi)[v]
         : kind == DTC_IK2 ? (int64_t) static_cast<const int16_t*>(phi)[v]
                           : (int64_t) static_cast<const int32_t*>(phi)[v];
}

int check_device() {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return set_error(DTC_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
    return DTC_OK;
}

// ---- load-case tables (init_displ_hexe8 / init_loadcase_el_order_lin)
// vv(dof, lc) with dof = d*8 + l (f-src/mod_struct_calcmat.f90:1684-1770); the
// entries are the same (d, l, m) list as the boundary table.
void vv_lin(double rve_strain, double vv[576]) {
    std::memset(vv, 0, sizeof(double) * 576);
    for (int lc = 0; lc < R; ++lc) {
        int cnt;
        const UUEntry* e = uu_entries(lc, &cnt);
        for (int k = 0; k < cnt; ++k) vv[(e[k].d * 8 + e[k].l) + R * lc] = (double)e[k].m * rve_strain;
    }
}

// HEX20 macro load-case matrix vv(dof, lc) = uu(l, d, lc) * RVE_STRAIN / 1e-6, dof = d*20 + l.
// The reference's own table (init_loadcase_el_order_quad, f-src/mod_struct_calcmat.f90:
// 1780-2028) leaves row 49 unassigned and holds unscaled literals, so it is
// singular; K_M of the HEX20 macro element is formed with this matrix instead.
void vv_quad(double rve_strain, double* vv) {
    std::memset(vv, 0, sizeof(double) * 3600);
    for (int lc = 0; lc < 60; ++lc) {
        int cnt;
        const UU20Entry* e = uu20_entries(lc, &cnt);
        for (int k = 0; k < cnt; ++k) vv[(e[k].d * 20 + e[k].l) + 60 * lc] = e[k].m * rve_strain;
    }
}

int invert(double* A, int n) {
    std::vector<double> M(A, A + n * n), I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i + n * i] = 1.0;
    for (int k = 0; k < n; ++k) {
        int pr = k;
        double pv = std::fabs(M[k + n * k]);
        for (int i = k + 1; i < n; ++i)
            if (std::fabs(M[i + n * k]) > pv) { pv = std::fabs(M[i + n * k]); pr = i; }
        if (pv == 0.0) return set_error(DTC_ERR_BAD_ARGUMENT, "singular load-case matrix");
        if (pr != k)
            for (int j = 0; j < n; ++j) { std::swap(M[k + n * j], M[pr + n * j]); std::swap(I[k + n * j], I[pr + n * j]); }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            double f = M[i + n * k] / M[k + n * k];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) { M[i + n * j] -= f * M[k + n * j]; I[i + n * j] -= f * I[k + n * j]; }
        }
    }
    for (int k = 0; k < n; ++k) {
        double d = M[k + n * k];
        for (int j = 0; j < n; ++j) I[k + n * j] /= d;
    }
    std::memcpy(A, I.data(), sizeof(double) * n * n);
    return DTC_OK;
}

// fv = W ff(:, 1:6), HEX08 macro (f-src/mod_struct_calcmat.f90:345-359)
void fv_weights(const double xdp[3], double W[6 * 24]) {
    std::memset(W, 0, sizeof(double) * 6 * 24);
    const double A23 = xdp[1] * xdp[2], A13 = xdp[0] * xdp[2], A12 = xdp[0] * xdp[1];
    auto set = [&](int i, std::initializer_list<int> rows, double w) { for (int r : rows) W[i + 6 * (r - 1)] += w; };
    set(0, {2, 3, 6, 7}, 1.0 / A23);
    set(1, {11, 12, 15, 16}, 1.0 / A13);
    set(2, {21, 22, 23, 24}, 1.0 / A12);
    set(3, {3, 4, 7, 8}, 0.5 / A13);
    set(3, {10, 11, 14, 15}, 0.5 / A23);
    set(4, {5, 6, 7, 8}, 0.5 / A12);
    set(4, {18, 19, 22, 23}, 0.5 / A23);
    set(5, {13, 14, 15, 16}, 0.5 / A12);
    set(5, {19, 20, 23, 24}, 0.5 / A13);
}

// the same for the HEX20 macro element (f-src/mod_struct_calcmat.f90:365-401): 6 x 60
void fv_weights_quad(const double xdp[3], double W[6 * 60]) {
    std::memset(W, 0, sizeof(double) * 6 * 60);
    const double A23 = xdp[1] * xdp[2], A13 = xdp[0] * xdp[2], A12 = xdp[0] * xdp[1];
    auto set = [&](int i, std::initializer_list<int> rows, double w) { for (int r : rows) W[i + 6 * (r - 1)] += w; };
    set(0, {2, 3, 6, 7, 10, 14, 18, 19}, 1.0 / A23);
    set(1, {23, 24, 27, 28, 31, 35, 39, 40}, 1.0 / A13);
    set(2, {45, 46, 47, 48, 53, 54, 55, 56}, 1.0 / A12);
    set(3, {3, 4, 7, 8, 11, 15, 19, 20}, 0.5 / A13);
    set(3, {22, 23, 26, 27, 30, 34, 38, 39}, 0.5 / A23);
    set(4, {5, 6, 7, 8, 13, 14, 15, 16}, 0.5 / A12);
    set(4, {42, 43, 46, 47, 50, 54, 58, 59}, 0.5 / A23);
    set(5, {25, 26, 27, 28, 33, 34, 35, 36}, 0.5 / A12);
    set(5, {43, 44, 47, 48, 51, 55, 59, 60}, 0.5 / A13);
}

constexpr double MEPS[6] = {1.e6, 1.e6, 1.e6, -1.e6, -1.e6, 1.e6};   // :434-441

// ff, K_M: column-major RM x RM, RM = 24 (macro order 1) or 60 (order 2)
int effective(const double* ff, int order, const double xdp[3], double rve_strain, double* K_M, double* C, double* EE) {
    const int RM = macro_lcs(order);
    std::vector<double> vv(3600);
    if (order == 1) vv_lin(rve_strain, vv.data());
    else vv_quad(rve_strain, vv.data());
    DTC_TRY(invert(vv.data(), RM));
    for (int i = 0; i < RM; ++i)
        for (int j = 0; j < RM; ++j) {
            double s = 0.0;
            for (int k = 0; k < RM; ++k) s += ff[i + RM * k] * vv[k + RM * j];
            K_M[i + RM * j] = s;
        }
    double W[6 * 60], fv[36];
    if (order == 1) fv_weights(xdp, W);
    else fv_weights_quad(xdp, W);
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int a = 0; a < RM; ++a) s += W[i + 6 * a] * ff[a + RM * j];
            fv[i + 6 * j] = s;
        }
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) C[i + 6 * j] = fv[i + 6 * j] * MEPS[j];
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) EE[i + 6 * j] = 0.5 * (C[i + 6 * j] + C[j + 6 * i]);
    return DTC_OK;
}

// HEX20 element stiffness (3x3x3 Gauss) in gen_quadmesh node order
int element_stiffness20(double E, double nu, double a, double* K) {
    static const int OFF[20][3] = {
        {0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {2, 1, 0}, {2, 2, 0}, {1, 2, 0}, {0, 2, 0}, {0, 1, 0},
        {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
        {0, 0, 2}, {1, 0, 2}, {2, 0, 2}, {2, 1, 2}, {2, 2, 2}, {1, 2, 2}, {0, 2, 2}, {0, 1, 2}};
    const int nd = 60;
    double D[6][6] = {};
    const double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D[i][j] = lam;
    for (int i = 0; i < 3; ++i) D[i][i] += 2.0 * mu;
    D[3][3] = D[4][4] = D[5][5] = mu;
    std::memset(K, 0, sizeof(double) * nd * nd);
    const double gp[3] = {-std::sqrt(0.6), 0.0, std::sqrt(0.6)}, gw[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    const double jac = 2.0 / a, detJ = (a / 2.0) * (a / 2.0) * (a / 2.0);
    static thread_local double B[6][60];
    for (int qi = 0; qi < 27; ++qi) {
        const double q[3] = {gp[qi % 3], gp[(qi / 3) % 3], gp[qi / 9]};
        const double w = gw[qi % 3] * gw[(qi / 3) % 3] * gw[qi / 9] * detJ;
        std::memset(B, 0, sizeof(B));
        for (int l = 0; l < 20; ++l) {
            double s[3] = {OFF[l][0] - 1.0, OFF[l][1] - 1.0, OFF[l][2] - 1.0};
            int zero = -1;
            for (int d = 0; d < 3; ++d) if (s[d] == 0.0) zero = d;
            double dN[3];
            if (zero < 0) {
                double f0 = 1 + s[0] * q[0], f1 = 1 + s[1] * q[1], f2 = 1 + s[2] * q[2];
                double t = s[0] * q[0] + s[1] * q[1] + s[2] * q[2] - 2.0;
                dN[0] = (s[0] * f1 * f2 * t + f0 * f1 * f2 * s[0]) / 8.0;
                dN[1] = (f0 * s[1] * f2 * t + f0 * f1 * f2 * s[1]) / 8.0;
                dN[2] = (f0 * f1 * s[2] * t + f0 * f1 * f2 * s[2]) / 8.0;
            } else {
                double f[3], df[3];
                for (int d = 0; d < 3; ++d) {
                    if (d == zero) { f[d] = 1 - q[d] * q[d]; df[d] = -2 * q[d]; }
                    else { f[d] = 1 + s[d] * q[d]; df[d] = s[d]; }
                }
                dN[0] = df[0] * f[1] * f[2] / 4.0;
                dN[1] = f[0] * df[1] * f[2] / 4.0;
                dN[2] = f[0] * f[1] * df[2] / 4.0;
            }
            const double dx = dN[0] * jac, dy = dN[1] * jac, dz = dN[2] * jac;
            B[0][3 * l] = dx; B[1][3 * l + 1] = dy; B[2][3 * l + 2] = dz;
            B[3][3 * l] = dy; B[3][3 * l + 1] = dx;
            B[4][3 * l] = dz; B[4][3 * l + 2] = dx;
            B[5][3 * l + 1] = dz; B[5][3 * l + 2] = dy;
        }
        for (int r = 0; r < nd; ++r)
            for (int m = 0; m < 6; ++m) {
                if (B[m][r] == 0.0) continue;
                for (int p = 0; p < 6; ++p) {
                    const double f = w * B[m][r] * D[m][p];
                    if (f == 0.0) continue;
                    for (int c = 0; c < nd; ++c) K[r + nd * c] += f * B[p][c];
                }
            }
    }
    return DTC_OK;
}

int validate_desc(const dtc_domain_desc& d) {
    if (!d.phi) return set_error(DTC_ERR_BAD_ARGUMENT, "domain voxels are NULL");
    if (!kind_size(d.phi_kind)) return set_error(DTC_ERR_BAD_ARGUMENT, "voxel kind %d (expected 1, 2 or 4)", d.phi_kind);
    for (int k = 0; k < 3; ++k) {
        if (d.n[k] <= 0 || d.n[k] > 4096) return set_error(DTC_ERR_BAD_ARGUMENT, "domain extent n[%d] = %lld", k, (long long)d.n[k]);
        if (!(d.delta[k] > 0.0)) return set_error(DTC_ERR_BAD_ARGUMENT, "grid spacing delta[%d] = %g", k, d.delta[k]);
    }
    if (d.micro_nen != 8 && d.micro_nen != 20)
        return set_error(DTC_ERR_BAD_ARGUMENT, "micro element with %d nodes (expected 8 or 20)", d.micro_nen);
    if (d.macro_order != 1 && d.macro_order != 2) return set_error(DTC_ERR_BAD_ARGUMENT, "macro element order %d", d.macro_order);
    if (!(d.nu > -1.0 && d.nu < 0.5) || !(d.E > 0.0) || !(d.size_domain > 0.0))
        return set_error(DTC_ERR_BAD_ARGUMENT, "material E=%g nu=%g size_domain=%g", d.E, d.nu, d.size_domain);
    return DTC_OK;
}

}  // namespace
}  // namespace dtc

using namespace dtc;

// ---------------------------------------------------------------- batch object
struct dtc_batch {
    std::vector<dtc_domain_desc> d;
    dtc_solve_opts o{};
    Stream st;
    std::vector<std::unique_ptr<DevMem>> phi, mod;
    std::vector<double> escale;
    std::vector<double> contrast;  // per domain: modulus contrast (max / min) over its solid voxels
    // Domains are solved in sub-batches of consecutive domains whose solver
    // nodes stay below NODE_CAP (32-bit vector indexing); each sub-batch keeps
    // its topology and solutions for the sensitivity pass.
    struct Sub {
        Stream own;                // the sub-batch's stream
        SolverBatch sb;
        SolverBatch sb08;          // HEX20 micro elements: the HEX08 twin of the p-multigrid
        std::vector<double> ff;
        std::vector<int64_t> iters, outer;
        std::vector<double> relres;
    };
    std::vector<std::unique_ptr<Sub>> subs;
    int n_sub = 0;                 // sub-batches used by the last run
    int n_main = 0;                // of which solved in the requested precision (the rest: FP64 repeats)
    bool timing = false;           // kernel timing (dtc_batch_timing)
    std::vector<int> sbi, g0, ng;  // sub-batch, first solver group and group count of each domain
    double nu = 0.3, a0 = 1.0;
    std::vector<double> khat20;    // HEX20 micro: E = 1 element stiffness
    MeshDev mesh;
    std::vector<double> EE;
    double run_ms = 0.0;
    bool ran = false;
};

namespace {
// solver nodes of one sub-batch: below 2^31 / 72 (FP32 vectors [node][3][24])
constexpr int64_t NODE_CAP = 24000000;
}

extern "C" {

const char* dtc_last_error(void) { return g_err.c_str(); }
int dtc_version(void) { return 10000; }

int dtc_ddc_params(const double size_domain[3], const double delta[3], const int64_t vdim[3], int64_t x_D[3],
                   int64_t nn_D[3], double x_D_phy[3]) {
    if (!size_domain || !delta || !vdim || !x_D || !nn_D || !x_D_phy) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    int64_t xd[3], nd[3];
    double xp[3];
    for (int k = 0; k < 3; ++k) {
        if (!(delta[k] > 0.0)) return set_error(DTC_ERR_BAD_ARGUMENT, "delta[%d] = %g", k, delta[k]);
        xd[k] = (int64_t)std::llround(size_domain[k] / delta[k]);   // NINT
        if (xd[k] <= 0) return set_error(DTC_ERR_BAD_ARGUMENT, "SIZE_DOMAIN/SPACING rounds to %lld", (long long)xd[k]);
        xp[k] = (double)xd[k] * delta[k];
        nd[k] = (vdim[k] - 2) / xd[k];                              // bpoints = 1
    }
    for (int k = 0; k < 3; ++k) { x_D[k] = xd[k]; nn_D[k] = nd[k]; x_D_phy[k] = xp[k]; }
    return DTC_OK;
}

int dtc_domain_bounds(int64_t domain, const int64_t x_D[3], const int64_t nn_D[3], int64_t nn_i[3], int64_t xa_n[3],
                      int64_t xe_n[3]) {
    if (!x_D || !nn_D || !nn_i || !xa_n || !xe_n) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    if (domain < 0 || nn_D[0] <= 0 || nn_D[1] <= 0 || nn_D[2] <= 0 || domain >= nn_D[0] * nn_D[1] * nn_D[2])
        return set_error(DTC_ERR_BAD_ARGUMENT, "domain %lld outside the decomposition", (long long)domain);
    const int64_t n3 = domain / (nn_D[0] * nn_D[1]);
    const int64_t n2 = (domain - nn_D[0] * nn_D[1] * n3) / nn_D[0];
    const int64_t n1 = domain - nn_D[0] * nn_D[1] * n3 - nn_D[0] * n2;
    const int64_t nn[3] = {n1, n2, n3};
    for (int k = 0; k < 3; ++k) {
        nn_i[k] = nn[k];
        xa_n[k] = x_D[k] * nn[k] + 2;
        xe_n[k] = x_D[k] * (nn[k] + 1) + 1;
    }
    return DTC_OK;
}

int dtc_extract_block(const void* vol, int kind, const int64_t vdim[3], const int64_t xa_n[3], const int64_t xe_n[3],
                      void* out) {
    if (!vol || !vdim || !xa_n || !xe_n || !out || !kind_size(kind)) return set_error(DTC_ERR_BAD_ARGUMENT, "bad argument");
    DTC_TRY(check_device());
    const int es = kind_size(kind);
    int64_t n[3];
    for (int k = 0; k < 3; ++k) {
        n[k] = xe_n[k] - xa_n[k] + 1;
        if (n[k] <= 0 || xa_n[k] < 1 || xe_n[k] > vdim[k]) return set_error(DTC_ERR_BAD_ARGUMENT, "block outside the volume");
    }
    Stream st;
    DTC_TRY(st.create());
    DevMem dv, dout;
    const size_t vb = (size_t)vdim[0] * vdim[1] * vdim[2] * es, ob = (size_t)n[0] * n[1] * n[2] * es;
    DTC_TRY(dv.alloc(vb));
    DTC_TRY(dout.alloc(ob));
    DTC_HIP(hipMemcpyAsync(dv.p, vol, vb, hipMemcpyHostToDevice, st.s));
    DTC_TRY(extract_block_device(dv.p, kind, vdim, xa_n, xe_n, dout.p, st.s));
    DTC_HIP(hipMemcpyAsync(out, dout.p, ob, hipMemcpyDeviceToHost, st.s));
    DTC_HIP(hipStreamSynchronize(st.s));
    return DTC_OK;
}

static int mesh_common(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen, MeshDev& m, Stream& st,
                       DevMem& dphi) {
    if (!phi || !n || !kind_size(kind)) return set_error(DTC_ERR_BAD_ARGUMENT, "bad argument");
    for (int k = 0; k < 3; ++k)
        if (n[k] <= 0 || n[k] > 4096) return set_error(DTC_ERR_BAD_ARGUMENT, "n[%d] = %lld", k, (long long)n[k]);
    DTC_TRY(check_device());
    DTC_TRY(st.create());
    const size_t bytes = (size_t)n[0] * n[1] * n[2] * kind_size(kind);
    DTC_TRY(dphi.alloc(bytes));
    DTC_HIP(hipMemcpyAsync(dphi.p, phi, bytes, hipMemcpyHostToDevice, st.s));
    return mesh_build_device(dphi.p, kind, n, llimit, nen, m, st.s);
}

int dtc_mesh_count(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen, int64_t* n_elems,
                   int64_t* n_nodes) {
    if (!n_elems || !n_nodes) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL output");
    MeshDev m;
    Stream st;
    DevMem dphi;
    int rc = mesh_common(phi, kind, n, llimit, nen, m, st, dphi);
    if (rc == DTC_ERR_EMPTY_MESH) { *n_elems = 0; *n_nodes = 0; return rc; }
    DTC_TRY(rc);
    *n_elems = m.n_elems;
    *n_nodes = m.n_nodes;
    return DTC_OK;
}

int dtc_mesh_build(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen, const double delta[3],
                   const int64_t xa_n[3], int64_t elems_cap, int64_t nodes_cap, double* nodes, int64_t* elems,
                   int64_t* n_elems, int64_t* n_nodes) {
    if (!delta || !xa_n || !nodes || !elems || !n_elems || !n_nodes) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    MeshDev m;
    Stream st;
    DevMem dphi;
    DTC_TRY(mesh_common(phi, kind, n, llimit, nen, m, st, dphi));
    if (m.n_elems > elems_cap || m.n_nodes > nodes_cap)
        return set_error(DTC_ERR_BAD_ARGUMENT, "mesh has %lld elements / %lld nodes, capacity %lld / %lld",
                         (long long)m.n_elems, (long long)m.n_nodes, (long long)elems_cap, (long long)nodes_cap);
    DevMem dn, de;
    DTC_TRY(dn.alloc((size_t)m.n_nodes * 3 * 8));
    DTC_TRY(de.alloc((size_t)m.n_elems * nen * 8));
    DTC_TRY(mesh_write_outputs(m, delta, xa_n, dn.as<double>(), de.as<int64_t>(), st.s));
    DTC_HIP(hipMemcpyAsync(nodes, dn.p, (size_t)m.n_nodes * 3 * 8, hipMemcpyDeviceToHost, st.s));
    DTC_HIP(hipMemcpyAsync(elems, de.p, (size_t)m.n_elems * nen * 8, hipMemcpyDeviceToHost, st.s));
    DTC_HIP(hipStreamSynchronize(st.s));
    *n_elems = m.n_elems;
    *n_nodes = m.n_nodes;
    return DTC_OK;
}

int dtc_boundaries(const double* nodes, int64_t n_nodes, const double min_c[3], const double max_c[3],
                   const double dim_c[3], int macro_order, int64_t* n_b, int64_t* bnode_ids, double* bvals) {
    if (!nodes || !min_c || !max_c || !dim_c || !n_b || n_nodes < 0) return set_error(DTC_ERR_BAD_ARGUMENT, "bad argument");
    if (macro_order != 1 && macro_order != 2) return set_error(DTC_ERR_BAD_ARGUMENT, "macro order %d", macro_order);
    const int RM = macro_lcs(macro_order);
    if ((bnode_ids == nullptr) != (bvals == nullptr)) return set_error(DTC_ERR_BAD_ARGUMENT, "ids and values go together");
    DTC_TRY(check_device());
    Stream st;
    DTC_TRY(st.create());
    DevMem dn, flag, pos, tmp, di, dv;
    DTC_TRY(dn.alloc((size_t)n_nodes * 3 * 8));
    DTC_HIP(hipMemcpyAsync(dn.p, nodes, (size_t)n_nodes * 3 * 8, hipMemcpyHostToDevice, st.s));
    int64_t nb = 0;
    DTC_TRY(boundaries_device(dn.as<double>(), n_nodes, min_c, max_c, dim_c, macro_order, &nb, nullptr, nullptr, flag,
                              pos, tmp, st.s));
    if (bnode_ids) {
        DTC_TRY(di.alloc((size_t)nb * 8));
        DTC_TRY(dv.alloc((size_t)nb * 3 * RM * 8));
        DTC_TRY(boundaries_device(dn.as<double>(), n_nodes, min_c, max_c, dim_c, macro_order, &nb, di.as<int64_t>(),
                                  dv.as<double>(), flag, pos, tmp, st.s));
        DTC_HIP(hipMemcpyAsync(bnode_ids, di.p, (size_t)nb * 8, hipMemcpyDeviceToHost, st.s));
        DTC_HIP(hipMemcpyAsync(bvals, dv.p, (size_t)nb * 3 * RM * 8, hipMemcpyDeviceToHost, st.s));
        DTC_HIP(hipStreamSynchronize(st.s));
    }
    *n_b = nb;
    return DTC_OK;
}

int dtc_element_stiffness(int nen, double E, double nu, double a, double* K) {
    if (!K) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL output");
    if (!(a > 0.0) || !(nu > -1.0 && nu < 0.5)) return set_error(DTC_ERR_BAD_ARGUMENT, "a=%g nu=%g", a, nu);
    if (nen == 8) return element_stiffness(8, E, nu, a, K);
    if (nen == 20) return element_stiffness20(E, nu, a, K);
    return set_error(DTC_ERR_BAD_ARGUMENT, "nen = %d", nen);
}

int dtc_batch_create(const dtc_domain_desc* d, int64_t n_dom, const dtc_solve_opts* o, dtc_batch** out) {
    if (!d || !o || !out || n_dom <= 0) return set_error(DTC_ERR_BAD_ARGUMENT, "bad argument");
    *out = nullptr;
    for (int64_t i = 0; i < n_dom; ++i) DTC_TRY(validate_desc(d[i]));
    for (int64_t i = 1; i < n_dom; ++i)
        if (d[i].nu != d[0].nu) return set_error(DTC_ERR_UNSUPPORTED, "one Poisson ratio per batch");
    for (int64_t i = 1; i < n_dom; ++i)
        if (d[i].micro_nen != d[0].micro_nen) return set_error(DTC_ERR_UNSUPPORTED, "one micro element type per batch");
    if (o->precision != DTC_FP64 && o->precision != DTC_FP32_REFINE) return set_error(DTC_ERR_BAD_ARGUMENT, "precision %d", o->precision);
    if (!(o->rel_tol > 0.0)) return set_error(DTC_ERR_BAD_ARGUMENT, "rel_tol %g", o->rel_tol);
    DTC_TRY(check_device());
    std::unique_ptr<dtc_batch> b(new dtc_batch);
    b->d.assign(d, d + n_dom);
    b->o = *o;
    if (!(b->o.inner_tol > 0.0 && b->o.inner_tol < 1.0)) b->o.inner_tol = 0.0;   // automatic (batch_solve)
    if (!(b->o.rot_step_deg > 0.0)) b->o.rot_step_deg = 1.0;
    if (!(b->o.num_zero > 0.0)) b->o.num_zero = 1e-9;
    DTC_TRY(b->st.create());
    const double a0 = d[0].size_domain * d[0].delta[0];   // a = phdsize * delta (f-src/mod_linfe.f90:180)
    for (int64_t i = 0; i < n_dom; ++i) {
        const size_t nv = (size_t)d[i].n[0] * d[i].n[1] * d[i].n[2];
        std::unique_ptr<DevMem> p(new DevMem), m(new DevMem);
        double ctr = 1.0;
        DTC_TRY(p->alloc(nv * kind_size(d[i].phi_kind)));
        DTC_HIP(hipMemcpyAsync(p->p, d[i].phi, nv * kind_size(d[i].phi_kind), hipMemcpyHostToDevice, b->st.s));
        if (d[i].modulus) {
            DTC_TRY(m->alloc(nv * 8));
            DTC_HIP(hipMemcpyAsync(m->p, d[i].modulus, nv * 8, hipMemcpyHostToDevice, b->st.s));
            // modulus contrast (max / min positive) over the solid voxels (phi >=
            // llimit), which bounds how deep an FP32 inner solve can go (the
            // automatic inner_tol of the domain)
            double lo = INFINITY, hi = 0.0;
            for (size_t v = 0; v < nv; ++v) {
                const double e = d[i].modulus[v];
                if (e > 0.0 && voxel_value(d[i].phi, d[i].phi_kind, v) >= d[i].llimit) {
                    lo = std::min(lo, e);
                    hi = std::max(hi, e);
                }
            }
            if (hi > 0.0) ctr = hi / lo;
        }
        b->contrast.push_back(ctr);
        b->phi.push_back(std::move(p));
        b->mod.push_back(std::move(m));
        b->escale.push_back(d[i].size_domain * d[i].delta[0] / a0);
        b->d[i].phi = nullptr;       // the host copy is not used after this call
        b->d[i].modulus = nullptr;
    }
    DTC_HIP(hipStreamSynchronize(b->st.s));
    b->nu = d[0].nu;
    b->a0 = a0;
    if (d[0].micro_nen == 20) {
        b->khat20.assign(3600, 0.0);
        DTC_TRY(element_stiffness20(1.0, d[0].nu, a0, b->khat20.data()));
    }
    *out = b.release();
    return DTC_OK;
}

int dtc_batch_run(dtc_batch* b, dtc_domain_result* res) {
    if (!b || !res) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    const int D = (int)b->d.size();
    hipStream_t s = b->st.s;
    DTC_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    // each domain is solved as groups of 24 macro load cases (1 group for the
    // HEX08 macro element, 3 for the 60 load cases of the HEX20 one); domains are
    // dealt to sub-batches by an upper bound of their solver nodes (lattice size)
    b->sbi.assign(D, 0);
    b->g0.assign(D, 0);
    b->ng.assign(D, 0);
    std::vector<int> first(1, 0);
    bool concurrent = false;
    int conc = 1;
    {
        int64_t cap = NODE_CAP;   // DTC_SUBBATCH_NODES lowers it (smaller sub-batches)
        if (const char* e = std::getenv("DTC_SUBBATCH_NODES")) cap = std::max<int64_t>(1, std::min<int64_t>(cap, std::atoll(e)));
        int64_t acc = 0;
        for (int i = 0; i < D; ++i) {
            const dtc_domain_desc& d = b->d[i];
            b->ng[i] = (macro_lcs(d.macro_order) + R - 1) / R;
            const int64_t m = d.micro_nen == 20 ? 2 : 1;
            const int64_t est = (m * d.n[0] + 1) * (m * d.n[1] + 1) * (m * d.n[2] + 1) * b->ng[i];
            if (acc > 0 && acc + est > cap) {
                first.push_back(i);
                acc = 0;
            }
            acc += est;
        }
        first.push_back(D);
        // A batch that fits one sub-batch is split into CONC_DEFAULT parts of about
        // equal size whose solves run at the same time on their own streams: the
        // iterations of one part fill the gaps of the others (the ramp-down tail of
        // each kernel, reduction latency, and the end of a solve where few domains
        // are still active); measured on 16 64^3 domains: 1 part 19.2, 2 parts 20.3,
        // 3 parts 20.5-20.8, 4 parts 19.0 domains/s.  With kernel timing on, one part
        // (the kernels of the parts would overlap, and per-kernel timings would no
        // longer measure one kernel alone).  DTC_CONCURRENT=k sets the parts.
        // A batch of several sub-batches solves them CONC_DEFAULT at a time (all are
        // set up, and hold their memory, before the first solve starts).
        constexpr int CONC_DEFAULT = 3;
        conc = b->timing ? 1 : CONC_DEFAULT;
        if (const char* e = std::getenv("DTC_CONCURRENT")) conc = std::max(1, std::atoi(e));
        conc = std::min(conc, D);
        if (first.size() > 2 && conc > 1) concurrent = true;
        if (first.size() == 2 && conc > 1) {
            int64_t total = 0;
            std::vector<int64_t> est(D);
            for (int i = 0; i < D; ++i) {
                const dtc_domain_desc& d = b->d[i];
                const int64_t m = d.micro_nen == 20 ? 2 : 1;
                est[i] = (m * d.n[0] + 1) * (m * d.n[1] + 1) * (m * d.n[2] + 1) * b->ng[i];
                total += est[i];
            }
            first.assign(1, 0);
            int64_t acc = 0;
            for (int i = 0; i < D; ++i) {
                if (acc * conc >= total * (int64_t)first.size() && (int)first.size() < conc) first.push_back(i);
                acc += est[i];
            }
            first.push_back(D);
            concurrent = first.size() > 2;
        }
    }
    const int NS = (int)first.size() - 1;
    int device = 0;
    DTC_HIP(hipGetDevice(&device));
    // one sub-batch, set-up: mesh -> topology for the listed domains (sequential:
    // the sub-batches share the mesh work space)
    auto setup_sub = [&](int k, const std::vector<int>& list, int precision) -> int {
        while ((int)b->subs.size() <= k) b->subs.emplace_back(new dtc_batch::Sub);
        dtc_batch::Sub& S = *b->subs[k];
        if (!S.own.s) DTC_TRY(S.own.create());
        SolverBatch& sb = S.sb;
        const hipStream_t s = S.own.s;
        sb.stream = s;
        sb.nu = b->nu;
        sb.a = b->a0;
        sb.micro20 = b->d[0].micro_nen == 20;
        sb.khat20 = b->khat20;
        sb.timing = b->timing;
        std::vector<dtc_domain_desc> gd;
        for (int i : list) {
            b->sbi[i] = k;
            b->g0[i] = (int)gd.size();
            for (int g = 0; g < b->ng[i]; ++g) gd.push_back(b->d[i]);
        }
        DTC_TRY(batch_begin(sb, gd.data(), (int64_t)gd.size()));
        for (int i : list)
            for (int g = 0; g < b->ng[i]; ++g) sb.h_contrast.push_back(b->contrast[i]);
        for (int i : list) {
            const dtc_domain_desc& d = b->d[i];
            int rc = mesh_build_device(b->phi[i]->p, d.phi_kind, d.n, d.llimit, d.micro_nen, b->mesh, s);
            if (rc == DTC_ERR_EMPTY_MESH) {
                for (int g = 0; g < b->ng[i]; ++g) DTC_TRY(batch_add_empty(sb, d, DTC_ERR_EMPTY_MESH, R * g));
                continue;
            }
            DTC_TRY(rc);
            const double* mod = b->mod[i]->p ? b->mod[i]->as<double>() : nullptr;
            for (int g = 0; g < b->ng[i]; ++g) {
                if (sb.micro20) DTC_TRY(batch_add_domain20(sb, b->mesh, d, mod, b->escale[i], R * g, s));
                else DTC_TRY(batch_add_domain(sb, b->mesh, d, mod, b->escale[i], R * g, s));
            }
        }
        DTC_TRY(batch_finalize(sb, precision));
        // the multigrid preconditioner of the mixed solve: the geometric V-cycle over
        // the voxel lattice (HEX08 micro), or for HEX20 micro elements a p-multigrid
        // whose coarse space is the HEX08 twin of the same voxels and domain groups
        sb.precond = (precision == DTC_FP32_REFINE && b->o.precond == DTC_PC_MG) ? DTC_PC_MG : DTC_PC_JACOBI;
        sb.mg_degree = b->o.mg_degree > 0 ? b->o.mg_degree : 2;
        sb.p08 = nullptr;
        if (sb.micro20 && sb.precond == DTC_PC_MG) {
            SolverBatch& tw = S.sb08;
            tw.stream = s;
            tw.nu = b->nu;
            tw.a = b->a0;
            tw.micro20 = false;
            tw.timing = false;
            DTC_TRY(batch_begin(tw, gd.data(), (int64_t)gd.size()));
            for (int i : list) {
                const dtc_domain_desc& d = b->d[i];
                int rc = mesh_build_device(b->phi[i]->p, d.phi_kind, d.n, d.llimit, 8, b->mesh, s);
                if (rc == DTC_ERR_EMPTY_MESH) {
                    for (int g = 0; g < b->ng[i]; ++g) DTC_TRY(batch_add_empty(tw, d, DTC_ERR_EMPTY_MESH, R * g));
                    continue;
                }
                DTC_TRY(rc);
                const double* mod = b->mod[i]->p ? b->mod[i]->as<double>() : nullptr;
                for (int g = 0; g < b->ng[i]; ++g) DTC_TRY(batch_add_domain(tw, b->mesh, d, mod, b->escale[i], R * g, s));
            }
            DTC_TRY(batch_finalize(tw, DTC_FP32_REFINE));
        }
        return DTC_OK;
    };
    // one sub-batch, solve: multigrid hierarchy (unless built by prep_sub) -> solve -> reactions
    auto prep_sub = [&](int k) -> int {
        dtc_batch::Sub& S = *b->subs[k];
        SolverBatch& sb = S.sb;
        if (sb.precond == DTC_PC_MG) {
            if (sb.micro20) DTC_TRY(batch_build_pmg(sb, S.sb08));
            else DTC_TRY(batch_build_mg(sb, 8));
        }
        return DTC_OK;
    };
    auto solve_sub = [&](int k, bool prepared) -> int {
        dtc_batch::Sub& S = *b->subs[k];
        SolverBatch& sb = S.sb;
        if (!prepared) DTC_TRY(prep_sub(k));
        DTC_TRY(batch_solve(sb, b->o, S.iters, S.outer, S.relres));
        DTC_TRY(batch_forces(sb, S.ff));
        return DTC_OK;
    };
    auto run_sub = [&](int k, const std::vector<int>& list, int precision) -> int {
        DTC_TRY(setup_sub(k, list, precision));
        return solve_sub(k, false);
    };
    for (int k = 0; k < NS; ++k) {
        std::vector<int> list;
        for (int i = first[k]; i < first[k + 1]; ++i) list.push_back(i);
        DTC_TRY(setup_sub(k, list, b->o.precision));
    }
    DTC_HIP(hipDeviceSynchronize());
    const auto t_setup = std::chrono::steady_clock::now();
    if (!concurrent) {
        for (int k = 0; k < NS; ++k) DTC_TRY(solve_sub(k, false));
    } else {
        // multigrid hierarchies and launch-shape tuning one sub-batch after another
        // (the tuning launches are timed: none may overlap another sub-batch's work),
        // then concurrent solves, conc at a time (one host thread per sub-batch, each
        // on its stream); the error message of a failed solve is carried back to this thread
        for (int k = 0; k < NS; ++k) {
            DTC_TRY(prep_sub(k));
            DTC_TRY(batch_autotune(b->subs[k]->sb));
            DTC_HIP(hipStreamSynchronize(b->subs[k]->sb.stream));
        }
        std::vector<int> rc(NS, DTC_OK);
        std::vector<std::string> msg(NS);
        const int wave = std::max(conc, 2);
        for (int k0 = 0; k0 < NS; k0 += wave) {
            std::vector<std::thread> th;
            for (int k = k0; k < std::min(NS, k0 + wave); ++k)
                th.emplace_back([&, k] {
                    if (hipSetDevice(device) != hipSuccess) {
                        rc[k] = DTC_ERR_DEVICE;
                        msg[k] = "hipSetDevice failed in a solve thread";
                        return;
                    }
                    rc[k] = solve_sub(k, true);
                    if (rc[k] != DTC_OK) msg[k] = dtc_last_error();
                });
            for (std::thread& t : th) t.join();
        }
        for (int k = 0; k < NS; ++k)
            if (rc[k] != DTC_OK) return set_error(rc[k], "%s", msg[k].c_str());
    }
    // A mixed-precision solve that did not converge (a near-singular mode, e.g. parts
    // hinged on one voxel edge, can grow without bound in the FP32 inner solve) is
    // repeated for that domain in FP64 with Jacobi-PCG, the reference's own scheme.
    int nsub = NS;
    b->n_main = NS;
    if (b->o.precision == DTC_FP32_REFINE) {
        std::vector<int> retry;
        for (int i = 0; i < D; ++i) {
            const dtc_batch::Sub& S = *b->subs[b->sbi[i]];
            bool bad = false;
            for (int g = 0; g < b->ng[i]; ++g) bad = bad || S.sb.h_status[b->g0[i] + g] == DTC_ERR_NO_CONVERGENCE;
            if (bad) retry.push_back(i);
        }
        if (!retry.empty()) {
            DTC_TRY(run_sub(NS, retry, DTC_FP64));
            nsub = NS + 1;
        }
    }
    b->n_sub = nsub;
    b->EE.assign((size_t)D * 36, 0.0);
    std::vector<dtc_domain_result> out(D);
    std::vector<int> solved;
    constexpr int FFS = 60;   // macro DOF slots per load case in the force partials
    for (int i = 0; i < D; ++i) {
        dtc_domain_result& r = out[i];
        std::memset(&r, 0, sizeof(r));
        const dtc_domain_desc& d = b->d[i];
        const dtc_batch::Sub& S = *b->subs[b->sbi[i]];
        const int gi = b->g0[i], RM = macro_lcs(d.macro_order), FD = 3 * macro_nodes(d.macro_order);
        r.status = S.sb.h_status[gi];
        r.precision_used = S.sb.precision;
        for (int g = 0; g < b->ng[i]; ++g) {
            const int st = S.sb.h_status[gi + g];
            if (r.status == DTC_OK && st != DTC_OK) r.status = st;
            r.iterations = std::max<int64_t>(r.iterations, S.iters[gi + g]);
            r.outer_iterations = std::max<int64_t>(r.outer_iterations, S.outer[gi + g]);
            const double rr = S.relres[gi + g];
            r.relres_max = (std::isnan(rr) || std::isnan(r.relres_max)) ? NAN : std::max(r.relres_max, rr);
        }
        r.n_elems = S.sb.h_elems[gi];
        r.n_nodes = S.sb.h_nodes[gi];
        r.n_bnodes = S.sb.h_bnodes[gi];
        double xdp[3], v = 1.0;
        for (int k = 0; k < 3; ++k) {
            xdp[k] = (double)d.n[k] * d.delta[k];
            v *= std::round(xdp[k] / d.delta[k]);
        }
        r.density = (double)r.n_elems / v;   // f-src/mod_struct_calcmat.f90:1642-1646
        if (r.status != DTC_OK && r.status != DTC_ERR_NO_CONVERGENCE) continue;
        // ff(dof, lc), column-major RM x RM
        for (int g = 0; g < b->ng[i]; ++g)
            for (int l = 0; l < R && R * g + l < RM; ++l)
                for (int k = 0; k < FD; ++k)
                    r.ff[k + RM * (R * g + l)] = S.ff[((size_t)(gi + g) * R + l) * FFS + k];
        DTC_TRY(effective(r.ff, d.macro_order, xdp, d.rve_strain, r.K_macro, r.C, r.EE));
        std::memcpy(&b->EE[(size_t)i * 36], r.EE, sizeof(r.EE));
        solved.push_back(i);
    }
    if (b->o.rotation && !solved.empty()) {
        std::vector<double> EEs;
        for (int i : solved) EEs.insert(EEs.end(), &b->EE[(size_t)i * 36], &b->EE[(size_t)i * 36] + 36);
        std::vector<RotResult> rr;
        DTC_TRY(rotation_search_batch(EEs.data(), (int)solved.size(), b->o.rot_step_deg, b->o.num_zero, s, rr));
        for (size_t k = 0; k < solved.size(); ++k) {
            dtc_domain_result& r = out[solved[k]];
            for (int c = 0; c < 2; ++c) {
                const RotResult& q = rr[2 * k + c];
                r.rot_alpha[c] = q.alpha;
                std::memcpy(&r.rot_n[3 * c], q.n, sizeof(q.n));
                std::memcpy(&r.rot_aa[9 * c], q.aa, sizeof(q.aa));
                std::memcpy(&r.rot_EE[36 * c], q.EE_opt, sizeof(q.EE_opt));
                r.rot_jj[c] = q.jj;
                r.rot_crit[c] = q.crit;
            }
        }
    }
    DTC_HIP(hipStreamSynchronize(s));
    b->run_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    // DTC_TRACE=1: where the run's time went (set-up = meshing, tables and the batch
    // data of every sub-batch; the rest = multigrid, solves, reactions, fit, rotation)
    if (const char* te = std::getenv("DTC_TRACE"); te && te[0] == '1')
        std::fprintf(stderr, "[dtc] run: %d sub-batch(es)%s, set-up %.1f ms, solve and results %.1f ms\n", NS,
                     concurrent ? " at the same time" : "",
                     std::chrono::duration<double, std::milli>(t_setup - t0).count(),
                     std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_setup).count());
    std::memcpy(res, out.data(), sizeof(dtc_domain_result) * D);
    b->ran = true;
    return DTC_OK;
}

int dtc_batch_grad(dtc_batch* b, int64_t i, const double dL_dC[36], double* dL_dE) {
    if (!b || !dL_dC || !dL_dE) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    if (!b->ran) return set_error(DTC_ERR_BAD_ARGUMENT, "dtc_batch_grad before dtc_batch_run");
    if (i < 0 || i >= (int64_t)b->d.size()) return set_error(DTC_ERR_BAD_ARGUMENT, "domain index %lld", (long long)i);
    const dtc_domain_desc& d = b->d[i];
    const int gi = b->g0[i];
    SolverBatch& sb = b->subs[b->sbi[i]]->sb;
    const size_t nv = (size_t)d.n[0] * d.n[1] * d.n[2];
    if (sb.h_status[gi] != DTC_OK && sb.h_status[gi] != DTC_ERR_NO_CONVERGENCE) {
        std::memset(dL_dE, 0, nv * 8);
        return DTC_OK;
    }
    // M(b,j) = meps_j sum_i G_ij sum_a W_ia vv_eps^-1(b,a) over the RM macro load
    // cases (24 or 60, the vv and fit weights of dtc's effective()); vv at eps =
    // 1e-6, the value the boundary tables use (f-src/mod_write_deck.f90:68)
    const int RM = macro_lcs(d.macro_order), ng = b->ng[i];
    std::vector<double> vv((size_t)RM * RM), W(6 * 60), M((size_t)ng * R * 6, 0.0);
    if (d.macro_order == 1) vv_lin(1.e-6, vv.data());
    else vv_quad(1.e-6, vv.data());
    DTC_TRY(invert(vv.data(), RM));
    double xdp[3] = {(double)d.n[0] * d.delta[0], (double)d.n[1] * d.delta[1], (double)d.n[2] * d.delta[2]};
    if (d.macro_order == 1) fv_weights(xdp, W.data());
    else fv_weights_quad(xdp, W.data());
    if (ng * R < RM) return set_error(DTC_ERR_DEVICE, "%d load-case groups for %d load cases", ng, RM);
    for (int bb = 0; bb < RM; ++bb)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int ii = 0; ii < 6; ++ii) {
                double t = 0.0;
                for (int a = 0; a < RM; ++a) t += W[ii + 6 * a] * vv[bb + RM * a];
                s += dL_dC[ii + 6 * j] * t;
            }
            M[bb * 6 + j] = s * MEPS[j];
        }
    std::vector<double> g(nv);
    DTC_TRY(batch_grad(sb, gi, ng, M.data(), g.data()));
    const double sc = b->escale[i];   // d(emod)/d(E_v)
    for (size_t v = 0; v < nv; ++v) dL_dE[v] = g[v] * sc;
    return DTC_OK;
}

int dtc_batch_stats(dtc_batch* b, double* mv_avg_ms, int64_t* mv_launches, double* mv_bytes, double* mv_flops,
                    double* run_ms) {
    if (!b) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL batch");
    double avg = 0.0, bytes = 0.0;
    int64_t n = 0;
    DTC_TRY(dtc_batch_kernel_stats(b, DTC_KSTAT_MV, &avg, &n, &bytes));
    // dense-equivalent flops F_mv = 1152 R N_e per domain group, counted in as
    // many launches as the group took PCG iterations, averaged over the launches
    double flops = 0.0;
    for (int k = 0; k < b->n_main && k < (int)b->subs.size(); ++k) {
        const auto& S = b->subs[k];
        for (size_t i = 0; i < S->sb.h_elems.size() && i < S->iters.size(); ++i)
            if (S->sb.h_status[i] == DTC_OK || S->sb.h_status[i] == DTC_ERR_NO_CONVERGENCE)
                flops += (double)S->iters[i] * 1152.0 * R * (double)S->sb.h_elems[i];
    }
    if (mv_avg_ms) *mv_avg_ms = avg;
    if (mv_launches) *mv_launches = n;
    if (mv_bytes) *mv_bytes = bytes;
    if (mv_flops) *mv_flops = n ? flops / (double)n : 0.0;
    if (run_ms) *run_ms = b->run_ms;
    return DTC_OK;
}

int dtc_batch_timing(dtc_batch* b, int enable) {
    if (!b) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL batch");
    b->timing = enable != 0;
    return DTC_OK;
}

int dtc_batch_kernel_stats(dtc_batch* b, int kernel, double* avg_ms, int64_t* launches, double* avg_bytes) {
    if (!b) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL batch");
    if (kernel < 0 || kernel >= DTC_KSTAT_COUNT) return set_error(DTC_ERR_BAD_ARGUMENT, "kernel %d", kernel);
    double ms = 0.0, bytes = 0.0;
    int64_t n = 0;
    // the sub-batches of the requested precision (FP64 repeats of stalled domains excluded)
    for (int k = 0; k < b->n_main && k < (int)b->subs.size(); ++k) {
        const KernelStat& q = b->subs[k]->sb.kstat[kernel];
        ms += q.ms_sum;
        bytes += q.bytes_sum;
        n += q.count;
    }
    if (avg_ms) *avg_ms = n ? ms / (double)n : 0.0;
    if (launches) *launches = n;
    if (avg_bytes) *avg_bytes = n ? bytes / (double)n : 0.0;
    return DTC_OK;
}

int dtc_batch_kernel_flops(dtc_batch* b, int kernel, double* avg_flops) {
    if (!b) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL batch");
    if (kernel < 0 || kernel >= DTC_KSTAT_COUNT) return set_error(DTC_ERR_BAD_ARGUMENT, "kernel %d", kernel);
    double fl = 0.0;
    int64_t n = 0;
    for (int k = 0; k < b->n_main && k < (int)b->subs.size(); ++k) {
        const KernelStat& q = b->subs[k]->sb.kstat[kernel];
        fl += q.flops_sum;
        n += q.count;
    }
    if (avg_flops) *avg_flops = n ? fl / (double)n : 0.0;
    return DTC_OK;
}

int dtc_batch_destroy(dtc_batch* b) {
    if (!b) return DTC_OK;
    if (b->st.s) (void)hipStreamSynchronize(b->st.s);
    for (const auto& S : b->subs)
        for (hipEvent_t e : S->sb.ev) (void)hipEventDestroy(e);
    delete b;
    return DTC_OK;
}

int dtc_solve_batch(const dtc_domain_desc* d, int64_t n_dom, const dtc_solve_opts* o, dtc_domain_result* res) {
    dtc_batch* b = nullptr;
    DTC_TRY(dtc_batch_create(d, n_dom, o, &b));
    int rc = dtc_batch_run(b, res);
    dtc_batch_destroy(b);
    return rc;
}

int dtc_solve_domain(const dtc_domain_desc* d, const dtc_solve_opts* o, dtc_domain_result* res, double* u_out,
                     double* r_out) {
    dtc_batch* b = nullptr;
    DTC_TRY(dtc_batch_create(d, 1, o, &b));
    int rc = dtc_batch_run(b, res);
    if (rc == DTC_OK && (u_out || r_out) && (res->status == DTC_OK || res->status == DTC_ERR_NO_CONVERGENCE)) {
        const int RM = macro_lcs(d->macro_order);
        const size_t col = (size_t)3 * res->n_nodes;
        for (int g = 0; g < b->ng[0] && rc == DTC_OK; ++g)
            rc = batch_fields(b->subs[b->sbi[0]]->sb, b->g0[0] + g, std::min(R, RM - R * g), u_out ? u_out + col * R * g : nullptr,
                              r_out ? r_out + col * R * g : nullptr);
    }
    dtc_batch_destroy(b);
    return rc;
}

int dtc_rotation_search(const double EE[36], int criterion, double step_deg, double num_zero, double* alpha, double n[3],
                        double aa[9], double EE_opt[36], int64_t* jj, double* crit) {
    if (!EE || !alpha || !n || !aa || !EE_opt) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    if (criterion != 1 && criterion != 2) return set_error(DTC_ERR_BAD_ARGUMENT, "criterion %d", criterion);
    DTC_TRY(check_device());
    Stream st;
    DTC_TRY(st.create());
    std::vector<RotResult> rr;
    DTC_TRY(rotation_search_batch(EE, 1, step_deg, num_zero > 0.0 ? num_zero : 1e-9, st.s, rr));
    const RotResult& q = rr[criterion - 1];
    *alpha = q.alpha;
    std::memcpy(n, q.n, sizeof(q.n));
    std::memcpy(aa, q.aa, sizeof(q.aa));
    std::memcpy(EE_opt, q.EE_opt, sizeof(q.EE_opt));
    if (jj) *jj = q.jj;
    if (crit) *crit = q.crit;
    return DTC_OK;
}

int dtc_rotation_search_noalloc(const double EE[36], int criterion, double step_deg, double num_zero, double* alpha,
                                double n[3], double aa[9], double EE_opt[36], int64_t* jj, double* crit) {
    if (!EE || !alpha || !n || !aa || !EE_opt) return set_error(DTC_ERR_BAD_ARGUMENT, "NULL argument");
    if (criterion != 1 && criterion != 2) return set_error(DTC_ERR_BAD_ARGUMENT, "criterion %d", criterion);
    RotResult rr[2];
    DTC_TRY(rotation_search_noalloc(EE, step_deg, num_zero > 0.0 ? num_zero : 1e-9, rr));
    const RotResult& q = rr[criterion - 1];
    *alpha = q.alpha;
    std::memcpy(n, q.n, sizeof(q.n));
    std::memcpy(aa, q.aa, sizeof(q.aa));
    std::memcpy(EE_opt, q.EE_opt, sizeof(q.EE_opt));
    if (jj) *jj = q.jj;
    if (crit) *crit = q.crit;
    return DTC_OK;
}

int64_t dtc_rotation_noalloc_max_points(void) { return rotation_noalloc_max_points(); }

}  // extern "C"
