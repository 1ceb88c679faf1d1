// rotation.hip -- CR_1 / CR_2 rotation search of the 6x6 effective stiffness on gfx950.
//
// Follows f-src/mod_struct_calcmat.f90:533-1603: a sweep over the axis
// n(phi, eta) = (cos phi sin eta, sin phi sin eta, cos eta) and the angle
// alpha on a 1-degree grid (alpha, phi in [0,180], eta in [0,90]), then up to
// 16 refinements of +-1 grid step zoomed by 10, stopping when the minimum
// changes by less than num_zero.  Each grid point is a Rodrigues rotation
// (rot_alg), its 6x6 Voigt form with sqrt(2) shear terms (tra_R6, inlined at
// :556-589) and the congruence T = B^T EE B; CR_1 = sum T(1:4,5:6)^2,
// CR_2 = sum T(1:3,4:6)^2 + T45^2 + T46^2 + T56^2 (:762-779).  Ties resolve to
// the first grid point in Fortran column-major order (alpha fastest), like
// MINLOC.  The sweep evaluates both criteria per point (the reference does
// the same in one loop); ~3M points per domain are spread over the GPU and
// reduced per block, then per domain.
#include <cmath>
#include <mutex>

#include "dtc_internal.hpp"
#include "rotation.hpp"

namespace dtc {

namespace {

constexpr int NT = 256;
constexpr int PPT = 16;   // points per thread in the sweep

__device__ inline void rot_point(const double* __restrict__ EE, double alpha, double phi, double eta, bool normalise,
                                 double T12[12]) {
    double n0 = cos(phi) * sin(eta), n1 = sin(phi) * sin(eta), n2 = cos(eta);
    if (normalise) {
        double nn = sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        n0 /= nn; n1 /= nn; n2 /= nn;
    }
    const double c = cos(alpha), s = sin(alpha), omc = 1.0 - c;
    double a[3][3];
    a[0][0] = c + n0 * n0 * omc;
    a[1][1] = c + n1 * n1 * omc;
    a[2][2] = c + n2 * n2 * omc;
    a[0][1] = n0 * n1 * omc - n2 * s;
    a[1][0] = n0 * n1 * omc + n2 * s;
    a[0][2] = n0 * n2 * omc + n1 * s;
    a[2][0] = n0 * n2 * omc - n1 * s;
    a[1][2] = n1 * n2 * omc - n0 * s;
    a[2][1] = n1 * n2 * omc + n0 * s;
    const double sq2 = 1.4142135623730951;
    double B[6][6];   // B[row][col]
    for (int k = 0; k < 3; ++k) {
        B[0][k] = a[0][k] * a[0][k];
        B[1][k] = a[1][k] * a[1][k];
        B[2][k] = a[2][k] * a[2][k];
        B[3][k] = sq2 * a[1][k] * a[0][k];
        B[4][k] = sq2 * a[0][k] * a[2][k];
        B[5][k] = sq2 * a[1][k] * a[2][k];
    }
    const int P[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    for (int t = 0; t < 3; ++t) {
        const int i = P[t][0], j = P[t][1], k = 3 + t;
        B[0][k] = sq2 * a[0][i] * a[0][j];
        B[1][k] = sq2 * a[1][i] * a[1][j];
        B[2][k] = sq2 * a[2][i] * a[2][j];
        B[3][k] = a[1][i] * a[0][j] + a[1][j] * a[0][i];
        B[4][k] = a[0][i] * a[2][j] + a[0][j] * a[2][i];
        B[5][k] = a[1][i] * a[2][j] + a[1][j] * a[2][i];
    }
    // M = EE B for columns 3..5 and the 12 entries T(i, j), i < j, j >= 3
    double M[6][3];
    for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 3; ++k) {
            double v = 0.0;
            for (int l = 0; l < 6; ++l) v += EE[r + 6 * l] * B[l][3 + k];
            M[r][k] = v;
        }
    // tmp_r12 order: T14 T15 T16 T24 T25 T26 T34 T35 T36 T45 T46 T56
    const int I[12] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 4};
    const int J[12] = {3, 4, 5, 3, 4, 5, 3, 4, 5, 4, 5, 5};
    for (int q = 0; q < 12; ++q) {
        double v = 0.0;
        for (int r = 0; r < 6; ++r) v += B[r][I[q]] * M[r][J[q] - 3];
        T12[q] = v;
    }
}

__device__ inline double crit1(const double t[12]) {
    return t[1] * t[1] + t[2] * t[2] + t[4] * t[4] + t[5] * t[5] + t[7] * t[7] + t[8] * t[8] + t[9] * t[9] + t[10] * t[10];
}
__device__ inline double crit2(const double t[12]) {
    double s = 0.0;
    for (int q = 0; q < 12; ++q) s += t[q] * t[q];
    return s;
}

__device__ inline bool better(double v, int64_t i, double bv, int64_t bi) {
    return v < bv || (v == bv && i < bi);
}

__device__ inline void block_argmin(double& v, int64_t& i, double* sv, int64_t* si) {
    const int t = threadIdx.x;
    sv[t] = v; si[t] = i;
    __syncthreads();
    for (int w = NT / 2; w > 0; w >>= 1) {
        if (t < w && better(sv[t + w], si[t + w], sv[t], si[t])) { sv[t] = sv[t + w]; si[t] = si[t + w]; }
        __syncthreads();
    }
    v = sv[0]; i = si[0];
}

// initial sweep: grid (na, na, ne), point index = ia + na*(ip + na*ie)
__global__ __launch_bounds__(NT) void k_sweep(const double* __restrict__ EE, int64_t na, int64_t ne, double ustep,
                                              int blocks_per_dom, double* __restrict__ pv, int64_t* __restrict__ pi) {
    __shared__ double sv[NT];
    __shared__ int64_t si[NT];
    const int dom = blockIdx.x / blocks_per_dom, blk = blockIdx.x % blocks_per_dom;
    const double* E = EE + 36 * dom;
    const int64_t npts = na * na * ne;
    double b1 = INFINITY, b2 = INFINITY;
    int64_t i1 = INT64_MAX, i2 = INT64_MAX;
    for (int k = 0; k < PPT; ++k) {
        int64_t idx = ((int64_t)blk * PPT + k) * NT + threadIdx.x;
        if (idx >= npts) break;
        int64_t ia = idx % na, ip = (idx / na) % na, ie = idx / (na * na);
        double t[12];
        rot_point(E, (double)ia * ustep, (double)ip * ustep, (double)ie * ustep, true, t);
        double c1 = crit1(t), c2 = crit2(t);
        if (better(c1, idx, b1, i1)) { b1 = c1; i1 = idx; }
        if (better(c2, idx, b2, i2)) { b2 = c2; i2 = idx; }
    }
    block_argmin(b1, i1, sv, si);
    __syncthreads();
    block_argmin(b2, i2, sv, si);
    if (threadIdx.x == 0) {
        pv[2 * blockIdx.x] = b1; pi[2 * blockIdx.x] = i1;
        pv[2 * blockIdx.x + 1] = b2; pi[2 * blockIdx.x + 1] = i2;
    }
}

__global__ __launch_bounds__(NT) void k_sweep_reduce(const double* __restrict__ pv, const int64_t* __restrict__ pi,
                                                     int blocks_per_dom, double* __restrict__ best,
                                                     int64_t* __restrict__ besti) {
    __shared__ double sv[NT];
    __shared__ int64_t si[NT];
    const int dom = blockIdx.x >> 1, c = blockIdx.x & 1;
    double v = INFINITY;
    int64_t i = INT64_MAX;
    for (int b = threadIdx.x; b < blocks_per_dom; b += NT) {
        int64_t k = 2 * ((int64_t)dom * blocks_per_dom + b) + c;
        if (better(pv[k], pi[k], v, i)) { v = pv[k]; i = pi[k]; }
    }
    block_argmin(v, i, sv, si);
    if (threadIdx.x == 0) { best[blockIdx.x] = v; besti[blockIdx.x] = i; }
}

// one refinement step for each active (domain, criterion) job: 21^3 points
// from start s0 with spacing div, criterion 1 normalises n, criterion 2 not
__global__ __launch_bounds__(NT) void k_refine(const double* __restrict__ EE, const RotJob* __restrict__ jobs,
                                               double* __restrict__ best, int64_t* __restrict__ besti) {
    __shared__ double sv[NT];
    __shared__ int64_t si[NT];
    const RotJob J = jobs[blockIdx.x];
    if (!J.active) return;
    const double* E = EE + 36 * J.dom;
    double bv = INFINITY;
    int64_t bi = INT64_MAX;
    for (int idx = threadIdx.x; idx < 21 * 21 * 21; idx += NT) {
        int ia = idx % 21, ip = (idx / 21) % 21, ie = idx / 441;
        double t[12];
        rot_point(E, (double)(J.s0[0] + ia) * J.div, (double)(J.s0[1] + ip) * J.div, (double)(J.s0[2] + ie) * J.div,
                  J.criterion == 1, t);
        double c = J.criterion == 1 ? crit1(t) : crit2(t);
        if (better(c, idx, bv, bi)) { bv = c; bi = idx; }
    }
    block_argmin(bv, bi, sv, si);
    if (threadIdx.x == 0) { best[blockIdx.x] = bv; besti[blockIdx.x] = bi; }
}

// host copies of rot_alg / tra_R6 / congruence for the final outputs
void h_rot_alg(const double n[3], double alpha, double aa[9]) {
    double c = std::cos(alpha), s = std::sin(alpha), omc = 1.0 - c;
    double n12 = n[0] * n[1], n13 = n[0] * n[2], n23 = n[1] * n[2];
    aa[0] = c + n[0] * n[0] * omc; aa[4] = c + n[1] * n[1] * omc; aa[8] = c + n[2] * n[2] * omc;
    aa[3] = n12 * omc - n[2] * s;  aa[1] = n12 * omc + n[2] * s;
    aa[6] = n13 * omc + n[1] * s;  aa[2] = n13 * omc - n[1] * s;
    aa[7] = n23 * omc - n[0] * s;  aa[5] = n23 * omc + n[0] * s;
}
void h_tra_R6(const double aa[9], double BB[36]) {
    const double sq2 = std::sqrt(2.0);
    auto A = [&](int i, int j) { return aa[(i - 1) + 3 * (j - 1)]; };
    auto B = [&](int i, int j) -> double& { return BB[(i - 1) + 6 * (j - 1)]; };
    for (int c = 1; c <= 3; ++c) {
        B(1, c) = A(1, c) * A(1, c); B(2, c) = A(2, c) * A(2, c); B(3, c) = A(3, c) * A(3, c);
        B(4, c) = sq2 * A(2, c) * A(1, c); B(5, c) = sq2 * A(1, c) * A(3, c); B(6, c) = sq2 * A(2, c) * A(3, c);
    }
    const int P[3][2] = {{1, 2}, {1, 3}, {2, 3}};
    for (int t = 0; t < 3; ++t) {
        int a = P[t][0], b = P[t][1], c = 4 + t;
        B(1, c) = sq2 * A(1, a) * A(1, b); B(2, c) = sq2 * A(2, a) * A(2, b); B(3, c) = sq2 * A(3, a) * A(3, b);
        B(4, c) = A(2, a) * A(1, b) + A(2, b) * A(1, a);
        B(5, c) = A(1, a) * A(3, b) + A(1, b) * A(3, a);
        B(6, c) = A(2, a) * A(3, b) + A(2, b) * A(3, a);
    }
}
void h_congruence(const double BB[36], const double EE[36], double T[36]) {
    double tmp[36];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += EE[i + 6 * k] * BB[k + 6 * j];
            tmp[i + 6 * j] = s;
        }
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += BB[k + 6 * i] * tmp[k + 6 * j];
            T[i + 6 * j] = s;
        }
}

}  // namespace

// final angle, axis and axis permutation (f-src/mod_struct_calcmat.f90:1015-1137)
void rotation_finish(const double EE[36], const int64_t ang[3], int64_t jj, double step_deg, RotResult& out) {
    const double pi = std::acos(-1.0);
    const double den = 180.0 * (std::pow(10.0, (double)jj) - 1.0);   // the reference's 10**jj - 1
    double alpha = (double)ang[0] * step_deg * pi / den;
    double phi = (double)ang[1] * step_deg * pi / den;
    double eta = (double)ang[2] * step_deg * pi / den;
    double n[3] = {std::cos(phi) * std::sin(eta), std::sin(phi) * std::sin(eta), std::cos(eta)};
    double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (double& v : n) v /= nn;
    out.alpha = alpha;
    for (int k = 0; k < 3; ++k) out.n[k] = n[k];
    double aa[9], BB[36], T[36];
    h_rot_alg(n, alpha, aa);
    h_tra_R6(aa, BB);
    h_congruence(BB, EE, T);
    const double e11 = T[0], e22 = T[7], e33 = T[14];
    int seq[2] = {-1, -1};
    if (e11 < e22 && e11 < e33 && e22 < e33) {}
    else if (e11 < e22 && e11 < e33 && e22 > e33) { seq[0] = 0; }
    else if (e11 < e22 && e11 > e33 && e22 > e33) { seq[0] = 1; seq[1] = 0; }
    else if (e11 > e22 && e11 < e33 && e22 < e33) { seq[0] = 2; }
    else if (e11 > e22 && e11 > e33 && e22 < e33) { seq[0] = 2; seq[1] = 0; }
    else if (e11 > e22 && e11 > e33 && e22 > e33) { seq[0] = 1; }
    for (int s = 0; s < 2 && seq[s] >= 0; ++s) {
        double ax[3] = {aa[3 * seq[s]], aa[3 * seq[s] + 1], aa[3 * seq[s] + 2]};
        double rr[9], prod[9];
        h_rot_alg(ax, pi / 2.0, rr);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double t = 0.0;
                for (int k = 0; k < 3; ++k) t += rr[i + 3 * k] * aa[k + 3 * j];
                prod[i + 3 * j] = t;
            }
        for (int k = 0; k < 9; ++k) aa[k] = prod[k];
    }
    h_tra_R6(aa, BB);
    h_congruence(BB, EE, out.EE_opt);
    for (int k = 0; k < 9; ++k) out.aa[k] = aa[k];
}

// Device buffers of one search: n_dom 6x6 tensors, per-block sweep minima
// (bpd per domain and criterion), per-job minima and the refinement jobs.
struct RotBufs {
    double* EE;
    double* pv;
    int64_t* pi;
    double* best;
    int64_t* besti;
    RotJob* jobs;
};

// Grid sizes of the initial sweep for an initial step of step_deg degrees.
void rotation_grid(double step_deg, int64_t* na, int64_t* ne, int* bpd) {
    *na = (int64_t)std::floor(180.0 / step_deg + 1e-9) + 1;
    *ne = (int64_t)std::floor(90.0 / step_deg + 1e-9) + 1;
    const int64_t npts = *na * *na * *ne;
    *bpd = (int)((npts + (int64_t)NT * PPT - 1) / ((int64_t)NT * PPT));
}

// The search on caller-provided device buffers and host state (hb, prev: 2 n_dom
// doubles; hbi, jjv: 2 n_dom int64; ang: 6 n_dom int64; J: 2 n_dom jobs), so that
// a caller without device or host allocations can run it.
int rotation_search_core(const double* EE_host, int n_dom, double step_deg, double num_zero, hipStream_t s,
                         const RotBufs& d, double* hb, int64_t* hbi, double* prev, int64_t* jjv, int64_t* ang,
                         RotJob* J, RotResult* out) {
    const double pi = std::acos(-1.0), pi_div_180 = pi / 180.0;
    int64_t na, ne;
    int bpd;
    rotation_grid(step_deg, &na, &ne, &bpd);
    DTC_HIP(hipMemcpyAsync(d.EE, EE_host, (size_t)n_dom * 36 * 8, hipMemcpyHostToDevice, s));
    k_sweep<<<(unsigned)(n_dom * bpd), NT, 0, s>>>(d.EE, na, ne, step_deg * pi_div_180, bpd, d.pv, d.pi);
    k_sweep_reduce<<<(unsigned)(n_dom * 2), NT, 0, s>>>(d.pv, d.pi, bpd, d.best, d.besti);
    DTC_HIP(hipGetLastError());
    DTC_HIP(hipMemcpyAsync(hb, d.best, (size_t)n_dom * 2 * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipMemcpyAsync(hbi, d.besti, (size_t)n_dom * 2 * 8, hipMemcpyDeviceToHost, s));
    DTC_HIP(hipStreamSynchronize(s));
    // refinement state per job (dom, criterion)
    for (int k = 0; k < n_dom * 2; ++k) {
        const int64_t idx = hbi[k];
        ang[3 * k] = idx % na; ang[3 * k + 1] = (idx / na) % na; ang[3 * k + 2] = idx / (na * na);
        prev[k] = hb[k];
        jjv[k] = 1;
        out[k] = RotResult{};
        J[k] = RotJob{};
        J[k].dom = k / 2; J[k].criterion = 1 + (k & 1); J[k].active = 1;
    }
    double div = pi_div_180 * step_deg;
    for (int jj = 1; jj <= 16; ++jj) {
        div = div * 0.1;
        bool any = false;
        for (int k = 0; k < n_dom * 2; ++k) {
            if (!J[k].active) continue;
            any = true;
            for (int c = 0; c < 3; ++c) J[k].s0[c] = (ang[3 * k + c] - 1) * 10;
            J[k].div = div;
        }
        if (!any) break;
        DTC_HIP(hipMemcpyAsync(d.jobs, J, (size_t)n_dom * 2 * sizeof(RotJob), hipMemcpyHostToDevice, s));
        k_refine<<<(unsigned)(n_dom * 2), NT, 0, s>>>(d.EE, d.jobs, d.best, d.besti);
        DTC_HIP(hipGetLastError());
        DTC_HIP(hipMemcpyAsync(hb, d.best, (size_t)n_dom * 2 * 8, hipMemcpyDeviceToHost, s));
        DTC_HIP(hipMemcpyAsync(hbi, d.besti, (size_t)n_dom * 2 * 8, hipMemcpyDeviceToHost, s));
        DTC_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < n_dom * 2; ++k) {
            if (!J[k].active) continue;
            const int64_t i = hbi[k];
            ang[3 * k] = J[k].s0[0] + i % 21;
            ang[3 * k + 1] = J[k].s0[1] + (i / 21) % 21;
            ang[3 * k + 2] = J[k].s0[2] + i / 441;
            jjv[k] = jj;
            out[k].crit = hb[k];
            if (std::fabs(prev[k] - hb[k]) < num_zero || jj >= 16) J[k].active = 0;
            prev[k] = hb[k];
        }
    }
    for (int k = 0; k < n_dom * 2; ++k) {
        out[k].jj = jjv[k];
        rotation_finish(EE_host + 36 * (k / 2), &ang[3 * k], jjv[k], step_deg, out[k]);
    }
    return DTC_OK;
}

int rotation_search_batch(const double* EE_host, int n_dom, double step_deg, double num_zero, hipStream_t s,
                          std::vector<RotResult>& out) {
    out.assign((size_t)n_dom * 2, RotResult{});
    if (n_dom == 0) return DTC_OK;
    if (!(step_deg > 0.0)) return set_error(DTC_ERR_BAD_ARGUMENT, "rotation step %g", step_deg);
    int64_t na, ne;
    int bpd;
    rotation_grid(step_deg, &na, &ne, &bpd);
    DevMem dEE, pv, pi_, best, besti, jobs;
    DTC_TRY(dEE.alloc((size_t)n_dom * 36 * 8));
    DTC_TRY(pv.alloc((size_t)n_dom * bpd * 2 * 8));
    DTC_TRY(pi_.alloc((size_t)n_dom * bpd * 2 * 8));
    DTC_TRY(best.alloc((size_t)n_dom * 2 * 8));
    DTC_TRY(besti.alloc((size_t)n_dom * 2 * 8));
    DTC_TRY(jobs.alloc((size_t)n_dom * 2 * sizeof(RotJob)));
    const RotBufs d{dEE.as<double>(), pv.as<double>(), pi_.as<int64_t>(), best.as<double>(), besti.as<int64_t>(),
                    jobs.as<RotJob>()};
    std::vector<double> hb(n_dom * 2), prev(n_dom * 2);
    std::vector<int64_t> hbi(n_dom * 2), jjv(n_dom * 2), ang(n_dom * 2 * 3);
    std::vector<RotJob> J(n_dom * 2);
    return rotation_search_core(EE_host, n_dom, step_deg, num_zero, s, d, hb.data(), hbi.data(), prev.data(),
                                jjv.data(), ang.data(), J.data(), out.data());
}

// ---------------------------------------------------------------- no-allocation search
// Work space of the search that allocates nothing: one per GPU, held until the
// process ends.  Host code never uses the array's name as an address (in host
// code it names a host-side stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) unsigned char scratch[ROT_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[MAX_DEVICES] = {};

// byte offsets of the buffers in the work space (one domain, 256-byte aligned)
constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
size_t rot_scratch_layout(int bpd, size_t off[6]) {
    const size_t sz[6] = {36 * 8, (size_t)bpd * 2 * 8, (size_t)bpd * 2 * 8, 2 * 8, 2 * 8, 2 * sizeof(RotJob)};
    size_t o = 0;
    for (int k = 0; k < 6; ++k) {
        off[k] = o;
        o = align256(o + sz[k]);
    }
    return o;
}
}  // namespace

// device address of this GPU's work space (looked up once per device, with the
// device current)
int rotation_scratch(void** ptr) {
    int dev = 0;
    DTC_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= MAX_DEVICES) return set_error(DTC_ERR_DEVICE, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(DTC_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return DTC_OK;
}

int64_t rotation_noalloc_max_points() {
    // largest sweep whose per-block minima fit the work space
    size_t off[6];
    int bpd = 1;
    while (rot_scratch_layout(bpd + 1, off) <= ROT_SCRATCH_BYTES) ++bpd;
    return (int64_t)bpd * NT * PPT;
}

int rotation_search_noalloc(const double EE[36], double step_deg, double num_zero, RotResult out[2]) {
    if (!(step_deg > 0.0)) return set_error(DTC_ERR_BAD_ARGUMENT, "rotation step %g", step_deg);
    int64_t na, ne;
    int bpd;
    rotation_grid(step_deg, &na, &ne, &bpd);
    size_t off[6];
    const size_t need = rot_scratch_layout(bpd, off);
    if (need > ROT_SCRATCH_BYTES)
        return set_error(DTC_ERR_BAD_ARGUMENT,
                         "rotation step %g deg: %lld grid points need %zu bytes of work space, the fixed "
                         "work space holds %zu (at most %lld points)",
                         step_deg, (long long)(na * na * ne), need, (size_t)ROT_SCRATCH_BYTES,
                         (long long)rotation_noalloc_max_points());
    void* base = nullptr;
    DTC_TRY(rotation_scratch(&base));
    unsigned char* b = static_cast<unsigned char*>(base);
    const RotBufs d{reinterpret_cast<double*>(b + off[0]), reinterpret_cast<double*>(b + off[1]),
                    reinterpret_cast<int64_t*>(b + off[2]), reinterpret_cast<double*>(b + off[3]),
                    reinterpret_cast<int64_t*>(b + off[4]), reinterpret_cast<RotJob*>(b + off[5])};
    double hb[2], prev[2];
    int64_t hbi[2], jjv[2], ang[6];
    RotJob J[2];
    // the null stream: no stream object is created
    return rotation_search_core(EE, 1, step_deg, num_zero, nullptr, d, hb, hbi, prev, jjv, ang, J, out);
}

}  // namespace dtc
