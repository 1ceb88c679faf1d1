// solver.hpp -- batched matrix-free PCG over HEX08 voxel domains (internal).
#pragma once
#include "dtc_internal.hpp"

namespace dtc {

// Accumulated timing of one timed kernel variant.
struct KernelStat {
    double ms_sum = 0.0, bytes_sum = 0.0, flops_sum = 0.0;
    int64_t count = 0;
};

// One coarse level of the multigrid preconditioner: voxels of 2^l fine voxels,
// same brick layout and matvec tables as the fine level, FP32 vectors.
struct Level {
    int n_bricks = 0;
    int64_t n_nodes = 0, n_lat = 0, n_vox = 0, n_tile_elems = 0;
    std::vector<DomInfo> h_dom;
    std::vector<BrickInfo> h_brick;
    DevMem dom, brick, lat2node, emod, halo, nmask, einfo, eE64, eE32, dinv64, dinv32;
    DevMem blist;                  // bricks for the FP32 matvecs: n_sparse sparse ones first, then the dense ones
    int n_sparse = 0;
    DevMem b, x, t, d;             // right-hand side, iterate, matvec output, Chebyshev direction
};

// Mesh of one HEX20 micro-element domain (owning) and its device view.
struct Dom20 {
    DevMem coords;                 // double [node][3], mesh order
    DevMem elems;                  // int [elem][20]: 0-based local node ids
    DevMem emod;                   // double per element
    DevMem evox;                   // int per element: its voxel (x fastest) in the domain
    int64_t nvox = 0;
    DevMem adj_off, adj;           // node -> sorted (element << 5 | local node)
    DevMem bcf;                    // uint8 per node: Dirichlet node
    DevMem scan_tmp;
    DevMem lat20;                  // int [(2nx+1)(2ny+1)(2nz+1)]: local node at a half-voxel lattice point, -1 none
    int Lx = 0, Ly = 0, Lz = 0;    // half-voxel lattice extents
    int64_t node_begin = 0, n_nodes = 0;
};
struct Dom20Dev {
    const double* coords;
    const int* elems;
    const double* emod;
    const int* adj_off;
    const int* adj;
    const unsigned char* bcf;
    const int* lat20;
    int64_t node_begin;
    int n_nodes;
    int Lx, Ly, Lz;
};

// An instantiated HIP graph and the captured graph it was made from (both destroyed
// with their owner; the captured graph is kept while its instance lives, since
// tools that follow graph launches, such as a kernel-trace profiler, look at it)
struct GraphExec {
    hipGraphExec_t e = nullptr;
    hipGraph_t g = nullptr;
    GraphExec() = default;
    GraphExec(const GraphExec&) = delete;
    GraphExec& operator=(const GraphExec&) = delete;
    ~GraphExec() { reset(); }
    void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
    }
};

// Page-locked host words the device writes (hipHostMalloc, freed with the owner).
struct HostFlags {
    int* p = nullptr;
    HostFlags() = default;
    HostFlags(const HostFlags&) = delete;
    HostFlags& operator=(const HostFlags&) = delete;
    ~HostFlags() {
        if (p) (void)hipHostFree(p);
    }
};

// One batch of domains resident on the device: topology, vectors, scalars.
struct SolverBatch {
    int n_dom = 0;
    int n_bricks = 0;
    int64_t n_nodes = 0;           // solver nodes over all domains
    int64_t n_lat = 0, n_vox = 0;  // lattice map / voxel entries over all domains
    std::vector<DomInfo> h_dom;
    std::vector<BrickInfo> h_brick;
    std::vector<int64_t> h_elems, h_nodes, h_bnodes;   // per domain counts
    std::vector<int> h_status;
    DevMem dom, brick;             // DomInfo[], BrickInfo[]
    DevMem lat2node;               // int32 per lattice node (-1 = no solver node)
    DevMem sol2ref;                // int32 per solver node: mesh node id (1-based)
    DevMem emod;                   // double per voxel (0 = no element)
    // per-brick matvec tables, built once per solve
    DevMem halo;                   // int [brick][HX*HY*HZ]: solver node | BC_BIT, -1 = none
    DevMem nmask;                  // uint8 [brick][BX*BY*BZ]: solid tile elements around each owned node
    DevMem einfo;                  // int per tile element: halo base | (owned base + 41) << 9 | owned corners << 24
    DevMem eE64, eE32;             // element modulus per tile element
    int64_t n_tile_elems = 0;
    DevMem blist;                  // bricks for the FP32 matvecs: n_sparse sparse ones first, then the dense ones
    int n_sparse = 0;
    DevMem dinv64, dinv32;         // Jacobi: 1 / diag(K) per DOF [node][3] (0 on Dirichlet DOFs)
    // vectors, layout [node][3][24]
    DevMem x64, r64, p64, q64;
    DevMem x32, r32, p32, q32;
    // per-brick partial sums
    DevMem part_a, part_b, part_c; // [brick][24]
    DevMem part_ff;                // [brick][24 lc][60 macro DOF slots]
    // per-domain scalars [dom][24]
    DevMem alpha, beta, rz, rr, rr0, thr2, scale;
    DevMem active, active64;       // int [dom][24]
    DevMem active_prev;            // inner activity when the residual was last replaced
    DevMem done;                   // int [dom]: inner iteration finished
    DevMem done64;                 // int [dom]: outer (FP64 residual) converged
    DevMem iters;                  // int64 [dom]
    DevMem ff;                     // double [dom][24 lc][60 macro DOF slots]
    DevMem gtmp;                   // gradient work
    // multigrid preconditioner of the FP32 PCG (empty when Jacobi is used)
    std::vector<Level> mg;
    DevMem z32, t32, d32;          // fine-level V-cycle output, matvec output, Chebyshev direction
    // block PCG of the mixed multigrid solve (the 24 load cases of a domain share one Krylov block)
    int n_slices = 0;
    DevMem slices;                 // int4 [slice]: domain, first node, end node
    DevMem dslice;                 // int [dom + 1]: first slice of each domain
    DevMem gpart;                  // double [slice][2][24 x 24]: Gram partial sums
    DevMem gsum;                   // double [dom][2][24 x 24]: Gram sums
    DevMem lfac;                   // double [dom][24 x 24]: Cholesky factor of P^T A P
    DevMem bcoef;                  // float [dom][24 x 24]: alpha, then beta, of the current step
    bool block_q64 = false;        // Q = A P in FP64 arithmetic
    bool block64 = false;          // block PCG on FP64 vectors (x, r, Q) with the FP32 V-cycle
    int precond = 0;               // DTC_PC_JACOBI / DTC_PC_MG
    int mg_degree = 2;             // Chebyshev degree of the fine and intermediate levels
    double cheb_lmax = 0.0;        // element-by-element bound of lambda_max(D^-1 K)
    // HEX20 micro elements: node-gather operator over pseudo-bricks of mesh nodes
    bool micro20 = false;
    // p-multigrid of HEX20 micro domains: the HEX08 discretisation of the same
    // voxels (a batch of its own, same domain groups, its geometric V-cycle) as
    // the coarse space; the fine level is smoothed with Chebyshev-Jacobi on the
    // interval [cheb_lmax / ratio, cheb_lmax] of D^-1 K (HEX20 element bound)
    SolverBatch* p08 = nullptr;
    const int* done_ext = nullptr;  // done flags to use instead of `done` (the coarse batch of a p-multigrid)
    std::vector<Dom20> d20;        // per domain group (empty entries for failed groups)
    DevMem d20dev;                 // Dom20Dev[]
    std::vector<Dom20Dev> h_d20dev;   // host copy
    DevMem ye20;                   // [element][60][24] element forces (FP32 or FP64) of the two-pass operator
    DevMem ebase20;                // int64 [group]: first element of each domain group in ye20
    std::vector<int64_t> h_ebase20;
    std::vector<double> khat20;    // E = 1 element stiffness, 60 x 60
    DevMem k20_64, k20_32;
    WalshK wk64{};
    // element products on the matrix cores (FP32 matvecs): B operands of the
    // normalised element stiffness, its normalisation, and which kernels use them
    DevMem kmf;                    // uint4 [tile 2][term 2][lane 64]
    float kmf_max = 0.f;
    int mf_cg = 0;                 // Q = K P (and q = K p): 0 Walsh VALU kernel, 1 MFMA (3 FP16 terms)
    int mf_smooth = 0;             // smoothing and restriction residual: 0 Walsh, 1 or 3 MFMA terms
    double kdiag = 0.0;            // diagonal entry of the E = 1 element stiffness
    double nu = 0.3, a = 1.0;
    std::vector<double> h_contrast;   // per domain group: modulus contrast (max / min) over its solid voxels
    DevMem itol2;                  // double [dom]: squared inner (FP32 phase) reduction of each domain group
    int precision = DTC_FP64;
    bool reactions_valid = false;
    bool capturing = false;
    hipStream_t stream = nullptr;
    // kernel timing (off unless enabled): HIP events around the fine-level
    // launches of each timed kernel, and the algorithmic bytes of each launch
    // summed on the device over the domains still iterating (k_launch_bytes)
    std::vector<hipEvent_t> ev, ev_free;
    struct Timed { hipEvent_t e0, e1; int kind; int slot; };
    std::vector<Timed> ev_used;
    DevMem dom_bytes;              // double [dom][5]: vector bytes V, modulus bytes, Jacobi bytes, FP64 V,
                                   // dense element-product flops
    DevMem launch_bytes;           // double [slot][2]: algorithmic bytes and flops of a timed launch
    int n_slots = 0;
    KernelStat kstat[DTC_KSTAT_COUNT];
    bool timing = false;
    // the block PCG's iterations captured as one HIP graph (replayed until every
    // domain is done), the arguments it was captured with, and the device word /
    // page-locked host words of its "all domains done" summary
    GraphExec block_graph;
    const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    int64_t graph_iters = -1;
    int graph_bricks = -1;
    int graph_mode = -1;           // block64 of the captured iterations
    DevMem all_done_dev;           // int: 1 when every domain's done flag is set
    HostFlags all_done_host;       // int [2]: copies of all_done_dev, one per launch parity
    int graph_launches = 0;        // graph launches of the last solve (statistics)
};

// Walsh-basis block coefficients of the E = 1 HEX08 element stiffness.
int walsh_blocks(const double* Khat, WalshK* out);
// Host element stiffness by Gauss quadrature (column-major 3nen x 3nen).
int element_stiffness(int nen, double E, double nu, double a, double* K);

// Reserve the lattice map and voxel modulus arrays of n_dom domains.
int batch_begin(SolverBatch& b, const dtc_domain_desc* d, int64_t n_dom);
// Append one meshed domain (HEX08 micro) to the batch topology as the group of
// 24 macro load cases starting at lc_base (load cases past the macro element's
// count are zero).  modulus_dev: optional per-voxel moduli on the device (n^3),
// otherwise d.E; escale multiplies every modulus.
int batch_add_domain(SolverBatch& b, const MeshDev& m, const dtc_domain_desc& d, const double* modulus_dev,
                     double escale, int lc_base, hipStream_t s);
// Append a domain that failed before the solve (empty mesh, ...).
int batch_add_empty(SolverBatch& b, const dtc_domain_desc& d, int status, int lc_base);
// Allocate vectors and scalars once all domains are added.
int batch_finalize(SolverBatch& b, int precision);
// Build the coarse levels of the multigrid preconditioner (after batch_finalize).
int batch_build_mg(SolverBatch& b, int max_levels);
// p-multigrid of a HEX20 micro batch b over c, the HEX08 discretisation of the same
// voxels and domain groups (both finalised): builds c's levels and links c to b.
int batch_build_pmg(SolverBatch& b, SolverBatch& c);
int batch_autotune(SolverBatch& b);
// DTC_HEX20_GATHER=1: the HEX20 operator as the node-gather kernel (the reference form;
// no Chebyshev epilogue), else element pass + node sums
inline bool hex20_gather_form() {
    const char* e = std::getenv("DTC_HEX20_GATHER");
    return e && e[0] == '1';
}
// HEX20 FP32 operator with one Chebyshev-Jacobi step in its epilogue: on free rows
// d = c1 d + c2 dinv (bv - K x), x += d (x updated in place; Dirichlet rows: K x = 0)
int mv20_cheb(SolverBatch& b, float* x, const int* done, const float* bv, const float* dinv, float* d, float c1,
              float c2);
// Full solve: returns per-domain status in b.h_status; fills iteration counts.
int batch_solve(SolverBatch& b, const dtc_solve_opts& o, std::vector<int64_t>& iters,
                std::vector<int64_t>& outer, std::vector<double>& relres);
// ff per domain group, [24 lc][60 macro DOF slots] (first 3 x nodes slots used), after a solve.
int batch_forces(SolverBatch& b, std::vector<double>& ff);
// displacements / reactions of domain group i in mesh numbering, column-major
// (3N, ncol): the first ncol load cases of the group.
int batch_fields(SolverBatch& b, int i, int ncol, double* u_out, double* r_out);
// HEX20 micro elements (solver20.hip): append a domain group, finish the tables
// after batch_finalize's allocations, the operator in the HEX08 kernel's modes,
// the force projection and the fields of one group.
int batch_add_domain20(SolverBatch& b, const MeshDev& m, const dtc_domain_desc& d, const double* modulus_dev,
                       double escale, int lc_base, hipStream_t s);
int finalize20(SolverBatch& b);
int mv20(SolverBatch& b, bool fp64, int in, int out, const void* xin, void* yout, double* part, const int* done);
// Q = K P of the block PCG on a HEX20 batch: FP32 directions, FP64 arithmetic
// (masked rows); Q is written to q (FP32), or to q64 (FP64) when that is set
int mv20_q64(SolverBatch& b, const float* p, float* q, const int* done, double* q64 = nullptr);
int project20(SolverBatch& b);
int fields20(SolverBatch& b, int i, int ncol, double* u_dev, double* r_dev);
// dL/dE per voxel of the domain whose load cases are groups i..i+ng-1 of a HEX20
// batch (M: 24 ng x 6, rows of absent load cases zero)
int grad20(SolverBatch& b, int i, int ng, const double* M, double* dL_dE_host);
// sensitivity dL/dE per voxel of the domain whose load cases are groups i..i+ng-1
// (M: 24 ng x 6, rows of absent load cases zero)
int batch_grad(SolverBatch& b, int i, int ng, const double* M, double* dL_dE_host);

}  // namespace dtc
