/*
 * dtc_capi.h -- C ABI of the MI355X (gfx950) implementation of DTC's
 * per-control-volume micro-FE path: voxel -> hex meshing, the 24 load-case
 * solves, the effective-stiffness fit and the rotation search.
 *
 * Conventions
 *  - plain C types only; int64_t indices; FP64 at the boundary;
 *  - arrays are column-major (Fortran order) where they mirror a Fortran
 *    array of the reference, ids are 1-based where the reference's are;
 *  - the caller owns every output buffer; device memory behind a dtc_batch
 *    is owned by the library;
 *  - return value 0 on success, else a DTC_ERR_* code; on failure the output
 *    buffers are left unmodified and dtc_last_error() describes the failure.
 *  - one host thread per dtc_batch; calls return when the results are on
 *    the host (work runs on a stream the batch owns).
 *
 * Reference citations are relative to the DTC repository root.
 */
#ifndef DTC_CAPI_H
#define DTC_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTC_OK                   0
#define DTC_ERR_EMPTY_MESH       1  /* max(phi) < llimit (f-src/mod_gen_quadmesh.f90:142-164) */
#define DTC_ERR_FEW_CONSTRAINTS  2  /* < 6 constrained DOFs (f-src/mod_struct_preprocess.f90:471-473) */
#define DTC_ERR_OOM              3  /* device allocation failed (the reference's mem_critical < 0) */
#define DTC_ERR_NO_CONVERGENCE   4  /* iteration cap reached before rel_tol */
#define DTC_ERR_BAD_ARGUMENT     5
#define DTC_ERR_DEVICE           6  /* HIP runtime error */
#define DTC_ERR_UNSUPPORTED      7  /* combination not implemented (e.g. HEX20 micro solve) */

/* voxel kinds: signed Integer(1/2/4) as typeraw ik1/ik2/ik4
 * (f-src/mod_struct_preprocess.f90:55-57, Appendix quirk: uint8 > 127 is negative under ik1) */
#define DTC_IK1 1
#define DTC_IK2 2
#define DTC_IK4 4

/* precision modes of the solve */
#define DTC_FP64        0   /* Jacobi-PCG in FP64 throughout */
#define DTC_FP32_REFINE 1   /* FP32 PCG inner solves wrapped in FP64 iterative refinement */

/* preconditioners of the PCG (DTC_FP64 always uses Jacobi, as PETSc's PCJACOBI) */
#define DTC_PC_JACOBI   0   /* diagonal scaling (the reference's choice) */
#define DTC_PC_MG       1   /* geometric multigrid V-cycle over 2x2x2 voxel coarsening, Chebyshev-Jacobi smoothing */

const char* dtc_last_error(void);
int dtc_version(void);

/* ---- a1: domain decomposition arithmetic --------------------------------
 * replaces calc_general_ddc_params (f-src/mod_decomp.f90:247-326) */
int dtc_ddc_params(const double size_domain[3], const double delta[3], const int64_t vdim[3],
                   int64_t x_D[3], int64_t nn_D[3], double x_D_phy[3]);
/* replaces the domain index math of generate_geometry (f-src/mod_struct_preprocess.f90:85-98) */
int dtc_domain_bounds(int64_t domain, const int64_t x_D[3], const int64_t nn_D[3],
                      int64_t nn_i[3], int64_t xa_n[3], int64_t xe_n[3]);

/* ---- a2: voxel block of one domain from a whole volume ------------------
 * replaces the stream read of f-src/mod_struct_preprocess.f90:230-244; vol
 * and out are host arrays (x fastest), the copy runs on the GPU. */
int dtc_extract_block(const void* vol, int kind, const int64_t vdim[3], const int64_t xa_n[3],
                      const int64_t xe_n[3], void* out);

/* ---- a3: voxel -> hex mesh, bit-exact to gen_quadmesh_from_phi ----------
 * replaces the call at f-src/mod_struct_preprocess.f90:275-288
 * (f-src/mod_gen_quadmesh.f90:22-679).  phi: n[0]*n[1]*n[2] voxels, x fastest.
 * nen = 8 (HEX08) or 20 (HEX20).  Two-phase like METIS: count, then build
 * with nodes (3 x N_n, column-major) and elems (nen x N_e, 1-based ids). */
int dtc_mesh_count(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen,
                   int64_t* n_elems, int64_t* n_nodes);
int dtc_mesh_build(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen,
                   const double delta[3], const int64_t xa_n[3], int64_t elems_cap,
                   int64_t nodes_cap, double* nodes, int64_t* elems, int64_t* n_elems,
                   int64_t* n_nodes);

/* ---- a5/a6: boundary nodes and prescribed displacements -----------------
 * replaces generate_boundaries (f-src/mod_struct_preprocess.f90:490-738) with
 * determine_prescribed_displ (f-src/mod_write_deck.f90:22-53), HEX08 macro.
 * bnode_ids == NULL: only *n_b is returned.  bvals: column-major (24, 3*n_b),
 * i.e. the reference's bnode_vals(lc, 3*(b-1)+d). */
int dtc_boundaries(const double* nodes, int64_t n_nodes, const double min_c[3],
                   const double max_c[3], const double dim_c[3], int macro_order,
                   int64_t* n_b, int64_t* bnode_ids, double* bvals);

/* ---- a7: element stiffness (column-major 3nen x 3nen) -------------------
 * equals Hexe08 / Hexe20 (f-src/mod_linfe.f90:161-766, :890-4509) with a = phdsize*delta */
int dtc_element_stiffness(int nen, double E, double nu, double a, double* K);

/* ---- a8-a11: whole-domain solve ------------------------------------------ */
typedef struct {
    const void* phi;        /* voxel block, host, n[0]*n[1]*n[2], x fastest */
    int phi_kind;           /* DTC_IK1 / DTC_IK2 / DTC_IK4 */
    int micro_nen;          /* 8 (solve supports HEX08 micro elements) */
    int macro_order;        /* 1 = HEX08 macro element, 24 load cases */
    int _pad;
    int64_t n[3];           /* voxels per edge (x_D) */
    int64_t llimit;         /* Lower limit of iso value (BINARIZE_LO) */
    int64_t xa_n[3];        /* first voxel of the domain (1-based) */
    double delta[3];        /* SPACING */
    double size_domain;     /* SIZE_DOMAIN (phdsize); K_e uses a = phdsize*delta */
    double E, nu;           /* YOUNG_MODULUS, POISSON_RATIO */
    double rve_strain;      /* RVE_STRAIN (used in vv) */
    const double* modulus;  /* optional per-voxel modulus E_v (host, n^3); NULL = uniform E */
} dtc_domain_desc;

typedef struct {
    int precision;          /* DTC_FP64 or DTC_FP32_REFINE */
    int rotation;           /* 0: skip; 1: CR_1 and CR_2 searches */
    double rel_tol;         /* true ||f - K u||_2 / ||f||_2 per load case (e.g. 1e-8) */
    double inner_tol;       /* FP32_REFINE: relative reduction per inner solve; 0 = automatic
                               (1e-6 for the multigrid block PCG when the per-voxel moduli
                               differ by at most 100x, 1e-4 otherwise) */
    int64_t max_iter;       /* cap on the total PCG iterations */
    double rot_step_deg;    /* initial grid step of the rotation search (1 = reference) */
    double num_zero;        /* refinement stop |dmin| < num_zero (1e-9 in mod_tensor_opt.f90:61) */
    int precond;            /* DTC_PC_JACOBI or DTC_PC_MG (FP32_REFINE only) */
    int mg_degree;          /* Chebyshev degree per smoothing pass (0 = default 2) */
} dtc_solve_opts;

typedef struct {
    int status;                 /* DTC_OK or DTC_ERR_* for this domain */
    int precision_used;         /* DTC_FP32_REFINE, or DTC_FP64 when the solve ran in FP64 (requested,
                                   or the mixed solve stalled and the domain was repeated in FP64) */
    int64_t n_elems, n_nodes, n_bnodes;
    int64_t iterations;         /* PCG iterations (inner, summed) */
    int64_t outer_iterations;   /* FP64 residual evaluations */
    double relres_max;          /* max over load cases of the true relative residual */
    double ff[3600];            /* "Domain forces" ff(dof, lc), column-major R x R, R = 24 (HEX08
                                   macro element) or 60 (HEX20), dof = (d-1)*nodes + l */
    double K_macro[3600];       /* "Effective numerical stiffness" ff vv^-1, R x R (HEX20: vv from
                                   the boundary table, the reference's own vv being singular) */
    double C[36];               /* "Effective stiffness" fv meps */
    double EE[36];              /* "Averaged Effective stiffness" (C + C^T)/2 */
    double density;             /* "Effective density" */
    double rot_alpha[2];        /* "Rotation Angle CR_1/CR_2" */
    double rot_n[6];            /* "Rotation Vector CR_1/CR_2" (3 each) */
    double rot_aa[18];          /* "Final coordinate system CR_1/CR_2" (9 each, column-major) */
    double rot_EE[72];          /* "Optimized Effective stiffness CR_1/CR_2" (36 each) */
    int64_t rot_jj[2];          /* refinement iterations */
    double rot_crit[2];         /* final criterion values */
} dtc_domain_result;

/* One domain; replaces f-src/mod_dtc_main_subroutines.f90:426-1238 (assembly,
 * Dirichlet lifting, the KSPSolve loop, reactions) plus the ff/C part of
 * calc_effective_material_parameters (f-src/mod_struct_calcmat.f90:252-473)
 * and, with opts->rotation, its rotation search (:533-1603).
 * u_out / r_out (optional, NULL to skip): displacements / reaction forces,
 * column-major (3*N_n, 24) in the mesh numbering of dtc_mesh_build. */
int dtc_solve_domain(const dtc_domain_desc* d, const dtc_solve_opts* o, dtc_domain_result* res,
                     double* u_out, double* r_out);
/* Many domains in one batched solve (the master/worker dispatch of
 * f-src_main/struct_process.f90:963-1094 becomes one launch sequence). */
int dtc_solve_batch(const dtc_domain_desc* d, int64_t n_dom, const dtc_solve_opts* o,
                    dtc_domain_result* res);

/* ---- device-resident batch (voxels resident in HBM) ---------------------- */
typedef struct dtc_batch dtc_batch;
int dtc_batch_create(const dtc_domain_desc* d, int64_t n_dom, const dtc_solve_opts* o,
                     dtc_batch** out);
/* mesh -> boundary conditions -> solve -> fit -> rotation search, all on the GPU */
int dtc_batch_run(dtc_batch* b, dtc_domain_result* res);
/* sensitivity of L = sum_ij dL_dC[ij] C_ij to every voxel modulus of domain i
 * after dtc_batch_run (self-adjoint: uses the stored load-case solutions,
 * no extra solve).  dL_dE: n^3 per voxel (0 where no element).  HEX08 or HEX20
 * micro elements, HEX08 or HEX20 macro element. */
int dtc_batch_grad(dtc_batch* b, int64_t i, const double dL_dC[36], double* dL_dE);
/* timing of the last run: average duration (ms) of the CG matvec kernel
 * launches measured with HIP events on the batch's stream, the number of
 * timed launches, and the algorithmic bytes and flops of an average launch
 * (a domain counts in the launches of the PCG iterations it took; finished
 * domains are skipped by the kernels and not counted). */
int dtc_batch_stats(dtc_batch* b, double* mv_avg_ms, int64_t* mv_launches, double* mv_bytes,
                    double* mv_flops, double* run_ms);
/* kernel timing of the following runs on (1) or off (0, the default): HIP
 * events on the solve stream around the fine-level launches of the timed
 * kernels below.  No reference counterpart (the reference's timers,
 * f-src/mod_times.f90:25,99, time whole program sections). */
int dtc_batch_timing(dtc_batch* b, int enable);
#define DTC_KSTAT_MV    0   /* CG matvec: Q = K P of the block PCG / q = K p of the PCG (FP32) */
#define DTC_KSTAT_CHEB  1   /* Chebyshev-Jacobi smoothing step fused into the matvec (fine level) */
#define DTC_KSTAT_BRES  2   /* b - K x before restriction (fine level) */
#define DTC_KSTAT_RESID 3   /* FP64 true residual f - K u */
#define DTC_KSTAT_COUNT 4
/* average duration (ms) and average algorithmic bytes of one launch of timed
 * kernel `kernel` (DTC_KSTAT_*) over the last timed run, and the number of
 * launches.  Bytes count each vector the launch reads or writes once (3 N_n R s
 * per vector), the element moduli (4 N_e) and the Jacobi scaling (12 N_n) of
 * the domains still iterating in that launch.  The FP64 repeat of domains whose
 * mixed solve stalled is not included. */
int dtc_batch_kernel_stats(dtc_batch* b, int kernel, double* avg_ms, int64_t* launches, double* avg_bytes);
/* the same launches' average dense element-product flops, 2 (3 nen)^2 R N_e over
 * the domains still iterating (the flop count of an assembled element-by-element
 * product; the Walsh-basis kernel executes fewer), for the flop-side roof of the
 * operator (the counterpart of the per-launch work the reference's
 * KSPSolve matvec does in f-src/mod_dtc_main_subroutines.f90:995-1194). */
int dtc_batch_kernel_flops(dtc_batch* b, int kernel, double* avg_flops);
int dtc_batch_destroy(dtc_batch* b);

/* ---- a12: rotation search ------------------------------------------------
 * replaces the CR_1 (criterion 1) / CR_2 (criterion 2) search of
 * f-src/mod_struct_calcmat.f90:533-1603 (library twin
 * f-src/mod_tensor_opt.f90:40-601).  EE, EE_opt column-major 6x6. */
int dtc_rotation_search(const double EE[36], int criterion, double step_deg, double num_zero,
                        double* alpha, double n[3], double aa[9], double EE_opt[36],
                        int64_t* jj, double* crit);

/* Rotation search that allocates nothing (no hipMalloc/hipFree, no stream
 * creation): no counterpart in the reference.  For callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid it).
 * Same arguments, results and errors as dtc_rotation_search; it runs on the
 * null stream with its buffers in a fixed per-GPU work space of 256 KB, held
 * until the process ends, which bounds the initial sweep to
 * dtc_rotation_noalloc_max_points() grid points (about 33 M: an initial step
 * of 0.45 degrees or more; 1 degree needs 25 KB).  A larger sweep is
 * rejected with DTC_ERR_BAD_ARGUMENT.  Calls that overlap on one GPU share the
 * work space and race: serialise them per GPU. */
int dtc_rotation_search_noalloc(const double EE[36], int criterion, double step_deg, double num_zero,
                                double* alpha, double n[3], double aa[9], double EE_opt[36],
                                int64_t* jj, double* crit);
int64_t dtc_rotation_noalloc_max_points(void);

#ifdef __cplusplus
}
#endif
#endif
