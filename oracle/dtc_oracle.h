/*
 * dtc_oracle.h -- CPU FP64 restatement of DTC's per-control-volume micro-FE path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * library (libdtc_amd.so).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product never links or calls it.
 *
 * Every function names the reference file:line it restates.  Reference paths
 * are relative to the DTC repository root.
 */
#ifndef DTC_ORACLE_H
#define DTC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* voxel kinds: signed Integer(1), Integer(2), Integer(4)
 * (f-src/mod_struct_preprocess.f90:55-57) */
enum { DTCO_IK1 = 1, DTCO_IK2 = 2, DTCO_IK4 = 4 };

void dtco_free(void* p);

/* a1: f-src/mod_decomp.f90:247-326 and f-src/mod_struct_preprocess.f90:85-98 */
int dtco_ddc_params(const double size_domain[3], const double delta[3], const int64_t vdim[3],
                    int64_t x_D[3], int64_t nn_D[3], double x_D_phy[3]);
int dtco_domain_bounds(int64_t domain, const int64_t x_D[3], const int64_t nn_D[3],
                       int64_t nn_i[3], int64_t xa_n[3], int64_t xe_n[3]);
/* a2: f-src/mod_struct_preprocess.f90:165-251 (x fastest) */
int dtco_extract_block(const void* vol, int kind, const int64_t vdim[3], const int64_t xa_n[3],
                       const int64_t xe_n[3], void* out);

/* a3: f-src/mod_gen_quadmesh.f90:22-679.  Allocates *nodes (3 x N_n) and
 * *elems (nen x N_e, 1-based ids); free with dtco_free.  Returns 1 for an
 * empty domain (max(phi) < llimit). */
int dtco_mesh(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen,
              const double delta[3], const int64_t xa_n[3],
              int64_t* n_nodes, int64_t* n_elems, double** nodes, int64_t** elems);

/* a5/a6: f-src/mod_struct_preprocess.f90:490-738, f-src/mod_write_deck.f90:22-279.
 * *ids: boundary node ids (1-based, node order); *vals: column-major (R, 3*n_b). */
int dtco_boundaries(const double* nodes, int64_t n_nodes, const double min_c[3],
                    const double max_c[3], const double dim_c[3], int macro_order,
                    int64_t* n_b, int64_t** ids, double** vals);

/* macro shape functions: f-src/mod_linfe.f90:51-84 (HEX08), :827-888 (HEX20, ANSYS order) */
int dtco_macro_shape(int macro_order, const double x[3], const double min_c[3],
                     const double max_c[3], double* N);
/* macro unit displacement tables uu(l,d,lc), eps = 1e-6: f-src/mod_write_deck.f90:63-107 */
int dtco_uu(int macro_order, double* uu);
/* load-case matrix vv(dof,lc) for HEX08 macro: f-src/mod_struct_calcmat.f90:1684-1770 */
int dtco_vv_lin(double rve_strain, double* vv);
/* HEX20 macro: vv(dof, lc) = uu * RVE_STRAIN / 1e-6, 60 x 60 (see dtco_effective) */
int dtco_vv_quad_consistent(double rve_strain, double* vv);

/* a7: element stiffness by Gauss quadrature (2x2x2 HEX08, 3x3x3 HEX20) in
 * gen_quadmesh node order; equals f-src/mod_linfe.f90:161-766 / :890-4509. */
int dtco_ke(int nen, double E, double nu, double a, double* K);

/* a8/a9/a10: assembled operator (3x3 block CSR), Dirichlet lifting and
 * MatZeroRowsColumns semantics (f-src/mod_dtc_main_subroutines.f90:426-951),
 * Jacobi-preconditioned CG per load case with x0 = 0 (:995-1194,
 * f-src/mod_parameters.f90:22-32), reactions r = A_org u, and the force
 * projection ff (f-src/mod_struct_calcmat.f90:252-280).
 * emod: per-element modulus (NULL = uniform E).  K_e uses a = phdsize*delta.
 * norm_type 0: stop on true ||r||_2/||f||_2 <= tol; 1: PETSc default
 * (preconditioned norm ||D^-1 r|| / ||D^-1 f||).
 * u, r (optional, may be NULL): column-major (3*N_n, R).
 * ff: column-major (R, R).  iters, relres: per load case (true residual). */
int dtco_solve_domain(const double* nodes, int64_t n_nodes, const int64_t* elems,
                      int64_t n_elems, int nen, const double* emod, double E, double nu,
                      double a, const double min_c[3], const double max_c[3],
                      const double dim_c[3], int macro_order, double tol, int64_t max_it,
                      int norm_type, int lc_first, int lc_count,
                      double* u, double* r, double* ff, int64_t* iters, double* relres);

/* a10/a11: K_M = ff vv^-1, fv, C = fv meps, EE = (C+C^T)/2
 * (f-src/mod_struct_calcmat.f90:216-303, :340-473), R = 24 (macro_order 1) or
 * 60 (macro_order 2).  K_M may be NULL.  For the HEX20 macro element vv is
 * dtco_vv_quad_consistent (the reference's own table is singular). */
int dtco_effective(const double* ff, int macro_order, double rve_strain,
                   const double x_D_phy[3], double* K_M, double* C, double* EE);
/* density: f-src/mod_struct_calcmat.f90:1642-1646 */
double dtco_density(int64_t n_elems, const double x_D_phy[3], const double delta[3]);

/* a12: CR_1 (criterion 1) / CR_2 (criterion 2) grid search + refinement +
 * axis permutation, f-src/mod_struct_calcmat.f90:533-1603 (step_deg = 1
 * reproduces the reference grid; mod_tensor_opt.f90:40-601 is the library twin).
 * Outputs: alpha, n (pre-permutation, as written to the result leaves), aa
 * (final coordinate system, column-major 3x3), EE_opt (6x6), jj (refinement
 * iterations), crit (final minimum). */
int dtco_rotation_search(const double EE[36], int criterion, double step_deg, double num_zero,
                         double* alpha, double n[3], double aa[9], double EE_opt[36],
                         int64_t* jj, double* crit);
/* T = B(R)^T EE B(R) for the rotation of angle alpha about n (rot_alg, tra_R6
 * as inlined at f-src/mod_struct_calcmat.f90:556-589). */
int dtco_rotate66(const double EE[36], const double n[3], double alpha, double T[36]);

/* Sensitivity of the 6x6 effective stiffness: given the 24 load-case
 * displacements u (3N_n x 24, or x 60 for the HEX20 macro element), returns
 * dL/dE_e for L = sum G_ij C_ij with G = dL/dC (6x6 column-major).  HEX08 or
 * HEX20 micro (nen), HEX08 or HEX20 macro (macro_order). */
int dtco_grad_modulus(const double* nodes, int64_t n_nodes, const int64_t* elems,
                      int64_t n_elems, int nen, int macro_order, double nu, double a,
                      const double min_c[3], const double max_c[3], const double x_D_phy[3],
                      double rve_strain, const double* u, const double G[36], double* dL_dE);

#ifdef __cplusplus
}
#endif
#endif
