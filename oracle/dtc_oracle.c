/*
 * dtc_oracle.c -- CPU FP64 restatement of DTC's per-control-volume micro-FE
 * path, used as the parity checker for the HIP library.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Never linked into the product.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the DTC repository root).  Build with -ffp-contract=off so that FP64 results
 * that the reference computes with plain IEEE operations (coordinates,
 * boundary values) come out bit-identical to the GPU kernels.
 */
#include "dtc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

void dtco_free(void* p) { free(p); }

static int64_t fortran_nint(double x) { return (int64_t)llround(x); }

/* ------------------------------------------------------------------------ */
/* a1  f-src/mod_decomp.f90:247-326                                          */
/* ------------------------------------------------------------------------ */
int dtco_ddc_params(const double size_domain[3], const double delta[3], const int64_t vdim[3],
                    int64_t x_D[3], int64_t nn_D[3], double x_D_phy[3]) {
    for (int d = 0; d < 3; ++d) {
        if (!(delta[d] > 0.0)) return 5;
        x_D[d] = fortran_nint(size_domain[d] / delta[d]);          /* :283 */
        if (x_D[d] <= 0) return 5;
        x_D_phy[d] = (double)x_D[d] * delta[d];                     /* :288 */
        nn_D[d] = (vdim[d] - 2) / x_D[d];                           /* :294, bpoints = 1 */
    }
    return 0;
}

/* f-src/mod_struct_preprocess.f90:85-98 */
int dtco_domain_bounds(int64_t domain, const int64_t x_D[3], const int64_t nn_D[3],
                       int64_t nn_i[3], int64_t xa_n[3], int64_t xe_n[3]) {
    if (domain < 0 || nn_D[0] <= 0 || nn_D[1] <= 0 || nn_D[2] <= 0) return 5;
    int64_t n3 = domain / (nn_D[0] * nn_D[1]);
    int64_t n2 = (domain - nn_D[0] * nn_D[1] * n3) / nn_D[0];
    int64_t n1 = domain - nn_D[0] * nn_D[1] * n3 - nn_D[0] * n2;
    nn_i[0] = n1; nn_i[1] = n2; nn_i[2] = n3;
    for (int d = 0; d < 3; ++d) {
        xa_n[d] = x_D[d] * nn_i[d] + 1 + 1;
        xe_n[d] = x_D[d] * (nn_i[d] + 1) + 1;
    }
    return 0;
}

static int64_t phi_at(const void* phi, int kind, int64_t idx) {
    switch (kind) {
    case DTCO_IK1: return (int64_t)((const int8_t*)phi)[idx];
    case DTCO_IK2: return (int64_t)((const int16_t*)phi)[idx];
    default: return (int64_t)((const int32_t*)phi)[idx];
    }
}

/* a2  f-src/mod_struct_preprocess.f90:230-244: rows of x read at
 * pos = vdim1*vdim2*(k-1) + vdim1*(j-1) + (xa_n1-1) (elements, 1-based k,j). */
int dtco_extract_block(const void* vol, int kind, const int64_t vdim[3], const int64_t xa_n[3],
                       const int64_t xe_n[3], void* out) {
    int64_t nx = xe_n[0] - xa_n[0] + 1, ny = xe_n[1] - xa_n[1] + 1, nz = xe_n[2] - xa_n[2] + 1;
    if (nx <= 0 || ny <= 0 || nz <= 0) return 5;
    for (int d = 0; d < 3; ++d)
        if (xa_n[d] < 1 || xe_n[d] > vdim[d]) return 5;
    size_t es = (size_t)(kind == DTCO_IK1 ? 1 : kind == DTCO_IK2 ? 2 : 4);
    for (int64_t k = 0; k < nz; ++k)
        for (int64_t j = 0; j < ny; ++j) {
            int64_t src = vdim[0] * vdim[1] * (xa_n[2] - 1 + k) + vdim[0] * (xa_n[1] - 1 + j) +
                          (xa_n[0] - 1);
            int64_t dst = (k * ny + j) * nx;
            memcpy((char*)out + dst * es, (const char*)vol + src * es, (size_t)nx * es);
        }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a3  gen_quadmesh_from_phi, f-src/mod_gen_quadmesh.f90:22-679              */
/* ------------------------------------------------------------------------ */

/* HEX20 local node offsets on the doubled lattice, f-src/mod_gen_quadmesh.f90:244-265 */
static const int H20_OFF[20][3] = {
    {0, 0, 0}, {1, 0, 0}, {2, 0, 0}, {2, 1, 0}, {2, 2, 0}, {1, 2, 0}, {0, 2, 0}, {0, 1, 0},
    {0, 0, 1}, {2, 0, 1}, {2, 2, 1}, {0, 2, 1},
    {0, 0, 2}, {1, 0, 2}, {2, 0, 2}, {2, 1, 2}, {2, 2, 2}, {1, 2, 2}, {0, 2, 2}, {0, 1, 2}};
/* HEX08 local node offsets, f-src/mod_gen_quadmesh.f90:298-306 */
static const int H08_OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

int dtco_mesh(const void* phi, int kind, const int64_t n[3], int64_t llimit, int nen,
              const double delta[3], const int64_t xa_n[3],
              int64_t* n_nodes_out, int64_t* n_elems_out, double** nodes_out, int64_t** elems_out) {
    *n_nodes_out = 0; *n_elems_out = 0; *nodes_out = NULL; *elems_out = NULL;
    if (nen != 8 && nen != 20) return 5;
    const int64_t nx = n[0], ny = n[1], nz = n[2];
    const int64_t nvox = nx * ny * nz;
    if (nvox <= 0) return 5;

    /* :142-164: empty domain when max(phi) < llimit */
    int64_t maxv = phi_at(phi, kind, 0);
    for (int64_t v = 1; v < nvox; ++v) { int64_t p = phi_at(phi, kind, v); if (p > maxv) maxv = p; }
    if (maxv < llimit) return 1;

    const int mul = (nen == 8) ? 1 : 2;
    const int64_t Lx = nx * mul + 1, Ly = ny * mul + 1, Lz = nz * mul + 1;
    const int64_t nlat = Lx * Ly * Lz;

    /* :226-318: element list, scan kk,jj,ii (ii fastest) */
    int64_t* elems = (int64_t*)malloc(sizeof(int64_t) * (size_t)nen * (size_t)nvox);
    int64_t* nodes_no = (int64_t*)calloc((size_t)nlat + 1, sizeof(int64_t));
    if (!elems || !nodes_no) { free(elems); free(nodes_no); return 3; }
    int64_t ne = 0;
    for (int64_t kk = 0; kk < nz; ++kk)
        for (int64_t jj = 0; jj < ny; ++jj)
            for (int64_t ii = 0; ii < nx; ++ii) {
                if (phi_at(phi, kind, (kk * ny + jj) * nx + ii) < llimit) continue;
                for (int l = 0; l < nen; ++l) {
                    const int* o = (nen == 8) ? H08_OFF[l] : H20_OFF[l];
                    int64_t lat = (ii * mul + o[0]) + (jj * mul + o[1]) * Lx + (kk * mul + o[2]) * Lx * Ly;
                    nodes_no[lat] = -1;
                    elems[ne * nen + l] = lat;
                }
                ++ne;
            }

    /* :326-341: first-touch renumbering */
    int64_t nn = 0;
    for (int64_t e = 0; e < ne; ++e)
        for (int l = 0; l < nen; ++l) {
            int64_t lat = elems[e * nen + l];
            if (nodes_no[lat] == -1) nodes_no[lat] = ++nn;
        }
    for (int64_t e = 0; e < ne * nen; ++e) elems[e] = nodes_no[elems[e]];

    /* :346-384: coordinates */
    double* nodes = (double*)malloc(sizeof(double) * 3 * (size_t)nn);
    int64_t* node_col = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    int64_t* elem_col = (int64_t*)malloc(sizeof(int64_t) * (size_t)ne);
    int64_t* node_cref = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nn + 1));
    if (!nodes || !node_col || !elem_col || !node_cref) {
        free(elems); free(nodes_no); free(nodes); free(node_col); free(elem_col); free(node_cref);
        return 3;
    }
    for (int64_t ll = 0; ll < nlat; ++ll) {
        int64_t id = nodes_no[ll];
        if (id == 0) continue;
        int64_t kk = ll / (Lx * Ly);
        int64_t jj = (ll - kk * Lx * Ly) / Lx;
        int64_t ii = ll - kk * Lx * Ly - jj * Lx;
        double* c = nodes + 3 * (id - 1);
        if (nen == 20) {
            c[2] = ((double)kk / 2.0 + (double)(xa_n[2] - 1)) * delta[2];
            c[1] = ((double)jj / 2.0 + (double)(xa_n[1] - 1)) * delta[1];
            c[0] = ((double)ii / 2.0 + (double)(xa_n[0] - 1)) * delta[0];
        } else {
            c[2] = (double)(kk + xa_n[2] - 1) * delta[2];
            c[1] = (double)(jj + xa_n[1] - 1) * delta[1];
            c[0] = (double)(ii + xa_n[0] - 1) * delta[0];
        }
    }
    free(nodes_no);

    /* :408-433: colouring, max label propagation until no change */
    for (int64_t i = 0; i < nn; ++i) node_col[i] = i + 1;
    int change = 1;
    while (change) {
        change = 0;
        for (int64_t e = 0; e < ne; ++e) {
            const int64_t* en = elems + e * nen;
            int64_t mn = node_col[en[0] - 1], mx = mn;
            for (int l = 1; l < nen; ++l) {
                int64_t c = node_col[en[l] - 1];
                if (c < mn) mn = c;
                if (c > mx) mx = c;
            }
            if (mn != mx) {
                for (int l = 0; l < nen; ++l) node_col[en[l] - 1] = mx;
                change = 1;
            }
        }
    }
    /* :438-442 */
    for (int64_t e = 0; e < ne; ++e) elem_col[e] = node_col[elems[e * nen] - 1];

    /* :447-480: distinct colours in order of first appearance */
    int64_t* cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    unsigned char* tb = (unsigned char*)calloc((size_t)nn, 1);
    int64_t ncol = 0;
    {
        /* a colour's list position is found through a lookup keyed by colour value */
        int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nn + 1));
        for (int64_t i = 0; i <= nn; ++i) pos[i] = -1;
        for (int64_t i = 0; i < nn; ++i) {
            int64_t c = node_col[i];
            if (pos[c] < 0) { pos[c] = ncol; cols[ncol++] = c; }
        }
        /* :482-553: x_min/x_max over all nodes; first colour uses the test
         * without abs() (always true through the x_max branch), the others use abs() */
        double xmin[3], xmax[3];
        for (int d = 0; d < 3; ++d) { xmin[d] = nodes[d]; xmax[d] = nodes[d]; }
        for (int64_t i = 1; i < nn; ++i)
            for (int d = 0; d < 3; ++d) {
                double v = nodes[3 * i + d];
                if (v < xmin[d]) xmin[d] = v;
                if (v > xmax[d]) xmax[d] = v;
            }
        for (int64_t i = 0; i < nn; ++i) {
            int64_t ci = pos[node_col[i]];
            if (tb[ci]) continue;
            const double* c = nodes + 3 * i;
            int hit = 0;
            if (ci == 0) {
                for (int d = 0; d < 3 && !hit; ++d)
                    if ((c[d] - xmin[d]) < 0.1 * delta[d] || (c[d] - xmax[d]) < 0.1 * delta[d]) hit = 1;
            } else {
                for (int d = 0; d < 3 && !hit; ++d)
                    if (fabs(c[d] - xmin[d]) < 0.1 * delta[d] || fabs(c[d] - xmax[d]) < 0.1 * delta[d]) hit = 1;
            }
            if (hit) tb[ci] = 1;
        }
        /* :570-592: colours with tb become -1 */
        for (int64_t i = 0; i < nn; ++i) if (tb[pos[node_col[i]]]) node_col[i] = -1;
        for (int64_t e = 0; e < ne; ++e) if (elem_col[e] != -1 && tb[pos[elem_col[e]]]) elem_col[e] = -1;
        free(pos);
    }
    free(cols); free(tb);

    /* :597-614: delete unconnected elements (last element always kept) */
    int64_t jj = 1, kk = 1;   /* 1-based as in the reference */
    while (jj < ne) {
        if (elem_col[jj - 1] != -1) { jj = jj + 1; }
        else { jj = jj + 1; kk = kk + 1; }
        elem_col[kk - 1] = elem_col[jj - 1];
        memmove(elems + (kk - 1) * nen, elems + (jj - 1) * nen, sizeof(int64_t) * (size_t)nen);
    }
    int64_t ne_kept = (ne >= 1) ? kk : 0;

    /* :619-651: delete unconnected nodes (last node always kept) */
    int64_t nn_kept = 0;
    if (nn == 1) {
        /* the reference indexes past the end here (Appendix A3); keep the node */
        node_cref[1] = 1; nn_kept = 1;
    } else if (nn > 1) {
        jj = 1; kk = 1;
        if (node_col[0] == -1) { node_cref[1] = 1; }
        for (;;) {
            if (node_col[jj - 1] != -1) {
                node_cref[jj] = 0;
                node_col[jj - 1] = 0;
                jj = jj + 1;
            } else {
                jj = jj + 1;
                kk = kk + 1;
            }
            node_col[kk - 1] = node_col[jj - 1];
            memmove(nodes + 3 * (kk - 1), nodes + 3 * (jj - 1), sizeof(double) * 3);
            node_cref[jj] = kk;
            if (jj == nn) break;
        }
        nn_kept = kk;
    }
    /* :656-660: renumber element node ids (0 for deleted nodes) */
    for (int64_t e = 0; e < ne_kept * nen; ++e) elems[e] = node_cref[elems[e]];

    free(node_col); free(elem_col); free(node_cref);
    *n_nodes_out = nn_kept;
    *n_elems_out = ne_kept;
    *nodes_out = (double*)realloc(nodes, sizeof(double) * 3 * (size_t)(nn_kept > 0 ? nn_kept : 1));
    *elems_out = (int64_t*)realloc(elems, sizeof(int64_t) * (size_t)nen * (size_t)(ne_kept > 0 ? ne_kept : 1));
    return 0;
}

/* ------------------------------------------------------------------------ */
/* macro shape functions and tables                                          */
/* ------------------------------------------------------------------------ */
/* f-src/mod_linfe.f90:77-84 */
static void t_geom_xi(const double x[3], const double xa[3], const double xe[3], double xi[3]) {
    for (int d = 0; d < 3; ++d) xi[d] = 2.0 / (xe[d] - xa[d]) * x[d] - (xe[d] + xa[d]) / (xe[d] - xa[d]);
}
static double g1(double xi) { return 0.5 * (1.0 - xi); }
static double g2(double xi) { return 0.5 * (1.0 + xi); }
/* f-src/mod_linfe.f90:51-67 */
static void phi_hexe8(const double xi[3], double N[8]) {
    N[0] = g1(xi[0]) * g1(xi[1]) * g1(xi[2]);
    N[1] = g2(xi[0]) * g1(xi[1]) * g1(xi[2]);
    N[2] = g2(xi[0]) * g2(xi[1]) * g1(xi[2]);
    N[3] = g1(xi[0]) * g2(xi[1]) * g1(xi[2]);
    N[4] = g1(xi[0]) * g1(xi[1]) * g2(xi[2]);
    N[5] = g2(xi[0]) * g1(xi[1]) * g2(xi[2]);
    N[6] = g2(xi[0]) * g2(xi[1]) * g2(xi[2]);
    N[7] = g1(xi[0]) * g2(xi[1]) * g2(xi[2]);
}
/* f-src/mod_linfe.f90:827-888 (ANSYS 20-node brick, corners then mid-edges).
 * The reference's constants f = 1./8. and ff = 1./4. are default-REAL
 * literals; both are exact in single precision. */
static void phi_hexe20(const double xi[3], double N[20]) {
    const double f = 0.125, ff = 0.25, e = 1.0, z = 2.0;
    const double x = xi[0], y = xi[1], w = xi[2];
    N[0] = f * (e - x) * (e - y) * (e - w) * (-x - y - w - z);
    N[1] = f * (e + x) * (e - y) * (e - w) * (x - y - w - z);
    N[2] = f * (e + x) * (e + y) * (e - w) * (x + y - w - z);
    N[3] = f * (e - x) * (e + y) * (e - w) * (-x + y - w - z);
    N[4] = f * (e - x) * (e - y) * (e + w) * (-x - y + w - z);
    N[5] = f * (e + x) * (e - y) * (e + w) * (x - y + w - z);
    N[6] = f * (e + x) * (e + y) * (e + w) * (x + y + w - z);
    N[7] = f * (e - x) * (e + y) * (e + w) * (-x + y + w - z);
    N[8] = ff * (e - x * x) * (e - y) * (e - w);
    N[9] = ff * (e + x) * (e - y * y) * (e - w);
    N[10] = ff * (e - x * x) * (e + y) * (e - w);
    N[11] = ff * (e - x) * (e - y * y) * (e - w);
    N[12] = ff * (e - x * x) * (e - y) * (e + w);
    N[13] = ff * (e + x) * (e - y * y) * (e + w);
    N[14] = ff * (e - x * x) * (e + y) * (e + w);
    N[15] = ff * (e - x) * (e - y * y) * (e + w);
    N[16] = ff * (e - x) * (e - y) * (e - w * w);
    N[17] = ff * (e + x) * (e - y) * (e - w * w);
    N[18] = ff * (e + x) * (e + y) * (e - w * w);
    N[19] = ff * (e - x) * (e + y) * (e - w * w);
}

int dtco_macro_shape(int macro_order, const double x[3], const double min_c[3],
                     const double max_c[3], double* N) {
    double xi[3];
    t_geom_xi(x, min_c, max_c, xi);
    if (macro_order == 1) { phi_hexe8(xi, N); return 0; }
    if (macro_order == 2) { phi_hexe20(xi, N); return 0; }
    return 5;
}

/* init_displ_hexe8, f-src/mod_write_deck.f90:63-107, written as
 * (load case, direction, node, multiple of eps) entries. */
static const struct { int lc, d, node; double m; } UU8[] = {
    {1, 1, 2, 1}, {1, 1, 3, 1}, {1, 1, 6, 1}, {1, 1, 7, 1},
    {2, 2, 3, 1}, {2, 2, 4, 1}, {2, 2, 7, 1}, {2, 2, 8, 1},
    {3, 3, 5, 1}, {3, 3, 6, 1}, {3, 3, 7, 1}, {3, 3, 8, 1},
    {4, 1, 1, 1}, {4, 1, 2, 1}, {4, 1, 5, 1}, {4, 1, 6, 1},
    {5, 1, 1, 1}, {5, 1, 2, 1}, {5, 1, 3, 1}, {5, 1, 4, 1},
    {6, 2, 5, 1}, {6, 2, 6, 1}, {6, 2, 7, 1}, {6, 2, 8, 1},
    {7, 1, 8, 1}, {8, 2, 1, 1}, {9, 2, 2, 1}, {10, 3, 1, 1}, {11, 3, 2, 1}, {12, 3, 3, 1}, {13, 3, 4, 1},
    {14, 1, 4, 1}, {15, 1, 5, 1}, {16, 1, 6, 1}, {17, 1, 7, 1},
    {18, 1, 8, 1}, {18, 3, 6, -2},
    {19, 2, 1, 1}, {19, 3, 7, -3},
    {20, 2, 2, 1}, {20, 3, 8, -4},
    {21, 2, 3, 1}, {22, 2, 4, 1}, {23, 2, 5, 1},
    {24, 2, 6, 1}, {24, 2, 7, -1},
};

/* init_displ_hexe20, f-src/mod_write_deck.f90:118-279 (same entry layout; the
 * table is generated from the parsed reference literals, see the include) */
#include "uu20_table.inc"

int dtco_uu(int macro_order, double* uu) {
    const double eps = 1.e-6;
    if (macro_order == 1) {
        memset(uu, 0, sizeof(double) * 8 * 3 * 24);
        for (size_t i = 0; i < sizeof(UU8) / sizeof(UU8[0]); ++i)
            uu[(UU8[i].node - 1) + 8 * ((UU8[i].d - 1) + 3 * (UU8[i].lc - 1))] = UU8[i].m * eps;
        return 0;
    }
    if (macro_order == 2) {
        memset(uu, 0, sizeof(double) * 20 * 3 * 60);
        for (size_t i = 0; i < sizeof(UU20) / sizeof(UU20[0]); ++i)
            uu[(UU20[i].node - 1) + 20 * ((UU20[i].d - 1) + 3 * (UU20[i].lc - 1))] = UU20[i].m * eps;
        return 0;
    }
    return 5;
}

/* init_loadcase_el_order_lin, f-src/mod_struct_calcmat.f90:1684-1770: the
 * same 24 load cases as macro nodal displacements, row = (d-1)*8 + node. */
int dtco_vv_lin(double rve_strain, double* vv) {
    memset(vv, 0, sizeof(double) * 24 * 24);
    for (size_t i = 0; i < sizeof(UU8) / sizeof(UU8[0]); ++i) {
        int row = (UU8[i].d - 1) * 8 + (UU8[i].node - 1);
        vv[row + 24 * (UU8[i].lc - 1)] = UU8[i].m * rve_strain;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a5/a6  generate_boundaries + determine_prescribed_displ                   */
/* ------------------------------------------------------------------------ */
static int on_face(const double c[3], const double min_c[3], const double max_c[3], const double dim_c[3]) {
    const double delta_b = 1.e-9;      /* f-src/mod_write_deck.f90:9 */
    if ((c[2] - min_c[2]) <= (dim_c[2] * delta_b)) return 1;
    if ((max_c[2] - c[2]) <= (dim_c[2] * delta_b)) return 1;
    if ((c[1] - min_c[1]) <= (dim_c[1] * delta_b)) return 1;
    if ((max_c[1] - c[1]) <= (dim_c[1] * delta_b)) return 1;
    if ((c[0] - min_c[0]) <= (dim_c[0] * delta_b)) return 1;
    if ((max_c[0] - c[0]) <= (dim_c[0] * delta_b)) return 1;
    return 0;
}

int dtco_boundaries(const double* nodes, int64_t n_nodes, const double min_c[3],
                    const double max_c[3], const double dim_c[3], int macro_order,
                    int64_t* n_b_out, int64_t** ids_out, double** vals_out) {
    *n_b_out = 0; *ids_out = NULL; *vals_out = NULL;
    if (macro_order != 1 && macro_order != 2) return 5;
    const int nenM = macro_order == 1 ? 8 : 20, R = 3 * nenM;
    double uu[20 * 3 * 60];
    dtco_uu(macro_order, uu);
    int64_t nb = 0;
    for (int64_t i = 0; i < n_nodes; ++i) nb += on_face(nodes + 3 * i, min_c, max_c, dim_c);
    int64_t* ids = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb > 0 ? nb : 1));
    double* vals = (double*)malloc(sizeof(double) * (size_t)R * 3 * (size_t)(nb > 0 ? nb : 1));
    if (!ids || !vals) { free(ids); free(vals); return 3; }
    int64_t b = 0;
    for (int64_t i = 0; i < n_nodes; ++i) {
        const double* c = nodes + 3 * i;
        if (!on_face(c, min_c, max_c, dim_c)) continue;
        ids[b] = i + 1;
        double N[20];
        dtco_macro_shape(macro_order, c, min_c, max_c, N);
        for (int lc = 0; lc < R; ++lc)
            for (int d = 0; d < 3; ++d) {
                double s = 0.0;                   /* Fortran SUM, left to right */
                for (int l = 0; l < nenM; ++l) s += N[l] * uu[l + nenM * (d + 3 * lc)];
                vals[lc + (size_t)R * (3 * b + d)] = s;
            }
        ++b;
    }
    *n_b_out = nb; *ids_out = ids; *vals_out = vals;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a7  element stiffness by quadrature                                       */
/* ------------------------------------------------------------------------ */
static void dmat_iso(double E, double nu, double D[6][6]) {
    double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = E / (2.0 * (1.0 + nu));
    memset(D, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) D[i][j] = lam;
    for (int i = 0; i < 3; ++i) D[i][i] += 2.0 * mu;
    D[3][3] = D[4][4] = D[5][5] = mu;
}

/* derivative of the micro shape functions w.r.t. xi in [-1,1]^3 */
static void micro_dshape(int nen, const double q[3], double dN[20][3]) {
    if (nen == 8) {
        for (int l = 0; l < 8; ++l) {
            double sx = 2.0 * H08_OFF[l][0] - 1, sy = 2.0 * H08_OFF[l][1] - 1, sz = 2.0 * H08_OFF[l][2] - 1;
            dN[l][0] = sx * (1 + sy * q[1]) * (1 + sz * q[2]) / 8.0;
            dN[l][1] = (1 + sx * q[0]) * sy * (1 + sz * q[2]) / 8.0;
            dN[l][2] = (1 + sx * q[0]) * (1 + sy * q[1]) * sz / 8.0;
        }
        return;
    }
    for (int l = 0; l < 20; ++l) {
        double s[3] = {H20_OFF[l][0] - 1.0, H20_OFF[l][1] - 1.0, H20_OFF[l][2] - 1.0};
        int zero = -1;
        for (int d = 0; d < 3; ++d) if (s[d] == 0.0) zero = d;
        if (zero < 0) {   /* corner: (1+s.x)(1+s.y)(1+s.z)(s.x+s.y+s.z-2)/8 */
            double a = 1 + s[0] * q[0], b = 1 + s[1] * q[1], c = 1 + s[2] * q[2];
            double t = s[0] * q[0] + s[1] * q[1] + s[2] * q[2] - 2.0;
            dN[l][0] = (s[0] * b * c * t + a * b * c * s[0]) / 8.0;
            dN[l][1] = (a * s[1] * c * t + a * b * c * s[1]) / 8.0;
            dN[l][2] = (a * b * s[2] * t + a * b * c * s[2]) / 8.0;
        } else {          /* mid-edge along axis `zero`: (1-q_z^2) prod_{d!=z}(1+s_d q_d)/4 */
            double f[3], df[3];
            for (int d = 0; d < 3; ++d) {
                if (d == zero) { f[d] = 1 - q[d] * q[d]; df[d] = -2 * q[d]; }
                else { f[d] = 1 + s[d] * q[d]; df[d] = s[d]; }
            }
            dN[l][0] = df[0] * f[1] * f[2] / 4.0;
            dN[l][1] = f[0] * df[1] * f[2] / 4.0;
            dN[l][2] = f[0] * f[1] * df[2] / 4.0;
        }
    }
}

int dtco_ke(int nen, double E, double nu, double a, double* K) {
    if (nen != 8 && nen != 20) return 5;
    const int nd = 3 * nen;
    double D[6][6];
    dmat_iso(E, nu, D);
    memset(K, 0, sizeof(double) * (size_t)nd * nd);
    double gp[3], gw[3];
    int ng;
    if (nen == 8) { ng = 2; gp[0] = -1.0 / sqrt(3.0); gp[1] = -gp[0]; gw[0] = gw[1] = 1.0; }
    else { ng = 3; gp[0] = -sqrt(0.6); gp[1] = 0.0; gp[2] = sqrt(0.6); gw[0] = gw[2] = 5.0 / 9.0; gw[1] = 8.0 / 9.0; }
    const double jac = 2.0 / a;               /* d xi / dx for edge length a */
    const double detJ = (a / 2.0) * (a / 2.0) * (a / 2.0);
    double B[6][60];
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j)
            for (int k = 0; k < ng; ++k) {
                double q[3] = {gp[i], gp[j], gp[k]};
                double w = gw[i] * gw[j] * gw[k] * detJ;
                double dN[20][3];
                micro_dshape(nen, q, dN);
                memset(B, 0, sizeof(B));
                for (int l = 0; l < nen; ++l) {
                    double dx = dN[l][0] * jac, dy = dN[l][1] * jac, dz = dN[l][2] * jac;
                    B[0][3 * l] = dx; B[1][3 * l + 1] = dy; B[2][3 * l + 2] = dz;
                    B[3][3 * l] = dy; B[3][3 * l + 1] = dx;
                    B[4][3 * l] = dz; B[4][3 * l + 2] = dx;
                    B[5][3 * l + 1] = dz; B[5][3 * l + 2] = dy;
                }
                for (int r = 0; r < nd; ++r)
                    for (int c = 0; c < nd; ++c) {
                        double s = 0.0;
                        for (int m = 0; m < 6; ++m) {
                            if (B[m][r] == 0.0) continue;
                            double t = 0.0;
                            for (int p = 0; p < 6; ++p) t += D[m][p] * B[p][c];
                            s += B[m][r] * t;
                        }
                        K[r + nd * c] += w * s;
                    }
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a8-a10  assembled operator, Jacobi-PCG, reactions, force projection      */
/* ------------------------------------------------------------------------ */
typedef struct {
    int64_t nn;         /* nodes */
    int64_t* rowptr;    /* nn+1 */
    int64_t* col;       /* node columns */
    double* val;        /* 9 per block, row-major 3x3 */
} bsr_t;

static void bsr_free(bsr_t* A) { free(A->rowptr); free(A->col); free(A->val); }

/* MatSetValues(ADD_VALUES) of K_e per element, f-src/mod_dtc_main_subroutines.f90:518-577.
 * Node id 0 (a deleted node in the kept last element) maps to a negative
 * global index, which PETSc ignores; those entries are skipped here. */
static int bsr_assemble(bsr_t* A, int64_t nn, const int64_t* elems, int64_t ne, int nen,
                        const double* Ke, const double* emod, double E) {
    const int cap = 81;
    int64_t* nbr = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn * cap);
    int* cnt = (int*)calloc((size_t)nn, sizeof(int));
    if (!nbr || !cnt) { free(nbr); free(cnt); return 3; }
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t* en = elems + e * nen;
        for (int a = 0; a < nen; ++a) {
            int64_t i = en[a] - 1;
            if (i < 0) continue;
            for (int b = 0; b < nen; ++b) {
                int64_t j = en[b] - 1;
                if (j < 0) continue;
                int found = 0;
                for (int t = 0; t < cnt[i]; ++t) if (nbr[i * cap + t] == j) { found = 1; break; }
                if (!found) nbr[i * cap + cnt[i]++] = j;
            }
        }
    }
    A->nn = nn;
    A->rowptr = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nn + 1));
    A->rowptr[0] = 0;
    for (int64_t i = 0; i < nn; ++i) A->rowptr[i + 1] = A->rowptr[i] + cnt[i];
    int64_t nnzb = A->rowptr[nn];
    A->col = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnzb > 0 ? nnzb : 1));
    A->val = (double*)calloc((size_t)(nnzb > 0 ? nnzb : 1) * 9, sizeof(double));
    for (int64_t i = 0; i < nn; ++i) {
        /* sorted columns */
        int64_t* src = nbr + i * cap;
        for (int t = 1; t < cnt[i]; ++t) {
            int64_t v = src[t]; int s = t - 1;
            while (s >= 0 && src[s] > v) { src[s + 1] = src[s]; --s; }
            src[s + 1] = v;
        }
        memcpy(A->col + A->rowptr[i], src, sizeof(int64_t) * (size_t)cnt[i]);
    }
    free(nbr); free(cnt);
    const int nd = 3 * nen;
    for (int64_t e = 0; e < ne; ++e) {
        const int64_t* en = elems + e * nen;
        double s = emod ? emod[e] / E : 1.0;
        for (int a = 0; a < nen; ++a) {
            int64_t i = en[a] - 1;
            if (i < 0) continue;
            for (int b = 0; b < nen; ++b) {
                int64_t j = en[b] - 1;
                if (j < 0) continue;
                int64_t lo = A->rowptr[i], hi = A->rowptr[i + 1] - 1;
                while (lo < hi) { int64_t m = (lo + hi) / 2; if (A->col[m] < j) lo = m + 1; else hi = m; }
                double* blk = A->val + 9 * lo;
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 3; ++c)
                        blk[3 * r + c] += s * Ke[(3 * a + r) + nd * (3 * b + c)];
            }
        }
    }
    return 0;
}

/* y = A x; with mask (1 = constrained DOF) the constrained rows/columns are
 * zero (MatZeroRowsColumns with diag 0.0). */
static void bsr_mv(const bsr_t* A, const double* x, double* y, const unsigned char* mask) {
    for (int64_t i = 0; i < A->nn; ++i) {
        double s0 = 0, s1 = 0, s2 = 0;
        for (int64_t t = A->rowptr[i]; t < A->rowptr[i + 1]; ++t) {
            const double* b = A->val + 9 * t;
            const double* xv = x + 3 * A->col[t];
            double x0 = xv[0], x1 = xv[1], x2 = xv[2];
            if (mask) {
                const unsigned char* m = mask + 3 * A->col[t];
                if (m[0]) x0 = 0;
                if (m[1]) x1 = 0;
                if (m[2]) x2 = 0;
            }
            s0 += b[0] * x0 + b[1] * x1 + b[2] * x2;
            s1 += b[3] * x0 + b[4] * x1 + b[5] * x2;
            s2 += b[6] * x0 + b[7] * x1 + b[8] * x2;
        }
        y[3 * i] = s0; y[3 * i + 1] = s1; y[3 * i + 2] = s2;
        if (mask) for (int d = 0; d < 3; ++d) if (mask[3 * i + d]) y[3 * i + d] = 0.0;
    }
}

static double dot(const double* a, const double* b, int64_t n) {
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

int dtco_solve_domain(const double* nodes, int64_t n_nodes, const int64_t* elems,
                      int64_t n_elems, int nen, const double* emod, double E, double nu,
                      double a, const double min_c[3], const double max_c[3],
                      const double dim_c[3], int macro_order, double tol, int64_t max_it,
                      int norm_type, int lc_first, int lc_count,
                      double* u_out, double* r_out, double* ff, int64_t* iters, double* relres) {
    if (macro_order != 1 && macro_order != 2) return 5;
    if (n_nodes <= 0 || n_elems <= 0) return 1;
    const int nenM = macro_order == 1 ? 8 : 20, R = 3 * nenM;
    if (lc_first < 0 || lc_count <= 0 || lc_first + lc_count > R) return 5;
    const int64_t m = 3 * n_nodes;
    double* Ke = (double*)malloc(sizeof(double) * 3600);
    dtco_ke(nen, E, nu, a, Ke);
    bsr_t A;
    int rc = bsr_assemble(&A, n_nodes, elems, n_elems, nen, Ke, emod, E);
    free(Ke);
    if (rc) return rc;

    int64_t nb; int64_t* bid; double* bval;
    rc = dtco_boundaries(nodes, n_nodes, min_c, max_c, dim_c, macro_order, &nb, &bid, &bval);
    if (rc) { bsr_free(&A); return rc; }
    if (3 * nb < 6) { bsr_free(&A); free(bid); free(bval); return 2; }   /* :471-473 */

    unsigned char* mask = (unsigned char*)calloc((size_t)m, 1);
    for (int64_t b = 0; b < nb; ++b) for (int d = 0; d < 3; ++d) mask[3 * (bid[b] - 1) + d] = 1;
    /* Jacobi diagonal of the constrained operator; PCJACOBI uses 1 for a zero diagonal */
    double* dinv = (double*)malloc(sizeof(double) * (size_t)m);
    for (int64_t i = 0; i < n_nodes; ++i)
        for (int64_t t = A.rowptr[i]; t < A.rowptr[i + 1]; ++t)
            if (A.col[t] == i)
                for (int d = 0; d < 3; ++d) {
                    double dg = mask[3 * i + d] ? 0.0 : A.val[9 * t + 4 * d];
                    dinv[3 * i + d] = dg != 0.0 ? 1.0 / dg : 1.0;
                }
    double *xbc = (double*)calloc((size_t)m, sizeof(double)), *f = malloc(sizeof(double) * m),
           *x = malloc(sizeof(double) * m), *r = malloc(sizeof(double) * m),
           *z = malloc(sizeof(double) * m), *p = malloc(sizeof(double) * m),
           *q = malloc(sizeof(double) * m), *uu = malloc(sizeof(double) * m),
           *rf = malloc(sizeof(double) * m);
    if (ff) memset(ff, 0, sizeof(double) * R * R);
    for (int lc = lc_first; lc < lc_first + lc_count; ++lc) {
        /* x_bc and the lifted RHS f = -A x_bc with f[bc] = 0 (:735-951) */
        memset(xbc, 0, sizeof(double) * (size_t)m);
        for (int64_t b = 0; b < nb; ++b)
            for (int d = 0; d < 3; ++d) xbc[3 * (bid[b] - 1) + d] = bval[lc + (size_t)R * (3 * b + d)];
        bsr_mv(&A, xbc, f, NULL);
        for (int64_t i = 0; i < m; ++i) f[i] = mask[i] ? 0.0 : -f[i];
        /* CG + Jacobi, x0 = 0 (:1072-1075) */
        memset(x, 0, sizeof(double) * (size_t)m);
        memcpy(r, f, sizeof(double) * (size_t)m);
        for (int64_t i = 0; i < m; ++i) z[i] = dinv[i] * r[i];
        memcpy(p, z, sizeof(double) * (size_t)m);
        double rz = dot(r, z, m);
        double fnorm = sqrt(dot(f, f, m));
        double bnorm = norm_type == 1 ? sqrt(dot(z, z, m)) : fnorm;
        int64_t it = 0;
        if (bnorm > 0.0) {
            for (it = 1; it <= max_it; ++it) {
                bsr_mv(&A, p, q, mask);
                double pq = dot(p, q, m);
                if (pq == 0.0) break;
                double alpha = rz / pq;
                for (int64_t i = 0; i < m; ++i) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
                for (int64_t i = 0; i < m; ++i) z[i] = dinv[i] * r[i];
                double rz1 = dot(r, z, m);
                double nrm = norm_type == 1 ? sqrt(dot(z, z, m)) : sqrt(dot(r, r, m));
                if (nrm <= tol * bnorm) break;
                double beta = rz1 / rz;
                rz = rz1;
                for (int64_t i = 0; i < m; ++i) p[i] = z[i] + beta * p[i];
            }
            if (it > max_it) it = max_it;
        }
        /* u = x with prescribed values re-inserted (:1092-1100) */
        for (int64_t i = 0; i < m; ++i) uu[i] = mask[i] ? xbc[i] : x[i];
        /* true residual of the constrained system */
        bsr_mv(&A, uu, rf, NULL);
        double rr = 0.0;
        for (int64_t i = 0; i < m; ++i) if (!mask[i]) rr += rf[i] * rf[i];
        if (iters) iters[lc] = it;
        if (relres) relres[lc] = fnorm > 0.0 ? sqrt(rr) / fnorm : 0.0;
        /* reactions on the unmodified matrix (:1105) */
        if (u_out) memcpy(u_out + (size_t)m * lc, uu, sizeof(double) * (size_t)m);
        if (r_out) memcpy(r_out + (size_t)m * lc, rf, sizeof(double) * (size_t)m);
        /* ff((k-1)*nenM + l, lc) = sum over all nodes r_k N_l (f-src/mod_struct_calcmat.f90:254-280) */
        if (ff) {
            for (int64_t nd = 0; nd < n_nodes; ++nd) {
                double N[20];
                dtco_macro_shape(macro_order, nodes + 3 * nd, min_c, max_c, N);
                for (int k = 0; k < 3; ++k)
                    for (int l = 0; l < nenM; ++l) ff[(k * nenM + l) + R * lc] += rf[3 * nd + k] * N[l];
            }
        }
    }
    free(xbc); free(f); free(x); free(r); free(z); free(p); free(q); free(uu); free(rf);
    free(mask); free(dinv); free(bid); free(bval); bsr_free(&A);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a10/a11  effective stiffness                                              */
/* ------------------------------------------------------------------------ */
/* LU with partial pivoting inverse (the reference uses LINPACK dgefa/dgedi,
 * linpack/mod_linpack.f90:10-61, also partial-pivoting LU). */
static int invert(double* A, int n) {
    int* piv = (int*)malloc(sizeof(int) * n);
    double* inv = (double*)calloc((size_t)n * n, sizeof(double));
    double* M = (double*)malloc(sizeof(double) * n * n);
    memcpy(M, A, sizeof(double) * n * n);
    for (int i = 0; i < n; ++i) inv[i + n * i] = 1.0;
    int rc = 0;
    for (int k = 0; k < n; ++k) {
        int pr = k; double pv = fabs(M[k + n * k]);
        for (int i = k + 1; i < n; ++i) if (fabs(M[i + n * k]) > pv) { pv = fabs(M[i + n * k]); pr = i; }
        if (pv == 0.0) { rc = 4; break; }
        if (pr != k)
            for (int j = 0; j < n; ++j) {
                double t = M[k + n * j]; M[k + n * j] = M[pr + n * j]; M[pr + n * j] = t;
                t = inv[k + n * j]; inv[k + n * j] = inv[pr + n * j]; inv[pr + n * j] = t;
            }
        for (int i = 0; i < n; ++i) {
            if (i == k) continue;
            double f = M[i + n * k] / M[k + n * k];
            if (f == 0.0) continue;
            for (int j = 0; j < n; ++j) { M[i + n * j] -= f * M[k + n * j]; inv[i + n * j] -= f * inv[k + n * j]; }
        }
    }
    if (!rc) {
        for (int k = 0; k < n; ++k) { double d = M[k + n * k]; for (int j = 0; j < n; ++j) inv[k + n * j] /= d; }
        memcpy(A, inv, sizeof(double) * n * n);
    }
    free(piv); free(inv); free(M);
    return rc;
}

/* fv = W ff(:,1:6): face-averaged force weights of
 * f-src/mod_struct_calcmat.f90:345-359 (HEX08 macro). W is 6 x 24, column-major. */
static void fv_weights_lin(const double xdp[3], double W[6 * 24]) {
    memset(W, 0, sizeof(double) * 6 * 24);
    const double A23 = xdp[1] * xdp[2], A13 = xdp[0] * xdp[2], A12 = xdp[0] * xdp[1];
#define SETW(i, list, val) do { const int L_[] = list; for (size_t t_ = 0; t_ < sizeof(L_)/sizeof(int); ++t_) W[(i) + 6 * (L_[t_] - 1)] += (val); } while (0)
#define LIST(...) {__VA_ARGS__}
    SETW(0, LIST(2, 3, 6, 7), 1.0 / A23);
    SETW(1, LIST(11, 12, 15, 16), 1.0 / A13);
    SETW(2, LIST(21, 22, 23, 24), 1.0 / A12);
    SETW(3, LIST(3, 4, 7, 8), 0.5 / A13);
    SETW(3, LIST(10, 11, 14, 15), 0.5 / A23);
    SETW(4, LIST(5, 6, 7, 8), 0.5 / A12);
    SETW(4, LIST(18, 19, 22, 23), 0.5 / A23);
    SETW(5, LIST(13, 14, 15, 16), 0.5 / A12);
    SETW(5, LIST(19, 20, 23, 24), 0.5 / A13);
#undef SETW
#undef LIST
}

/* the same for the HEX20 macro element, f-src/mod_struct_calcmat.f90:365-401
 * (20 nodes per direction block; W is 6 x 60) */
static void fv_weights_quad(const double xdp[3], double W[6 * 60]) {
    memset(W, 0, sizeof(double) * 6 * 60);
    const double A23 = xdp[1] * xdp[2], A13 = xdp[0] * xdp[2], A12 = xdp[0] * xdp[1];
#define SETW(i, list, val) do { const int L_[] = list; for (size_t t_ = 0; t_ < sizeof(L_)/sizeof(int); ++t_) W[(i) + 6 * (L_[t_] - 1)] += (val); } while (0)
#define LIST(...) {__VA_ARGS__}
    SETW(0, LIST(2, 3, 6, 7, 10, 14, 18, 19), 1.0 / A23);
    SETW(1, LIST(23, 24, 27, 28, 31, 35, 39, 40), 1.0 / A13);
    SETW(2, LIST(45, 46, 47, 48, 53, 54, 55, 56), 1.0 / A12);
    SETW(3, LIST(3, 4, 7, 8, 11, 15, 19, 20), 0.5 / A13);
    SETW(3, LIST(22, 23, 26, 27, 30, 34, 38, 39), 0.5 / A23);
    SETW(4, LIST(5, 6, 7, 8, 13, 14, 15, 16), 0.5 / A12);
    SETW(4, LIST(42, 43, 46, 47, 50, 54, 58, 59), 0.5 / A23);
    SETW(5, LIST(25, 26, 27, 28, 33, 34, 35, 36), 0.5 / A12);
    SETW(5, LIST(43, 44, 47, 48, 51, 55, 59, 60), 0.5 / A13);
#undef SETW
#undef LIST
}

static const double MEPS[6] = {1.e6, 1.e6, 1.e6, -1.e6, -1.e6, 1.e6};   /* :434-441 */

/* HEX20 load-case matrix vv(dof, lc) = uu(l, d, lc) * RVE_STRAIN / 1e-6 with dof =
 * (d-1)*20 + l.  The reference's own table (init_loadcase_el_order_quad,
 * f-src/mod_struct_calcmat.f90:1780-2028) leaves row 49 unassigned and holds
 * unscaled literals, so it is singular; K_M of the HEX20 macro element is formed
 * with this matrix instead (documented deviation). */
int dtco_vv_quad_consistent(double rve_strain, double* vv) {
    double uu[20 * 3 * 60];
    dtco_uu(2, uu);
    for (int lc = 0; lc < 60; ++lc)
        for (int d = 0; d < 3; ++d)
            for (int l = 0; l < 20; ++l) vv[(d * 20 + l) + 60 * lc] = uu[l + 20 * (d + 3 * lc)] * rve_strain / 1.e-6;
    return 0;
}

int dtco_effective(const double* ff, int macro_order, double rve_strain,
                   const double x_D_phy[3], double* K_M, double* C, double* EE) {
    if (macro_order != 1 && macro_order != 2) return 5;
    const int R = macro_order == 1 ? 24 : 60;
    int rc = 0;
    if (K_M) {
        double vv[60 * 60];
        if (macro_order == 1) dtco_vv_lin(rve_strain, vv);
        else dtco_vv_quad_consistent(rve_strain, vv);
        rc = invert(vv, R);
        if (!rc)
            for (int i = 0; i < R; ++i)
                for (int j = 0; j < R; ++j) {
                    double s = 0.0;
                    for (int k = 0; k < R; ++k) s += ff[i + R * k] * vv[k + R * j];
                    K_M[i + R * j] = s;
                }
    }
    /* fv(i, lc) for lc = 1..6 */
    double W[6 * 60], fv[36];
    if (macro_order == 1) fv_weights_lin(x_D_phy, W);
    else fv_weights_quad(x_D_phy, W);
    /* the reference sums each group left to right before dividing; this is a
     * different FP order of the same expression (tolerance parity) */
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int a = 0; a < R; ++a) s += W[i + 6 * a] * ff[a + R * j];
            fv[i + 6 * j] = s;
        }
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) C[i + 6 * j] = fv[i + 6 * j] * MEPS[j];
    for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) EE[i + 6 * j] = 0.5 * (C[i + 6 * j] + C[j + 6 * i]);
    return rc;
}

double dtco_density(int64_t n_elems, const double x_D_phy[3], const double delta[3]) {
    double v = 1.0;
    for (int d = 0; d < 3; ++d) v *= round(x_D_phy[d] / delta[d]);
    return (double)n_elems / v;
}

/* ------------------------------------------------------------------------ */
/* a12  rotation search                                                      */
/* ------------------------------------------------------------------------ */
/* rot_alg / tra_R6 as inlined at f-src/mod_struct_calcmat.f90:556-589; aa and
 * BB are column-major. */
static void rot_alg(const double n[3], double alpha, double aa[9]) {
    double c = cos(alpha), s = sin(alpha), omc = 1.0 - c;
    double n12 = n[0] * n[1], n13 = n[0] * n[2], n23 = n[1] * n[2];
#define A(i, j) aa[((i) - 1) + 3 * ((j) - 1)]
    A(1, 1) = c + n[0] * n[0] * omc;
    A(2, 2) = c + n[1] * n[1] * omc;
    A(3, 3) = c + n[2] * n[2] * omc;
    A(1, 2) = n12 * omc - n[2] * s;
    A(2, 1) = n12 * omc + n[2] * s;
    A(1, 3) = n13 * omc + n[1] * s;
    A(3, 1) = n13 * omc - n[1] * s;
    A(2, 3) = n23 * omc - n[0] * s;
    A(3, 2) = n23 * omc + n[0] * s;
#undef A
}

static void tra_R6(const double aa[9], double BB[36]) {
    const double sq2 = sqrt(2.0);
#define A(i, j) aa[((i) - 1) + 3 * ((j) - 1)]
#define B(i, j) BB[((i) - 1) + 6 * ((j) - 1)]
    for (int c = 1; c <= 3; ++c) {
        B(1, c) = A(1, c) * A(1, c); B(2, c) = A(2, c) * A(2, c); B(3, c) = A(3, c) * A(3, c);
        B(4, c) = sq2 * A(2, c) * A(1, c); B(5, c) = sq2 * A(1, c) * A(3, c); B(6, c) = sq2 * A(2, c) * A(3, c);
    }
    const int P[3][2] = {{1, 2}, {1, 3}, {2, 3}};
    for (int t = 0; t < 3; ++t) {
        int a = P[t][0], b = P[t][1], c = 4 + t;
        B(1, c) = sq2 * A(1, a) * A(1, b);
        B(2, c) = sq2 * A(2, a) * A(2, b);
        B(3, c) = sq2 * A(3, a) * A(3, b);
        B(4, c) = A(2, a) * A(1, b) + A(2, b) * A(1, a);
        B(5, c) = A(1, a) * A(3, b) + A(1, b) * A(3, a);
        B(6, c) = A(2, a) * A(3, b) + A(2, b) * A(3, a);
    }
#undef A
#undef B
}

static void congruence(const double BB[36], const double EE[36], double T[36]) {
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

int dtco_rotate66(const double EE[36], const double n[3], double alpha, double T[36]) {
    double aa[9], BB[36];
    rot_alg(n, alpha, aa);
    tra_R6(aa, BB);
    congruence(BB, EE, T);
    return 0;
}

/* CR_1: sum T(1:4,5:6)^2; CR_2: sum T(1:3,4:6)^2 + T45^2 + T46^2 + T56^2 (:762-779) */
static double crit_of(const double T[36], int criterion) {
#define t(i, j) T[((i) - 1) + 6 * ((j) - 1)]
    double s = 0.0;
    if (criterion == 1) {
        for (int i = 1; i <= 4; ++i) for (int j = 5; j <= 6; ++j) s += t(i, j) * t(i, j);
    } else {
        for (int i = 1; i <= 3; ++i) for (int j = 4; j <= 6; ++j) s += t(i, j) * t(i, j);
        s += t(4, 5) * t(4, 5) + t(4, 6) * t(4, 6) + t(5, 6) * t(5, 6);
    }
#undef t
    return s;
}

static double eval_point(const double EE[36], int criterion, double alpha, double phi, double eta, int normalise) {
    double n[3] = {cos(phi) * sin(eta), sin(phi) * sin(eta), cos(eta)};
    if (normalise) {
        double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        n[0] /= nn; n[1] /= nn; n[2] /= nn;
    }
    double T[36];
    dtco_rotate66(EE, n, alpha, T);
    return crit_of(T, criterion);
}

int dtco_rotation_search(const double EE[36], int criterion, double step_deg, double num_zero,
                         double* alpha_out, double n_out[3], double aa_out[9], double EE_opt[36],
                         int64_t* jj_out, double* crit_out) {
    if (criterion != 1 && criterion != 2) return 5;
    if (!(step_deg > 0.0)) return 5;
    const double pi = acos(-1.0);
    const double pi_div_180 = pi / 180.0;
    const int64_t na = (int64_t)floor(180.0 / step_deg + 1e-9) + 1;
    const int64_t ne = (int64_t)floor(90.0 / step_deg + 1e-9) + 1;
    /* initial sweep :533-785, alpha fastest, minloc = first minimum */
    double best = INFINITY;
    int64_t ang[3] = {0, 0, 0};
    for (int64_t ie = 0; ie < ne; ++ie)
        for (int64_t ip = 0; ip < na; ++ip)
            for (int64_t ia = 0; ia < na; ++ia) {
                double c = eval_point(EE, criterion, (double)ia * step_deg * pi_div_180,
                                      (double)ip * step_deg * pi_div_180,
                                      (double)ie * step_deg * pi_div_180, 1);
                if (c < best) { best = c; ang[0] = ia; ang[1] = ip; ang[2] = ie; }
            }
    double crit_min_prev = best;
    /* refinement :790-1012 (CR_2: :1169-1300, n not renormalised) */
    int64_t jj = 1;
    double div = pi_div_180 * step_deg;
    double cur = best;
    for (;;) {
        int64_t s0[3], e0[3];
        for (int d = 0; d < 3; ++d) { s0[d] = (ang[d] - 1) * 10; e0[d] = (ang[d] + 1) * 10; }
        div = div * 0.1;
        double b = INFINITY;
        int64_t na2[3] = {ang[0], ang[1], ang[2]};
        for (int64_t ie = s0[2]; ie <= e0[2]; ++ie)
            for (int64_t ip = s0[1]; ip <= e0[1]; ++ip)
                for (int64_t ia = s0[0]; ia <= e0[0]; ++ia) {
                    double c = eval_point(EE, criterion, (double)ia * div, (double)ip * div,
                                          (double)ie * div, criterion == 1);
                    if (c < b) { b = c; na2[0] = ia; na2[1] = ip; na2[2] = ie; }
                }
        ang[0] = na2[0]; ang[1] = na2[1]; ang[2] = na2[2];
        cur = b;
        if (fabs(crit_min_prev - cur) < num_zero || jj >= 16) break;
        crit_min_prev = cur;
        jj = jj + 1;
    }
    /* final angles with the reference's /(10**jj-1) scale (:1017-1019) */
    const double den = 180.0 * (pow(10.0, (double)jj) - 1.0);
    double alpha = (double)ang[0] * step_deg * pi / den;
    double phi = (double)ang[1] * step_deg * pi / den;
    double eta = (double)ang[2] * step_deg * pi / den;
    double n[3] = {cos(phi) * sin(eta), sin(phi) * sin(eta), cos(eta)};
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    n[0] /= nn; n[1] /= nn; n[2] /= nn;
    *alpha_out = alpha;
    n_out[0] = n[0]; n_out[1] = n[1]; n_out[2] = n[2];
    /* inlining + axis permutation (:1054-1137) */
    double aa[9], BB[36], T[36];
    rot_alg(n, alpha, aa);
    tra_R6(aa, BB);
    congruence(BB, EE, T);
    double e11 = T[0], e22 = T[7], e33 = T[14];
    int seq[2] = {-1, -1};   /* columns of aa used as rotation axes, in order */
    if (e11 < e22 && e11 < e33 && e22 < e33) { }
    else if (e11 < e22 && e11 < e33 && e22 > e33) { seq[0] = 0; }
    else if (e11 < e22 && e11 > e33 && e22 > e33) { seq[0] = 1; seq[1] = 0; }
    else if (e11 > e22 && e11 < e33 && e22 < e33) { seq[0] = 2; }
    else if (e11 > e22 && e11 > e33 && e22 < e33) { seq[0] = 2; seq[1] = 0; }
    else if (e11 > e22 && e11 > e33 && e22 > e33) { seq[0] = 1; }
    for (int s = 0; s < 2; ++s) {
        if (seq[s] < 0) break;
        double ax[3] = {aa[3 * seq[s]], aa[3 * seq[s] + 1], aa[3 * seq[s] + 2]};
        double rr[9], prod[9];
        rot_alg(ax, pi / 2.0, rr);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double t = 0.0;
                for (int k = 0; k < 3; ++k) t += rr[i + 3 * k] * aa[k + 3 * j];
                prod[i + 3 * j] = t;
            }
        memcpy(aa, prod, sizeof(aa));
    }
    tra_R6(aa, BB);
    congruence(BB, EE, EE_opt);
    memcpy(aa_out, aa, sizeof(aa));
    if (jj_out) *jj_out = jj;
    if (crit_out) *crit_out = cur;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* sensitivity dL/dE_e (self-adjoint: no extra solve)                         */
/* ------------------------------------------------------------------------ */
/* ff(a,b) = g_a^T K u_b where g_a is the macro unit field of DOF a; at
 * convergence this equals w_a^T K u_b with w_a = sum_b u_b vv_eps^-1(b,a) the
 * harmonic extension of g_a (vv_eps: load-case matrix at eps = 1e-6, the
 * value init_displ_hexe8 uses).  Hence dC_ij/dE_e = meps_j sum_a W_ia
 * w_a,e^T Khat u_j,e and dL/dE_e = sum_j h_j,e^T Khat u_j,e, for HEX08 and HEX20
 * micro elements (nen nodes, Khat of the element type) under the HEX08 macro
 * element's 24 load cases or the HEX20 macro element's 60 (its consistent vv,
 * as in dtco_effective). */
int dtco_grad_modulus(const double* nodes, int64_t n_nodes, const int64_t* elems,
                      int64_t n_elems, int nen, int macro_order, double nu, double a,
                      const double min_c[3], const double max_c[3], const double x_D_phy[3],
                      double rve_strain, const double* u, const double G[36], double* dL_dE) {
    (void)nodes; (void)min_c; (void)max_c; (void)rve_strain;
    if (nen != 8 && nen != 20) return 5;
    if (macro_order != 1 && macro_order != 2) return 5;
    const int R = macro_order == 1 ? 24 : 60, nd = 3 * nen;
    const int64_t m = 3 * n_nodes;
    double vv[3600];
    double W[6 * 60], M[60 * 6], Khat[3600];
    if (macro_order == 1) dtco_vv_lin(1.e-6, vv);
    else dtco_vv_quad_consistent(1.e-6, vv);
    if (invert(vv, R)) return 4;
    if (macro_order == 1) fv_weights_lin(x_D_phy, W);
    else fv_weights_quad(x_D_phy, W);
    /* M(b,j) = meps_j sum_i G_ij sum_a W_ia vvinv(b,a) */
    for (int b = 0; b < R; ++b)
        for (int j = 0; j < 6; ++j) {
            double s = 0.0;
            for (int i = 0; i < 6; ++i) {
                double t = 0.0;
                for (int aidx = 0; aidx < R; ++aidx) t += W[i + 6 * aidx] * vv[b + R * aidx];
                s += G[i + 6 * j] * t;
            }
            M[b + R * j] = s * MEPS[j];
        }
    dtco_ke(nen, 1.0, nu, a, Khat);
    for (int64_t e = 0; e < n_elems; ++e) {
        const int64_t* en = elems + (int64_t)nen * e;
        double ue[60][60];   /* [lc][local dof] */
        for (int lc = 0; lc < R; ++lc)
            for (int l = 0; l < nen; ++l)
                for (int d = 0; d < 3; ++d)
                    ue[lc][3 * l + d] = en[l] > 0 ? u[(size_t)m * lc + 3 * (en[l] - 1) + d] : 0.0;
        double g = 0.0;
        for (int j = 0; j < 6; ++j) {
            double h[60], ku[60];
            for (int t = 0; t < nd; ++t) {
                double s = 0.0;
                for (int b = 0; b < R; ++b) s += M[b + R * j] * ue[b][t];
                h[t] = s;
            }
            for (int r = 0; r < nd; ++r) {
                double s = 0.0;
                for (int c = 0; c < nd; ++c) s += Khat[r + nd * c] * ue[j][c];
                ku[r] = s;
            }
            for (int t = 0; t < nd; ++t) g += h[t] * ku[t];
        }
        dL_dE[e] = g;
    }
    return 0;
}
