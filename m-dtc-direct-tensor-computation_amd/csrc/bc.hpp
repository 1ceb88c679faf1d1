// bc.hpp -- macro-element boundary values (device + host), internal.
//
// Boundary displacement of load case lc at a node with coordinate x:
//   u_d = sum_l N_l(t_geom_xi(x, min_c, max_c)) * uu(l, d, lc)
// (f-src/mod_write_deck.f90:22-53) with the HEX08 macro shape functions
// (f-src/mod_linfe.f90:51-84) and the 24 unit load cases of init_displ_hexe8
// (f-src/mod_write_deck.f90:63-107).  The table below lists the nonzero
// entries of uu in the reference's order; the sum over l runs in increasing
// l, so skipping the zero terms changes no bit of the result.
//
// The HEX20 macro element (60 load cases) uses the ANSYS 20-node shape
// functions (f-src/mod_linfe.f90:827-888) and init_displ_hexe20
// (f-src/mod_write_deck.f90:118-279, table in uu20_table.hpp).
#pragma once
#include <hip/hip_runtime.h>

#include "uu20_table.hpp"

namespace dtc {

struct UUEntry {
    signed char d, l, m;   // direction, macro node (0-based), multiple of eps
};

// [lc][k] nonzero entries (d, l, m), m * 1e-6 = value; count in UU_N
__host__ __device__ inline const UUEntry* uu_entries(int lc, int* count) {
    static constexpr UUEntry T[24][4] = {
        {{0, 1, 1}, {0, 2, 1}, {0, 5, 1}, {0, 6, 1}},   // LC 1: x of nodes 2,3,6,7
        {{1, 2, 1}, {1, 3, 1}, {1, 6, 1}, {1, 7, 1}},   // LC 2: y of nodes 3,4,7,8
        {{2, 4, 1}, {2, 5, 1}, {2, 6, 1}, {2, 7, 1}},   // LC 3: z of nodes 5..8
        {{0, 0, 1}, {0, 1, 1}, {0, 4, 1}, {0, 5, 1}},   // LC 4: x of nodes 1,2,5,6
        {{0, 0, 1}, {0, 1, 1}, {0, 2, 1}, {0, 3, 1}},   // LC 5: x of nodes 1..4
        {{1, 4, 1}, {1, 5, 1}, {1, 6, 1}, {1, 7, 1}},   // LC 6: y of nodes 5..8
        {{0, 7, 1}}, {{1, 0, 1}}, {{1, 1, 1}}, {{2, 0, 1}}, {{2, 1, 1}}, {{2, 2, 1}}, {{2, 3, 1}},
        {{0, 3, 1}}, {{0, 4, 1}}, {{0, 5, 1}}, {{0, 6, 1}},
        {{0, 7, 1}, {2, 5, -2}},                         // LC 18
        {{1, 0, 1}, {2, 6, -3}},                         // LC 19
        {{1, 1, 1}, {2, 7, -4}},                         // LC 20
        {{1, 2, 1}}, {{1, 3, 1}}, {{1, 4, 1}},
        {{1, 5, 1}, {1, 6, -1}},                         // LC 24
    };
    static constexpr int N[24] = {4, 4, 4, 4, 4, 4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1, 1, 2};
    *count = N[lc];
    return T[lc];
}

// N(xi(x)) for the HEX08 macro element, same IEEE operations as the
// reference (no contraction into FMA).
__host__ __device__ inline void macro_shape8(const double x[3], const double mn[3], const double mx[3],
                                             double N[8]) {
#pragma clang fp contract(off)
    double xi[3];
    for (int d = 0; d < 3; ++d) xi[d] = 2.0 / (mx[d] - mn[d]) * x[d] - (mx[d] + mn[d]) / (mx[d] - mn[d]);
    const double a0 = 0.5 * (1.0 - xi[0]), b0 = 0.5 * (1.0 + xi[0]);
    const double a1 = 0.5 * (1.0 - xi[1]), b1 = 0.5 * (1.0 + xi[1]);
    const double a2 = 0.5 * (1.0 - xi[2]), b2 = 0.5 * (1.0 + xi[2]);
    N[0] = a0 * a1 * a2;
    N[1] = b0 * a1 * a2;
    N[2] = b0 * b1 * a2;
    N[3] = a0 * b1 * a2;
    N[4] = a0 * a1 * b2;
    N[5] = b0 * a1 * b2;
    N[6] = b0 * b1 * b2;
    N[7] = a0 * b1 * b2;
}

// boundary value of direction d, load case lc from precomputed N
__host__ __device__ inline double bc_value(const double N[8], int d, int lc) {
#pragma clang fp contract(off)
    int cnt;
    const UUEntry* e = uu_entries(lc, &cnt);
    double s = 0.0;
    for (int k = 0; k < cnt; ++k)
        if (e[k].d == d) s += N[e[k].l] * ((double)e[k].m * 1.e-6);
    return s;
}

// N(xi(x)) for the HEX20 macro element in the reference's operation order
// (products left to right; f = 1/8 and ff = 1/4 are exact in single precision)
__host__ __device__ inline void macro_shape20(const double x[3], const double mn[3], const double mx[3],
                                              double N[20]) {
#pragma clang fp contract(off)
    double xi[3];
    for (int d = 0; d < 3; ++d) xi[d] = 2.0 / (mx[d] - mn[d]) * x[d] - (mx[d] + mn[d]) / (mx[d] - mn[d]);
    const double f = 0.125, ff = 0.25, e = 1.0, z = 2.0;
    const double a = xi[0], b = xi[1], c = xi[2];
    N[0] = f * (e - a) * (e - b) * (e - c) * (-a - b - c - z);
    N[1] = f * (e + a) * (e - b) * (e - c) * (a - b - c - z);
    N[2] = f * (e + a) * (e + b) * (e - c) * (a + b - c - z);
    N[3] = f * (e - a) * (e + b) * (e - c) * (-a + b - c - z);
    N[4] = f * (e - a) * (e - b) * (e + c) * (-a - b + c - z);
    N[5] = f * (e + a) * (e - b) * (e + c) * (a - b + c - z);
    N[6] = f * (e + a) * (e + b) * (e + c) * (a + b + c - z);
    N[7] = f * (e - a) * (e + b) * (e + c) * (-a + b + c - z);
    N[8] = ff * (e - a * a) * (e - b) * (e - c);
    N[9] = ff * (e + a) * (e - b * b) * (e - c);
    N[10] = ff * (e - a * a) * (e + b) * (e - c);
    N[11] = ff * (e - a) * (e - b * b) * (e - c);
    N[12] = ff * (e - a * a) * (e - b) * (e + c);
    N[13] = ff * (e + a) * (e - b * b) * (e + c);
    N[14] = ff * (e - a * a) * (e + b) * (e + c);
    N[15] = ff * (e - a) * (e - b * b) * (e + c);
    N[16] = ff * (e - a) * (e - b) * (e - c * c);
    N[17] = ff * (e + a) * (e - b) * (e - c * c);
    N[18] = ff * (e + a) * (e + b) * (e - c * c);
    N[19] = ff * (e - a) * (e + b) * (e - c * c);
}

// HEX20 macro boundary value of direction d, load case lc (0..59)
__host__ __device__ inline double bc_value20(const double N[20], int d, int lc) {
#pragma clang fp contract(off)
    int cnt;
    const UU20Entry* e = uu20_entries(lc, &cnt);
    double s = 0.0;
    for (int k = 0; k < cnt; ++k)
        if (e[k].d == d) s += N[e[k].l] * (e[k].m * 1.e-6);
    return s;
}

// load cases and macro nodes of a macro element order (1: HEX08, 2: HEX20)
__host__ __device__ constexpr int macro_lcs(int order) { return order == 2 ? 60 : 24; }
__host__ __device__ constexpr int macro_nodes(int order) { return order == 2 ? 20 : 8; }

__host__ __device__ inline void macro_shape(int order, const double x[3], const double mn[3], const double mx[3],
                                            double N[20]) {
    if (order == 2) macro_shape20(x, mn, mx, N);
    else macro_shape8(x, mn, mx, N);
}

// boundary value of global load case lc; 0 beyond the element's load cases
__host__ __device__ inline double bc_value_m(int order, const double N[20], int d, int lc) {
    if (lc >= macro_lcs(order)) return 0.0;
    return order == 2 ? bc_value20(N, d, lc) : bc_value(N, d, lc);
}

// face test of generate_boundaries (f-src/mod_struct_preprocess.f90:564-592)
__host__ __device__ inline bool on_box_face(const double c[3], const double mn[3], const double mx[3],
                                            const double dim[3]) {
#pragma clang fp contract(off)
    const double delta_b = 1.e-9;   // f-src/mod_write_deck.f90:9
    if ((c[2] - mn[2]) <= (dim[2] * delta_b)) return true;
    if ((mx[2] - c[2]) <= (dim[2] * delta_b)) return true;
    if ((c[1] - mn[1]) <= (dim[1] * delta_b)) return true;
    if ((mx[1] - c[1]) <= (dim[1] * delta_b)) return true;
    if ((c[0] - mn[0]) <= (dim[0] * delta_b)) return true;
    if ((mx[0] - c[0]) <= (dim[0] * delta_b)) return true;
    return false;
}

}  // namespace dtc
