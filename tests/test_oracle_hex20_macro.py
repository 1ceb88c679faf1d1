"""The oracle's HEX20 macro element (60 load cases): load-case table against the
reference's literal text (f-src/mod_write_deck.f90:118-279), boundary values
against closed forms, solves and the 6x6 fit (f-src/mod_struct_calcmat.f90:365-401)
against the isotropic closed form and the HEX08 macro element."""
import os

import numpy as np
import pytest

from helpers import domain_box, iso_voigt, random_mask
from oracle import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))

# ANSYS 20-node brick local coordinates (f-src/mod_linfe.f90:839-859)
XI20 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1], [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1],
                 [0, -1, -1], [1, 0, -1], [0, 1, -1], [-1, 0, -1], [0, -1, 1], [1, 0, 1], [0, 1, 1], [-1, 0, 1],
                 [-1, -1, 0], [1, -1, 0], [1, 1, 0], [-1, 1, 0]], float)


def _cube(n, delta=0.01, xa=(2, 2, 2)):
    phi = np.ones((n, n, n), np.int8)
    nodes, elems = O.mesh(phi, 1, 8, (delta,) * 3, xa)
    mn, mx, dim = domain_box((n, n, n), delta, xa)
    return nodes, elems, mn, mx, dim, n * delta * delta


def test_uu20_table_equals_the_reference_literals():
    z = np.load(os.path.join(HERE, "golden", "reference_tables.npz"))
    assert np.array_equal(O.uu(2), z["uu20"])


def test_uu20_default_real_literals_are_single_precision():
    """`.9*eps` and `-.7*eps` carry no kind suffix: the literal is a default
    (single-precision) real promoted to double before the product."""
    u = O.uu(2)
    assert u[18, 2, 11] == float(np.float32(0.9)) * 1e-6 != 0.9e-6
    assert u[5, 2, 17] == -float(np.float32(0.7)) * 1e-6


def test_hex20_shape_functions_are_nodal_and_partition_unity():
    mn, mx = [0.0, 0.0, 0.0], [2.0, 2.0, 2.0]
    for m, xi in enumerate(XI20):
        N = O.macro_shape(xi + 1.0, mn, mx, 2)
        e = np.zeros(20); e[m] = 1.0
        assert np.abs(N - e).max() < 1e-15
    rng = np.random.default_rng(0)
    for x in rng.uniform(0, 2, (20, 3)):
        assert abs(O.macro_shape(x, mn, mx, 2).sum() - 1.0) < 1e-14


def test_hex20_strain_load_cases_equal_the_hex08_ones():
    """Load cases 1-6 are affine strain fields in both tables; the serendipity
    interpolation reproduces affine fields, so the boundary values agree."""
    nodes, elems, mn, mx, dim, a = _cube(5)
    i8, v8 = O.boundaries(nodes, mn, mx, dim, 1)
    i20, v20 = O.boundaries(nodes, mn, mx, dim, 2)
    assert np.array_equal(i8, i20)
    assert v20.shape == (60, 3 * len(i20))
    assert np.abs(v20[:6] - v8[:6]).max() < 1e-21


def test_hex20_boundary_values_at_macro_nodes_are_the_table():
    n = 4
    nodes, elems, mn, mx, dim, a = _cube(n)
    ids, vals = O.boundaries(nodes, mn, mx, dim, 2)
    uu = O.uu(2)
    X = nodes[ids - 1]
    for m, xi in enumerate(XI20):
        target = np.array(mn) + (xi + 1.0) / 2.0 * (np.array(mx) - np.array(mn))
        b = int(np.argmin(np.abs(X - target).sum(axis=1)))
        assert np.abs(X[b] - target).max() < 1e-12
        for d in range(3):
            assert np.allclose(vals[:, 3 * b + d], uu[m, d, :], rtol=1e-13, atol=1e-21)


def test_solid_cube_hex20_macro_gives_isotropic_voigt_and_symmetric_macro_stiffness():
    nodes, elems, mn, mx, dim, a = _cube(4)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-12, macro_order=2)
    assert r["ff"].shape == (60, 60) and r["relres"].max() < 1e-11
    KM, Cm, EE = O.effective(r["ff"], dim, macro_order=2)
    assert np.abs(Cm - iso_voigt(5600.0, 0.3)).max() < 1e-8 * 7538.5
    s = np.abs(KM).max()
    assert KM.shape == (60, 60) and np.abs(KM - KM.T).max() < 1e-7 * s


def test_hex20_macro_strain_load_cases_match_hex08_solves():
    """Same boundary values for load cases 1-6 -> same reactions; the force
    projections differ only through the macro shape functions."""
    phi = random_mask((6, 6, 6), 0.65, 17)
    phi[0] = 1
    nodes, elems = O.mesh(phi, 1, 8, (0.01,) * 3, (2, 2, 2))
    mn, mx, dim = domain_box((6, 6, 6), 0.01, (2, 2, 2))
    a = 0.06 * 0.01
    r1 = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, lc_count=6, want_r=True)
    r2 = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, lc_count=6, want_r=True,
                        macro_order=2)
    assert np.abs(r1["r"][:6] - r2["r"][:6]).max() < 1e-9 * np.abs(r1["r"][:6]).max()
    # ff rows of HEX08 macro node l equal the HEX20 corner rows plus half of the adjacent mid-edge rows
    f8, f20 = r1["ff"][:, :6], r2["ff"][:, :6]
    mids = {0: (8, 11, 16), 1: (8, 9, 17), 2: (9, 10, 18), 3: (10, 11, 19),
            4: (12, 15, 16), 5: (12, 13, 17), 6: (13, 14, 18), 7: (14, 15, 19)}
    for d in range(3):
        for l in range(8):
            want = f20[20 * d + l] + 0.5 * sum(f20[20 * d + m] for m in mids[l])
            assert np.allclose(f8[8 * d + l], want, rtol=1e-9, atol=1e-9 * np.abs(f8).max())
