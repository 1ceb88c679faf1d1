"""The oracle's element stiffness and load-case tables against the reference's
own literal tables (tests/golden/reference_tables.npz, written by
tests/golden/make_fixtures.py from f-src/mod_linfe.f90, f-src/mod_write_deck.f90
and f-src/mod_struct_calcmat.f90), plus the invariants that fix them."""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_tables.npz"))


@pytest.mark.parametrize("nu", [0.3, 0.2])
@pytest.mark.parametrize("nen,den", [(8, 144.0), (20, 256.0)])
def test_ke_matches_reference_table(nen, den, nu):
    E, a = 5600.0, 0.0024
    K = O.ke(nen, E, nu, a)
    # factor of Hexe08 / Hexe20 (f-src/mod_linfe.f90:182, :904)
    T = GOLD[f"ke{nen:02d}_nu{nu}"] * E * a / (den * (2 * nu ** 2 + nu - 1))
    assert np.abs(K - T).max() <= 1e-12 * np.abs(T).max()


def _rigid_modes(nen):
    if nen == 8:
        off = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    else:
        off = np.array([[0, 0, 0], [1, 0, 0], [2, 0, 0], [2, 1, 0], [2, 2, 0], [1, 2, 0], [0, 2, 0], [0, 1, 0],
                        [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1], [0, 0, 2], [1, 0, 2], [2, 0, 2], [2, 1, 2],
                        [2, 2, 2], [1, 2, 2], [0, 2, 2], [0, 1, 2]], float) / 2.0
    modes = []
    for d in range(3):
        v = np.zeros((nen, 3)); v[:, d] = 1; modes.append(v.ravel())
    for ax in range(3):
        w = np.zeros(3); w[ax] = 1
        modes.append(np.cross(w, off).ravel())
    return np.array(modes).T


@pytest.mark.parametrize("nen", [8, 20])
def test_ke_invariants(nen):
    K = O.ke(nen, 1.0, 0.3, 1.0)
    assert np.abs(K - K.T).max() < 1e-13 * np.abs(K).max()
    ev = np.linalg.eigvalsh(K)
    assert ev.min() > -1e-12 * ev.max()
    assert np.sum(ev > 1e-10 * ev.max()) == 3 * nen - 6          # rank = DOFs - 6 rigid modes
    assert np.abs(K @ _rigid_modes(nen)).max() < 1e-12 * np.abs(K).max()


def test_ke_linear_in_edge_length():
    # K_e scales with a = phdsize * delta (Appendix A1 of the length cancellation)
    assert np.allclose(O.ke(8, 5600, 0.3, 0.5) * 2, O.ke(8, 5600, 0.3, 1.0), rtol=1e-14, atol=0)


def test_uu_and_vv_match_reference():
    assert np.array_equal(O.uu(1), GOLD["uu08"])
    assert np.array_equal(O.vv_lin(1.0), GOLD["vv_lin_unit"])


def test_vv_is_uu_in_macro_dof_order():
    # vv((d-1)*8 + l, lc) == uu(l, d, lc) * RVE_STRAIN / eps (the two tables of
    # f-src/mod_write_deck.f90:63-107 and f-src/mod_struct_calcmat.f90:1684-1770 agree)
    uu = GOLD["uu08"]
    vv = GOLD["vv_lin_unit"]
    for d in range(3):
        for l in range(8):
            assert np.array_equal(vv[d * 8 + l, :], uu[l, d, :] / 1e-6)
    assert abs(np.linalg.det(vv)) > 0.5     # invertible load-case matrix


def test_quad_vv_table_is_singular_in_the_reference():
    # row 49 is never assigned and row 54 holds 61 literals (Appendix A5)
    assert list(np.where(~GOLD["vv_quad_assigned"])[0] + 1) == [49, 54]
    assert GOLD["vv_quad_bad_rows"].tolist() == [[54, 61]]


def test_walsh_basis_block_diagonal():
    """The HEX08 element stiffness is block diagonal in the signed Walsh basis of
    its 8 nodes; the GPU matvec relies on this (45 coefficients)."""
    K = GOLD["ke08_nu0.3"]
    off = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    H = np.array([[np.prod([(2 * o[i] - 1) if (m >> i) & 1 else 1 for i in range(3)]) for o in off] for m in range(8)])
    T = np.kron(H, np.eye(3))
    Kw = T @ K @ T.T / 64.0
    groups = [[(1, 0), (2, 1), (4, 2)], [(1, 1), (2, 0)], [(1, 2), (4, 0)], [(2, 2), (4, 1)], [(3, 0), (6, 2)],
              [(3, 1), (5, 2)], [(3, 2), (5, 1), (6, 0)], [(5, 0), (6, 1)], [(7, 0)], [(7, 1)], [(7, 2)]]
    mask = np.zeros((24, 24), bool)
    for g in groups:
        idx = [3 * m + d for m, d in g]
        mask[np.ix_(idx, idx)] = True
    assert sum(len(g) ** 2 for g in groups) == 45
    assert np.abs(Kw[~mask]).max() < 1e-12 * np.abs(Kw).max()
    assert np.allclose(T.T @ Kw @ T / 64.0 * 64.0, K)
