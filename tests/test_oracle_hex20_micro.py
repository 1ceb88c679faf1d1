"""The oracle's solve on HEX20 micro meshes (gen_quadmesh with 20-node elements,
f-src/mod_gen_quadmesh.f90:244-265; solve f-src/mod_dtc_main_subroutines.f90:426-1194)
against closed forms and a direct sparse solve of the assembled system."""
import numpy as np
import pytest

from helpers import domain_box, iso_voigt, random_mask
from oracle import oracle as O


def _mesh20(phi, delta=0.01, xa=(2, 2, 2)):
    nodes, elems = O.mesh(phi, 1, 20, (delta,) * 3, xa)
    mn, mx, dim = domain_box(phi.shape[::-1], delta, xa)
    return nodes, elems, mn, mx, dim, phi.shape[2] * delta * delta


@pytest.mark.parametrize("n,macro_order", [(3, 1), (4, 1), (3, 2)])
def test_solid_cube_hex20_micro_gives_isotropic_voigt(n, macro_order):
    """Serendipity elements reproduce affine fields exactly: the solid cube's
    C is the analytic isotropic stiffness."""
    nodes, elems, mn, mx, dim, a = _mesh20(np.ones((n, n, n), np.int8))
    assert elems.shape == (n ** 3, 20)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-12, macro_order=macro_order)
    KM, Cm, EE = O.effective(r["ff"], dim, macro_order=macro_order)
    assert np.abs(Cm - iso_voigt(5600.0, 0.3)).max() < 1e-8 * 7538.5
    assert r["relres"].max() < 1e-11


def test_hex20_micro_affine_load_case_is_exact_at_mid_edge_nodes():
    nodes, elems, mn, mx, dim, a = _mesh20(np.ones((3, 3, 3), np.int8))
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, want_u=True)
    xi = 2.0 * (nodes - np.array(mn)) / (np.array(mx) - np.array(mn)) - 1.0
    u = r["u"][1].reshape(-1, 3)      # LC 2: u_y = eps (1 + xi_y) / 2
    assert np.allclose(u[:, 1], 1e-6 * (1 + xi[:, 1]) / 2, atol=1e-17)
    assert np.abs(u[:, [0, 2]]).max() < 1e-17


def test_hex20_micro_pcg_matches_direct_sparse_solve():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    phi = random_mask((4, 4, 4), 0.6, 21)
    phi[0, :, :] = 1
    nodes, elems, mn, mx, dim, a = _mesh20(phi)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, want_u=True, want_r=True)
    K = O.ke(20, 5600.0, 0.3, a)
    m = 3 * nodes.shape[0]
    rows, cols, vals = [], [], []
    for e in elems:
        dof = (3 * (e[:, None] - 1) + np.arange(3)).ravel()
        rows.append(np.repeat(dof, 60)); cols.append(np.tile(dof, 60)); vals.append(K.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, m))
    ids, bv = O.boundaries(nodes, mn, mx, dim)
    bc = (3 * (ids[:, None] - 1) + np.arange(3)).ravel()
    free = np.setdiff1d(np.arange(m), bc)
    for lc in (0, 4, 13, 22):
        ub = bv[lc]
        u = np.zeros(m); u[bc] = ub
        u[free] = spl.spsolve(A[free][:, free].tocsc(), -A[free][:, bc] @ ub)
        assert np.abs(u - r["u"][lc]).max() < 1e-9 * np.abs(u).max()
        assert np.abs(A @ u - r["r"][lc]).max() < 1e-8 * np.abs(A @ u).max()


def test_hex20_micro_strain_energy_below_hex08():
    """Minimum potential energy: with the same affine boundary trace (load cases
    1-6) the trilinear field space is contained in the serendipity one, so the
    HEX20 solution's strain energy u^T K u is at most the HEX08 one."""
    phi = random_mask((5, 5, 5), 0.7, 5)
    phi[0, :, :] = 1
    n8, e8 = O.mesh(phi, 1, 8, (0.01,) * 3, (2, 2, 2))
    n20, e20, mn, mx, dim, a = _mesh20(phi)
    assert e8.shape[0] == e20.shape[0]
    kw = dict(tol=1e-13, want_u=True, want_r=True, lc_count=6)
    r8 = O.solve_domain(n8, e8, 5600.0, 0.3, a, mn, mx, dim, **kw)
    r20 = O.solve_domain(n20, e20, 5600.0, 0.3, a, mn, mx, dim, **kw)
    for lc in range(6):
        w8 = r8["u"][lc] @ r8["r"][lc]
        w20 = r20["u"][lc] @ r20["r"][lc]
        assert 0.0 < w20 <= w8 * (1 + 1e-9)
        assert w20 > 0.5 * w8
