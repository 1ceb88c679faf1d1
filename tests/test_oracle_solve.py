"""The oracle's solve and fit against closed forms, a direct sparse solve and
invariants (f-src/mod_dtc_main_subroutines.f90:426-1194,
f-src/mod_struct_calcmat.f90:216-473)."""
import numpy as np
import pytest

from helpers import domain_box, iso_voigt, random_mask
from oracle import oracle as O


def _cube(n, delta, xa=(2, 2, 2), E=5600.0, nu=0.3, size=None):
    phi = np.ones((n, n, n), np.int8)
    nodes, elems = O.mesh(phi, 1, 8, (delta,) * 3, xa)
    mn, mx, dim = domain_box((n, n, n), delta, xa)
    a = (size if size is not None else n * delta) * delta
    return nodes, elems, mn, mx, dim, a


@pytest.mark.parametrize("n,delta,size", [(4, 0.01, None), (5, 0.02, None), (3, 0.5, 7.0)])
def test_solid_cube_gives_isotropic_voigt(n, delta, size):
    """Trilinear FE reproduces affine fields: C is the analytic isotropic
    stiffness for any n, delta and SIZE_DOMAIN (Appendix A1)."""
    nodes, elems, mn, mx, dim, a = _cube(n, delta, size=size)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-12)
    KM, Cm, EE = O.effective(r["ff"], dim)
    if size is not None:
        # K_e uses a = SIZE_DOMAIN * delta; C is exact only when SIZE_DOMAIN = x_D * delta
        Cm = Cm * (n * delta) / size
    assert np.abs(Cm - iso_voigt(5600.0, 0.3)).max() < 1e-8 * 7538.5
    assert r["relres"].max() < 1e-11


def test_affine_load_case_solution_is_exact():
    n, delta = 4, 0.01
    nodes, elems, mn, mx, dim, a = _cube(n, delta)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, want_u=True)
    xi = 2.0 * (nodes - np.array(mn)) / (np.array(mx) - np.array(mn)) - 1.0
    u = r["u"][0].reshape(-1, 3)      # LC 1: u_x = eps (1 + xi_x) / 2 on the box
    assert np.allclose(u[:, 0], 1e-6 * (1 + xi[:, 0]) / 2, atol=1e-17)
    assert np.abs(u[:, 1:]).max() < 1e-17


def _porous(seed=5, shape=(6, 6, 6), p=0.6):
    phi = random_mask(shape, p, seed)
    phi[0, :, :] = 1                   # one connected component touching the box
    nodes, elems = O.mesh(phi, 1, 8, (0.01,) * 3, (2, 2, 2))
    mn, mx, dim = domain_box(shape[::-1], 0.01, (2, 2, 2))
    return phi, nodes, elems, mn, mx, dim, shape[2] * 0.01 * 0.01


def test_pcg_matches_direct_sparse_solve():
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    phi, nodes, elems, mn, mx, dim, a = _porous()
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, want_u=True, want_r=True)
    K = O.ke(8, 5600.0, 0.3, a)
    m = 3 * nodes.shape[0]
    rows, cols, vals = [], [], []
    for e in elems:
        dof = (3 * (e[:, None] - 1) + np.arange(3)).ravel()
        rows.append(np.repeat(dof, 24)); cols.append(np.tile(dof, 24)); vals.append(K.ravel())
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, m))
    ids, bv = O.boundaries(nodes, mn, mx, dim)
    bc = (3 * (ids[:, None] - 1) + np.arange(3)).ravel()
    free = np.setdiff1d(np.arange(m), bc)
    for lc in (0, 5, 17, 23):
        ub = bv[lc]
        u = np.zeros(m); u[bc] = ub
        u[free] = spl.spsolve(A[free][:, free].tocsc(), -A[free][:, bc] @ ub)
        assert np.abs(u - r["u"][lc]).max() < 1e-9 * np.abs(u).max()
        assert np.abs(A @ u - r["r"][lc]).max() < 1e-8 * np.abs(A @ u).max()


def test_macro_stiffness_symmetric_with_rigid_null_space():
    phi, nodes, elems, mn, mx, dim, a = _porous(seed=9)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13)
    KM, Cm, EE = O.effective(r["ff"], dim)
    s = np.abs(KM).max()
    assert np.abs(KM - KM.T).max() < 1e-7 * s
    # macro nodal translations and rotations carry no force
    X = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    for d in range(3):
        v = np.zeros((3, 8)); v[d] = 1
        assert np.abs(KM @ v.ravel()).max() < 1e-7 * s
    for ax in range(3):
        w = np.zeros(3); w[ax] = 1
        v = np.cross(w, X).T
        assert np.abs(KM @ v.ravel()).max() < 1e-7 * s


def test_energy_form_equals_force_projection():
    """ff(a, b) = g_a^T K u_b: the projected reactions equal the energy
    product of the load-case displacements through vv (self-adjointness)."""
    phi, nodes, elems, mn, mx, dim, a = _porous(seed=11)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-13, want_u=True, want_r=True)
    U, F = r["u"], r["r"]                       # (24, 3N)
    energy = U @ F.T                            # u_a^T K u_b
    vv = O.vv_lin(1e-6)
    # u_a^T K u_b = sum_c vv(c, a) ff(c, b)
    assert np.allclose(energy, vv.T @ r["ff"], rtol=1e-8, atol=1e-8 * np.abs(energy).max())


def test_petsc_norm_option_and_few_constraints():
    phi, nodes, elems, mn, mx, dim, a = _porous(seed=12)
    r = O.solve_domain(nodes, elems, 5600.0, 0.3, a, mn, mx, dim, tol=1e-3, norm_type=1)
    assert r["iters"].max() > 0
    # a voxel that does not reach the domain box: no constrained DOF
    phi = np.zeros((5, 5, 5), np.int8); phi[2, 2, 2] = 1
    nodes, elems = O.mesh(phi, 1, 8, (0.01,) * 3, (2, 2, 2))
    mn, mx, dim = domain_box((5, 5, 5), 0.01, (2, 2, 2))
    with pytest.raises(O.OracleError) as e:
        O.solve_domain(nodes, elems, 5600.0, 0.3, 0.05 * 0.01, mn, mx, dim)
    assert e.value.rc == 2


def test_domain_decomposition_of_the_regression_case():
    # 100^3 image, spacing 0.01, SIZE_DOMAIN 0.5 -> x_D = 50, one domain per axis,
    # domain 0 spans voxels 2..51 (regression/TC01-1_*.meta)
    x_D, nn_D, xdp = O.ddc_params((0.5, 0.5, 0.5), (0.01, 0.01, 0.01), (100, 100, 100))
    assert x_D.tolist() == [50, 50, 50] and nn_D.tolist() == [1, 1, 1] and np.allclose(xdp, 0.5)
    nn_i, xa, xe = O.domain_bounds(0, x_D, nn_D)
    assert xa.tolist() == [2, 2, 2] and xe.tolist() == [51, 51, 51]
    x_D, nn_D, _ = O.ddc_params((0.24, 0.24, 0.24), (0.01,) * 3, (100, 100, 100))
    assert x_D.tolist() == [24] * 3 and nn_D.tolist() == [4] * 3
    nn_i, xa, xe = O.domain_bounds(1 + 4 * 2 + 16 * 3, x_D, nn_D)
    assert nn_i.tolist() == [1, 2, 3] and xa.tolist() == [26, 50, 74] and xe.tolist() == [49, 73, 97]
    vol = np.arange(10 * 9 * 8, dtype=np.int32).reshape(8, 9, 10)   # (z, y, x)
    blk = O.extract_block(vol, (2, 3, 4), (5, 6, 7))
    assert np.array_equal(blk, vol[3:7, 2:6, 1:5])
