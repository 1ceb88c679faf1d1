"""The oracle's rotation search (f-src/mod_struct_calcmat.f90:533-1603) against
tensor algebra and closed forms, and its sensitivity dL/dE against finite
differences of full solves."""
import numpy as np
import pytest

from helpers import domain_box, random_mask
from oracle import oracle as O


def _rot(n, alpha):
    n = np.asarray(n, float) / np.linalg.norm(n)
    K = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
    return np.eye(3) + np.sin(alpha) * K + (1 - np.cos(alpha)) * K @ K


def _mandel_rotation(R):
    """6x6 Mandel matrix Q with Q s = mandel(R^T sigma R) built from rank-2 algebra."""
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    w = [1, 1, 1, np.sqrt(2), np.sqrt(2), np.sqrt(2)]
    Q = np.zeros((6, 6))
    for c, (k, l) in enumerate(pairs):
        S = np.zeros((3, 3))
        S[k, l] = S[l, k] = 1.0 / w[c]
        T = R.T @ S @ R
        Q[:, c] = [w[r] * T[i, j] for r, (i, j) in enumerate(pairs)]
    return Q


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tra_R6_is_the_mandel_rotation(seed):
    rng = np.random.default_rng(seed)
    n, alpha = rng.standard_normal(3), rng.uniform(0, np.pi)
    n /= np.linalg.norm(n)
    A = rng.standard_normal((6, 6)); A = A + A.T
    T = O.rotate66(A, n, alpha)
    R = _rot(n, alpha)
    Q = _mandel_rotation(R.T)
    assert np.allclose(T, Q.T @ A @ Q, atol=1e-12)
    assert np.allclose(O.rotate66(np.eye(6), n, alpha), np.eye(6), atol=1e-14)   # B^T B = I


def _ortho():
    rng = np.random.default_rng(3)
    from dtc_amd import synth
    return synth.orthotropic_C(rng)


def test_cr2_recovers_a_rotated_orthotropic_tensor():
    C0 = _ortho()
    n = np.array([0.3, -0.5, 0.8]); n /= np.linalg.norm(n)
    T = O.rotate66(C0, n, 0.7)
    r = O.rotation_search(T, 2)
    assert r["crit"] < 1e-6 * np.sum(C0 ** 2)
    E = r["EE_opt"]
    off = np.concatenate([E[:3, 3:].ravel(), [E[3, 4], E[3, 5], E[4, 5]]])
    assert np.abs(off).max() < 1e-2 * np.abs(E).max()
    assert E[0, 0] <= E[1, 1] <= E[2, 2]                  # axis permutation to 11 < 22 < 33
    assert np.allclose(r["aa"].T @ r["aa"], np.eye(3), atol=1e-12)


def test_final_angle_uses_the_reference_scale():
    """An orthotropic tensor rotated by exactly 30 degrees about z has its CR_2
    minimum on the grid point (30, 0, 0); the refinement converges at jj = 1
    and the reference scales the angle by 1/(10**jj - 1) (Appendix A7)."""
    C0 = _ortho()
    T = O.rotate66(C0, [0, 0, 1], np.deg2rad(-30.0))
    r = O.rotation_search(T, 2)
    assert r["jj"] == 1
    assert abs(r["alpha"] - 300 * np.pi / (180 * 9)) < 1e-12
    assert np.allclose(r["n"], [0, 0, 1])


def test_cr1_monotropic_criterion():
    C0 = _ortho()
    n = np.array([1.0, 1.0, 0.5]); n /= np.linalg.norm(n)
    T = O.rotate66(C0, n, 1.1)
    r = O.rotation_search(T, 1)
    E = r["EE_opt"]
    assert r["crit"] < 1e-6 * np.sum(C0 ** 2)
    assert np.abs(E[:4, 4:]).max() < 1e-2 * np.abs(E).max()


@pytest.mark.parametrize("nen,shape,mo", [(8, (5, 5, 5), 1), (20, (4, 4, 4), 1), (8, (4, 4, 4), 2), (20, (3, 3, 4), 2)])
def test_gradient_matches_finite_differences(nen, shape, mo):
    """dL/dE_e of the oracle (HEX08 or HEX20 micro elements, HEX08 or HEX20 macro
    element) against central differences of full oracle solves at relative step 1e-4."""
    phi = random_mask(shape, 0.65, 21)
    phi[0] = 1
    delta, xa = 0.01, (2, 2, 2)
    nodes, elems = O.mesh(phi, 1, nen, (delta,) * 3, xa)
    mn, mx, dim = domain_box(shape[::-1], delta, xa)
    a = shape[2] * delta * delta
    rng = np.random.default_rng(4)
    emod = rng.uniform(2000, 8000, elems.shape[0])
    G = rng.standard_normal((6, 6))

    def loss(em):
        r = O.solve_domain(nodes, elems, 1.0, 0.3, a, mn, mx, dim, tol=1e-13, emod=em, macro_order=mo)
        return np.sum(G * O.effective(r["ff"], dim, macro_order=mo)[1]), r

    L0, r0 = loss(emod)
    r0 = O.solve_domain(nodes, elems, 1.0, 0.3, a, mn, mx, dim, tol=1e-13, emod=emod, want_u=True, macro_order=mo)
    g = O.grad_modulus(nodes, elems, 0.3, a, mn, mx, dim, r0["u"], G, macro_order=mo)
    for e in rng.choice(elems.shape[0], 6, replace=False):
        h = 1e-4 * emod[e]
        ep, em = emod.copy(), emod.copy()
        ep[e] += h; em[e] -= h
        fd = (loss(ep)[0] - loss(em)[0]) / (2 * h)
        assert abs(fd - g[e]) < 1e-5 * max(abs(fd), 1e-3 * np.abs(g).max())
