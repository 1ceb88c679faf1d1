"""First GPU parity checks: mesher bit-exactness and the solid-cube closed form."""
import numpy as np
import pytest

from helpers import domain_box, iso_voigt, random_mask

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,shape,p,nen", [(1, (5, 6, 7), 0.5, 8), (2, (8, 8, 8), 0.3, 8), (3, (4, 5, 3), 0.6, 20),
                                              (4, (12, 10, 9), 0.2, 8)])
def test_mesh_bit_exact(seed, shape, p, nen):
    import dtc_amd
    from oracle import oracle as O
    phi = random_mask(shape, p, seed)
    delta = (0.01, 0.02, 0.015)
    xa = (3, 5, 7)
    ne, ee = O.mesh(phi, 1, nen, delta, xa)
    nn, eg = dtc_amd.mesh(phi, 1, nen, delta, xa)
    assert eg.shape == ee.shape and nn.shape == ne.shape
    assert np.array_equal(eg, ee)
    assert np.array_equal(nn.view(np.int64), ne.view(np.int64))   # bit-exact coordinates


def test_solid_cube_fp64():
    import dtc_amd
    n = 6
    d = dtc_amd.Domain(np.ones((n, n, n), np.int8), llimit=1, delta=(0.01,) * 3, xa_n=(2, 2, 2))
    r = dtc_amd.solve_domain(d, precision=dtc_amd.DTC_FP64, rel_tol=1e-10, rotation=False)
    assert r["status"] == 0, r
    C = iso_voigt(5600.0, 0.3)
    assert np.abs(r["C"] - C).max() / np.abs(C).max() < 1e-8
