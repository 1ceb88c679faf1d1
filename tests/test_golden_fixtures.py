"""The oracle fixtures of whole domains (tests/golden/trabecular64_seed2.npz, the
metric's 64^3 domain, and tests/golden/uneven3_d4/d5.npz, the two smallest domains
of the uneven batch) checked against what the mathematics fixes, so that a fixture
written wrongly (a column of ff lost, a transposed block) fails here on the CPU:
every load case converged to 1e-10, K_M = ff vv^-1 symmetric (it equals the energy
form u_a^T K u_b at convergence), C positive definite with EE its symmetric part,
every column of ff in equilibrium (the reactions of a load case sum to zero per
direction), and the density N_e / n^3."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_trabecular64 as G64  # noqa: E402
import make_uneven3 as GU  # noqa: E402

FIXTURES = [(G64.OUT, 64)] + [(GU.out_path(i), None) for i in GU.DOMAINS]


def _load(path, n):
    fx = np.load(path, allow_pickle=False)
    return fx, (n if n is not None else int(fx["edge"]))


@pytest.mark.parametrize("path,n", FIXTURES, ids=[os.path.basename(f[0]) for f in FIXTURES])
def test_fixture_converged_and_sized(path, n):
    fx, n = _load(path, n)
    assert float(fx["tol"]) == 1e-10 and (fx["relres"] <= 1e-10).all() and (fx["iters"] > 0).all()
    assert int(fx["n_elems"]) > 0.1 * n ** 3 and int(fx["n_nodes"]) > int(fx["n_elems"])
    assert abs(float(fx["density"]) - int(fx["n_elems"]) / n ** 3) < 1e-15


@pytest.mark.parametrize("path,n", FIXTURES, ids=[os.path.basename(f[0]) for f in FIXTURES])
def test_fixture_macro_stiffness_symmetric_and_tensor_spd(path, n):
    fx, n = _load(path, n)
    KM, C, EE = fx["K_M"], fx["C"], fx["EE"]
    assert np.abs(KM - KM.T).max() < 1e-6 * np.abs(KM).max()
    assert np.abs(C - C.T).max() < 1e-6 * np.abs(C).max()
    assert np.allclose(EE, 0.5 * (C + C.T), rtol=0, atol=1e-12 * np.abs(C).max())
    assert np.linalg.eigvalsh(EE).min() > 0.0
    # every load case carries force: no column of ff is empty
    assert (np.abs(fx["ff"]).max(axis=0) > 0).all()


@pytest.mark.parametrize("path,n", FIXTURES, ids=[os.path.basename(f[0]) for f in FIXTURES])
def test_fixture_reactions_in_equilibrium(path, n):
    """ff(dof, lc), dof = d * 8 + l (direction-major over the 8 macro nodes): the
    projected reactions of one load case sum to zero in each direction, since the
    shape functions are a partition of unity and the internal forces balance."""
    ff = _load(path, n)[0]["ff"]
    scale = np.abs(ff).max()
    for d in range(3):
        assert np.abs(ff[8 * d:8 * d + 8, :].sum(axis=0)).max() < 1e-7 * scale
