"""torch.autograd binding: C(modulus) and its backward against the oracle's
gradient and central finite differences of the GPU forward."""
import numpy as np
import pytest

from helpers import random_mask

pytestmark = pytest.mark.gpu


def test_autograd_backward_matches_oracle_and_fd():
    import torch
    import dtc_amd
    from dtc_amd.autograd import effective_stiffness
    from oracle import oracle as O

    rng = np.random.default_rng(12)
    phi = random_mask((6, 5, 7), 0.7, 31)
    phi[0] = 1
    mod = rng.uniform(2000, 8000, phi.shape)
    G = rng.standard_normal((6, 6))
    dom = dtc_amd.Domain(phi, delta=(0.01,) * 3)
    kw = dict(precision=0, rel_tol=1e-11)

    m = torch.tensor(mod, dtype=torch.float64, requires_grad=True)
    C = effective_stiffness(m, dom, **kw)
    loss = (torch.as_tensor(G) * C).sum()
    loss.backward()
    g = m.grad.numpy()
    assert np.all(g[phi == 0] == 0)

    # oracle gradient on the same mesh
    nodes, elems = O.mesh(phi, 1, 8, dom.delta, dom.xa_n)
    mn, mx, dim = dom.box()
    ijk = np.rint(nodes[elems[:, 0] - 1] / 0.01 - 1).astype(int)
    emod = mod[ijk[:, 2], ijk[:, 1], ijk[:, 0]]
    r = O.solve_domain(nodes, elems, 1.0, dom.nu, dom.a(), mn, mx, dim, tol=1e-13, emod=emod, want_u=True)
    go = O.grad_modulus(nodes, elems, dom.nu, dom.a(), mn, mx, dim, r["u"], G)
    gv = g[ijk[:, 2], ijk[:, 1], ijk[:, 0]]
    assert np.abs(gv - go).max() < 1e-6 * np.abs(go).max()

    # central differences through the same autograd forward
    for k in rng.choice(np.argwhere(phi.ravel() > 0).ravel(), 3, replace=False):
        h = 1e-4 * mod.ravel()[k]
        vals = []
        for s in (1, -1):
            m2 = mod.copy().ravel()
            m2[k] += s * h
            with torch.no_grad():
                vals.append(float((torch.as_tensor(G) * effective_stiffness(
                    torch.tensor(m2.reshape(phi.shape)), dom, **kw)).sum()))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g.ravel()[k]) < 1e-4 * np.abs(g).max()


def test_autograd_mixed_precision_gradient_close_to_fp64():
    import torch
    import dtc_amd
    from dtc_amd.autograd import effective_stiffness

    rng = np.random.default_rng(13)
    phi = random_mask((8, 8, 8), 0.6, 41)
    phi[0] = 1
    mod = rng.uniform(2000, 8000, phi.shape)
    G = torch.as_tensor(rng.standard_normal((6, 6)))
    dom = dtc_amd.Domain(phi, delta=(0.01,) * 3)
    grads = []
    for prec in (0, 1):
        m = torch.tensor(mod, requires_grad=True)
        (G * effective_stiffness(m, dom, precision=prec, rel_tol=1e-8)).sum().backward()
        grads.append(m.grad.numpy())
    assert np.abs(grads[1] - grads[0]).max() < 1e-5 * np.abs(grads[0]).max()
