"""GPU parity of HEX20 micro-element domains (node-gather operator) against the
oracle: closed forms, porous domains, both macro elements, failures in a batch.
Bars as for HEX08: relative 1e-6 (FP64) / 1e-5 (FP32 + FP64 refinement)."""
import numpy as np
import pytest

from helpers import iso_voigt

pytestmark = pytest.mark.gpu


def _dtc():
    import dtc_amd
    return dtc_amd


def _O():
    from oracle import oracle as O
    return O


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _oracle(d, macro_order=1, want=False):
    O = _O()
    nodes, elems = O.mesh(d.phi, d.llimit, 20, d.delta, d.xa_n)
    mn, mx, dim = d.box()
    r = O.solve_domain(nodes, elems, d.E, d.nu, d.a(), mn, mx, dim, tol=1e-12, want_u=want, want_r=want,
                       macro_order=macro_order)
    KM, Cm, EE = O.effective(r["ff"], dim, d.rve_strain, macro_order=macro_order)
    return dict(r, KM=KM, C=Cm, EE=EE, nodes=nodes, elems=elems)


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-5)])
def test_hex20_micro_solid_cube_closed_form(prec, tol):
    dtc = _dtc()
    d = dtc.Domain(np.ones((5, 5, 5), np.int8), delta=(0.02,) * 3, xa_n=(3, 3, 3), micro_nen=20)
    r = dtc.solve_domain(d, precision=prec, rel_tol=1e-8, rotation=False)
    assert r["status"] == 0 and r["relres"] <= 1e-8
    assert r["n_elems"] == 125 and r["n_nodes"] == 6 ** 3 + 3 * 5 * 36
    assert _rel(r["C"], iso_voigt(5600.0, 0.3)) < tol


@pytest.mark.parametrize("prec,tol", [(0, 1e-6), (1, 1e-5)])
def test_hex20_micro_porous_domain_matches_oracle(prec, tol):
    dtc = _dtc()
    from dtc_amd import synth
    phi = synth.trabecular(9, seed=81, sigma=1.4, bvtv=0.45)
    d = dtc.Domain(phi, delta=(0.01,) * 3, xa_n=(2, 2, 2), micro_nen=20)
    r = dtc.solve_domain(d, precision=prec, rel_tol=1e-8, rotation=False, want_u=True, want_r=True)
    o = _oracle(d, want=True)
    assert r["status"] == 0 and r["relres"] <= 1e-8
    assert r["n_elems"] == o["elems"].shape[0] and r["n_nodes"] == o["nodes"].shape[0]
    for key in ("ff", "K_macro", "C", "EE"):
        ref = o["KM"] if key == "K_macro" else o[key]
        assert _rel(r[key], ref) < tol, key
    assert r["u"].shape == o["u"].shape == (24, 3 * o["nodes"].shape[0])
    assert _rel(r["u"], o["u"]) < 10 * tol
    assert np.abs(r["r"] - o["r"]).max() < 10 * tol * np.abs(o["r"]).max()


def test_hex20_micro_with_hex20_macro_matches_oracle():
    dtc = _dtc()
    from dtc_amd import synth
    phi = synth.trabecular(8, seed=82, sigma=1.4, bvtv=0.5)
    d = dtc.Domain(phi, delta=(0.01,) * 3, micro_nen=20, macro_order=2)
    r = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False)
    o = _oracle(d, macro_order=2)
    assert r["status"] == 0 and r["ff"].shape == (60, 60)
    assert _rel(r["ff"], o["ff"]) < 1e-5
    assert _rel(r["C"], o["C"]) < 1e-5 and _rel(r["EE"], o["EE"]) < 1e-5


def test_hex20_micro_batch_with_failures_and_per_voxel_modulus():
    dtc = _dtc()
    from dtc_amd import synth
    O = _O()
    grey, modulus = synth.greyscale(8, seed=6, bvtv=0.5)
    d1 = dtc.Domain(grey, llimit=1, delta=(0.01,) * 3, modulus=modulus, E=10000.0, micro_nen=20)
    d2 = dtc.Domain(np.zeros((4, 4, 4), np.int8), micro_nen=20)                 # empty
    few = np.zeros((5, 5, 5), np.int8); few[2, 2, 2] = 1                          # no node on the box
    d3 = dtc.Domain(few, micro_nen=20)
    phi = synth.trabecular(7, seed=83, sigma=1.3, bvtv=0.5)
    d4 = dtc.Domain(phi, delta=(0.02,) * 3, xa_n=(3, 4, 5), micro_nen=20)
    rb = dtc.solve_batch([d1, d2, d3, d4], precision=1, rel_tol=1e-8, rotation=True, precond=1)
    assert [x["status"] for x in rb] == [0, 1, 2, 0]
    # per-voxel modulus: element e of the mesh takes the modulus of its voxel
    nodes, elems = O.mesh(grey, 1, 20, d1.delta, d1.xa_n)
    base = nodes[elems[:, 0] - 1]
    ijk = np.rint(base / np.array(d1.delta) - (np.array(d1.xa_n) - 1)).astype(int)
    mn, mx, dim = d1.box()
    o1 = O.solve_domain(nodes, elems, d1.E, d1.nu, d1.a(), mn, mx, dim, tol=1e-12,
                        emod=modulus[ijk[:, 2], ijk[:, 1], ijk[:, 0]])
    _, C1, _ = O.effective(o1["ff"], dim)
    assert _rel(rb[0]["C"], C1) < 1e-5
    assert _rel(rb[3]["C"], _oracle(d4)["C"]) < 1e-5
    s4 = dtc.solve_domain(d4, precision=1, rel_tol=1e-8, rotation=True)
    assert _rel(rb[3]["C"], s4["C"]) < 1e-6
    for c in range(2):
        assert rb[3]["rot"][c]["alpha"] == pytest.approx(s4["rot"][c]["alpha"], abs=1e-6)


def test_hex20_micro_rejects_mixed_element_types():
    dtc = _dtc()
    phi = np.ones((4, 4, 4), np.int8)
    with pytest.raises(Exception):
        dtc.solve_batch([dtc.Domain(phi), dtc.Domain(phi, micro_nen=20)], precision=1)


@pytest.mark.parametrize("nen,macro_order", [(20, 1), (8, 2), (20, 2)])
def test_hex20_gradient_matches_oracle_and_finite_differences(nen, macro_order):
    """dL/dE per voxel with HEX20 micro and/or HEX20 macro elements (the 60 load
    cases as 3 groups) from the stored solutions against the oracle's sensitivity
    (1e-6 of max |g|, zero on void voxels) and against central differences of full
    GPU solves (relative step 1e-4)."""
    dtc, O = _dtc(), _O()
    from helpers import random_mask
    rng = np.random.default_rng(19)
    phi = random_mask((5, 5, 5), 0.7, 91)
    phi[0] = 1
    modulus = rng.uniform(2000, 8000, phi.shape)
    G = rng.standard_normal((6, 6))
    mk = lambda m: dtc.Domain(phi, delta=(0.01,) * 3, modulus=m, micro_nen=nen, macro_order=macro_order)
    d = mk(modulus)
    with dtc.Batch([d], precision=0, rel_tol=1e-11, rotation=False) as b:
        r = b.run()[0]
        g = b.grad(0, G)
    assert r["status"] == 0
    nodes, elems = O.mesh(d.phi, d.llimit, nen, d.delta, d.xa_n)
    ijk = np.rint(nodes[elems[:, 0] - 1] / 0.01 - 1).astype(int)
    mn, mx, dim = d.box()
    o = O.solve_domain(nodes, elems, d.E, d.nu, d.a(), mn, mx, dim, tol=1e-13, want_u=True,
                       emod=modulus[ijk[:, 2], ijk[:, 1], ijk[:, 0]], macro_order=macro_order)
    go = O.grad_modulus(nodes, elems, d.nu, d.a(), mn, mx, dim, o["u"], G, macro_order=macro_order)
    gv = g[ijk[:, 2], ijk[:, 1], ijk[:, 0]]
    assert _rel(gv, go) < 1e-6
    assert np.all(g[phi == 0] == 0)
    for k in rng.choice(np.argwhere(phi.ravel() > 0).ravel(), 3, replace=False):
        h = 1e-4 * modulus.ravel()[k]
        vals = []
        for sgn in (1, -1):
            m2 = modulus.copy().ravel()
            m2[k] += sgn * h
            x = dtc.solve_domain(mk(m2.reshape(phi.shape)), precision=0, rel_tol=1e-11, rotation=False)
            vals.append(np.sum(G * x["C"]))
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - g.ravel()[k]) < 1e-4 * np.abs(g).max()


@pytest.mark.parametrize("macro_order", [1, 2])
def test_hex20_pmultigrid_matches_oracle(macro_order):
    """The mixed solve of HEX20 micro domains with the p-multigrid (HEX20
    Chebyshev-Jacobi smoothing, coarse space the HEX08 discretisation of the same
    voxels with its geometric V-cycle, block PCG) against the oracle, both macro
    elements, within the mixed bar 1e-5."""
    dtc = _dtc()
    from dtc_amd import synth
    phi = synth.trabecular(9, seed=83, sigma=1.4, bvtv=0.45)
    d = dtc.Domain(phi, delta=(0.01,) * 3, xa_n=(2, 2, 2), micro_nen=20, macro_order=macro_order)
    r = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    o = _oracle(d, macro_order=macro_order)
    assert r["status"] == 0 and r["relres"] <= 1e-8 and r["precision_used"] == dtc.DTC_FP32_REFINE
    for key in ("ff", "C", "EE"):
        assert _rel(r[key], o[key]) < 1e-5, key
    assert _rel(r["K_macro"], o["KM"]) < (1e-5 if macro_order == 1 else 1e-4)


def test_hex20_pmultigrid_cuts_iterations_greyscale():
    """Greyscale 24^3 HEX20 domain (per-voxel moduli, contrast up to 6e4): the
    p-multigrid block PCG reaches 1e-8 with far fewer iterations than Jacobi-PCG
    and the same tensors (to what rel-res 1e-8 fixes)."""
    dtc = _dtc()
    from dtc_amd import synth
    grey, modulus = synth.greyscale(24, seed=24, bvtv=0.35)
    d = dtc.Domain(grey, llimit=1, delta=(0.01,) * 3, modulus=modulus, E=10000.0, micro_nen=20)
    rj = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_JACOBI)
    rm = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    assert rj["status"] == 0 == rm["status"] and rm["relres"] <= 1e-8
    assert _rel(rm["C"], rj["C"]) < 1e-6 and _rel(rm["K_macro"], rj["K_macro"]) < 1e-6
    assert rm["iterations"] < rj["iterations"] / 5, (rm["iterations"], rj["iterations"])


@pytest.mark.parametrize("mfma", ["1", "0"])
@pytest.mark.parametrize("precond", ["jacobi", "mg"])
def test_hex20_element_operator_equals_node_gather(monkeypatch, precond, mfma):
    """The FP32 HEX20 operator in two passes (element forces Khat x_e on the matrix
    cores or on the VALU, then a sum per node) against the node-gather form of
    the same product: they differ only in FP32 rounding (the order of the 60
    products of a row, E * (Khat x_e) rounded before the sum over a node's
    elements), so tensors agree to 1e-6 and the iteration counts within 5 %."""
    monkeypatch.setenv("DTC_HEX20_MFMA", mfma)
    dtc = _dtc()
    from dtc_amd import synth
    phi = synth.trabecular(12, seed=7, sigma=1.5, bvtv=0.4)
    d = dtc.Domain(phi, delta=(0.01,) * 3, xa_n=(2, 2, 2), micro_nen=20)
    pc = dtc.DTC_PC_MG if precond == "mg" else dtc.DTC_PC_JACOBI
    out = {}
    for gather in ("1", "0"):
        monkeypatch.setenv("DTC_HEX20_GATHER", gather)
        out[gather] = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=pc)
    a, b = out["1"], out["0"]
    assert a["status"] == 0 and b["status"] == 0 and b["relres"] <= 1e-8
    for key in ("ff", "K_macro", "C"):
        assert _rel(b[key], a[key]) < 1e-6, key
    assert abs(b["iterations"] - a["iterations"]) <= max(2, 0.05 * a["iterations"])


def test_hex20_fused_chebyshev_step_equals_the_separate_kernel(monkeypatch):
    """The HEX20 smoothing step formed in the epilogue of the operator's node sums
    (no t = K x vector) computes the same direction and update expression as the
    separate vector kernel (DTC_CHEB_FUSE=0): the solves agree bit for bit, on a
    binary and a greyscale HEX20 batch with the p-multigrid."""
    dtc = _dtc()
    from dtc_amd import synth
    grey, modulus = synth.greyscale(16, seed=31, bvtv=0.35)
    doms = [dtc.Domain(synth.trabecular(12, seed=7, sigma=1.5, bvtv=0.4), delta=(0.01,) * 3, xa_n=(2, 2, 2),
                       micro_nen=20),
            dtc.Domain(grey, llimit=1, delta=(0.01,) * 3, modulus=modulus, E=10000.0, micro_nen=20)]
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    for d in doms:
        monkeypatch.setenv("DTC_CHEB_FUSE", "0")
        ra = dtc.solve_domain(d, **kw)
        monkeypatch.delenv("DTC_CHEB_FUSE")
        rb = dtc.solve_domain(d, **kw)
        assert ra["status"] == rb["status"] == 0
        assert ra["iterations"] == rb["iterations"]
        for k in ("C", "K_macro", "ff"):
            assert np.array_equal(ra[k], rb[k]), k


def test_hex20_hinged_clusters_converge_in_the_mixed_solve():
    """A 48^3 trabecular HEX20 domain whose solid holds voxel clusters attached to
    the rest only through an edge or a corner (near-mechanisms): the block PCG
    with the p-multigrid converges in the mixed precision (no FP64 fallback).
    With the block's Q = K P in FP32 arithmetic (DTC_BLOCK_Q64=0) this domain
    diverged after the first residual replacement."""
    dtc = _dtc()
    from dtc_amd import synth
    from scipy import ndimage
    phi = synth.trabecular(48, seed=4, sigma=2.0, bvtv=0.2)
    lab6, n6 = ndimage.label(phi > 0)                          # face-connected parts
    lab26, n26 = ndimage.label(phi > 0, structure=np.ones((3, 3, 3)))
    assert n26 == 1 and n6 > 1                                  # hinged parts are present
    d = dtc.Domain(phi, delta=(0.01,) * 3, micro_nen=20)
    r = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    assert r["status"] == 0 and r["relres"] <= 1e-8
    assert r["precision_used"] == dtc.DTC_FP32_REFINE
    assert r["iterations"] <= 150, r["iterations"]
    K = r["K_macro"]
    assert np.abs(K - K.T).max() < 1e-6 * np.abs(K).max()
