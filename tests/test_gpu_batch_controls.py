"""Batch handle controls on the GPU: reuse of one handle under a different
sub-batch split, the per-domain automatic inner tolerance, the precision a
result was produced in, and the kernel timing statistics."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dtc():
    import dtc_amd
    return dtc_amd


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _binary(dtc, n, seed):
    from dtc_amd import synth
    return dtc.Domain(synth.trabecular(n, seed=seed, sigma=1.5, bvtv=0.3), delta=(0.01,) * 3)


def _grey(dtc, n, seed):
    from dtc_amd import synth
    g, m = synth.greyscale(n, seed=seed, bvtv=0.3)
    return dtc.Domain(g, llimit=1, delta=(0.01,) * 3, modulus=m, E=10000.0)


def test_one_handle_run_concurrent_then_sequential(monkeypatch):
    """The same Batch handle run first as two concurrent halves, then as one
    sub-batch, then cut into small sub-batches: each run equals, bit for bit, a
    fresh handle run with the same split (the block PCG's per-domain slices and
    every cached buffer follow the split), and the three splits agree to 1e-7,
    i.e. to what a relative residual of 1e-8 fixes (a domain's multigrid and
    launch shapes depend on its sub-batch, so the floating-point order, and an
    iteration count by one, can differ: measured 4.4e-9)."""
    dtc = _dtc()
    doms = [_binary(dtc, 16 + 2 * i, 40 + i) for i in range(4)]
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    splits = [dict(DTC_CONCURRENT="2"), dict(DTC_CONCURRENT="1"), dict(DTC_CONCURRENT="1", DTC_SUBBATCH_NODES="6000")]

    def env(sp):
        monkeypatch.delenv("DTC_SUBBATCH_NODES", raising=False)
        for k, v in sp.items():
            monkeypatch.setenv(k, v)

    reused = []
    with dtc.Batch(doms, **kw) as b:
        for sp in splits:
            env(sp)
            reused.append(b.run())
    for sp, rr in zip(splits, reused):
        env(sp)
        with dtc.Batch(doms, **kw) as f:
            rf = f.run()
        for a, c in zip(rr, rf):
            assert a["status"] == c["status"] == 0, sp
            assert a["iterations"] == c["iterations"] and a["outer_iterations"] == c["outer_iterations"], sp
            assert np.array_equal(a["C"], c["C"]) and np.array_equal(a["K_macro"], c["K_macro"]), sp
    for a, c, d in zip(*reused):
        assert abs(a["iterations"] - c["iterations"]) <= 2 and abs(d["iterations"] - c["iterations"]) <= 2
        assert _rel(a["C"], c["C"]) < 1e-7 and _rel(d["C"], c["C"]) < 1e-7
        assert _rel(a["K_macro"], c["K_macro"]) < 1e-7 and _rel(d["K_macro"], c["K_macro"]) < 1e-7


def test_inner_tolerance_is_chosen_per_domain():
    """Automatic inner tolerance of the FP32 block: 1e-6 for a binary domain (as
    with that tolerance given explicitly).  A batch that holds a greyscale domain
    (contrast > 100 over its solid voxels) iterates on FP64 vectors, where no
    inner tolerance applies: the greyscale domain iterates exactly as alone (and
    as with an inner tolerance given), the binary one in that batch reaches the
    same tensor as alone."""
    dtc = _dtc()
    bi, gr = _binary(dtc, 24, 51), _grey(dtc, 20, 52)
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    mix = dtc.solve_batch([bi, gr], **kw)
    b_alone = dtc.solve_domain(bi, **kw)
    g_alone = dtc.solve_domain(gr, **kw)
    b_6 = dtc.solve_domain(bi, inner_tol=1e-6, **kw)
    g_4 = dtc.solve_domain(gr, inner_tol=1e-4, **kw)
    for r in mix + [b_alone, g_alone, b_6, g_4]:
        assert r["status"] == 0 and r["relres"] <= 1e-8
        assert r["precision_used"] == dtc.DTC_FP32_REFINE
    key = lambda r: (r["iterations"], r["outer_iterations"])
    assert key(b_alone) == key(b_6)
    assert key(mix[1]) == key(g_alone) == key(g_4)
    # a domain's multigrid and launch shapes depend on its batch: equal to what rel-res 1e-8 fixes
    assert _rel(mix[0]["C"], b_alone["C"]) < 1e-7 and _rel(mix[1]["C"], g_alone["C"]) < 1e-7


def test_precision_used_flag():
    dtc = _dtc()
    d = _binary(dtc, 12, 61)
    r1 = dtc.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    r0 = dtc.solve_domain(d, precision=0, rel_tol=1e-8, rotation=False)
    assert r1["status"] == 0 == r0["status"]
    assert r1["precision_used"] == dtc.DTC_FP32_REFINE and r0["precision_used"] == dtc.DTC_FP64


def test_kernel_timing_statistics():
    """With timing on, the CG matvec, the fused Chebyshev step, the restriction
    residual and the FP64 residual are timed; the average algorithmic bytes and the
    average dense element-product flops of a launch lie between one domain's and the
    whole batch's, for the formula of each kernel.  With timing off nothing is
    timed."""
    dtc = _dtc()
    doms = [_binary(dtc, 20, 70 + i) for i in range(3)]
    with dtc.Batch(doms, precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG) as b:
        res = b.run()
        assert all(b.kernel_stats(k)["launches"] == 0 for k in dtc.KSTAT_NAMES)
        b.timing(True)
        res = b.run()
        st = {k: b.kernel_stats(k) for k in dtc.KSTAT_NAMES}
    assert all(r["status"] == 0 for r in res)
    V = [3 * r["n_nodes"] * 24 * 4 for r in res]
    E = [4 * r["n_elems"] for r in res]
    J = [3 * r["n_nodes"] * 4 for r in res]
    one = lambda vp, jac=False, s=1: min(vp * s * v + s * e + (j if jac else 0) for v, e, j in zip(V, E, J))
    full = lambda vp, jac=False, s=1: sum(vp * s * v + s * e + (j if jac else 0) for v, e, j in zip(V, E, J))
    for k, (lo, hi) in {dtc.DTC_KSTAT_MV: (one(2), full(2)), dtc.DTC_KSTAT_BRES: (one(3), full(3)),
                        dtc.DTC_KSTAT_CHEB: (one(3, True), full(5, True)),
                        dtc.DTC_KSTAT_RESID: (one(1, s=2), full(2, s=2))}.items():
        s = st[k]
        assert s["launches"] > 0 and s["avg_ms"] > 0.0, (k, s)
        assert lo * (1 - 1e-12) <= s["bytes"] <= hi * (1 + 1e-12), (k, s, lo, hi)
        # dense element-product flops 2 (3 nen)^2 R N_e of the domains still iterating
        F = [2 * 24 * 24 * 24 * r["n_elems"] for r in res]
        assert min(F) * (1 - 1e-12) <= s["flops"] <= sum(F) * (1 + 1e-12), (k, s, F)
    # the block PCG runs one Q = K P per iteration, the V-cycle three fused steps
    assert st[dtc.DTC_KSTAT_CHEB]["launches"] >= 2 * st[dtc.DTC_KSTAT_MV]["launches"]


def test_graph_replay_equals_launch_by_launch(monkeypatch):
    """The block PCG replayed as a captured HIP graph with the stop decided one
    replay late (the default) gives, bit for bit, the results of launching every
    kernel from the host with a synchronous check (DTC_GRAPH=0): a replay after
    every domain finished changes nothing, and each domain's iteration count is
    its own."""
    dtc = _dtc()
    doms = [_binary(dtc, 16 + 4 * i, 80 + i) for i in range(3)] + [_grey(dtc, 20, 83)]
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    monkeypatch.setenv("DTC_GRAPH", "0")
    with dtc.Batch(doms, **kw) as b:
        r0 = b.run()
        g0 = b.grad(1, np.eye(6))
    monkeypatch.delenv("DTC_GRAPH")
    with dtc.Batch(doms, **kw) as b:
        r1 = b.run()
        r2 = b.run()                                  # captured again for the rebuilt multigrid levels
        g1 = b.grad(1, np.eye(6))
    for a, c, e in zip(r0, r1, r2):
        assert a["status"] == c["status"] == e["status"] == 0
        assert a["iterations"] == c["iterations"] == e["iterations"]
        assert a["outer_iterations"] == c["outer_iterations"]
        for k in ("C", "K_macro", "ff"):
            assert np.array_equal(a[k], c[k]) and np.array_equal(a[k], e[k]), k
    assert np.array_equal(g0, g1)


@pytest.mark.parametrize("terms", ["1", "3"])
def test_mfma_element_products_match_the_walsh_kernel(monkeypatch, terms):
    """The FP32 matvecs with the element products on the matrix cores (FP16 pairs,
    DTC_MFMA=1) and the Walsh-basis VALU kernel (the default) solve to the same
    tensors, to what a relative residual of 1e-8 fixes, binary and greyscale, with
    1 or 3 FP16 terms in the multigrid smoother.  With 3 terms the iteration counts
    stay close; one FP16 term (2^-11) in the smoother costs iterations (measured
    35 -> 49 on the first domain, 101 -> 296 on the bench batch), which is why the
    Walsh kernel stays the default."""
    dtc = _dtc()
    doms = [_binary(dtc, 24, 90), _binary(dtc, 30, 91), _grey(dtc, 20, 92)]
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    monkeypatch.setenv("DTC_MFMA", "0")
    rw = dtc.solve_batch(doms, **kw)
    monkeypatch.setenv("DTC_MFMA", "1")
    monkeypatch.setenv("DTC_MG_TERMS", terms)
    rm = dtc.solve_batch(doms, **kw)
    for a, c in zip(rw, rm):
        assert a["status"] == c["status"] == 0 and c["relres"] <= 1e-8
        assert _rel(c["C"], a["C"]) < 1e-7 and _rel(c["K_macro"], a["K_macro"]) < 1e-7
        if terms == "3":
            assert c["iterations"] <= 1.25 * a["iterations"] + 4, (a["iterations"], c["iterations"])


def test_sparse_brick_kernel_equals_dense(monkeypatch):
    """DTC_SPARSE=1 runs bricks with at most 384 (owned node, element corner) pairs
    through the 256-thread matvec with compact corner slots; its corner sums skip
    the zero terms of the dense kernel and the compiler contracts its products
    differently, so the two solves are two different FP32 roundings of the same
    iteration: the tensors agree to what the relative residual of 1e-8 fixes
    (5e-7 of the largest entry; measured up to 1.3e-7 on K_macro), iteration counts
    within 5 % (the hundreds of FP32 Jacobi-PCG iterations drift apart by a few):
    binary and greyscale domains, multigrid block PCG and Jacobi PCG."""
    dtc = _dtc()
    doms = [_binary(dtc, 24, 95), _binary(dtc, 28, 96), _grey(dtc, 20, 97)]
    for pc in (dtc.DTC_PC_MG, dtc.DTC_PC_JACOBI):
        kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=pc)
        monkeypatch.setenv("DTC_SPARSE", "0")
        rd = dtc.solve_batch(doms, **kw)
        monkeypatch.setenv("DTC_SPARSE", "1")
        rs = dtc.solve_batch(doms, **kw)
        for a, c in zip(rd, rs):
            assert a["status"] == c["status"] == 0
            assert abs(a["iterations"] - c["iterations"]) <= 2 + 0.05 * a["iterations"]
            for k in ("C", "K_macro", "ff"):
                ref = np.abs(a[k]).max()
                assert np.abs(a[k] - c[k]).max() <= 5e-7 * ref, (pc, k)


def test_whole_segment_output_stores_equal_per_chunk_stores(monkeypatch):
    """The dense FP32 Chebyshev step keeps each item's outputs of all load-case
    chunks in registers and stores them after the last chunk; the values are the
    same as with one store per chunk (DTC_MV_DEFER=0), so the solves agree bit for
    bit (multigrid block PCG and Jacobi PCG, fresh handles with the same launch
    shapes)."""
    dtc = _dtc()
    doms = [_binary(dtc, 24, 71), _grey(dtc, 20, 72)]
    for pc in (dtc.DTC_PC_MG, dtc.DTC_PC_JACOBI):
        kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=pc)
        monkeypatch.setenv("DTC_MV_DEFER", "0")
        ra = dtc.solve_batch(doms, **kw)
        monkeypatch.delenv("DTC_MV_DEFER")
        rb = dtc.solve_batch(doms, **kw)
        for a, c in zip(ra, rb):
            assert a["status"] == c["status"] == 0
            assert a["iterations"] == c["iterations"]
            for k in ("C", "K_macro", "ff"):
                assert np.array_equal(a[k], c[k]), (pc, k)


def test_lds_staged_gram_sums_equal_the_global_form(monkeypatch):
    """The block PCG's Gram sums staged through LDS add the same rows in the same
    order per wave as the form that reads every operand from global memory
    (DTC_GRAM_LDS=0): the solves agree bit for bit, on FP32 and FP64-vector blocks."""
    dtc = _dtc()
    for doms in ([_binary(dtc, 24, 81), _binary(dtc, 20, 82)], [_grey(dtc, 20, 83)]):
        kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
        monkeypatch.setenv("DTC_GRAM_LDS", "0")
        ra = dtc.solve_batch(doms, **kw)
        monkeypatch.delenv("DTC_GRAM_LDS")
        rb = dtc.solve_batch(doms, **kw)
        for a, c in zip(ra, rb):
            assert a["status"] == c["status"] == 0
            assert a["iterations"] == c["iterations"]
            for k in ("C", "K_macro", "ff"):
                assert np.array_equal(a[k], c[k]), k


def test_fused_zero_start_smoothing_equals_the_stored_form(monkeypatch):
    """The first Chebyshev step of each pre-smoothing pass (from x = 0: x1 = c D^-1 b)
    is formed inside the next step's staging instead of being stored by a vector
    kernel; the values are the same products, so the solves agree bit for bit with
    DTC_CHEB_FUSE=0 (binary and greyscale batches)."""
    dtc = _dtc()
    for doms in ([_binary(dtc, 24, 91), _binary(dtc, 28, 92)], [_grey(dtc, 20, 93)]):
        kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
        monkeypatch.setenv("DTC_CHEB_FUSE", "0")
        ra = dtc.solve_batch(doms, **kw)
        monkeypatch.delenv("DTC_CHEB_FUSE")
        rb = dtc.solve_batch(doms, **kw)
        for a, c in zip(ra, rb):
            assert a["status"] == c["status"] == 0
            assert a["iterations"] == c["iterations"]
            for k in ("C", "K_macro", "ff"):
                assert np.array_equal(a[k], c[k]), k



def test_concurrent_sub_batches_equal_sequential(monkeypatch):
    """A batch that the solver-node cap splits into several sub-batches solves them
    several at a time (the default), each on its own stream after all are set up and
    tuned one after another; every sub-batch's arithmetic is that of the sequential
    run (the launch shapes only regroup work), so the results agree bit for bit."""
    dtc = _dtc()
    doms = [_binary(dtc, 16 + 2 * i, 110 + i) for i in range(5)] + [_grey(dtc, 18, 115)]
    kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
    monkeypatch.setenv("DTC_SUBBATCH_NODES", "9000")
    out = []
    for conc in ("1", "3"):
        monkeypatch.setenv("DTC_CONCURRENT", conc)
        out.append(dtc.solve_batch(doms, **kw))
    for a, c in zip(*out):
        assert a["status"] == c["status"] == 0
        assert a["iterations"] == c["iterations"] and a["outer_iterations"] == c["outer_iterations"]
        for k in ("C", "K_macro", "ff"):
            assert np.array_equal(a[k], c[k]), k


def test_fused_orthonormalisation_gram_agrees_with_the_separate_kernel(monkeypatch):
    """The Gram of the new direction block formed in the direction kernel (default)
    adds the same products as the separate Gram kernel (DTC_ORTH_FUSE=0) in another
    order, so the two solves are two roundings of one iteration: the tensors agree
    to what rel-res 1e-8 fixes and the iteration counts within 2 (binary batch on FP32
    vectors, greyscale batch on FP64 vectors)."""
    dtc = _dtc()
    for doms in ([_binary(dtc, 24, 121), _binary(dtc, 28, 122)], [_grey(dtc, 20, 123)]):
        kw = dict(precision=1, rel_tol=1e-8, rotation=False, precond=dtc.DTC_PC_MG)
        monkeypatch.setenv("DTC_ORTH_FUSE", "0")
        ra = dtc.solve_batch(doms, **kw)
        monkeypatch.delenv("DTC_ORTH_FUSE")
        rb = dtc.solve_batch(doms, **kw)
        for a, c in zip(ra, rb):
            assert a["status"] == c["status"] == 0
            assert abs(a["iterations"] - c["iterations"]) <= 2
            for k in ("C", "K_macro", "ff"):
                assert np.abs(a[k] - c[k]).max() <= 1e-7 * np.abs(a[k]).max(), k
