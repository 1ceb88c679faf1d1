"""The Fortran drop-in driver (fortran/dtc_gpu_driver.f90) end to end on the GPU:
.meta + raw image -> domain decomposition (f-src/mod_decomp.f90:247-326) ->
batches through the iso_c_binding module -> the puredat tree of the reference's
output (f-src_main/struct_process.f90:393-414,700-741) with its typed streams
and the .status restart tracker, checked against the same domains solved
through the Python binding; the reference's own TC01-1 regression inputs
(isotropic 100^3 image, HEX08 macro ME1 and HEX20 macro ME2) against the
isotropic closed form; and the crawler's tables from those streams."""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import iso_voigt, read_puredat_head, read_puredat_leaf

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "m-dtc-direct-tensor-computation_amd", "lib")
DRIVER = os.path.join(LIB, "dtc_gpu_driver")
CRAWLER = os.path.join(LIB, "dtc_crawl_tensors")


def _run(*args, rc=0):
    p = subprocess.run(list(args), capture_output=True, text=True, timeout=600)
    assert p.returncode == rc, p.stdout + p.stderr
    return p.stdout


def _leaves(stem):
    tree = read_puredat_head(stem + ".head")
    return tree, {l["desc"]: l for b in tree for l in b["leaves"]}


def _write_meta(path, meta):
    with open(path, "w") as f:
        for k, v in meta.items():
            f.write(f"r {k} {v}\n")


def _check_tree(tree, nd, nlc):
    """Structure and bounds: root 'results' with 3 branches; leaves numbered per
    stream in tree order (f-src/mod_puredat.f90:2007-2035); stream sizes."""
    assert [b["desc"] for b in tree] == ["results", "Input parameters", "Global domain decomposition",
                                         "Averaged Material Properties"]
    assert tree[0]["no_branches"] == 3 and tree[0]["no_leaves"] == 0 and tree[0]["streams"] is not None
    assert [len(b["leaves"]) for b in tree[1:]] == [18, 8, 24]
    assert all(b["no_leaves"] == len(b["leaves"]) and b["streams"] is None for b in tree[1:])
    nxt = {}
    for b in tree[1:]:
        for l in b["leaves"]:
            assert l["lb"] == nxt.get(l["type"], 1) and l["ub"] == l["lb"] + l["no"] - 1, l
            nxt[l["type"]] = l["ub"] + 1
    st = tree[0]["streams"]
    assert st["size_int8_stream"] == nxt[4] - 1 and st["size_real_stream"] == nxt[5] - 1
    assert st["size_char_stream"] == nxt[6] - 1 and st["size_int4_stream"] == 0
    res = {l["desc"]: l for l in tree[3]["leaves"]}
    assert res["Effective numerical stiffness"]["no"] == 2 * nd * nlc * nlc      # domains_per_comm = 2 nd
    assert res["Effective stiffness"]["no"] == 2 * nd * 36


@pytest.mark.parametrize("macro_order", [1, 2])
def test_fortran_driver_writes_the_puredat_tree(tmp_path, macro_order):
    assert os.path.exists(DRIVER), f"{DRIVER} is not built (run __graft_entry__.build())"
    import dtc_amd
    from dtc_amd import synth
    xd, vd = 10, 22                                   # x_D = SIZE_DOMAIN / SPACING, 2 x 2 x 2 domains
    vol = np.zeros((vd, vd, vd), np.int8)             # (z, y, x), x fastest on disk
    vol[1:21, 1:21, 1:21] = synth.trabecular(20, seed=5, sigma=1.5, bvtv=0.45)
    base = str(tmp_path / "vol")
    vol.tofile(base + ".raw")
    _write_meta(base + ".meta", {
        "TYPE_RAW": "ik1", "DIMENSIONS": f"{vd} {vd} {vd}", "SPACING": "0.01 0.01 0.01", "SIZE_DOMAIN": "0.1 0.1 0.1",
        "LO_BNDS_DMN_RANGE": "0 0 0", "UP_BNDS_DMN_RANGE": "1 0 1", "BINARIZE_LO": "1", "RVE_STRAIN": "1.e-6",
        "YOUNG_MODULUS": "5600.", "POISSON_RATIO": "0.3", "MACRO_ELMNT_ORDER": str(macro_order),
        "MICRO_ELMNT_TYPE": "HEX08"})
    env = dict(os.environ, DTC_DRIVER_GROUP="3")      # two groups: 3 + 1 domains
    p = subprocess.run([DRIVER, base + ".meta", "1"], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr

    stem = base + "_results"
    tree, L = _leaves(stem)
    nlc = 24 if macro_order == 1 else 60
    _check_tree(tree, 4, nlc)
    get = lambda name: read_puredat_leaf(stem, L[name])
    doms = [0, 1, 4, 5]                               # i + 2 j + 4 k for i, k in 0..1, j = 0
    assert list(get("Domain number")[:4]) == doms and not get("Domain number")[4:].any()
    assert list(np.fromfile(base + ".status", np.int64)) == doms
    assert get("muCT puredat pro_name").rstrip() == "vol" and list(get("Number of voxels per direction")) == [vd] * 3
    assert list(get("x_D")) == [xd] * 3 and list(get("nn_D")) == [2] * 3 and list(get("bpoints")) == [1] * 3
    assert np.allclose(get("Grid spacings"), 0.01) and get("Element type  on micro scale").rstrip() == "HEX08"
    C = get("Effective stiffness")[:4 * 36].reshape(4, 6, 6).transpose(0, 2, 1)
    rho = get("Effective density")
    for n, dom in enumerate(doms):
        i, k = dom % 2, dom // 4
        xa = (xd * i + 2, 2, xd * k + 2)              # 1-based first voxel (x, y, z)
        blk = vol[xa[2] - 1:xa[2] - 1 + xd, xa[1] - 1:xa[1] - 1 + xd, xa[0] - 1:xa[0] - 1 + xd]
        d = dtc_amd.Domain(np.ascontiguousarray(blk), delta=(0.01,) * 3, xa_n=xa, macro_order=macro_order)
        r = dtc_amd.solve_domain(d, precision=1, rel_tol=1e-8, rotation=False)
        assert r["status"] == 0
        assert np.abs(C[n] - r["C"]).max() <= 1e-6 * np.abs(r["C"]).max()
        assert get("Number of Elements")[n] == r["n_elems"]
        assert rho[n] == r["density"]
        # symmetry deviation of C (leaf 13): max |C - C^T| / max |C|
        sd = get("Symmetry deviation - effective stiffness")[n]
        assert abs(sd - np.abs(C[n] - C[n].T).max() / np.abs(C[n]).max()) < 1e-15


def test_fortran_driver_restarts_from_the_status_file(tmp_path):
    from dtc_amd import synth
    vd = 22
    vol = np.zeros((vd, vd, vd), np.int8)
    vol[1:21, 1:21, 1:21] = synth.trabecular(20, seed=6, sigma=1.5, bvtv=0.45)
    base = str(tmp_path / "rs")
    vol.tofile(base + ".raw")
    meta = {"TYPE_RAW": "ik1", "DIMENSIONS": f"{vd} {vd} {vd}", "SPACING": "0.01 0.01 0.01",
            "SIZE_DOMAIN": "0.1 0.1 0.1", "LO_BNDS_DMN_RANGE": "0 0 0", "UP_BNDS_DMN_RANGE": "1 1 0",
            "BINARIZE_LO": "1", "YOUNG_MODULUS": "5600.", "POISSON_RATIO": "0.3", "MACRO_ELMNT_ORDER": "1"}
    _write_meta(base + ".meta", meta)
    out = _run(DRIVER, base + ".meta", "1")
    assert "0 already computed" in out
    stem = base + "_results"
    _, L = _leaves(stem)
    C0 = read_puredat_leaf(stem, L["Effective stiffness"]).copy()
    # finished job: nothing left to compute, the streams stay as they are
    out = _run(DRIVER, base + ".meta", "1")
    assert "4 already computed" in out and "status" not in out
    assert np.array_equal(read_puredat_leaf(stem, L["Effective stiffness"]), C0)
    # interrupted job: domain 2 (slot 3) not in the tracker and its tensor lost -> only it is recomputed
    st = np.fromfile(base + ".status", np.int64)
    st[2] = -999
    st.tofile(base + ".status")
    with open(stem + ".real8.st", "r+b") as f:
        f.seek(8 * (L["Effective stiffness"]["lb"] - 1 + 2 * 36))
        f.write(np.zeros(36).tobytes())
    out = _run(DRIVER, base + ".meta", "1")
    assert "3 already computed" in out and out.count("domain ") == 1
    C1 = read_puredat_leaf(stem, L["Effective stiffness"])
    assert np.array_equal(np.delete(C1, range(72, 108)), np.delete(C0, range(72, 108)))
    # solved alone instead of in a batch of 4: equal to what rel-res 1e-8 fixes
    assert np.abs(C1[72:108] - C0[72:108]).max() < 1e-7 * np.abs(C0[72:108]).max()
    assert list(np.fromfile(base + ".status", np.int64)) == [0, 1, 2, 3]
    # a tracker of another size (the domain range changed): refused
    meta["UP_BNDS_DMN_RANGE"] = "1 1 1"
    _write_meta(base + ".meta", meta)
    out = _run(DRIVER, base + ".meta", "1", rc=2)
    assert "do not match" in out


def _tc01(tmp_path, variant):
    """The reference's TC01-1 regression case: its .meta keywords
    (tests/golden/regression_meta.json) and the isotropic 100^3 ik4 image of
    regression/f-src/create_testcase.f90 (every voxel 14250)."""
    metas = json.load(open(os.path.join(HERE, "golden", "regression_meta.json")))
    m = dict(metas[variant])
    base = str(tmp_path / variant)
    np.full((100, 100, 100), 14250, np.int32).tofile(base + ".raw")
    _write_meta(base + ".meta", m)
    return base, m


@pytest.mark.parametrize("variant", ["isotropic-std-HEX08-ME1", "isotropic-std-HEX08-ME2"])
def test_tc01_isotropic_regression_case(tmp_path, variant):
    """FP64 solve of the one 50^3 solid domain: C is the isotropic Voigt stiffness
    (E 5600, nu 0.3) to 1e-6, density 1, every leaf in place; the crawler writes
    the domain's tensor into .covo and its numerical stiffness into .covo.num."""
    base, m = _tc01(tmp_path, variant)
    _run(DRIVER, base + ".meta", "0")
    stem = base + "_results"
    tree, L = _leaves(stem)
    nlc = 24 if m["MACRO_ELMNT_ORDER"] == "1" else 60
    _check_tree(tree, 1, nlc)
    get = lambda name: read_puredat_leaf(stem, L[name])
    assert get("Domain number")[0] == 0 and get("Number of Elements")[0] == 50 ** 3
    assert get("Number of Nodes")[0] == 51 ** 3 and get("Effective density")[0] == 1.0
    C = get("Effective stiffness")[:36].reshape(6, 6).T
    ref = iso_voigt(5600.0, 0.3)
    assert np.abs(C - ref).max() < 1e-6 * np.abs(ref).max()
    KM = get("Effective numerical stiffness")[:nlc * nlc].reshape(nlc, nlc).T
    assert np.abs(KM - KM.T).max() < 1e-6 * np.abs(KM).max()
    assert list(np.fromfile(base + ".status", np.int64)) == [0]
    # crawler tables (f-src_main/mrd_crawl_tensors.f90)
    out = _run(CRAWLER, base + ".meta")
    assert "crawled 1 domains" in out
    for suf, leaf in (("covo", "Effective stiffness"), ("cr1", "Optimized Effective stiffness CR_1"),
                      ("cr2", "Optimized Effective stiffness CR_2")):
        lines = open(f"{base}.{suf}").read().splitlines()
        head, row = lines[0].split(","), lines[1].split(",")
        assert len(lines) == 2 and len(head) == len(row) == 23 + 24 + 36
        assert int(row[0]) == 0 and row[head.index("Micro element")] == "HEX08"
        t = np.array(row[-36:], float).reshape(6, 6)
        T = get(leaf)[:36].reshape(6, 6)            # read row index outer, as the reference's crawler
        assert np.array_equal(t, T)
    num = open(base + ".covo.num").read().splitlines()
    assert len(num) == 2 and len(num[1].split(",")) == 1 + nlc * nlc
    assert np.array_equal(np.array(num[1].split(",")[1:], float), get("Effective numerical stiffness")[:nlc * nlc])
