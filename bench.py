"""Throughput of the per-control-volume micro-FE path on one MI355X.

Workload (the metric's configuration): synthetic trabecular control volumes
of 64^3 voxels (seeded Gaussian-smoothed noise, sigma = 2 voxels, threshold
at BV/TV = 0.20, largest 26-connected component), HEX08 micro elements, HEX08
macro element with 24 load cases, E = 5600 MPa, nu = 0.3.  One step = one
pass of the whole path over a batch of such domains resident in HBM: mesh ->
boundary conditions -> 24 load-case solves to a true relative residual of
1e-8 -> force projection and effective stiffness -> CR_1/CR_2 rotation search.

value  = domains per second of the batched FP32 + FP64-refine solve (block PCG:
the 24 load cases of a domain share one Krylov block, multigrid V-cycle
preconditioner).
Also reported: the plain FP64 GPU form (one domain per launch sequence), the
tensor error of the mixed solve against it, the EbE matvec roofline and the
CPU FP64 restatement (the oracle) timed on a bounded sample on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--n 64]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "m-dtc-direct-tensor-computation_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
MFMA_PEAK_TFLOPS = {"f32": 157.3, "f64": 78.6}   # MI355X dense matrix-core peaks
# the timed kernels (fine-level launches): kernel_stats name -> kernel
KERNEL_NAMES = {"mv": "k_matvec<float, IN_V, OUT_MASKED> (Q = K P of the block PCG)",
                "cheb": "k_matvec<float, IN_V|IN_JS, OUT_CHEB> (Chebyshev-Jacobi smoothing step)",
                "bres": "k_matvec<float, IN_V, OUT_BRES> (b - K x before restriction)",
                "resid": "k_matvec<double, IN_U, OUT_RESID> (FP64 true residual)"}


def hbm_copy_probe(gib=1.0, iters=10):
    """Device-to-device copy of a gib-GiB buffer on device 0 through the HIP runtime the
    library uses (hipMemcpyAsync, timed with HIP events): the bandwidth an HBM-bound
    kernel can reach here (read + write bytes per second, GB/s)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    n = int(gib * (1 << 30))
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what}: hip error {rc}")
    ok(hip.hipMalloc(C.byref(a), C.c_size_t(n)), "hipMalloc")
    try:
        ok(hip.hipMalloc(C.byref(b), C.c_size_t(n)), "hipMalloc")
        ok(hip.hipMemset(a, 0, C.c_size_t(n)), "hipMemset")
        ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
        ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
        d2d = 3   # hipMemcpyDeviceToDevice
        for _ in range(2):
            ok(hip.hipMemcpyAsync(b, a, C.c_size_t(n), d2d, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e0, None), "hipEventRecord")
        for _ in range(iters):
            ok(hip.hipMemcpyAsync(b, a, C.c_size_t(n), d2d, None), "hipMemcpyAsync")
        ok(hip.hipEventRecord(e1, None), "hipEventRecord")
        ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
        ms = C.c_float()
        ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
        return 2.0 * n * iters / (ms.value * 1e-3) / 1e9
    finally:
        for e in (e0, e1):
            if e.value:
                hip.hipEventDestroy(e)
        for p in (a, b):
            if p.value:
                hip.hipFree(p)


def _sync():
    """Device synchronisation through the HIP runtime the library uses (the library's
    own calls return after their stream work; this brackets the timed regions)."""
    import ctypes as C
    rc = C.CDLL("libamdhip64.so.7").hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize: hip error {rc}")


def make_domains(count, n, seed0):
    import dtc_amd
    from dtc_amd import synth
    return [dtc_amd.Domain(synth.trabecular(n, seed=seed0 + i, sigma=2.0, bvtv=0.20), llimit=1,
                           delta=(0.01, 0.01, 0.01), xa_n=(2, 2, 2), E=5600.0, nu=0.3) for i in range(count)]


def _cpu_sample(job):
    """One core: the oracle (optimised build) on its own domain, bounded sample."""
    seed, n, iters_per_lc, sample_iters = job
    os.environ["DTC_ORACLE_FAST"] = "1"
    from oracle import oracle as O
    from dtc_amd import synth
    import dtc_amd
    dom = dtc_amd.Domain(synth.trabecular(n, seed=seed, sigma=2.0, bvtv=0.20), llimit=1, delta=(0.01, 0.01, 0.01),
                         xa_n=(2, 2, 2), E=5600.0, nu=0.3)
    t0 = time.perf_counter()
    nodes, elems = O.mesh(dom.phi, dom.llimit, 8, dom.delta, dom.xa_n)
    t_mesh = time.perf_counter() - t0
    mn, mx, dim = dom.box()
    t0 = time.perf_counter()
    r = O.solve_domain(nodes, elems, dom.E, dom.nu, dom.a(), mn, mx, dim, tol=1e-30, max_it=sample_iters,
                       lc_first=0, lc_count=1)
    t_solve = time.perf_counter() - t0
    it = int(r["iters"][0])
    t0 = time.perf_counter()
    O.rotation_search(np.diag([3000.0, 3100, 3300, 900, 950, 1000]) + 50.0, 1)
    t_rot = time.perf_counter() - t0
    per_iter = t_solve / max(it, 1)
    return t_mesh, per_iter, t_rot, it


def cpu_baseline(n, gpu_fp64_iters_per_lc, sample_iters, workers, log):
    """The oracle (assembled 3x3-block CSR, Jacobi-PCG per load case, FP64; the
    optimised build of oracle/Makefile) with one domain per host core, all cores
    at once (the reference runs one domain per MPI worker), on a bounded sample:
    each core meshes its domain, runs load case 1 for `sample_iters` PCG
    iterations and one CR_1 search; a domain's time is extrapolated with the
    FP64 iteration count per load case x 24 load cases + CR_1 and CR_2, and the
    host's rate is the sum of the per-core rates."""
    import multiprocessing as mp
    import platform
    t0 = time.perf_counter()
    jobs = [(2 + i % 16, n, gpu_fp64_iters_per_lc, sample_iters) for i in range(workers)]
    with mp.get_context("spawn").Pool(workers) as pool:
        out = pool.map(_cpu_sample, jobs)
    wall = time.perf_counter() - t0
    t_dom = [tm + pi * gpu_fp64_iters_per_lc * 24 + 2 * tr for tm, pi, tr, _ in out]
    rate = sum(1.0 / t for t in t_dom)
    per_iter = float(np.mean([o[1] for o in out]))
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        cpu = platform.processor() or cpu
    log(f"cpu sample: {workers} cores ({cpu}), {out[0][3]} PCG iterations each, {per_iter * 1e3:.2f} ms per "
        f"iteration, {np.mean(t_dom):.1f} s per domain extrapolated, {wall:.1f} s wall")
    return {"value": rate, "unit": "domains/s", "cores": workers, "kind": "port", "cpu_model": cpu,
            "host_cpus": os.cpu_count(),
            "sample": f"oracle (gcc -O3 -march=x86-64-v3) on {workers} cores at once, one 64^3 domain each (seeds "
                      f"2..): mesh, load case 1 for {out[0][3]} Jacobi-PCG iterations ({per_iter * 1e3:.2f} ms "
                      f"each on average), one CR_1 search; each domain extrapolated to 24 load cases x "
                      f"{gpu_fp64_iters_per_lc:.0f} iterations (the FP64 GPU iteration count) + CR_1 and CR_2; "
                      f"{wall:.1f} s of sampling"}


def uneven(args, log):
    """Uneven synthetic domains solved in runs of 128 (each one dtc_batch_run, split
    into sub-batches by the library), domains/s over all runs."""
    import dtc_amd
    from dtc_amd import synth
    t0 = time.perf_counter()
    phis = synth.uneven_batch(args.uneven, seed=3)
    log(f"generated {len(phis)} domains in {time.perf_counter() - t0:.1f} s; "
        f"{sum(p.size for p in phis) / 1e6:.1f} M voxels")
    doms = [dtc_amd.Domain(p, llimit=1, delta=(0.01, 0.01, 0.01), xa_n=(2, 2, 2)) for p in phis]
    res, dt, G = [], 0.0, 128      # runs of 128 domains (each split into sub-batches by the library)
    for g0 in range(0, len(doms), G):
        with dtc_amd.Batch(doms[g0:g0 + G], precision=dtc_amd.DTC_FP32_REFINE, rel_tol=1e-8,
                           inner_tol=args.inner_tol, rotation=True,
                           rot_step_deg=5.0, precond=dtc_amd.DTC_PC_MG, max_iter=args.max_iter) as b:
            _sync()
            t0 = time.perf_counter()
            res += b.run()
            _sync()
            dt += time.perf_counter() - t0
        log(f"domains {g0}..{min(g0 + G, len(doms)) - 1}: {dt:.1f} s so far")
    ok = sum(1 for r in res if r["status"] == 0 and r["relres"] <= 1e-8)
    st = {}
    for i, r in enumerate(res):
        st.setdefault(int(r["status"]), []).append(i)
    log(f"{ok}/{len(res)} converged in {dt:.1f} s; domains by status: " +
        ", ".join(f"{k}: {len(v)}" + (f" {v[:8]}" if k else "") for k, v in sorted(st.items())))
    print(json.dumps({"metric": "converged uneven domains/sec @rel-res 1e-8 (edges 32..96, BV/TV 0.08..0.35)",
                      "value": round(ok / dt, 3), "unit": "domains/s", "n_gpus": 1, "domains": len(res),
                      "converged": ok, "seconds": round(dt, 2), "voxels": int(sum(p.size for p in phis)),
                      "pcg_iters_mean": float(np.mean([r["iterations"] for r in res])),
                      "data": "synthetic", "dtype": "f32+f64"}))


def config4(args, log):
    """Configuration 4: one 96^3 greyscale domain (uint8 values of the smoothed field,
    void 0, E = 10000 (g/255)^2 MPa), HEX20 micro + HEX20 macro elements (60 load cases),
    FP32 + FP64 refinement to 1e-8 with the p-multigrid block PCG: time, iterations and
    the properties of the 60 x 60 macro stiffness."""
    import dtc_amd
    from dtc_amd import synth
    n = args.n if args.n != 64 else 96
    grey, modulus = synth.greyscale(n, seed=4, bvtv=0.20)
    d = dtc_amd.Domain(grey, llimit=1, delta=(0.01,) * 3, modulus=modulus, E=10000.0, micro_nen=20, macro_order=2)
    with dtc_amd.Batch([d], precision=dtc_amd.DTC_FP32_REFINE, rel_tol=1e-8, rotation=True,
                       precond=dtc_amd.DTC_PC_MG, max_iter=args.max_iter) as b:
        _sync()
        t0 = time.perf_counter()
        r = b.run()[0]
        _sync()
        dt = time.perf_counter() - t0
    K = r["K_macro"]
    s = np.abs(K).max()
    null = max(np.abs(K @ np.repeat(np.eye(3), 20, axis=1)[k]).max() for k in range(3)) / s
    log(f"config 4: {n}^3 greyscale HEX20/HEX20: status {r['status']}, {r['n_elems']} elements, {r['n_nodes']} "
        f"nodes, {r['iterations']} block iterations, {r['outer_iterations']} FP64 replacements, relres "
        f"{r['relres']:.2e}, {dt:.2f} s")
    print(json.dumps({"metric": "config 4: 96^3 greyscale HEX20 micro + HEX20 macro, rel-res 1e-8", "value": round(dt, 3),
                      "unit": "s per domain", "higher_is_better": False, "n_gpus": 1, "status": r["status"],
                      "relres": r["relres"], "iterations": r["iterations"], "outer": r["outer_iterations"],
                      "n_elems": r["n_elems"], "n_nodes": r["n_nodes"],
                      "KM_symmetry": float(np.abs(K - K.T).max() / s), "KM_translation_residual": float(null),
                      "data": "synthetic", "dtype": "f32+f64"}))


def config5(args, log):
    """Configuration 5: 128^3 greyscale domain (HEX08 micro + macro, 24 load cases),
    loss = ||C - C_target||_F^2 (C_target a seeded orthotropic SPD tensor); the
    hand-written sensitivity (no second solve) against central differences of full
    solves at relative step 1e-4 on `--fd-voxels` seeded voxels among the stiffest
    tenth (the perturbed domains solved as one batch to rel-res 1e-11), and the
    gradient of the 1e-8 solve against the one of the FP64-refined 1e-11 solve.
    With `--fd-dirs K` the `--fd-voxels` voxels are perturbed together along K seeded
    sign vectors s (each voxel by s_k 1e-4 E_k): one central difference per direction
    checks g . (s h) over all of them (2 K solves instead of 2 per voxel)."""
    import dtc_amd
    from dtc_amd import synth
    n = args.n if args.n != 64 else 128
    rng = np.random.default_rng(55)
    C_t = synth.orthotropic_C(rng)
    grey, modulus = synth.greyscale(n, seed=5, bvtv=0.20)
    dom = lambda m: dtc_amd.Domain(grey, llimit=1, delta=(0.01,) * 3, modulus=m, E=10000.0)
    loss = lambda C: float(np.sum((C - C_t) ** 2))
    out = {}
    for tol in (1e-8, 1e-11):
        with dtc_amd.Batch([dom(modulus)], precision=dtc_amd.DTC_FP32_REFINE, rel_tol=tol, rotation=False,
                           precond=dtc_amd.DTC_PC_MG, max_iter=args.max_iter) as b:
            _sync()
            t0 = time.perf_counter()
            r = b.run()[0]
            _sync()
            t_f = time.perf_counter() - t0
            t0 = time.perf_counter()
            g = b.grad(0, 2.0 * (r["C"] - C_t))
            _sync()
            t_g = time.perf_counter() - t0
        out[tol] = (r, g, t_f, t_g)
        log(f"config 5 at rel-res {tol:g}: {r['iterations']} block iterations, forward {t_f:.2f} s, gradient "
            f"{t_g:.3f} s")
    g8, g11 = out[1e-8][1], out[1e-11][1]
    gerr = float(np.abs(g8 - g11).max() / np.abs(g11).max())
    solid = np.argwhere(grey.ravel() > 0).ravel()
    stiff = solid[modulus.ravel()[solid] >= np.quantile(modulus.ravel()[solid], 0.9)]
    vox = rng.choice(stiff, args.fd_voxels, replace=False)
    hv = 1e-4 * modulus.ravel()[vox]
    gk = g11.ravel()[vox]
    doms = []
    if args.fd_dirs > 0:
        # simultaneous perturbation: direction j moves voxel k by S[j, k] hv[k]
        S = rng.choice([-1.0, 1.0], size=(args.fd_dirs, len(vox)))
        for j in range(args.fd_dirs):
            for sgn in (1, -1):
                m2 = modulus.copy().ravel()
                m2[vox] += sgn * S[j] * hv
                doms.append(dom(m2.reshape(grey.shape)))
    else:
        for i in range(len(vox)):
            for sgn in (1, -1):
                m2 = modulus.copy().ravel()
                m2[vox[i]] += sgn * hv[i]
                doms.append(dom(m2.reshape(grey.shape)))
    t0 = time.perf_counter()
    rf = dtc_amd.solve_batch(doms, precision=dtc_amd.DTC_FP32_REFINE, rel_tol=1e-11, rotation=False,
                             precond=dtc_amd.DTC_PC_MG, max_iter=args.max_iter)
    t_fd = time.perf_counter() - t0
    if args.fd_dirs > 0:
        # (L(E + s h) - L(E - s h)) / 2 against g . (s h); scale: sum |g_k| h_k
        fd = np.array([(loss(rf[2 * j]["C"]) - loss(rf[2 * j + 1]["C"])) / 2.0 for j in range(args.fd_dirs)])
        pred = S @ (gk * hv)
        fd_err = float(np.abs(fd - pred).max() / np.sum(np.abs(gk) * hv))
        what = f"{args.fd_dirs} sign directions over {len(vox)} voxels"
    else:
        fd = np.array([(loss(rf[2 * i]["C"]) - loss(rf[2 * i + 1]["C"])) / (2 * hv[i]) for i in range(len(vox))])
        fd_err = float(np.abs(fd - gk).max() / np.abs(g11).max())
        what = f"{len(vox)} voxels"
    log(f"config 5: {what} by central differences ({len(doms)} solves, {t_fd:.1f} s): relative error "
        f"{fd_err:.2e}; gradient 1e-8 vs 1e-11: {gerr:.2e}")
    r8, _, tf8, tg8 = out[1e-8]
    print(json.dumps({"metric": "config 5: 128^3 greyscale gradient of ||C - C_target||_F^2", "n_gpus": 1,
                      "forward_s": round(tf8, 3), "gradient_s": round(tg8, 4), "grad_over_forward": round(tg8 / tf8, 4),
                      "iterations": r8["iterations"], "fd_voxels": len(vox), "fd_dirs": args.fd_dirs,
                      "fd_rel_err": fd_err,
                      "grad_1e-8_vs_1e-11": gerr, "fd_batch_s": round(t_fd, 2), "data": "synthetic",
                      "dtype": "f32+f64"}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--fp64-domains", type=int, default=2)
    ap.add_argument("--cpu-iters", type=int, default=400)
    ap.add_argument("--cpu-workers", type=int, default=min(16, os.cpu_count() or 1),
                    help="host cores of the CPU baseline (one domain each; 16 = one GPU's share of the host)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--max-iter", type=int, default=200000, help="PCG iteration cap (profiling runs)")
    ap.add_argument("--no-check", action="store_true", help="do not require convergence (profiling runs)")
    ap.add_argument("--precond", choices=["mg", "jacobi"], default="mg",
                    help="preconditioner of the batched mixed-precision PCG")
    ap.add_argument("--mg-degree", type=int, default=0, help="Chebyshev degree per smoothing pass (0 = library default)")
    ap.add_argument("--inner-tol", type=float, default=0.0,
                    help="FP32 inner reduction before an FP64 residual replacement (0 = library default)")
    ap.add_argument("--uneven", type=int, default=0,
                    help="instead: one batch of this many uneven domains (edges 32..96, BV/TV 0.08..0.35, "
                         "anisotropic frames; rotation search on a 5 degree grid)")
    ap.add_argument("--config4", action="store_true",
                    help="instead: configuration 4 (96^3 greyscale, HEX20 micro + HEX20 macro, p-multigrid)")
    ap.add_argument("--config5", action="store_true",
                    help="instead: configuration 5 (128^3 greyscale gradient, central differences, 1e-11 reference)")
    ap.add_argument("--fd-voxels", type=int, default=8, help="config 5: voxels checked by central differences")
    ap.add_argument("--fd-dirs", type=int, default=0,
                    help="config 5: check the voxels along this many seeded sign directions (0: one by one)")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.gpus != 1:
        print(f"bench.py: one GPU is measured; --gpus {args.gpus} is not supported (domains would shard "
              f"across GPUs, one process per GPU, with no exchange during the solves)", file=sys.stderr)
        sys.exit(2)

    import dtc_amd
    if args.uneven > 0:
        uneven(args, log)
        return
    if args.config4:
        config4(args, log)
        return
    if args.config5:
        config5(args, log)
        return
    doms = make_domains(args.batch, args.n, 2)
    log(f"batch of {len(doms)} domains of {args.n}^3, BV/TV " +
        ", ".join(f"{d.phi.mean():.3f}" for d in doms[:4]) + " ...")

    # ---- batched FP32 + FP64 refinement (kernel timing off in the timed steps)
    pc = dtc_amd.DTC_PC_MG if args.precond == "mg" else dtc_amd.DTC_PC_JACOBI
    b = dtc_amd.Batch(doms, precision=dtc_amd.DTC_FP32_REFINE, rel_tol=1e-8, inner_tol=args.inner_tol, rotation=True,
                      max_iter=args.max_iter, precond=pc, mg_degree=args.mg_degree)
    for w in range(args.warmup):
        res = b.run()
        log(f"warmup {w}: {b.stats()['run_ms']:.1f} ms")
    _sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        res = b.run()
        log(f"step {k}: {b.stats()['run_ms']:.1f} ms")
    _sync()
    dt = time.perf_counter() - t0
    bad = [i for i, r in enumerate(res) if r["status"] != 0 or r["relres"] > 1e-8]
    if bad and not args.no_check:
        raise SystemExit(f"domains {bad} did not converge to 1e-8: {[res[i]['status'] for i in bad]}")
    value = len(doms) * args.steps / dt
    mixed_iters = np.mean([r["iterations"] for r in res])
    fallbacks = sum(1 for r in res if r["precision_used"] != dtc_amd.DTC_FP32_REFINE)
    log("PCG iterations per domain: " + " ".join(str(r["iterations"]) for r in res) +
        "; FP64 residual replacements: " + " ".join(str(r["outer_iterations"]) for r in res) +
        f"; domains repeated in FP64: {fallbacks}")

    # ---- kernel timing: one more run (outside the timed steps) with HIP events on the
    # solve stream around the fine-level launches of each timed kernel
    b.timing(True)
    b.run()
    b.timing(False)
    kern = {}
    for k, name in dtc_amd.KSTAT_NAMES.items():
        ks = b.kernel_stats(k)
        if ks["launches"] == 0 or ks["avg_ms"] <= 0:
            continue
        gbs = ks["bytes"] / (ks["avg_ms"] * 1e-3) / 1e9
        tfl = ks["flops"] / (ks["avg_ms"] * 1e-3) / 1e12
        pk = MFMA_PEAK_TFLOPS["f64" if name == "resid" else "f32"]
        kern[name] = {"kernel": KERNEL_NAMES[name], "avg_launch_ms": round(ks["avg_ms"], 5),
                      "launches": ks["launches"], "total_ms": round(ks["avg_ms"] * ks["launches"], 2),
                      "algorithmic_bytes": round(ks["bytes"]), "achieved_GBs": round(gbs, 1),
                      "frac": round(gbs / HBM_PEAK_GBS, 4),
                      "dense_flops": round(ks["flops"]), "dense_TFLOPs": round(tfl, 2),
                      "flop_frac": round(tfl / pk, 4)}
    run_ms = b.stats()["run_ms"]
    for v in kern.values():
        v["share_of_run"] = round(v["total_ms"] / run_ms, 3)
    log("timed kernels: " + "; ".join(f"{k}: {v['avg_launch_ms'] * 1e3:.1f} us x {v['launches']}, "
                                      f"{v['achieved_GBs']:.0f} GB/s, {100 * v['share_of_run']:.1f}% of the run"
                                      for k, v in kern.items()))

    # ---- sensitivity dC/dE of every domain (reuses the stored load-case solutions)
    G = np.eye(6) + 0.1 * np.ones((6, 6))
    _sync()
    g0 = time.perf_counter()
    for i, r in enumerate(res):
        if r["status"] == 0:
            b.grad(i, G)
    _sync()
    t_grad = time.perf_counter() - g0
    log(f"gradient of all {len(doms)} domains: {t_grad * 1e3:.1f} ms ({t_grad / (dt / args.steps):.3f} x one step)")
    b.close()

    # ---- plain FP64 GPU form (Jacobi-PCG in FP64), one domain at a time; beside it
    # the mixed form with the same Jacobi preconditioner and one PCG per load case, so
    # that the gain of the precision alone is separated from the multigrid / block gain
    def one_by_one(prec, precond, count):
        t, Cs, its = 0.0, [], []
        for i in range(min(count, len(doms))):
            with dtc_amd.Batch([doms[i]], precision=prec, rel_tol=1e-8, rotation=True, max_iter=args.max_iter,
                               precond=precond) as b1:
                b1.run()
                _sync()
                s0 = time.perf_counter()
                r1 = b1.run()[0]
                _sync()
                t += time.perf_counter() - s0
            Cs.append(r1["C"])
            its.append(r1["iterations"])
            log(f"{'fp64' if prec == dtc_amd.DTC_FP64 else 'fp32+fp64 jacobi'} domain {i}: {r1['iterations']} "
                f"iterations, relres {r1['relres']:.2e}, {r1['status']}")
        return (len(Cs) / t if Cs else float("nan")), Cs, its
    fp64_rate, C64, it64 = one_by_one(dtc_amd.DTC_FP64, dtc_amd.DTC_PC_JACOBI, args.fp64_domains)
    jac_rate, _, itj = one_by_one(dtc_amd.DTC_FP32_REFINE, dtc_amd.DTC_PC_JACOBI, args.fp64_domains)
    terr = max((np.abs(res[i]["C"] - C64[i]).max() / np.abs(C64[i]).max() for i in range(len(C64))), default=float("nan"))

    # ---- measured copy bandwidth beside the HBM peak
    try:
        probe = round(hbm_copy_probe(), 1)
    except Exception as e:   # reported as missing, never guessed
        log(f"copy probe failed: {e}")
        probe = None

    # ---- roofline of the dominant kernel (largest share of the run)
    dom_k = max(kern, key=lambda k: kern[k]["total_ms"]) if kern else None
    roof = None
    if dom_k:
        kd = kern[dom_k]
        traffic = None
        tj = os.path.join(ROOT, "profiles", f"traffic_{dom_k}.json")
        if os.path.exists(tj):
            t = json.load(open(tj))
            if t.get("batch") == len(doms) and t.get("domain_edge") == args.n:
                traffic = t["traffic_bytes_per_dispatch"]
                kd["traffic"] = traffic
        roof = {"bound": "hbm", "achieved": kd["achieved_GBs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": kd["frac"], "traffic": traffic, "kernel": kd["kernel"],
                "avg_launch_ms": kd["avg_launch_ms"], "launches_timed": kd["launches"],
                "algorithmic_bytes": kd["algorithmic_bytes"], "share_of_run": kd["share_of_run"],
                "copy_probe_GBs": probe,
                "frac_of_copy_probe": round(kd["achieved_GBs"] / probe, 4) if probe else None,
                # the flop side: dense element-product count 2 (3 nen)^2 R N_e per launch
                # against the FP32 matrix-core peak (the Walsh-basis kernel executes fewer)
                "flop_side": {"achieved": kd["dense_TFLOPs"], "peak": MFMA_PEAK_TFLOPS["f32"], "unit": "TFLOP/s",
                              "frac": kd["flop_frac"], "dense_flops": kd["dense_flops"]}}
    for k, v in kern.items():
        tj = os.path.join(ROOT, "profiles", f"traffic_{k}.json")
        if "traffic" not in v and os.path.exists(tj):
            t = json.load(open(tj))
            if t.get("batch") == len(doms) and t.get("domain_edge") == args.n:
                v["traffic"] = t["traffic_bytes_per_dispatch"]

    cpu = None
    if not args.no_cpu and it64:
        cpu = cpu_baseline(args.n, float(np.mean(it64)), args.cpu_iters, args.cpu_workers, log)

    out = {
        "metric": "converged 64³ domains/sec @rel-res 1e-8: batched FP32+FP64-refine vs plain FP64",
        "value": round(value, 3), "unit": "domains/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(dt / args.steps * 1e3, 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"batch of {len(doms)} trabecular {args.n}^3 domains (BV/TV 0.20, sigma 2), HEX08 "
                               f"micro + HEX08 macro, 24 load cases, rel-res 1e-8, CR_1+CR_2 search",
                   "global_batch": len(doms), "domain_edge": args.n,
                   "preconditioner": "multigrid V-cycle (Chebyshev-Jacobi)" if args.precond == "mg" else "Jacobi",
                   "fp64_form": "Jacobi-PCG in FP64, one domain per batch"},
        "fp64_plain_domains_per_s": round(fp64_rate, 3), "speedup_vs_fp64": round(value / fp64_rate, 2),
        "fp32_refine_jacobi_domains_per_s": round(jac_rate, 3),
        "speedup_precision_only": round(jac_rate / fp64_rate, 2),
        "speedup_preconditioner_and_batch": round(value / jac_rate, 2),
        "tensor_rel_err_vs_fp64": float(f"{terr:.3e}"), "pcg_iters_mixed": float(mixed_iters),
        "pcg_iters_fp64": float(np.mean(it64)) if it64 else None,
        "pcg_iters_fp32_jacobi": float(np.mean(itj)) if itj else None,
        "fp64_fallback_domains": fallbacks,
        "grad_over_forward": round(t_grad / (dt / args.steps), 4),
        "roofline": roof, "kernels": kern, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
