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
FP32_PEAK_TFLOPS = 157.3       # MI355X dense FP32 matrix


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def make_domains(count, n, seed0):
    import dtc_amd
    from dtc_amd import synth
    return [dtc_amd.Domain(synth.trabecular(n, seed=seed0 + i, sigma=2.0, bvtv=0.20), llimit=1,
                           delta=(0.01, 0.01, 0.01), xa_n=(2, 2, 2), E=5600.0, nu=0.3) for i in range(count)]


def cpu_baseline(dom, gpu_fp64_iters_per_lc, sample_iters, log):
    """The oracle (assembled 3x3-block CSR, Jacobi-PCG per load case, FP64, one
    core) on a bounded sample: one load case for `sample_iters` PCG iterations
    plus one CR_1 search, extrapolated to a full domain with the iteration
    counts of the FP64 solve of the same domain."""
    from oracle import oracle as O
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
    t_domain = t_mesh + per_iter * gpu_fp64_iters_per_lc * 24 + 2 * t_rot
    log(f"cpu sample: mesh {t_mesh:.2f}s, {it} PCG iterations in {t_solve:.2f}s, CR_1 search {t_rot:.2f}s "
        f"-> {t_domain:.1f}s per domain")
    return {"value": 1.0 / t_domain, "unit": "domains/s", "cores": 1, "kind": "port",
            "sample": f"oracle on one 64^3 domain: mesh, load case 1 for {it} Jacobi-PCG iterations "
                      f"({per_iter * 1e3:.2f} ms each), one CR_1 search; extrapolated to 24 load cases x "
                      f"{gpu_fp64_iters_per_lc:.0f} iterations (the FP64 iteration count) + CR_1 and CR_2"}


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
        with dtc_amd.Batch(doms[g0:g0 + G], precision=dtc_amd.DTC_FP32_REFINE, rel_tol=1e-8, rotation=True,
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--fp64-domains", type=int, default=2)
    ap.add_argument("--cpu-iters", type=int, default=400)
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
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(domains shard across GPUs with no exchange, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import dtc_amd
    if args.uneven > 0:
        uneven(args, log)
        return
    doms = make_domains(args.batch, args.n, 2)
    log(f"batch of {len(doms)} domains of {args.n}^3, BV/TV " +
        ", ".join(f"{d.phi.mean():.3f}" for d in doms[:4]) + " ...")

    # ---- batched FP32 + FP64 refinement
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
    st = b.stats()
    bad = [i for i, r in enumerate(res) if r["status"] != 0 or r["relres"] > 1e-8]
    if bad and not args.no_check:
        raise SystemExit(f"domains {bad} did not converge to 1e-8: {[res[i]['status'] for i in bad]}")
    value = len(doms) * args.steps / dt
    mixed_iters = np.mean([r["iterations"] for r in res])
    log("PCG iterations per domain: " + " ".join(str(r["iterations"]) for r in res) +
        "; FP64 residual replacements: " + " ".join(str(r["outer_iterations"]) for r in res))

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

    # ---- plain FP64 GPU form, one domain at a time
    t64, C64, it64 = 0.0, [], []
    for i in range(min(args.fp64_domains, len(doms))):
        with dtc_amd.Batch([doms[i]], precision=dtc_amd.DTC_FP64, rel_tol=1e-8, rotation=True,
                           max_iter=args.max_iter) as b64:
            b64.run()
            _sync()
            s0 = time.perf_counter()
            r64 = b64.run()[0]
            _sync()
            t64 += time.perf_counter() - s0
        C64.append(r64["C"])
        it64.append(r64["iterations"])
        log(f"fp64 domain {i}: {r64['iterations']} iterations, relres {r64['relres']:.2e}")
    fp64_rate = len(C64) / t64 if C64 else float("nan")
    terr = max((np.abs(res[i]["C"] - C64[i]).max() / np.abs(C64[i]).max() for i in range(len(C64))), default=float("nan"))

    # ---- roofline of the dominant kernel (EbE matvec Q = A P of the FP32 block PCG)
    mv_s = st["mv_avg_ms"] * 1e-3
    achieved = st["mv_bytes"] / mv_s / 1e9 if mv_s > 0 else 0.0
    traffic = None
    tj = os.path.join(ROOT, "profiles", "k_matvec_traffic.json")
    if os.path.exists(tj):
        t = json.load(open(tj))
        if t.get("batch") == len(doms) and t.get("domain_edge") == args.n:
            traffic = t["traffic_bytes_per_dispatch"]
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "kernel": "k_matvec<float, IN_V, OUT_MASKED>", "avg_launch_ms": round(st["mv_avg_ms"], 5),
            "launches_timed": st["mv_launches"], "algorithmic_bytes": st["mv_bytes"],
            "dense_equiv_tflops": round(st["mv_flops"] / mv_s / 1e12, 2) if mv_s > 0 else 0.0}

    cpu = None
    if not args.no_cpu and it64:
        cpu = cpu_baseline(doms[0], float(np.mean(it64)), args.cpu_iters, log)

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
        "tensor_rel_err_vs_fp64": float(f"{terr:.3e}"), "pcg_iters_mixed": float(mixed_iters),
        "pcg_iters_fp64": float(np.mean(it64)) if it64 else None,
        "grad_over_forward": round(t_grad / (dt / args.steps), 4),
        "roofline": roof, "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
