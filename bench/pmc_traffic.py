"""HBM traffic of a timed bench kernel from rocprofv3 counters, calibrated on a copy.

    python bench/pmc_traffic.py [--kernel cheb|mv|bres|resid] [--batch 16] [--n 64]

Runs (as child processes, the program directly after `--`):
  1. rocprofv3 --pmc FETCH_SIZE on one bench.py step of the workload (the
     multigrid-preconditioned mixed solve, run as one part, DTC_CONCURRENT=1: the
     form of bench.py's kernel timing run, whose launches cover the whole batch),
  2. rocprofv3 --pmc WRITE_SIZE on the same run,
  3. both counters on a plain 1 GiB device-to-device copy.
FETCH_SIZE / WRITE_SIZE are reported in KiB per dispatch.  The copy reads and
writes exactly 1 GiB, so `factor = 1 GiB / counter(copy)` converts each counter
into bytes; traffic = FETCH * f_fetch + WRITE * f_write per dispatch, averaged
over the fine-level dispatches of the kernel (the most dispatched grid, the
larger on a tie: the coarse multigrid levels launch the same kernel on smaller
grids, the launch-shape autotuner a few launches on split grids), the launches
whose HIP-event average and algorithmic bytes bench.py reports.
Writes profiles/traffic_<kernel>.json, which bench.py reports as the kernel's
traffic (and as roofline.traffic for the dominant kernel) when its batch and
domain edge match.
"""
from __future__ import annotations

import argparse
import re
import glob
import json
import os
import sqlite3
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# bench kernel key -> (kernel name prefix as rocprofv3 writes it, the name bench.py reports).
# A fine-level FP32 matvec is one launch of the sparse-brick variant and one of the
# dense variant (template flags after the prefix): their per-dispatch averages add.
# name pattern (regular expression) of each timed kernel, and its label; the
# Chebyshev step runs with a vector input (IN_V = 0) or, after a zero start, with
# the scaled right-hand side (IN_JS = 2): both are that kernel's launches
KERNELS = {
    "mv": (r"k_matvec<float, 0, 3,", "k_matvec<float, IN_V, OUT_MASKED>"),
    "cheb": (r"k_matvec<float, [02], 4,", "k_matvec<float, IN_V|IN_JS, OUT_CHEB>"),
    "bres": (r"k_matvec<float, 0, 5,", "k_matvec<float, IN_V, OUT_BRES>"),
    "resid": (r"k_matvec<double, 1, 1,", "k_matvec<double, IN_U, OUT_RESID>"),
}
COPY = ("import torch; a = torch.ones(2**28, device='cuda'); b = torch.empty_like(a); torch.cuda.synchronize(); "
        "b.copy_(a); torch.cuda.synchronize(); print(float(b[-1]))")
GIB = float(1 << 30)


def _rows(dbdir: str, counter: str):
    """(kernel name, grid size or None, value) of every dispatch sample of `counter`."""
    out = []
    for db in glob.glob(os.path.join(dbdir, "**", "*_results.db"), recursive=True):
        c = sqlite3.connect(db)
        cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
        gcol = next((g for g in ("grid_size", "grid_x", "grid") if g in cols), None)
        q = f"select kernel_name, {gcol or 'null'}, value from counters_collection where counter_name = ?"
        out += [(n, g, float(v)) for n, g, v in c.execute(q, (counter,))]
        c.close()
    return out


def _counter(dbdir: str, counter: str, kname: str) -> tuple[float, int]:
    """(mean value per fine-level dispatch, dispatches) over every kernel variant
    whose name matches the pattern kname: for each variant its fine-level
    dispatches, and the mean over all of them (the variants are separate launches)."""
    pat = re.compile(kname)
    rows = [r for r in _rows(dbdir, counter) if pat.search(r[0].replace("(anonymous namespace)::", ""))]
    if not rows:
        raise RuntimeError(f"no {counter} samples of {kname} in {dbdir}")
    by = {}
    for r in rows:
        by.setdefault(r[0], []).append(r)
    vals = []
    for name, rs in by.items():
        if all(r[1] is not None for r in rs):
            cnt = {}
            for r in rs:
                cnt[r[1]] = cnt.get(r[1], 0) + 1
            g = max(cnt, key=lambda k: (cnt[k], k))   # most dispatched grid, the larger on a tie
            rs = [r for r in rs if r[1] == g]
        vals += [r[2] for r in rs]
    return sum(vals) / len(vals), len(vals)


def _pmc(counter: str, out: str, cmd: list[str]) -> str:
    full = ["rocprofv3", "--pmc", counter, "--kernel-include-regex", "k_matvec|copyBuffer", "-d", out, "-o", "run",
            "--"] + cmd
    print(" ".join(full), flush=True)
    p = subprocess.run(full, check=True, timeout=600, stdout=subprocess.PIPE, text=True,
                       env=dict(os.environ, TMPDIR="/tmp", DTC_CONCURRENT="1"))
    return p.stdout


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--kernel", choices=sorted(KERNELS), default="cheb")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=64)
    args = ap.parse_args()
    kname, label = KERNELS[args.kernel]
    out_path = os.path.join(ROOT, "profiles", f"traffic_{args.kernel}.json")
    bench = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "1", "--warmup", "0", "--batch",
             str(args.batch), "--n", str(args.n), "--fp64-domains", "0", "--no-cpu"]
    copy = [sys.executable, "-c", COPY]
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, "mv_" + counter)
            line = [l for l in _pmc(counter, d, bench).splitlines() if l.startswith("{")]
            if line:
                k = json.loads(line[-1]).get("kernels", {}).get(args.kernel)
                if k:
                    res["algorithmic"] = k["algorithmic_bytes"]
            res[counter] = _counter(d, counter, kname)
            d = os.path.join(tmp, "copy_" + counter)
            _pmc(counter, d, copy)
            # the copy is the dispatch with the largest counter value of that program
            res["copy_" + counter] = max(r[2] for r in _rows(d, counter))
    f_fetch = (GIB / 1024.0) / res["copy_FETCH_SIZE"]
    f_write = (GIB / 1024.0) / res["copy_WRITE_SIZE"]
    fetch_kib, n_f = res["FETCH_SIZE"]
    write_kib, n_w = res["WRITE_SIZE"]
    out = {
        "kernel": label,
        "workload": f"bench.py batch {args.batch}, {args.n}^3 trabecular domains, one step of the mixed solve, "
                    f"fine-level dispatches",
        "batch": args.batch, "domain_edge": args.n,
        "fetch_kib_per_dispatch": round(fetch_kib, 1), "write_kib_per_dispatch": round(write_kib, 1),
        "dispatches": [n_f, n_w],
        "copy_fetch_kib_per_gib": res["copy_FETCH_SIZE"], "copy_write_kib_per_gib": res["copy_WRITE_SIZE"],
        "fetch_factor": round(f_fetch, 4), "write_factor": round(f_write, 4),
        "traffic_bytes_per_dispatch": round(fetch_kib * 1024.0 * f_fetch + write_kib * 1024.0 * f_write),
        "algorithmic_bytes_per_dispatch": res.get("algorithmic"),
        "command": f"python bench/pmc_traffic.py --kernel {args.kernel} --batch {args.batch} --n {args.n}",
    }
    os.makedirs(os.path.dirname(out_path), exist_ok=True)
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
