"""Per-dispatch hardware counters of the timed bench kernels (rocprofv3 --pmc).

    python bench/pmc_counters.py --counters SQ_WAVES,SQ_INSTS_VALU,... [--batch 16] [--n 64]
                                 [--out profiles/<name>.json]

Runs one rocprofv3 --pmc pass (the counters of one pass must fit the hardware's
per-block limits: at most 8 SQ counters) over one bench.py step of the workload
(as a child process, the program directly after `--`; the step runs as one part,
DTC_CONCURRENT=1, the form of bench.py's kernel timing run), and averages every counter
over the fine-level dispatches (the most dispatched grid, the larger on a tie: the
coarse levels run smaller grids, the launch-shape autotuner a few split ones) of
each timed kernel: the CG matvec
Q = K P (mv), the fused Chebyshev smoothing step (cheb), the restriction
residual (bres) and the FP64 residual (resid).  Prints one JSON object.
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
sys.path.insert(0, os.path.join(ROOT, "bench"))
from pmc_traffic import KERNELS  # noqa: E402


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--counters", required=True)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    counters = [c.strip() for c in args.counters.split(",") if c.strip()]
    bench = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "1", "--warmup", "0", "--batch",
             str(args.batch), "--n", str(args.n), "--fp64-domains", "0", "--no-cpu"]
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        cmd = ["rocprofv3", "--pmc"] + counters + ["--kernel-include-regex", "k_matvec", "-d", tmp, "-o", "run",
                                                   "--"] + bench
        print(" ".join(cmd), file=sys.stderr, flush=True)
        subprocess.run(cmd, check=True, timeout=900, stdout=subprocess.DEVNULL, env=dict(os.environ, TMPDIR="/tmp", DTC_CONCURRENT="1"))
        rows = []
        for db in glob.glob(os.path.join(tmp, "**", "*_results.db"), recursive=True):
            c = sqlite3.connect(db)
            cols = [r[1] for r in c.execute("pragma table_info(counters_collection)")]
            gcol = next((g for g in ("grid_size", "grid_x", "grid") if g in cols), None)
            dcol = next((g for g in ("dispatch_id", "dispatch", "correlation_id") if g in cols), None)
            q = (f"select kernel_name, {gcol or 'null'}, {dcol or 'null'}, counter_name, value "
                 f"from counters_collection")
            rows += c.execute(q).fetchall()
            c.close()
    for key, (kname, label) in KERNELS.items():
        mine = [r for r in rows if re.search(kname, r[0].replace("(anonymous namespace)::", ""))]
        if not mine:
            continue
        variants = {}
        for r in mine:
            variants.setdefault(r[0], []).append(r)
        entry = {"kernel": label, "variants": {}}
        for name, rs in variants.items():
            if all(r[1] is not None for r in rs):
                # the solve's fine-level grid: the most dispatched grid, the larger on a tie
                # (the launch-shape autotuner runs a few launches on split grids first)
                cnt = {}
                for r in rs:
                    cnt.setdefault(r[1], set()).add(r[2])
                g = max(cnt, key=lambda k: (len(cnt[k]), k))
                rs = [r for r in rs if r[1] == g]
            per = {}
            for _, _, disp, cn, v in rs:
                per.setdefault(cn, {}).setdefault(disp, 0.0)
                per[cn][disp] += float(v)        # summed over the counter's instances per dispatch
            short = name.replace("(anonymous namespace)::", "").replace("dtc::", "").split("(")[0]
            entry["variants"][short] = {"dispatches": max(len(d) for d in per.values()),
                                        **{cn: sum(d.values()) / len(d) for cn, d in sorted(per.items())}}
        res[key] = entry
    out = {"workload": f"bench.py batch {args.batch}, {args.n}^3 trabecular domains, one step, fine-level dispatches",
           "counters": counters, "per_dispatch": res,
           "command": "python bench/pmc_counters.py --counters " + ",".join(counters) +
                      f" --batch {args.batch} --n {args.n}"}
    if args.out:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
