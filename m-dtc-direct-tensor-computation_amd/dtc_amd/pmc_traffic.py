"""HBM traffic of the CG matvec from rocprofv3 counters, calibrated on a copy.

    python -m dtc_amd.pmc_traffic [--batch 16] [--n 64] [--out profiles/k_matvec_traffic.json]

Runs (as child processes, the program directly after `--`):
  1. rocprofv3 --pmc FETCH_SIZE on one bench.py step of the workload (the
     multigrid-preconditioned mixed solve, as timed by bench.py),
  2. rocprofv3 --pmc WRITE_SIZE on the same run,
  3. both counters on a plain 1 GiB device-to-device copy.
FETCH_SIZE / WRITE_SIZE are reported in KiB per dispatch.  The copy reads and
writes exactly 1 GiB, so `factor = 1 GiB / counter(copy)` converts each counter
into bytes; traffic = FETCH * f_fetch + WRITE * f_write per dispatch of
k_matvec<float, IN_V, OUT_MASKED> (Q = A P of the block PCG), averaged over its
dispatches (the same launches whose HIP-event average and algorithmic bytes
bench.py reports).
bench.py reports the result as roofline.traffic when its batch and domain edge
match.
"""
from __future__ import annotations

import argparse
import glob
import json
import os
import sqlite3
import subprocess
import sys
import tempfile

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
KERNEL = "k_matvec<float, 0, 3, false>"
COPY = ("import torch; a = torch.ones(2**28, device='cuda'); b = torch.empty_like(a); torch.cuda.synchronize(); "
        "b.copy_(a); torch.cuda.synchronize(); print(float(b[-1]))")
GIB = float(1 << 30)


def _counter(dbdir: str, counter: str, match) -> tuple[float, int]:
    """(mean value per dispatch, dispatches) of the kernels whose name satisfies match()."""
    vals = []
    for db in glob.glob(os.path.join(dbdir, "**", "*_results.db"), recursive=True):
        c = sqlite3.connect(db)
        for name, v in c.execute("select kernel_name, value from counters_collection where counter_name = ?",
                                 (counter,)):
            if match(name):
                vals.append(float(v))
        c.close()
    if not vals:
        raise RuntimeError(f"no {counter} samples in {dbdir}")
    return sum(vals) / len(vals), len(vals)


def _pmc(counter: str, out: str, cmd: list[str]) -> str:
    full = ["rocprofv3", "--pmc", counter, "--kernel-include-regex", "k_matvec|copyBuffer", "-d", out, "-o", "run",
            "--"] + cmd
    print(" ".join(full), flush=True)
    p = subprocess.run(full, check=True, timeout=600, stdout=subprocess.PIPE, text=True,
                       env=dict(os.environ, TMPDIR="/tmp"))
    return p.stdout


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles", "k_matvec_traffic.json"))
    args = ap.parse_args()
    bench = [sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "1", "--warmup", "0", "--batch",
             str(args.batch), "--n", str(args.n), "--fp64-domains", "0", "--no-cpu"]
    copy = [sys.executable, "-c", COPY]
    res = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(tmp, "mv_" + counter)
            line = [l for l in _pmc(counter, d, bench).splitlines() if l.startswith("{")]
            if line:
                res["algorithmic"] = json.loads(line[-1])["roofline"]["algorithmic_bytes"]
            res[counter] = _counter(d, counter, lambda n: KERNEL in n.replace("(anonymous namespace)::", ""))
            d = os.path.join(tmp, "copy_" + counter)
            _pmc(counter, d, copy)
            # the copy is the kernel with the largest counter value of that program
            vals = []
            for db in glob.glob(os.path.join(d, "**", "*_results.db"), recursive=True):
                c = sqlite3.connect(db)
                vals += [float(x) for (x,) in c.execute(
                    "select value from counters_collection where counter_name = ?", (counter,))]
                c.close()
            res["copy_" + counter] = max(vals)
    f_fetch = (GIB / 1024.0) / res["copy_FETCH_SIZE"]
    f_write = (GIB / 1024.0) / res["copy_WRITE_SIZE"]
    fetch_kib, n_f = res["FETCH_SIZE"]
    write_kib, n_w = res["WRITE_SIZE"]
    out = {
        "kernel": "k_matvec<float, IN_V, OUT_MASKED>",
        "workload": f"bench.py batch {args.batch}, {args.n}^3 trabecular domains, one step of the mixed solve",
        "batch": args.batch, "domain_edge": args.n,
        "fetch_kib_per_dispatch": round(fetch_kib, 1), "write_kib_per_dispatch": round(write_kib, 1),
        "dispatches": [n_f, n_w],
        "copy_fetch_kib_per_gib": res["copy_FETCH_SIZE"], "copy_write_kib_per_gib": res["copy_WRITE_SIZE"],
        "fetch_factor": round(f_fetch, 4), "write_factor": round(f_write, 4),
        "traffic_bytes_per_dispatch": round(fetch_kib * 1024.0 * f_fetch + write_kib * 1024.0 * f_write),
        "algorithmic_bytes_per_dispatch": res.get("algorithmic"),
        "command": "python -m dtc_amd.pmc_traffic --batch %d --n %d" % (args.batch, args.n),
    }
    os.makedirs(os.path.dirname(args.out), exist_ok=True)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
