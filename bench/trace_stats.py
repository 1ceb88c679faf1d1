"""Per-kernel totals of a rocprofv3 kernel trace, grouped by kernel and grid.

    python bench/trace_stats.py <dir>/run_results.db [rows]

Reads the `kernels` view of the trace database written by
`rocprofv3 --kernel-trace -d <dir> -o run -- <program>` and prints calls,
total and average duration per (kernel, grid_x x grid_y), largest first.
"""
from __future__ import annotations

import sqlite3
import sys


def main() -> None:
    db = sys.argv[1]
    rows_max = int(sys.argv[2]) if len(sys.argv) > 2 else 40
    c = sqlite3.connect(db)
    q = ("select name, grid_x, grid_y, count(*), sum(end - start), avg(end - start) from kernels "
         "group by name, grid_x, grid_y order by sum(end - start) desc")
    rows = c.execute(q).fetchall()
    tot = sum(r[4] for r in rows)
    print(f"{'kernel':44s} {'grid':>14s} {'calls':>6s} {'total ms':>9s} {'avg us':>9s} {'%':>5s}")
    for r in rows[:rows_max]:
        nm = r[0].replace("(anonymous namespace)::", "").replace("dtc::", "").split("(")[0][:44]
        print(f"{nm:44s} {str(r[1]) + 'x' + str(r[2]):>14s} {r[3]:6d} {r[4] / 1e6:9.2f} {r[5] / 1e3:9.2f} "
              f"{100 * r[4] / tot:5.1f}")
    print("total ms", tot / 1e6)


if __name__ == "__main__":
    main()
