"""Summary of a rocprofv3 kernel trace by (kernel name, grid size).

rocprofv3 --stats aggregates by kernel name only, which folds launches of one kernel on
different grids (the 8192 x 2048 benchmark channel and the 512 x 128 adjoint case share the
FP32 two-step kernel) into one average.  This keeps them apart:

    python -m mrlbm_2d_amd.trace_summary <out>/run_kernel_trace.csv summary.csv

writes kernel, grid (work-items), calls, mean/min/max duration in ns, sorted by total time.
"""
from __future__ import annotations

import collections
import csv
import sys


def summarize(src: str, dst: str) -> None:
    agg = collections.defaultdict(lambda: [0, 0.0, float("inf"), 0.0])
    with open(src) as fh:
        for r in csv.DictReader(fh):
            d = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
            a = agg[(r["Kernel_Name"], r.get("Grid_Size") or r.get("Grid_Size_X", ""))]
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
    with open(dst, "w") as fh:
        fh.write("kernel,grid,calls,mean_ns,min_ns,max_ns\n")
        for (k, g), (n, s, lo, hi) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
            fh.write(f'"{k}",{g},{n},{s / n:.1f},{lo:.0f},{hi:.0f}\n')


if __name__ == "__main__":
    summarize(sys.argv[1], sys.argv[2])
