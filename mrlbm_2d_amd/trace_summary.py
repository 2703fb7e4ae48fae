"""Summary of a rocprofv3 kernel trace by (kernel name, grid size).

rocprofv3 --stats aggregates by kernel name only, which folds launches of one kernel on
different grids (the 8192 x 2048 benchmark channel and the 512 x 128 adjoint case share the
FP32 two-step kernel) into one average.  This keeps them apart:

    python -m mrlbm_2d_amd.trace_summary <out>/run_kernel_trace.csv summary.csv

writes kernel, grid (work-items), calls, mean/min/max duration in ns, sorted by total time.

    python -m mrlbm_2d_amd.trace_summary --timeline <out>/run_kernel_trace.csv t.csv PATTERN N

writes the dispatches from just before the N-th last launch of the kernel whose name holds
PATTERN to a few after its last, in start order, with start/end times in us from the first,
workgroups, workgroup size and queue (which stream ran them).
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


def timeline(src: str, dst: str, pattern: str, n: int) -> None:
    with open(src) as fh:
        rows = sorted(csv.DictReader(fh), key=lambda r: int(r["Start_Timestamp"]))
    idx = [i for i, r in enumerate(rows) if pattern in r["Kernel_Name"]]
    if len(idx) < n:
        raise SystemExit(f"{len(idx)} launches of a kernel named like {pattern!r}, {n} asked for")
    sel = rows[max(idx[-n] - 1, 0): idx[-1] + 6]
    t0 = int(sel[0]["Start_Timestamp"])
    with open(dst, "w") as fh:
        fh.write("start_us,end_us,dur_us,workgroups,wg_size,queue,kernel\n")
        for r in sel:
            s, e = (int(r["Start_Timestamp"]) - t0) / 1e3, (int(r["End_Timestamp"]) - t0) / 1e3
            wg = int(r.get("Workgroup_Size_X") or r.get("Workgroup_Size") or 1)
            g = int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0) // max(wg, 1)
            fh.write(f'{s:.1f},{e:.1f},{e - s:.1f},{g},{wg},{r["Queue_Id"]},"{r["Kernel_Name"][:90]}"\n')


if __name__ == "__main__":
    if sys.argv[1] == "--timeline":
        timeline(sys.argv[2], sys.argv[3], sys.argv[4], int(sys.argv[5]))
    else:
        summarize(sys.argv[1], sys.argv[2])
