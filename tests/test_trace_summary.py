"""Host-side check of the kernel-trace summary tool (by kernel and grid, and the timeline
around the last launches of a kernel) on a small synthetic rocprofv3 trace."""
import csv

from mrlbm_2d_amd import trace_summary

ROWS = [  # kernel, start ns, end ns, grid, workgroup, queue
    ("bulk", 1000, 5000, 1024, 256, 2),
    ("side", 1200, 2200, 128, 64, 1),
    ("side", 2200, 3000, 64, 64, 1),
    ("bulk", 6000, 9000, 1024, 256, 2),
    ("side", 6100, 7100, 128, 64, 1),
]


def _write(path):
    with open(path, "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X", "Workgroup_Size_X", "Queue_Id"])
        w.writerows(ROWS)


def test_summary_by_kernel_and_grid(tmp_path):
    src, dst = tmp_path / "t.csv", tmp_path / "s.csv"
    _write(src)
    trace_summary.summarize(str(src), str(dst))
    rows = list(csv.DictReader(open(dst)))
    assert [(r["kernel"], r["grid"], r["calls"]) for r in rows] == [
        ("bulk", "1024", "2"), ("side", "128", "2"), ("side", "64", "1")]  # by total time
    assert float(rows[0]["mean_ns"]) == 3500.0 and rows[0]["min_ns"] == "3000" and rows[0]["max_ns"] == "4000"


def test_timeline_around_last_launches(tmp_path):
    src, dst = tmp_path / "t.csv", tmp_path / "l.csv"
    _write(src)
    trace_summary.timeline(str(src), str(dst), "bulk", 1)
    rows = list(csv.DictReader(open(dst)))
    # from the dispatch before the last bulk launch to the end, times from the first
    assert [r["kernel"] for r in rows] == ["side", "bulk", "side"]
    assert [(r["start_us"], r["end_us"], r["workgroups"], r["queue"]) for r in rows] == [
        ("0.0", "0.8", "1", "1"), ("3.8", "6.8", "4", "2"), ("3.9", "4.9", "2", "1")]


def test_timeline_asks_for_more_launches_than_traced(tmp_path):
    src = tmp_path / "t.csv"
    _write(src)
    try:
        trace_summary.timeline(str(src), str(tmp_path / "l.csv"), "bulk", 3)
    except SystemExit as e:
        assert "2 launches" in str(e)
    else:
        raise AssertionError("expected SystemExit")
