"""bench.py's command-line contract and its CPU-side pieces (no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_roofline_fields_and_arithmetic():
    r = dict(achieved_gbs=7000.0, frac=7000.0 / bench.HBM_PEAK_GBS, steps_per_launch=2.0,
             bytes_per_step=805_306_368, launch_ms=0.23, launch_median_ms=0.229, launch_min_ms=0.22,
             launch_max_ms=0.25, timed_launches=10, bulk_launches=10, timed_steps=20)
    traffic, source, _ = bench.TRAFFIC[2]
    out = bench.roofline(r, "bulk2_kernel<float,float,10,false>", traffic["f32"], source)
    assert out["bound"] == "hbm" and out["unit"] == "GB/s" and out["peak"] == 8000.0
    assert out["frac"] == pytest.approx(0.875)
    assert out["algorithmic_bytes_per_launch"] == 2 * 805_306_368
    assert out["traffic"] == traffic["f32"]
    assert out["traffic_source"].startswith(source)
    assert out["traffic_gbs"] == pytest.approx(traffic["f32"] / 0.23e-3 / 1e9, rel=1e-4)
    assert out["traffic_frac"] == pytest.approx(out["traffic_gbs"] / 8000.0, rel=1e-3)
    for spp, (tr, src, _) in bench.TRAFFIC.items():
        assert os.path.exists(os.path.join(REPO, src)), src
        assert tr["f32"] > 0 and tr["f64"] > 0, spp
    none = bench.roofline(r, "k", None, None)
    assert none["traffic_source"] is None and none["traffic_frac"] is None
    assert none["valu_issue"] is None and none["valu_source"] is None
    shares, src, _ = bench.VALU_ISSUE[4]
    with_valu = bench.roofline(r, "k", traffic["f32"], src, shares["f32"])
    assert 0.5 < with_valu["valu_issue"] < 1.0 and with_valu["valu_source"].startswith(src)
    assert os.path.exists(os.path.join(REPO, src))
    json.dumps(out)


def test_cpu_baseline_on_a_small_budget():
    """The oracle baseline on the bench's grid, one timed step: the JSON fields the bench
    line carries."""
    out = bench.cpu_baseline(max_steps=1, budget_s=1e-9)
    assert out["unit"] == "MLUPS" and out["kind"] == "port" and out["value"] > 0
    assert out["cores"] >= 1
    assert "8192x2048" in out["sample"]
