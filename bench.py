"""Benchmark of the D2Q9 moments-based regularized LBM step on one MI355X.

Workload (the perf configuration of the metric): 8192 x 2048 channel, cylinder D = 128
centred at (1024, 1025), parabolic inlet u_max = 0.05, Re = 200, seeded +-1e-4 initial
velocity noise.  Both precision forms run: FP32 deviation storage (the headline `value`)
and plain FP64.  A "step" is one lattice update of every node (one launch).

Prints ONE JSON line.  `roofline` prices the step kernel against MI355X's 8 TB/s HBM:
achieved = algorithmic bytes per launch (12 moment words per node: 6 read + 6 written)
/ mean kernel time measured with HIP events on the library's stream.
`cpu_baseline` times the reference program itself (oracle/_ref/mrlbm_ref, built from
the reference sources with flang, OpenMP) on a bounded number of steps of the same grid.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
NX, NY = 8192, 2048
XC, YC, RC = 1024.0, 1025.0, 64.0
UMAX, RE = 0.05, 200.0


def make_lattice(precision, nx=NX, ny=NY, strip_rows=0, kernel=0):
    from mrlbm_2d_amd import capi
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = XC * nx / NX, YC * ny / NY, RC * ny / NY
    p.Re, p.uo = RE, UMAX
    p.inlet_profile = capi.INLET_PARABOLIC
    p.precision = precision
    p.strip_rows = strip_rows
    p.kernel = kernel
    lat = capi.Lattice(p)
    rng = np.random.default_rng(3)
    ux = rng.uniform(-1e-4, 1e-4, (nx, ny))
    uy = rng.uniform(-1e-4, 1e-4, (nx, ny))
    lat.set_moments(np.ones((nx, ny)), ux, uy)
    return lat


def time_form(precision, steps, warmup, **kw):
    lat = make_lattice(precision, **kw)
    lat.step(2, warmup)
    lat.synchronize()
    lat.set_timing(True)
    t0 = time.perf_counter()
    lat.step(2 + warmup, steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = lat.last_step_ms()
    assert lat.check_finite() == 0, "non-finite state after the timed steps"
    info = lat.info
    lat.close()
    n = info.nx * info.ny
    mlups = n * steps / (wall * 1e6)
    bytes_per_launch = int(info.bytes_per_step)
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    return dict(mlups=mlups, wall_s=wall, ms_per_step=1e3 * wall / steps, kernel_ms=kern_ms,
                bytes_per_launch=bytes_per_launch, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS)


def reference_input(nsteps, threads):
    # The reference's geometry for the same 8192 x 2048 grid: ncy = 128 (D = 128),
    # L_front + 1 + L_back = 64, L_top + 1 + L_bot = 16 (2d_mrlbm.f90:86-92).
    return f"""&DomainSize
 nx = 320
 ny = 32
 lattice_type = 9
/
&Cylinder
 ncy = 128
 L_front = 8
 L_back = 55
 L_top = 8
 L_bot = 7
/
&Numbers
 Re = {RE}
/
&Parallel
 num_threads={threads}
/
&Controls
 uo = {UMAX}
 rho_infty = 1.0
 iplot = 100000000
 max_iter = {1 + nsteps}
 isave = 100000000
 irestart = 0
 statsbegin = 100000000
 statsend = 100000001
 iplotstats = 1000
/
&LogicalControls
 post_process = .false.
 x_periodic = .false.
 y_periodic = .false.
 channel_with_cylinder = .true.
 channel_with_square = .false.
 incomp_bc = .false.
 incomp_bulk = .false.
 vel_interp = .true.
 mom_interp = .true.
 rotated_coordinate = .true.
 strong_mass_bc = .true.
 eq_mass_bc = .false.
 weak_mass_bc = .false.
/
"""


def cpu_baseline(threads, n1=1, n2=6):
    """Wall time of the reference program per step: two runs of n1 and n2 steps on the
    8192 x 2048 grid, differenced so that setup and file output cancel."""
    exe = os.path.join(REPO, "oracle", "_ref", "mrlbm_ref")
    if not os.path.exists(exe):
        return None
    times = []
    for n in (n1, n2):
        d = tempfile.mkdtemp(prefix="mrlbm_ref_")
        try:
            for sub in ("output", "restart", "data_probe", "data_geo", "data_mean"):
                os.makedirs(os.path.join(d, sub))
            with open(os.path.join(d, "input.dat"), "w") as fh:
                fh.write(reference_input(n, threads))
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
            t0 = time.perf_counter()
            r = subprocess.run([exe], cwd=d, env=env, capture_output=True, text=True, timeout=900)
            times.append(time.perf_counter() - t0)
            if r.returncode != 0:
                return None
        finally:
            shutil.rmtree(d, ignore_errors=True)
    per_step = (times[1] - times[0]) / (n2 - n1)
    return dict(value=NX * NY / (per_step * 1e6), unit="MLUPS", cores=threads, kind="reference",
                sample=f"reference program (flang -O3 -fopenmp, {threads} threads) on the same "
                       f"8192x2048 grid (ncy=128), {n2} minus {n1} steps, wall clock differenced "
                       f"({times[1]:.1f}s - {times[0]:.1f}s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--strip-rows", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the multi-GPU slab decomposition is "
              "design-only (see DESIGN.md). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    from mrlbm_2d_amd import capi
    print(f"[bench] FP32 deviation form {NX}x{NY}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)
    r32 = time_form(capi.MRLBM_FP32_DEV, args.steps, args.warmup, strip_rows=args.strip_rows)
    print(f"[bench] FP32: {r32}", file=sys.stderr, flush=True)
    print(f"[bench] FP64 plain form {NX}x{NY}", file=sys.stderr, flush=True)
    r64 = time_form(capi.MRLBM_FP64, args.steps, args.warmup, strip_rows=args.strip_rows)
    print(f"[bench] FP64: {r64}", file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline: reference program", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_threads)
        print(f"[bench] CPU: {cpu}", file=sys.stderr, flush=True)
    out = {
        "metric": "MLUPS of the D2Q9 regularized LBM step, FP32 deviation form (FP64 beside it)",
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded +-1e-4 velocity noise)",
        "config": {"workload": "cfg3: 8192x2048 channel, cylinder D=128 at (1024,1025), "
                               "parabolic inlet u_max=0.05, Re=200", "nx": NX, "ny": NY},
        "roofline": {"bound": "hbm", "achieved": round(r32["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r32["frac"], 4), "traffic": None},
        "fp64": {"value": round(r64["mlups"], 1), "ms_per_step": round(r64["ms_per_step"], 5),
                 "roofline": {"bound": "hbm", "achieved": round(r64["achieved_gbs"], 1),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(r64["frac"], 4)}},
        "fp32_over_fp64": round(r32["mlups"] / r64["mlups"], 3),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
