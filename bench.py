"""Benchmark of the D2Q9 moments-based regularized LBM step on one MI355X.

Workload (the perf configuration of the metric): 8192 x 2048 channel, cylinder D = 128
centred at (1024, 1025), parabolic inlet u_max = 0.05, Re = 200, seeded +-1e-4 initial
velocity noise.  Both precision forms run: FP32 deviation storage (the headline `value`)
and plain FP64.  A "step" is one lattice update of every node; the default path advances
four steps per launch of the bulk kernel (temporal blocking), bit-identical to single steps.

Prints ONE JSON line.  `value` is the wall clock of the K timed steps, synchronized on both
sides, with no per-launch events among them.  `roofline` prices the dominant kernel against
MI355X's 8 TB/s HBM over the next K steps of the same run: achieved = algorithmic bytes of
the timed launches (12 moment words per node and step: 6 read + 6 written; the library
reports how many steps the timed launches advanced) / their summed duration, from a HIP
event pair given to each timed launch on the bulk kernel's own stream (every launch when the
run is short, every 4th otherwise);
traffic = one launch's HBM bytes from rocprofv3 PMC counters (FETCH_SIZE, WRITE_SIZE) of the
committed profile named in `traffic_source`, not measured in this run; below the algorithmic
bytes because the intermediate steps never leave the chip, so `frac` (algorithmic) passes 1
and `traffic_frac` (those bytes over the same launch time, against 8 TB/s) is the share of HBM
bandwidth the launch actually uses; `valu_issue` (same profile) is the share of the SIMDs'
cycles the four-step launch spends issuing VALU instructions (instruction issue, not HBM,
bounds that kernel).
`statistics_window`: ms per step inside the statistics window (forces, surface pressure,
probes and u statistics recorded every step) against plain steps, both forms.
`fp32_vs_fp64` is the accuracy condition of the FP32 figure: both forms from the same seeded
state over 100,000 steps (mean Cd, rms Cl, spectral St over the last 50,000; velocity L2 at
the end), run before the timed legs.  Also: the 512x128 drag adjoint (adjoint / forward
time) and 256 batched runs (aggregate MLUPS).
`cpu_baseline` times the oracle (the CPU restatement of the reference step) on a bounded
number of steps of the same grid.  The reference program itself is not timed here: its
set-up writes every fluid node to formatted text files, which at this grid size runs for
many minutes before the first step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel at 8192 x 2048, from rocprofv3 FETCH_SIZE and
# WRITE_SIZE runs scaled by a plain copy of known size (FETCH_SIZE x 2, WRITE_SIZE x 1):
# the four-step bulk pass (default; strips of 20 rows in FP32, 40 in FP64) and the two-step
# one (--steps-per-pass 2; FP32 10 rows, FP64 12), with the profiles holding the summaries and
# the commands; the strip heights each was measured at
TRAFFIC = {
    4: ({"f32": 1_074_217_287, "f64": 1_851_616_351}, "profiles/bulkn_kernel_8192x2048_hbm_bytes.md", (20, 40)),
    2: ({"f32": 949_550_578, "f64": 2_289_678_504}, "profiles/bulk2_kernel_8192x2048_hbm_bytes.md", (10, 12)),
}
# Share of the SIMDs' cycles issuing VALU instructions during a four-step launch (instruction
# issue bounds that kernel), from the same profile's SQ_INSTS_VALU x 4 cycles over GRBM_GUI_ACTIVE
VALU_ISSUE = {4: ({"f32": 0.805, "f64": 0.743}, "profiles/bulkn_kernel_8192x2048_hbm_bytes.md", (20, 40))}
NX, NY = 8192, 2048
# HIP event pairs around every 4th bulk launch on long runs: the per-launch mean without the
# events' own cost (about 2% of the step when recorded around every launch); runs of up to
# TIME_ALL_UP_TO steps time every launch
TIMED_EVERY = 4
TIME_ALL_UP_TO = 400
XC, YC, RC = 1024.0, 1025.0, 64.0
# oracle MLUPS / reference MLUPS measured in the build container (8 cores) on the reference's
# own 400x100 (ncy = 20, L = 3/16/2/2) and 920x420 (its input.dat) cases, post_process off
REF_RATIO_8T, REF_RATIO_1T = (3.52, 3.73), (1.29, 2.15)
UMAX, RE = 0.05, 200.0


def make_lattice(precision, nx=NX, ny=NY, strip_rows=0, kernel=0, time_blocking=-1, steps_per_pass=0, quad_rows=0):
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
    p.time_blocking = time_blocking
    p.steps_per_pass = steps_per_pass
    p.quad_rows = quad_rows
    lat = capi.Lattice(p)
    rng = np.random.default_rng(3)
    ux = rng.uniform(-1e-4, 1e-4, (nx, ny))
    uy = rng.uniform(-1e-4, 1e-4, (nx, ny))
    lat.set_moments(np.ones((nx, ny)), ux, uy)
    return lat


def run_form(precision, steps, warmup, gate_steps=0, gate_window=0, **kw):
    """One 8192 x 2048 run in one precision form.  With gate_steps > 0 it first advances
    gate_steps steps recording the force series over the last gate_window (the FP32-vs-FP64
    accuracy comparison); then `warmup` untimed and `steps` timed steps continue the same run
    (the timed steps are steps of the run whose accuracy is checked); the velocity field at
    the end."""
    from mrlbm_2d_amd import capi
    lat = make_lattice(precision, **kw)
    it = 2
    fd, fl = [], []
    if gate_steps > 0:
        o = capi.default_step_opts()
        o.post_process = 1
        o.statsbegin = 2 + gate_steps - gate_window
        o.statsend = 1 + gate_steps
        chunk = 2000
        while it <= 1 + gate_steps:
            n = min(chunk, 2 + gate_steps - it)
            lat.step(it, n, o)
            _, a, b = lat.forces()
            fd.append(a)
            fl.append(b)
            it += n
            if (it - 2) % 20000 == 0:
                print(f"[bench] precision {precision}: {it - 2} of {gate_steps} accuracy steps", file=sys.stderr,
                      flush=True)
    lat.step(it, warmup)
    it += warmup
    first_timed = it
    # the timed steps run without per-launch events (a pair on every launch of a 20-step
    # window costs ~3% of its wall clock); the next `steps` steps of the same run carry them,
    # for the launch durations of the roofline
    lat.synchronize()
    lat.set_timing(0)
    t0 = time.perf_counter()
    lat.step(it, steps)
    lat.synchronize()
    wall = time.perf_counter() - t0
    it += steps
    lat.set_timing(1 if steps <= TIME_ALL_UP_TO else TIMED_EVERY)
    lat.step(it, steps)
    lat.synchronize()
    t = lat.last_timing()
    finite = lat.check_finite() == 0
    assert finite, "non-finite state after the timed steps"
    info = lat.info
    out = {}
    if gate_steps > 0:
        _, ux, uy = lat.fields()
        out = dict(fd=np.concatenate(fd), fl=np.concatenate(fl), ux=ux, uy=uy, finite=finite, r_cyl=info.r_cyl)
    lat.close()
    n = info.nx * info.ny
    mlups = n * steps / (wall * 1e6)
    # the timed launches advanced t.timed_steps steps (4, 2 or 1 per four-, two- or single-step
    # launch), t.timed_ms in all
    bytes_per_step = int(info.bytes_per_step)
    achieved = t.timed_steps * bytes_per_step / (t.timed_ms * 1e-3) / 1e9
    return dict(mlups=mlups, wall_s=wall, ms_per_step=1e3 * wall / steps, step_ms=t.step_ms,
                launch_ms=t.bulk_mean_ms, launch_median_ms=t.bulk_median_ms, launch_min_ms=t.bulk_min_ms,
                launch_max_ms=t.bulk_max_ms, timed_launches=t.timed_launches, bulk_launches=t.bulk_launches,
                steps_per_launch=t.bulk_steps / max(t.bulk_launches, 1), timed_steps=t.timed_steps,
                bytes_per_step=bytes_per_step, achieved_gbs=achieved, frac=achieved / HBM_PEAK_GBS,
                first_timed_step=first_timed, pair_rows=info.pair_rows, quad_rows=info.quad_rows), out


def roofline(r, kernel, traffic, source, valu=None):
    """Roofline of the dominant kernel: algorithmic bytes of the timed launches (12 moment
    words per node and step) over their summed duration, against 8 TB/s; traffic = HBM bytes
    of one launch from the committed PMC profile (None for a launch shape not profiled), and
    traffic_frac = those bytes over the mean launch time, against 8 TB/s; valu_issue = the
    profiled share of SIMD cycles issuing VALU instructions (the four-step kernel's bound)."""
    tgbs = traffic / (r["launch_ms"] * 1e-3) / 1e9 if traffic else None
    return {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(r["frac"], 4), "traffic": traffic,
            "traffic_source": (source + " (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE; not measured in "
                               "this run)") if traffic else None,
            "traffic_gbs": round(tgbs, 1) if tgbs else None,
            "traffic_frac": round(tgbs / HBM_PEAK_GBS, 4) if tgbs else None,
            "valu_issue": valu,
            "valu_source": (source + " (rocprofv3 --pmc SQ_INSTS_VALU, GRBM_GUI_ACTIVE; not measured in this "
                            "run)") if valu else None,
            "kernel": kernel, "steps_per_launch": r["steps_per_launch"],
            "algorithmic_bytes_per_launch": int(round(r["steps_per_launch"] * r["bytes_per_step"])),
            "launch_ms": round(r["launch_ms"], 5), "launch_median_ms": round(r["launch_median_ms"], 5),
            "launch_min_ms": round(r["launch_min_ms"], 5), "launch_max_ms": round(r["launch_max_ms"], 5),
            "timed_launches": r["timed_launches"], "bulk_launches": r["bulk_launches"],
            "timed_steps": r["timed_steps"]}


def precision_gate(run64, run32, steps, window):
    """FP32 deviation form against FP64 from the same seeded state (the two runs of run_form):
    mean Cd, rms Cl and the lift spectrum's St over the window, velocity L2 at the end."""
    from mrlbm_2d_amd import compare
    s64, s32 = compare.summarize(run64, UMAX), compare.summarize(run32, UMAX)
    du = np.sqrt(np.sum((run64["ux"] - run32["ux"]) ** 2 + (run64["uy"] - run32["uy"]) ** 2))
    nu = np.sqrt(np.sum(run64["ux"] ** 2 + run64["uy"] ** 2))
    out = {"steps": steps, "window": window,
           "cd_rel_err": abs(s32["cd_mean"] - s64["cd_mean"]) / abs(s64["cd_mean"]),
           "cl_rms_rel_err": abs(s32["cl_rms"] - s64["cl_rms"]) / abs(s64["cl_rms"]),
           "st_spectral_rel_err": abs(s32["st_spectral"] - s64["st_spectral"]) / s64["st_spectral"],
           "l2_u_rel": float(du / nu),
           "fd_series_max_rel": float(np.max(np.abs(run64["fd"] - run32["fd"])) / np.max(np.abs(run64["fd"]))),
           "cd_fp64": s64["cd_mean"], "cd_fp32": s32["cd_mean"],
           "cl_rms_fp64": s64["cl_rms"], "cl_rms_fp32": s32["cl_rms"],
           "st_spectral_fp64": s64["st_spectral"], "st_spectral_fp32": s32["st_spectral"],
           "finite": bool(run64["finite"] and run32["finite"]),
           "threshold": 1e-3,
           "velocity_field_at": "the end of the timed steps, both forms at the same step"}
    out["pass"] = bool(out["finite"] and out["cd_rel_err"] < 1e-3)
    return out


def cpu_baseline(max_steps=10, budget_s=8.0):
    """Wall time per step of the oracle (the CPU restatement of the reference step; its
    per-node sweeps on OpenMP threads, as the reference's) on the same 8192 x 2048 grid:
    one untimed step, then timed steps until `budget_s` seconds or `max_steps` steps.  The
    reference's own geometry (ncy = 128, L_front/back/top/bot = 8/55/8/7, 2d_mrlbm.f90:86-92)
    and its uniform inlet; the per-step work does not depend on the inlet profile."""
    from oracle import pyoracle as po
    p = po.OracleParams(nx=NX, ny=NY, ncy=128, L_front=8, L_back=55, L_top=8, L_bot=7, Re=RE, uo=UMAX,
                        post_process=0, iplot=10 ** 8, statsbegin=10 ** 8, statsend=10 ** 8 + 1)
    o = po.Oracle(p)
    assert (o.nx, o.ny) == (NX, NY)
    threads = po.num_threads()
    o.run(2, 1)
    print(f"[bench] CPU baseline: untimed step done ({threads} threads)", file=sys.stderr, flush=True)
    times = []
    while len(times) < max_steps and sum(times) < budget_s:
        t0 = time.perf_counter()
        o.run(3 + len(times), 1)
        times.append(time.perf_counter() - t0)
        print(f"[bench] CPU baseline: step {len(times)} {times[-1]:.2f}s", file=sys.stderr, flush=True)
    per_step = sum(times) / len(times)
    del o
    return dict(value=round(NX * NY / (per_step * 1e6), 3), unit="MLUPS", cores=threads, kind="port",
                sample=f"oracle (C restatement of the reference step, gcc -O2, OpenMP per-node sweeps on "
                       f"{threads} threads, boundary work serial as in the reference) on the same 8192x2048 "
                       f"grid, {len(times)} timed steps after one untimed step, {per_step:.3f} s/step; "
                       f"against the reference itself (flang -O3 -fopenmp build of 2d_mrlbm.f90, serial "
                       f"streaming) on 400x100 and 920x420 in the build container the oracle ran "
                       f"{REF_RATIO_8T[0]}-{REF_RATIO_8T[1]}x faster on 8 threads and "
                       f"{REF_RATIO_1T[0]}-{REF_RATIO_1T[1]}x on 1 thread")


def adjoint_bench():
    """The 512 x 128 drag-adjoint case in both precision forms: wall time of the checkpointed
    forward run and of the reverse sweep (recomputation + adjoint), and the gradients'
    agreement between the forms."""
    from mrlbm_2d_amd import brinkman as B, capi
    c = B.ADJ_CASE
    chi = B.adjoint_case_field()
    out, grads = {}, {}
    for name, prec in (("fp64", capi.MRLBM_FP64), ("fp32", capi.MRLBM_FP32_DEV)):
        B.run_adjoint_case(prec, chi)  # warm-up (allocations, code load)
        res, g, _ = B.run_adjoint_case(prec, chi)
        grads[name] = g
        out[name] = {"J": res.J, "dJ_duo": res.dJ_duo,
                     "forward_ms_per_step": round(res.forward_ms / c["nsteps"], 5),
                     "adjoint_ms_per_step": round(res.adjoint_ms / c["nsteps"], 5),
                     "adjoint_over_forward": round(res.adjoint_ms / res.forward_ms, 3)}
        print(f"[bench] adjoint {name}: {out[name]}", file=sys.stderr, flush=True)
    out["grad_rel_l2_fp32_vs_fp64"] = float(np.linalg.norm(grads["fp32"] - grads["fp64"]) / np.linalg.norm(grads["fp64"]))
    out["workload"] = ("512x128 channel, cylinder D=24 at (100,64.5), Re=100, u_max=0.05, 4000 steps, "
                       "J = mean Cd over the last 2000, checkpoints every 200, seeded Brinkman bumps")
    return out


def batch_bench(nruns=256, steps=20000, window=10000, seed=0, sample=(0, 1, 2)):
    """256 independent FP32 cylinder runs of uneven size in one batch: aggregate MLUPS over
    `steps` steps of every run, per-run mean Cd over the last `window` steps, and the same Cd
    from a few of the runs stepped alone."""
    from mrlbm_2d_amd import batch_case as BC, capi
    ps = BC.uneven_runs(nruns, seed=seed)
    t0 = time.perf_counter()
    b = capi.Batch(ps)
    setup_s = time.perf_counter() - t0
    info = b.info
    tune = [{"kernel": info.tune_kernel[i], "persistent": info.tune_persistent[i], "time_blocking": info.tune_tb[i],
             "quad_rows": info.tune_quad[i], "ms_per_step": round(info.tune_ms[i], 4)} for i in range(info.ntune)]
    print(f"[bench] batch: {nruns} runs, {info.nodes} nodes, kernel {info.kernel} persistent "
          f"{info.persistent} time_blocking {info.time_blocking} four-step rows {info.quad_rows}, "
          f"setup+autotune {setup_s:.1f} s", file=sys.stderr, flush=True)
    for r in range(nruns):
        BC.seed_state(b.run(r), seed, r)
    t0 = time.perf_counter()
    b.step(steps - window)
    b.step(window, record_forces=True)
    wall = time.perf_counter() - t0
    fd, fl, fl2, cnt = b.forces()
    r_cyl = [b.run(r).info.r_cyl for r in range(nruns)]
    omega = [b.run(r).info.omega for r in range(nruns)]
    cd = [BC.cd_of(ps[r], r_cyl[r], fd[r] / cnt[r]) for r in range(nruns)]
    b.close()
    print(f"[bench] batch: {steps} steps in {wall:.2f} s", file=sys.stderr, flush=True)
    cmp = []
    for r in sample:
        with capi.Lattice(ps[r]) as lat:
            BC.seed_state(lat, seed, r)
            lat.step(2, steps - window)
            o = capi.default_step_opts()
            o.post_process = 1
            o.statsbegin = 2 + steps - window
            o.statsend = 1 + steps
            lat.step(2 + steps - window, window, o)
            _, sfd, _ = lat.forces()
            s = 0.0
            for v in sfd:
                s += v
            cd1 = BC.cd_of(ps[r], lat.info.r_cyl, s / len(sfd))
        cmp.append({"run": r, "nx": ps[r].nx, "Re": round(ps[r].Re, 2), "cd_batch": cd[r], "cd_single": cd1})
    finite = [c for c in cd if np.isfinite(c)]
    diverged = [{"run": r, "nx": ps[r].nx, "D": round(2 * ps[r].r_c, 2), "Re": round(ps[r].Re, 1),
                 "omega": round(omega[r], 4)} for r in range(nruns) if not np.isfinite(cd[r])]
    return {"workload": f"{nruns} FP32 cylinder runs, nx in 256..2048 step 64, ny = nx/4, D in [ny/8, ny/4] "
                        f"at (nx/5, ny/2+0.5), Re in [40, 300], u_max = 0.05, seed {seed}; {steps} steps, "
                        f"Cd over the last {window}",
            "nodes": int(info.nodes), "value": round(info.nodes * steps / wall / 1e6, 1), "unit": "MLUPS",
            "wall_s": round(wall, 3), "kernel": info.kernel, "persistent": info.persistent,
            "time_blocking": info.time_blocking, "quad_rows": info.quad_rows,
            "graph_steps": info.graph_steps, "autotune": tune, "setup_s": round(setup_s, 2),
            "cd_finite_runs": len(finite), "cd_min": min(finite) if finite else None,
            "cd_max": max(finite) if finite else None, "diverged": diverged, "single_run_check": cmp}


def stats_window_bench(steps=400, warmup=40):
    """Cost of a step inside the statistics window at 8192 x 2048, as the drop-in driver runs
    it there: forces, surface pressure, four probes and the u statistics recorded every
    step, per precision form; beside it the same lattice stepped with nothing recorded."""
    from mrlbm_2d_amd import capi
    out = {"workload": f"8192x2048 benchmark channel, {steps} steps after {warmup}: forces, surface pressure, "
                       "4 probes (two on nodes the multi-step bulk writes) and u statistics every step"}
    for name, prec in (("fp64", capi.MRLBM_FP64), ("fp32", capi.MRLBM_FP32_DEV)):
        lat = make_lattice(prec)
        try:
            it = 2
            lat.step(it, warmup)
            it += warmup
            lat.synchronize()
            t0 = time.perf_counter()
            lat.step(it, steps)
            lat.synchronize()
            plain = (time.perf_counter() - t0) / steps
            it += steps
            o = capi.default_step_opts()
            o.post_process, o.statsbegin, o.statsend = 1, it, 10 ** 8
            o.record_surface_p, o.stats_interval = 1, 1
            o.nprobe = 4
            for k, (i, j) in enumerate(((1160, 1026), (4000, 1026), (2000, 3), (8192, 700))):
                o.probe_i[k], o.probe_j[k] = i, j
            lat.step(it, warmup, o)
            it += warmup
            lat.synchronize()
            t0 = time.perf_counter()
            lat.step(it, steps, o)
            lat.synchronize()
            win = (time.perf_counter() - t0) / steps
            assert lat.check_finite() == 0, "non-finite state in the statistics window"
        finally:
            lat.close()
        out[name] = {"ms_per_step_window": round(1e3 * win, 5), "ms_per_step_plain": round(1e3 * plain, 5),
                     "window_over_plain": round(win / plain, 3)}
        print(f"[bench] statistics window {name}: {out[name]}", file=sys.stderr, flush=True)
    return out


def copy_calibration(nbytes, reps):
    """A plain device copy of `nbytes` (torch's copy kernel), `reps` times: the known-size
    transfer that scales the HBM byte counters (FETCH_SIZE, WRITE_SIZE) of a profiled run,
    and the read + write rate a plain copy reaches (HIP events around the reps, after three
    untimed copies)."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        y.copy_(x)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps, "ms_per_copy": round(ms, 5),
                      "copy_gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-stats-window", action="store_true")
    ap.add_argument("--gate-steps", type=int, default=100000,
                    help="steps of the FP32-vs-FP64 accuracy run (0 skips it)")
    ap.add_argument("--gate-window", type=int, default=50000)
    ap.add_argument("--strip-rows", type=int, default=0)
    ap.add_argument("--time-blocking", type=int, default=-1,
                    help="-1 auto (multi-step passes on), 0 one step per pass, 2..16 two-step strip rows")
    ap.add_argument("--steps-per-pass", type=int, default=0, choices=(0, 2, 4),
                    help="0 or 4: four steps per bulk pass where they apply (default), 2: two")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device copy of BYTES (for profiler counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: the step runs on one GPU; a multi-GPU slab decomposition along x is "
              "designed but not built. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, max(args.steps, 1))
        return
    from mrlbm_2d_amd import capi
    gate_steps = max(args.gate_steps, 0)
    gate_window = min(args.gate_window, gate_steps)
    kw = dict(strip_rows=args.strip_rows, time_blocking=args.time_blocking, steps_per_pass=args.steps_per_pass)
    # FP64 first, then FP32 (the headline): each form's timed steps continue its own run
    # right after the accuracy steps
    print(f"[bench] FP64 plain form {NX}x{NY}: {gate_steps} accuracy steps (window {gate_window}), warmup "
          f"{args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    r64, run64 = run_form(capi.MRLBM_FP64, args.steps, args.warmup, gate_steps, gate_window, **kw)
    print(f"[bench] FP64: {r64}", file=sys.stderr, flush=True)
    print(f"[bench] FP32 deviation form {NX}x{NY}: same", file=sys.stderr, flush=True)
    r32, run32 = run_form(capi.MRLBM_FP32_DEV, args.steps, args.warmup, gate_steps, gate_window, **kw)
    print(f"[bench] FP32: {r32}", file=sys.stderr, flush=True)
    gate = precision_gate(run64, run32, gate_steps, gate_window) if gate_steps > 0 else None
    del run64, run32
    print(f"[bench] FP32 vs FP64: {gate}", file=sys.stderr, flush=True)
    window = None
    if not args.no_stats_window:
        print("[bench] statistics window", file=sys.stderr, flush=True)
        window = stats_window_bench()
    adj = None
    if not args.no_adjoint:
        print("[bench] drag adjoint, 512x128", file=sys.stderr, flush=True)
        adj = adjoint_bench()
    batch = None
    if not args.no_batch:
        print("[bench] batched runs", file=sys.stderr, flush=True)
        batch = batch_bench()
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline: oracle", file=sys.stderr, flush=True)
        cpu = cpu_baseline()
        print(f"[bench] CPU: {cpu}", file=sys.stderr, flush=True)
    tb, spp = args.time_blocking, (4 if args.steps_per_pass in (0, 4) else 2)
    if tb == 0:
        spp = 1
        k32, k64 = "bulk_kernel<float,float,false,304>", "bulk_kernel<double,double,false,8>"
    elif spp == 4:
        k32 = f"bulkn_kernel<float,float,4,{r32['quad_rows']},false>"
        k64 = f"bulkn_kernel<double,double,4,{r64['quad_rows']},false>"
    else:
        k32 = f"bulk2_kernel<float,float,{r32['pair_rows']},false>"
        k64 = f"bulk2_kernel<double,double,{r64['pair_rows']},false>"
    # measured traffic applies to the profiled launch shapes only
    traffic, source, shape = TRAFFIC.get(spp, ({"f32": None, "f64": None}, None, None))
    valu = VALU_ISSUE.get(spp, ({"f32": None, "f64": None}, None, None))[0]
    rows = (r32["quad_rows"], r64["quad_rows"]) if spp == 4 else (r32["pair_rows"], r64["pair_rows"])
    if args.strip_rows != 0 or tb not in (-1,) or rows != shape:
        traffic = valu = {"f32": None, "f64": None}
    out = {
        # the headline metric: MLUPS of the FP32 deviation form (value), its % of HBM peak
        # (roofline), the FP64 form beside it, the |dCd|/Cd condition (fp32_vs_fp64) and the
        # adjoint / forward time (adjoint)
        "metric": "MLUPS + % HBM peak, FP32 vs FP64 at |\u0394Cd|/Cd<1e-3; adjoint/forward time",
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
        "config": {"workload": "8192x2048 channel, cylinder D=128 at (1024,1025), "
                               "parabolic inlet u_max=0.05, Re=200", "nx": NX, "ny": NY},
        "roofline": roofline(r32, k32, traffic["f32"], source, valu["f32"]),
        "fp64": {"value": round(r64["mlups"], 1), "ms_per_step": round(r64["ms_per_step"], 5),
                 "roofline": roofline(r64, k64, traffic["f64"], source, valu["f64"])},
        "fp32_over_fp64": round(r32["mlups"] / r64["mlups"], 3),
        "fp32_vs_fp64": gate,
        # the timed steps continue each form's accuracy run from this step number
        "first_timed_step": r32["first_timed_step"],
        "statistics_window": window,
        "adjoint": adj,
        "batched": batch,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
