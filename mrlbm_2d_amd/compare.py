"""FP32 deviation form against FP64 on one cylinder configuration.

Runs the same case in both precision forms (same seeded initial state), records the
force series every step of the averaging window, and reports mean Cd, rms Cl and St
(2d_statistics algorithm where two lift peaks exist, plain window means otherwise; and St
from the lift spectrum),
the relative L2 difference of the final velocity field and both forms' throughput.

    python -m mrlbm_2d_amd.compare --nx 8192 --ny 2048 --steps 100000 --window 50000
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import capi
from . import statistics as S


def make_params(nx, ny, xc, yc, rc, re, umax, precision, parabolic=True):
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = xc, yc, rc
    p.Re, p.uo = re, umax
    p.inlet_profile = capi.INLET_PARABOLIC if parabolic else capi.INLET_UNIFORM
    p.precision = precision
    return p


def run(params, steps, window, seed, chunk=2000):
    lat = capi.Lattice(params)
    nx, ny = lat.nx, lat.ny
    rng = np.random.default_rng(seed)
    ux = rng.uniform(-1e-4, 1e-4, (nx, ny))
    uy = rng.uniform(-1e-4, 1e-4, (nx, ny))
    lat.set_moments(np.ones((nx, ny)), ux, uy)
    o = capi.default_step_opts()
    o.post_process = 1
    o.statsbegin = 2 + steps - window
    o.statsend = 1 + steps
    o.iplot = 0
    fd, fl = [], []
    t0 = time.perf_counter()
    it = 2
    while it <= 1 + steps:
        n = min(chunk, 2 + steps - it)
        lat.step(it, n, o)
        _, a, b = lat.forces()
        fd.append(a)
        fl.append(b)
        it += n
        if (it - 2) % (10 * chunk) == 0:
            print(f"[compare] precision {params.precision}: {it - 2} of {steps} steps", file=sys.stderr, flush=True)
    lat.synchronize()
    wall = time.perf_counter() - t0
    rho, ux, uy = lat.fields()
    finite = lat.check_finite() == 0
    info = lat.info
    lat.close()
    return dict(fd=np.concatenate(fd), fl=np.concatenate(fl), ux=ux, uy=uy, rho=rho, finite=finite,
                mlups=nx * ny * steps / (wall * 1e6), r_cyl=info.r_cyl)


def spectral_st(fl, D, uref):
    """St from the dominant frequency of the lift series: Hann-windowed spectrum, peak bin
    refined by a parabola through the log magnitudes of its neighbours.  Unlike the first-two-
    maxima rule of 2d_statistics.f90:183-214 it is not thrown off by small high-frequency
    ripple (FP32 rounding puts local maxima on consecutive steps)."""
    x = np.asarray(fl, dtype=np.float64)
    x = x - x.mean()
    n = len(x)
    mag = np.abs(np.fft.rfft(x * np.hanning(n)))
    k = int(np.argmax(mag[1:])) + 1
    delta = 0.0
    if k + 1 < len(mag):
        a, b, c = (np.log(mag[k + d] + 1e-300) for d in (-1, 0, 1))
        den = a - 2.0 * b + c
        delta = 0.5 * (a - c) / den if den != 0.0 else 0.0
    return float((k + delta) / n * D / uref)


def summarize(r, umax, rho_inf=1.0, parabolic=True):
    uref = (2.0 / 3.0) * umax if parabolic else umax
    out = dict(cd_mean=float(np.mean(r["fd"]) / (0.5 * rho_inf * uref * uref * 2 * r["r_cyl"])),
               cl_rms=float(np.sqrt(np.mean(r["fl"] ** 2)) / (0.5 * rho_inf * uref * uref * 2 * r["r_cyl"])),
               st_spectral=spectral_st(r["fl"], 2 * r["r_cyl"], uref))
    try:
        s = S.compute(r["fd"], r["fl"], uref, rho_inf, r["r_cyl"])
        out.update(st=s.st, cd_cycles=s.cd, cl_rms_cycles=s.cl_rms, period=s.period, n_cycle=s.n_cycle)
    except S.StatisticsError:
        out.update(st=None)
    return out


def compare(nx, ny, xc, yc, rc, re, umax, steps, window, seed=3, parabolic=True):
    res = {}
    runs = {}
    for name, prec in (("fp64", capi.MRLBM_FP64), ("fp32", capi.MRLBM_FP32_DEV)):
        r = run(make_params(nx, ny, xc, yc, rc, re, umax, prec, parabolic), steps, window, seed)
        runs[name] = r
        res[name] = dict(summarize(r, umax, parabolic=parabolic), mlups=r["mlups"], finite=r["finite"])
    a, b = runs["fp64"], runs["fp32"]
    du = np.sqrt(np.sum((a["ux"] - b["ux"]) ** 2 + (a["uy"] - b["uy"]) ** 2))
    nu = np.sqrt(np.sum(a["ux"] ** 2 + a["uy"] ** 2))
    res["l2_u_rel"] = float(du / nu)
    res["cd_rel_err"] = abs(res["fp32"]["cd_mean"] - res["fp64"]["cd_mean"]) / abs(res["fp64"]["cd_mean"])
    res["cl_rms_rel_err"] = (abs(res["fp32"]["cl_rms"] - res["fp64"]["cl_rms"]) / abs(res["fp64"]["cl_rms"])
                             if res["fp64"]["cl_rms"] else None)
    if res["fp64"].get("st") and res["fp32"].get("st"):
        res["st_rel_err"] = abs(res["fp32"]["st"] - res["fp64"]["st"]) / res["fp64"]["st"]
    res["st_spectral_rel_err"] = (abs(res["fp32"]["st_spectral"] - res["fp64"]["st_spectral"])
                                  / res["fp64"]["st_spectral"])
    res["fd_series_max_rel"] = float(np.max(np.abs(a["fd"] - b["fd"])) / np.max(np.abs(a["fd"])))
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nx", type=int, default=8192)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--xc", type=float, default=None)
    ap.add_argument("--yc", type=float, default=None)
    ap.add_argument("--rc", type=float, default=None)
    ap.add_argument("--re", type=float, default=200.0)
    ap.add_argument("--umax", type=float, default=0.05)
    ap.add_argument("--steps", type=int, default=100000)
    ap.add_argument("--window", type=int, default=50000)
    ap.add_argument("--seed", type=int, default=3)
    a = ap.parse_args()
    xc = a.xc if a.xc is not None else a.nx / 8.0
    yc = a.yc if a.yc is not None else a.ny / 2.0 + 1.0
    rc = a.rc if a.rc is not None else a.ny / 32.0
    print(json.dumps(compare(a.nx, a.ny, xc, yc, rc, a.re, a.umax, a.steps, a.window, a.seed)), flush=True)


if __name__ == "__main__":
    main()
