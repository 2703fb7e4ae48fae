"""Post-processing of a run's force and surface-pressure series: the algorithm of the
reference's separate program 2d_statistics.f90 (statistics_calculation).

* period and cycle start from the first two local maxima of the lift series
  (find_peaks, 2d_statistics.f90:183-214; rows are 1-based positions in the series);
* n_cycle whole periods after the first peak, St = D/(period*uo) (:92-98);
* running means over rows [cycle_start, cycle_start + n_cycle*period] of the surface
  pressure, Cd = Fd/f_norm and Cl^2, Cl_rms = sqrt(mean Cl^2) (:114-131), with
  f_norm = rho_infty uo^2 D / 2 and p_norm = rho_infty uo^2 / 2, D = 2 r_cy (:29-32);
* Cp(theta) of the averaged pressure, written for theta <= pi as |theta - 180 deg|
  (:145-165).

Raises StatisticsError where the reference prints "Could not detect two peaks." and stops.

Drop-in for the reference's post-processing program, run in a run directory:

    python -m mrlbm_2d_amd.statistics [run_dir]

reads input_stat.dat (or data_mean/input_stat.dat, which the drop-in driver writes from the
run's nbct, r_cyl, uo and inlet-density mean), data_mean/forces_time.dat, theta.dat and the unformatted
p_surface{,_smooth}_time.dat, and writes plots/cp.dat, cp_fit.dat, cd_time.dat,
cl_time.dat and drag_lift.dat with the same rows (2d_statistics.f90:29-178).
"""
from __future__ import annotations

import math
import os
import sys
from dataclasses import dataclass

import numpy as np


class StatisticsError(RuntimeError):
    pass


@dataclass
class Statistics:
    cycle_start: int
    period: int
    n_cycle: int
    st: float
    cd: float
    cl_rms: float
    cp: np.ndarray | None = None
    cp_fit: np.ndarray | None = None


def find_peaks(signal) -> tuple[int, int]:
    """(cycle_start, period) from the first two strict local maxima, 1-based rows."""
    s = np.asarray(signal, dtype=np.float64)
    p1 = p2 = -1
    for i in range(2, len(s)):          # rows 2 .. n-1
        if s[i - 1] > s[i - 2] and s[i - 1] > s[i]:
            if p1 == -1:
                p1 = i
            else:
                p2 = i
                break
    if p1 == -1 or p2 == -1:
        raise StatisticsError("Could not detect two peaks.")
    return p1, p2 - p1


def compute(fd, fl, uo: float, rho_infty: float, r_cy: float, ps=None, ps_fit=None) -> Statistics:
    fd = np.asarray(fd, dtype=np.float64)
    fl = np.asarray(fl, dtype=np.float64)
    D = 2.0 * r_cy
    p_infty = rho_infty / 3.0
    p_norm = (0.50 * rho_infty) * (uo * uo)
    f_norm = ((0.50 * rho_infty) * (uo * uo)) * D
    cs, period = find_peaks(fl)
    n_cycle = (len(fd) - cs) // period
    ce = cs + n_cycle * period
    st = D / (float(period) * uo)
    cd_avg = 0.0
    cl_avg = 0.0
    ps_avg = ps_fit_avg = None
    if ps is not None:
        ps = np.asarray(ps, dtype=np.float64)
        ps_fit = np.asarray(ps_fit, dtype=np.float64)
        ps_avg = np.zeros(ps.shape[1])
        ps_fit_avg = np.zeros(ps.shape[1])
    for i in range(cs, ce + 1):         # running means in the reference's form
        mc = float(i - cs)
        cd = fd[i - 1] / f_norm
        cl = fl[i - 1] / f_norm
        cd_avg = (mc * cd_avg + cd) / (mc + 1.0)
        cl_avg = (mc * cl_avg + cl * cl) / (mc + 1.0)
        if ps is not None:
            ps_avg = (mc * ps_avg + ps[i - 1]) / (mc + 1.0)
            ps_fit_avg = (mc * ps_fit_avg + ps_fit[i - 1]) / (mc + 1.0)
    out = Statistics(cs, period, n_cycle, st, cd_avg, math.sqrt(cl_avg))
    if ps is not None:
        out.cp = (ps_avg - p_infty) / p_norm
        out.cp_fit = (ps_fit_avg - p_infty) / p_norm
    return out


def cp_profile(theta, cp):
    """Rows of plots/cp.dat: (|theta*180/pi - 180|, Cp) for theta <= pi (2d_statistics.f90:153-158)."""
    theta = np.asarray(theta, dtype=np.float64)
    keep = theta <= math.pi
    return np.stack([np.abs((theta[keep] * 180.0 / math.pi) - 180.0), np.asarray(cp)[keep]], axis=1)


def read_unformatted_series(path, nrec):
    """`nrec` records of a Fortran sequential unformatted file whose records are
    (int32 iteration, float64 values...) between 4-byte length markers (:100-111)."""
    raw = open(path, "rb").read()
    iters, rows, pos = [], [], 0
    for _ in range(nrec):
        m = int(np.frombuffer(raw[pos:pos + 4], np.int32)[0])
        rec = raw[pos + 4:pos + 4 + m]
        iters.append(int(np.frombuffer(rec[:4], np.int32)[0]))
        rows.append(np.frombuffer(rec[4:], np.float64).copy())
        pos += 8 + m
    return np.array(iters, np.int32), np.array(rows)


def _fmt(v):
    return f" {v!r}" if isinstance(v, float) else f" {v}"


def run(directory="."):
    """The reference post-processor on the files of one run directory."""
    from .namelist import read_input_stat
    path = os.path.join(directory, "input_stat.dat")
    if not os.path.exists(path):  # the one the drop-in driver writes at the end of a run
        path = os.path.join(directory, "data_mean", "input_stat.dat")
    prm = read_input_stat(path)
    nbct, r_cy, uo, rho_infty = int(prm["nbct"]), float(prm["r_cy"]), float(prm["uo"]), float(prm["rho_infty"])
    dm = os.path.join(directory, "data_mean")
    rows = [l.split() for l in open(os.path.join(dm, "forces_time.dat")) if l.strip()]
    nstep = len(rows)
    fd = np.array([float(r[1]) for r in rows])
    fl = np.array([float(r[2]) for r in rows])
    theta = np.array([float(l.split()[0]) for l in open(os.path.join(dm, "theta.dat")) if l.strip()][:nbct])
    _, ps = read_unformatted_series(os.path.join(dm, "p_surface_time.dat"), nstep)
    _, pf = read_unformatted_series(os.path.join(dm, "p_surface_smooth_time.dat"), nstep)
    r = compute(fd, fl, uo, rho_infty, r_cy, ps=ps[:, :nbct], ps_fit=pf[:, :nbct])
    f_norm = ((0.50 * rho_infty) * (uo * uo)) * (2.0 * r_cy)
    out = os.path.join(directory, "plots")
    os.makedirs(out, exist_ok=True)
    for name, cp in (("cp.dat", r.cp), ("cp_fit.dat", r.cp_fit)):
        with open(os.path.join(out, name), "w") as fh:
            for a, b in cp_profile(theta, cp):
                fh.write(_fmt(float(a)) + _fmt(float(b)) + "\n")
    ce = r.cycle_start + r.n_cycle * r.period
    for name, sig in (("cd_time.dat", fd), ("cl_time.dat", fl)):
        with open(os.path.join(out, name), "w") as fh:
            for i in range(r.cycle_start, ce + 1):
                # real(i - cycle_start) / real(period) is a REAL(4) quotient
                x = float(np.float32(i - r.cycle_start) / np.float32(r.period))
                fh.write(_fmt(x) + _fmt(float(sig[i - 1] / f_norm)) + "\n")
    lines = [("Strouhal number (St)  ", float(r.st)), ("Drag coefficinet (C_D)  ", float(r.cd)),
             ("Lift coefficinet (C_L ", float(r.cl_rms))]
    with open(os.path.join(out, "drag_lift.dat"), "w") as fh:
        for k, v in lines:
            fh.write(f" {k}:{v!r}\n")
    print(f"Averging over {r.n_cycle} cycle(s)")
    for k, v in lines:
        print(f" {k}:{v!r}")
    return r


if __name__ == "__main__":
    try:
        run(sys.argv[1] if len(sys.argv) > 1 else ".")
    except StatisticsError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
