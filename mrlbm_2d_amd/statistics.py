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
"""
from __future__ import annotations

import math
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
