"""Seeded Brinkman solid-fraction fields for the drag adjoint, and the 512 x 128 adjoint case.

The field is zero in the fluid plus smooth Gaussian bumps whose centres lie within two
diameters of the cylinder surface; it is cut to zero farther out than that.
"""
from __future__ import annotations

import numpy as np

from . import capi


def gaussian_bumps(nx, ny, xc, yc, radius, amp=0.01, seed=0, nbumps=8, sigma=None):
    """(nx, ny) field of `nbumps` Gaussian bumps of peak `amp` (1-based lattice coordinates,
    as the cylinder centre)."""
    D = 2.0 * radius
    sigma = D / 6.0 if sigma is None else sigma
    rng = np.random.default_rng(seed)
    x = np.arange(1, nx + 1, dtype=np.float64)[:, None]
    y = np.arange(1, ny + 1, dtype=np.float64)[None, :]
    chi = np.zeros((nx, ny))
    for _ in range(nbumps):
        r = rng.uniform(radius + 2.0, radius + 2.0 * D)
        th = rng.uniform(0.0, 2.0 * np.pi)
        cx, cy = xc + r * np.cos(th), yc + r * np.sin(th)
        chi += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2.0 * sigma * sigma))
    dist = np.sqrt((x - xc) ** 2 + (y - yc) ** 2) - radius
    chi[dist > 2.0 * D] = 0.0
    return chi


# The adjoint case: 512 x 128 channel, D = 24 at (100, 64.5), Re = 100, u_max = 0.05,
# 4000 forward steps, J = mean Cd over the last 2000, checkpoints every 200 steps.
ADJ_CASE = dict(nx=512, ny=128, xc=100.0, yc=64.5, radius=12.0, Re=100.0, umax=0.05,
                nsteps=4000, window=2000, K=200, seed=5)


def adjoint_params(precision, case=ADJ_CASE, umax=None, omega=0.0):
    p = capi.default_params()
    p.nx, p.ny = case["nx"], case["ny"]
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = case["xc"], case["yc"], case["radius"]
    p.Re = case["Re"]
    p.uo = case["umax"] if umax is None else umax
    p.omega = omega
    p.inlet_profile = capi.INLET_PARABOLIC
    p.precision = precision
    # the half-integer centre gives wall nodes with labels the reference leaves unset
    p.label_policy = capi.LABEL_FLUID
    return p


def adjoint_case_field(case=ADJ_CASE):
    return gaussian_bumps(case["nx"], case["ny"], case["xc"], case["yc"], case["radius"], 0.01, case["seed"])


def run_adjoint_case(precision, chi=None, case=ADJ_CASE, umax=None, omega=0.0):
    """Forward + adjoint of the case; returns (result, dJ/dchi, omega)."""
    chi = adjoint_case_field(case) if chi is None else chi
    with capi.Lattice(adjoint_params(precision, case, umax, omega)) as lat:
        lat.set_brinkman(chi)
        res, g = lat.adjoint_drag(2, case["nsteps"], case["window"], case["K"])
        return res, g, lat.info.omega


def drag_objective(precision, chi, case=ADJ_CASE, umax=None, omega=0.0):
    """J = mean Cd over the window from a plain forward run of the case."""
    p = adjoint_params(precision, case, umax, omega)
    with capi.Lattice(p) as lat:
        lat.set_brinkman(chi)
        o = capi.default_step_opts()
        o.post_process = 1
        o.statsbegin = 2 + case["nsteps"] - case["window"]
        o.statsend = 1 + case["nsteps"]
        lat.step(2, case["nsteps"], o)
        _, fd, _ = lat.forces()
        uref = (2.0 / 3.0) * p.uo
        return float(np.sum(fd) / case["window"] / (0.5 * uref * uref * 2 * lat.info.r_cyl))
