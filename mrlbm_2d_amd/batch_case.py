"""Seeded sets of independent cylinder runs for the batched mode.

Each run: nx drawn from `sizes`, ny = nx/4, diameter D uniform in [ny/8, ny/4] centred at
(nx/5, ny/2 + 0.5), Re uniform in [40, 300], u_max = 0.05 parabolic inlet, walls at the
channel sides, initial velocity noise +-1e-4 seeded per run.
"""
from __future__ import annotations

import numpy as np

from . import capi


def uneven_runs(n=256, seed=0, precision=capi.MRLBM_FP32_DEV, sizes=tuple(range(256, 2049, 64)), umax=0.05):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        nx = int(rng.choice(sizes))
        ny = nx // 4
        D = float(rng.uniform(ny / 8.0, ny / 4.0))
        re = float(rng.uniform(40.0, 300.0))
        p = capi.default_params()
        p.nx, p.ny = nx, ny
        p.x_periodic, p.y_periodic = 0, 0
        p.channel_with_cylinder = 1
        p.x_c, p.y_c, p.r_c = nx / 5.0, ny / 2.0 + 0.5, D / 2.0
        p.Re, p.uo = re, umax
        p.inlet_profile = capi.INLET_PARABOLIC
        p.precision = precision
        p.label_policy = capi.LABEL_FLUID  # non-integer centres give labels the reference leaves unset
        out.append(p)
    return out


def seed_state(lat, seed, run):
    """rho = 1 and u uniform in +-1e-4 from a per-run stream."""
    rng = np.random.default_rng([seed, run])
    ux = rng.uniform(-1e-4, 1e-4, (lat.nx, lat.ny))
    uy = rng.uniform(-1e-4, 1e-4, (lat.nx, lat.ny))
    lat.set_moments(np.ones((lat.nx, lat.ny)), ux, uy)


def cd_of(params, r_cyl, fd_mean):
    uref = (2.0 / 3.0) * params.uo if params.inlet_profile == capi.INLET_PARABOLIC else params.uo
    return fd_mean / (0.5 * uref * uref * 2.0 * r_cyl)
