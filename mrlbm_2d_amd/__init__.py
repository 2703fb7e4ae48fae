"""mrlbm_2d_amd -- MI355X (gfx950) D2Q9 moments-based regularized LBM time step.

The hot path of sakthi-lbm/mrlbm_2d (the main-loop body of 2d_mrlbm.f90:229-535) as
hand-written HIP kernels behind a C-ABI (include/mrlbm.h, libmrlbm.so).  This package
holds the ctypes binding (capi), the reference's namelist reader (namelist) and the
helpers that turn the reference's run configuration into library parameters.
"""
from __future__ import annotations

from . import capi
from .capi import Lattice, MrlbmError, Params, StepOpts  # noqa: F401
from .namelist import RunConfig, read_input, read_input_stat  # noqa: F401

__all__ = ["capi", "Lattice", "MrlbmError", "Params", "StepOpts", "params_from_namelist",
           "step_opts_from_namelist", "read_input", "read_input_stat"]


def params_from_namelist(p: dict, precision: int = capi.MRLBM_FP64) -> Params:
    """Library parameters for a run described by the reference's namelist values
    (keys as in input.dat, case-insensitive): grid and centre follow 2d_mrlbm.f90:86-92
    and 946-948, the BC switches follow &LogicalControls."""
    q = {k.lower(): v for k, v in p.items()}
    cfg = RunConfig({"cylinder": {"ncy": q.get("ncy", 0), "l_front": q.get("l_front", 0),
                                  "l_back": q.get("l_back", 0), "l_top": q.get("l_top", 0),
                                  "l_bot": q.get("l_bot", 0)},
                     "domainsize": {"nx": q.get("nx", 0), "ny": q.get("ny", 0)},
                     "logicalcontrols": {"channel_with_cylinder": q.get("channel_with_cylinder", True)}})
    par = capi.default_params()
    par.nx, par.ny = cfg.grid()
    par.channel_with_cylinder = 1 if cfg.channel_with_cylinder else 0
    if cfg.channel_with_cylinder:
        par.x_c, par.y_c, par.r_c = cfg.centre()
    par.Re = float(q["re"])
    par.uo = float(q["uo"])
    par.x_periodic = int(bool(q.get("x_periodic", False)))
    par.y_periodic = int(bool(q.get("y_periodic", True)))
    par.rotated_coordinate = int(bool(q.get("rotated_coordinate", True)))
    par.vel_interp = int(bool(q.get("vel_interp", True)))
    par.mom_interp = int(bool(q.get("mom_interp", True)))
    par.incomp_bc = int(bool(q.get("incomp_bc", False)))
    par.incomp_bulk = int(bool(q.get("incomp_bulk", False)))
    if q.get("eq_mass_bc", False):
        par.mass_bc = capi.MASS_EQ
    elif q.get("weak_mass_bc", False):
        par.mass_bc = capi.MASS_WEAK
    else:
        par.mass_bc = capi.MASS_STRONG
    par.precision = precision
    return par


def step_opts_from_namelist(p: dict, record_surface_p: bool = True) -> StepOpts:
    q = {k.lower(): v for k, v in p.items()}
    o = capi.default_step_opts()
    o.post_process = int(bool(q.get("post_process", False)))
    o.statsbegin = int(q.get("statsbegin", 1 << 30))
    o.statsend = int(q.get("statsend", -1))
    o.iplot = int(q.get("iplot", 0))
    o.record_surface_p = 1 if record_surface_p else 0
    return o
