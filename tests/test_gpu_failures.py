"""Failure detection through the C-ABI and the drop-in driver on the GPU, and calls that must
leave a handle as they found it.

* NaN fed through mrlbm_set_moments is reported by mrlbm_check_finite (MRLBM_E_NONFINITE),
  before and after stepping; mrlbm_max_ux ignores it.
* The device scan behind mrlbm_check_finite / mrlbm_max_ux agrees with the host-side
  maximum of the fields it returns.
* The driver stops with exit status 3 at the first log step of a non-finite state (a NaN
  population in restart.dat), before writing a field or restart file of that state.
* mrlbm_adjoint_drag leaves the running inlet-density mean and the multi-step passes as they
  were (it installs a zero Brinkman field for its sweep and removes it again).
"""
import os
import subprocess

import numpy as np
import pytest

from mrlbm_2d_amd import brinkman as B
from mrlbm_2d_amd import build as BLD
from mrlbm_2d_amd import capi
from tests.golden import make_fixtures as mf

pytestmark = pytest.mark.gpu


def _lattice(precision=capi.MRLBM_FP64, nx=88, ny=40):
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic = p.y_periodic = 0
    p.x_c, p.y_c, p.r_c = 21.0, 20.5, 4.0
    p.Re, p.uo = 40.0, 0.05
    p.inlet_profile = capi.INLET_PARABOLIC
    p.label_policy = capi.LABEL_FLUID
    p.precision = precision
    return capi.Lattice(p)


@pytest.mark.parametrize("precision", [capi.MRLBM_FP64, capi.MRLBM_FP32_DEV])
def test_nan_through_set_moments_is_reported(precision):
    print(f"[gpu] NaN through set_moments, precision {precision}", flush=True)
    with _lattice(precision) as lat:
        nx, ny = lat.nx, lat.ny
        rng = np.random.default_rng(1)
        ux = rng.uniform(-1e-3, 1e-3, (nx, ny))
        lat.set_moments(np.ones((nx, ny)), ux, np.zeros((nx, ny)))
        lat.step(2, 4)
        assert lat.check_finite() == 0
        rho, ux, uy = lat.fields()
        assert lat.max_ux() == ux.max()
        ux[40, 10] = np.nan
        lat.set_moments(rho, ux, uy)
        lat.step(6, 1)
        assert lat.check_finite() == capi.E_NONFINITE
        assert np.isfinite(lat.max_ux())
        lat.step(7, 10)
        assert lat.check_finite() == capi.E_NONFINITE


def test_device_scan_matches_host_fields_at_8192x2048():
    import bench
    print("[gpu] device scan vs host fields, 8192x2048 FP32", flush=True)
    lat = bench.make_lattice(capi.MRLBM_FP32_DEV)
    try:
        lat.step(2, 40)
        assert lat.check_finite() == 0
        _, ux, _ = lat.fields()
        assert lat.max_ux() == ux.max()
    finally:
        lat.close()


def test_driver_stops_on_a_non_finite_state(tmp_path):
    d = str(tmp_path)
    for sub in ("output", "restart", "data_probe", "data_geo", "data_mean"):
        os.makedirs(os.path.join(d, sub))
    p = dict(mf.BASE, ncy=8, L_front=2, L_back=4, L_top=2, L_bot=2, post_process=False)
    nx, ny = mf.grid_size(p)
    f, ux, uy = mf.noisy_state(nx, ny, 3)
    f[30, 20, 0] = np.nan
    path = os.path.join(d, "restart", "restart.dat")
    mf.write_restart(path, 1, f, ux, uy, np.ones((nx, ny)))
    before = open(path, "rb").read()
    mf.write_input(os.path.join(d, "input.dat"), p, 301, 1)
    print("[gpu] driver on a NaN restart state", flush=True)
    r = subprocess.run([BLD.DRIVER], cwd=d, capture_output=True, text=True, timeout=120)
    assert r.returncode == 3, r.stdout + r.stderr
    assert "non-finite state" in r.stdout and "100" in r.stdout.split("non-finite")[0].splitlines()[-1]
    assert open(path, "rb").read() == before          # no restart of the NaN state
    assert not [n for n in os.listdir(os.path.join(d, "output")) if n.startswith("grid1")]


@pytest.mark.parametrize("precision", [capi.MRLBM_FP64, capi.MRLBM_FP32_DEV])
def test_adjoint_leaves_the_handle_as_it_was(precision):
    c = dict(B.ADJ_CASE, nsteps=400, window=200, K=100)
    print(f"[gpu] adjoint side effects, precision {precision}", flush=True)
    with capi.Lattice(B.adjoint_params(precision, c)) as lat:
        o = capi.default_step_opts()
        o.post_process, o.statsbegin, o.statsend = 1, 2, 10 ** 7
        lat.step(2, 300, o)
        lat.forces()
        rim = lat.rho_inf_mean()
        lat.adjoint_drag(302, c["nsteps"], c["window"], c["K"])
        assert lat.rho_inf_mean() == rim
        lat.set_timing(1)
        lat.step(702, 20)
        t = lat.last_timing()
        # four steps per bulk pass again (a Brinkman field would force single steps)
        assert t.bulk_launches == 5 and t.bulk_steps == 20
        assert lat.check_finite() == 0
