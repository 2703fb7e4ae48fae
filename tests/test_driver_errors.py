"""Failure handling of the drop-in driver that needs no GPU: a malformed input.dat and a
missing or truncated restart/restart.dat are refused with a message and a nonzero exit
status before any device work (the reference prints and stops on a bad namelist,
2d_mrlbm.f90:2235-2238, and fails inside the Fortran runtime on a short restart file,
:2208-2216).  The run directory is built the way the reference's Makefile builds it."""
import os
import subprocess

import numpy as np
import pytest

from mrlbm_2d_amd import build as B
from tests.golden import make_fixtures as mf

P = dict(mf.BASE, ncy=8, L_front=2, L_back=4, L_top=2, L_bot=2)


def _rundir(tmp_path, text=None, irestart=0):
    d = str(tmp_path)
    for sub in ("output", "restart", "data_probe", "data_geo", "data_mean"):
        os.makedirs(os.path.join(d, sub), exist_ok=True)
    mf.write_input(os.path.join(d, "input.dat"), P, 100, irestart)
    if text is not None:
        with open(os.path.join(d, "input.dat"), "w") as fh:
            fh.write(text)
    return d


def _run(d):
    if not os.path.exists(B.DRIVER):
        pytest.skip("driver not built")
    return subprocess.run([B.DRIVER], cwd=d, capture_output=True, text=True, timeout=120)


def test_malformed_namelist_is_refused(tmp_path):
    d = _rundir(tmp_path)
    txt = open(os.path.join(d, "input.dat")).read().replace("ncy = 8", "ncy = eight")
    with open(os.path.join(d, "input.dat"), "w") as fh:
        fh.write(txt)
    r = _run(d)
    assert r.returncode != 0
    assert "Error while reading input.dat" in r.stdout + r.stderr


def test_missing_input_is_refused(tmp_path):
    d = _rundir(tmp_path)
    os.remove(os.path.join(d, "input.dat"))
    r = _run(d)
    assert r.returncode != 0 and "input.dat" in r.stdout + r.stderr


def test_missing_restart_is_refused(tmp_path):
    d = _rundir(tmp_path, irestart=1)
    r = _run(d)
    assert r.returncode != 0 and "does not exist" in r.stdout


def test_truncated_restart_is_refused(tmp_path):
    d = _rundir(tmp_path, irestart=1)
    nx, ny = mf.grid_size(P)
    f, ux, uy = mf.noisy_state(nx, ny, 1)
    path = os.path.join(d, "restart", "restart.dat")
    mf.write_restart(path, 1, f, ux, uy, np.ones((nx, ny)))
    size = os.path.getsize(path)
    with open(path, "r+b") as fh:
        fh.truncate(size // 2)
    r = _run(d)
    assert r.returncode != 0
    assert "restart/restart.dat holds" in r.stdout and f"{nx} x {ny}" in r.stdout
    # nothing was written into the run directory's output folders
    assert not os.listdir(os.path.join(d, "output"))
