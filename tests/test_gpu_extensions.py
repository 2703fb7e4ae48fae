"""GPU against the oracle on the configurations the reference cannot express (the oracle's
extensions are pinned in tests/test_oracle_extensions.py): FP64 fields bit for bit, forces
within 1e-12 (tree against sequential sums).

* 400 x 100 channel as stated for the CPU-runnable case: cylinder r = 10 at (80, 50.5),
  parabolic inlet u_max = 0.05, Re = 20 on the mean velocity 2/3 u_max, walls, from rest,
  20,000 steps with the force series over the last 10,000; the half-integer centre's two
  label-15 wall nodes are fluid (label-fluid policy);
* the adjoint case's geometry, 512 x 128, D = 24 at (100, 64.5), Re = 100, with its seeded
  Brinkman field, 400 steps;
* a cylinder-free 48 x 24 channel with the parabolic inlet (nu on ny - 1), 2,000 steps.
"""
import numpy as np
import pytest

from mrlbm_2d_amd import brinkman as B
from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu


def _pair(po, nx, ny, re, uo, xc=None, yc=None, rc=None, statsbegin=10 ** 8, statsend=10 ** 8 + 1):
    cyl = xc is not None
    op = po.OracleParams(geometry_explicit=1, nx=nx, ny=ny, Re=re, uo=uo, inlet_parabolic=1, label_fluid=1,
                         x_periodic=0, y_periodic=0, channel_with_cylinder=1 if cyl else 0,
                         post_process=1, statsbegin=statsbegin, statsend=statsend, iplot=10 ** 8)
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic = p.y_periodic = 0
    p.channel_with_cylinder = 1 if cyl else 0
    p.Re, p.uo = re, uo
    p.inlet_profile = capi.INLET_PARABOLIC
    p.label_policy = capi.LABEL_FLUID
    if cyl:
        op.x_c, op.y_c, op.r_c = xc, yc, rc
        p.x_c, p.y_c, p.r_c = xc, yc, rc
    o = capi.default_step_opts()
    o.post_process, o.statsbegin, o.statsend = 1, statsbegin, statsend
    return po.Oracle(op), p, o


def _run_oracle(o, first, n, record_from):
    fd, fl = [], []
    for it in range(first, first + n):
        o.step(it)
        if it >= record_from:
            a, b = o.forces()
            fd.append(a)
            fl.append(b)
    return np.array(fd), np.array(fl)


def _check_fields(lat, o):
    rho, ux, uy, mxx, myy, mxy = lat.moments()
    orho, oux, ouy, omxx, omyy, omxy = o.moments()
    for a, b in ((rho, orho), (ux, oux), (uy, ouy), (mxx, omxx), (myy, omyy), (mxy, omxy)):
        np.testing.assert_array_equal(a, b)


def test_400x100_as_stated_20000_steps_matches_oracle(oracle_mod):
    o, p, opts = _pair(oracle_mod, 400, 100, 20.0, 0.05, 80.0, 50.5, 10.0, statsbegin=10002, statsend=10 ** 7)
    assert o.nbct == 78                      # 80 wall nodes less the two label-15 nodes
    with capi.Lattice(p) as lat:
        info = lat.info
        assert info.nbct == o.nbct and info.r_cyl == o.r_cyl and info.nu == o.nu and info.omega == o.omega
        b = lat.boundary()
        ob = o.boundary()
        np.testing.assert_array_equal(b["bou_i"], ob["bou_i"])
        np.testing.assert_array_equal(b["bou_j"], ob["bou_j"])
        np.testing.assert_array_equal(b["label"], ob["label"])
        print(f"[gpu] 400x100 (80,50.5) parabolic Re=20 x 20000 steps vs oracle ({oracle_mod.num_threads()} threads)",
              flush=True)
        lat.step(2, 20000, opts)
        it, gfd, gfl = lat.forces()
        fd, fl = _run_oracle(o, 2, 20000, 10002)
        _check_fields(lat, o)
        rim = lat.rho_inf_mean()
    assert rim == pytest.approx(o.rho_inf_mean(), rel=1e-13)
    np.testing.assert_array_equal(it, np.arange(10002, 20002, dtype=np.int32))
    scale = max(np.abs(fd).max(), np.abs(fl).max())
    np.testing.assert_allclose(gfd, fd, rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(gfl, fl, rtol=1e-12, atol=1e-12 * scale)
    # steady flow at Re = 20: the drag coefficient on 2/3 u_max and D = 2 r_cyl
    cd = fd[-1] / (0.5 * (2.0 / 3.0 * 0.05) ** 2 * 2 * o.r_cyl)
    print(f"[gpu] Cd = {cd:.6f}", flush=True)
    assert 2.0 < cd < 8.0


def test_adjoint_geometry_with_brinkman_field_matches_oracle(oracle_mod):
    c = B.ADJ_CASE
    o, p, opts = _pair(oracle_mod, c["nx"], c["ny"], c["Re"], c["umax"], c["xc"], c["yc"], c["radius"],
                       statsbegin=202, statsend=10 ** 7)
    chi = B.adjoint_case_field()
    assert chi.max() > 1e-3
    o.set_brinkman(chi)
    with capi.Lattice(p) as lat:
        lat.set_brinkman(chi)
        print("[gpu] 512x128 D=24 (100,64.5) with the Brinkman field x 400 steps vs oracle", flush=True)
        lat.step(2, 400, opts)
        it, gfd, gfl = lat.forces()
        fd, fl = _run_oracle(o, 2, 400, 202)
        _check_fields(lat, o)
    scale = max(np.abs(fd).max(), np.abs(fl).max())
    np.testing.assert_allclose(gfd, fd, rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(gfl, fl, rtol=1e-12, atol=1e-12 * scale)


def test_parabolic_channel_without_cylinder_matches_oracle(oracle_mod):
    o, p, opts = _pair(oracle_mod, 48, 24, 10.0, 0.05)
    with capi.Lattice(p) as lat:
        assert lat.info.nu == o.nu
        lat.step(2, 2000, opts)
        o.run(2, 2000)
        _check_fields(lat, o)


def test_8192x2048_fp64_matches_oracle(oracle_mod):
    """The benchmark grid itself: 8192 x 2048, D = 128 at (1024, 1025), parabolic inlet
    u_max = 0.05, Re = 200, from a seeded noisy population state (+-1e-4 velocity), 11 FP64
    steps (the first one single after the populations are set, then five two-step passes)
    with the forces of every step: fields bit for bit against the oracle."""
    from tests.golden import make_fixtures as MF
    nx, ny, n = 8192, 2048, 11
    o, p, opts = _pair(oracle_mod, nx, ny, 200.0, 0.05, 1024.0, 1025.0, 64.0, statsbegin=2, statsend=10 ** 7)
    f0, _, _ = MF.noisy_state(nx, ny, 3)
    print(f"[gpu] 8192x2048 FP64 x {n} steps vs oracle ({oracle_mod.num_threads()} threads)", flush=True)
    with capi.Lattice(p) as lat:
        lat.set_populations(f0)
        lat.step(2, n, opts)
        it, gfd, gfl = lat.forces()
        rho, ux, uy = lat.fields()
    o.set_state(f0)
    del f0
    fd, fl = _run_oracle(o, 2, n, 2)
    orho, oux, ouy, *_ = o.moments()
    np.testing.assert_array_equal(rho, orho)
    np.testing.assert_array_equal(ux, oux)
    np.testing.assert_array_equal(uy, ouy)
    np.testing.assert_array_equal(it, np.arange(2, 2 + n, dtype=np.int32))
    # Early in the run the net force (~1e-4) is a near-cancelling sum of 512 node forces of
    # size ~0.1: another summation order (a tree on the GPU) moves it by ~1e-15 per node.
    atol = 1e-15 * o.nbct
    np.testing.assert_allclose(gfd, fd, rtol=1e-12, atol=atol)
    np.testing.assert_allclose(gfl, fl, rtol=1e-12, atol=atol)
