"""The FP32 deviation form against the FP64 form from the same seeded state: mean Cd within
|dCd|/Cd < 1e-3 (the accuracy condition of the FP32 throughput figures), with rms Cl, the
lift spectrum's St and the final velocity field stated beside it.

* 1600 x 400 channel, D = 40 at (200, 201), parabolic inlet u_max = 0.06, Re = 100,
  seed 2, 200,000 steps, statistics over the last 100,000 (vortex shedding); also two
  steps per bulk pass against single steps over 2,000 recorded steps from the state
  reached, and St by both the reference's period rule and the spectrum; and the same
  FP64 run's mean / RMS velocity fields (sampled every 10th step) checked for their
  sample count, the imposed inlet and wall values and the RMS peak in the wake;
* 8192 x 2048, D = 128 at (1024, 1025), u_max = 0.05, Re = 200, seed 3: a bounded
  20,000 steps with Cd over the last 10,000 (the full 100,000-step comparison runs in
  the benchmark).
"""
import numpy as np
import pytest

from mrlbm_2d_amd import capi, compare

pytestmark = pytest.mark.gpu


def _check_gate(res, cd_tol=1e-3):
    for form in ("fp64", "fp32"):
        assert res[form]["finite"], form
    print(f"[gpu] |dCd|/Cd {res['cd_rel_err']:.3e}, rms Cl {res['cl_rms_rel_err']}, spectral St "
          f"{res['st_spectral_rel_err']:.3e}, L2(u) {res['l2_u_rel']:.3e}; MLUPS fp64 {res['fp64']['mlups']:.0f} "
          f"fp32 {res['fp32']['mlups']:.0f}", flush=True)
    assert res["cd_rel_err"] < cd_tol


def test_fp32_gate_shedding_1600x400():
    print("[gpu] precision gate 1600x400 Re=100, 200000 steps, window 100000", flush=True)
    res = compare.compare(1600, 400, 200.0, 201.0, 20.0, 100.0, 0.06, 200000, 100000, seed=2)
    _check_gate(res)
    assert res["cl_rms_rel_err"] < 1e-2
    assert res["st_spectral_rel_err"] < 1e-3
    # shedding: the spectral St of a D = 40 cylinder at Re = 100 in a channel of 10 D
    assert 0.2 < res["fp64"]["st_spectral"] < 0.35
    # the reference's rule (first two lift maxima) on the FP64 series gives the same period
    assert res["fp64"]["st"] is not None
    assert abs(res["fp64"]["st"] - res["fp64"]["st_spectral"]) < 0.02 * res["fp64"]["st_spectral"]


def test_pairs_equal_single_steps_from_shedding_state_1600x400():
    """From a state 30,000 steps into the 1600 x 400 run, 2,000 steps with forces, surface
    pressure and u statistics recorded: two steps per pass against one, bit for bit."""
    p = compare.make_params(1600, 400, 200.0, 201.0, 20.0, 100.0, 0.06, capi.MRLBM_FP64)
    with capi.Lattice(p) as lat:
        rng = np.random.default_rng(2)
        lat.set_moments(np.ones((1600, 400)), rng.uniform(-1e-4, 1e-4, (1600, 400)),
                        rng.uniform(-1e-4, 1e-4, (1600, 400)))
        lat.step(2, 30000)
        f = lat.populations()
    o = capi.default_step_opts()
    o.post_process, o.statsbegin, o.statsend, o.record_surface_p, o.stats_interval = 1, 30002, 10 ** 7, 1, 3
    out = []
    for tb in (-1, 0):
        p.time_blocking = tb
        with capi.Lattice(p) as lat:
            lat.set_populations(f)
            lat.step(30002, 2000, o)
            _, fd, fl = lat.forces()
            _, ps, pf = lat.surface_p_series(max_n=2000)
            out.append((lat.moments(), fd, fl, ps, pf, lat.stats()))
    (m0, *r0), (m1, *r1) = out
    for a, b in zip(m0, m1):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(r0[:4], r1[:4]):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(r0[4], r1[4]):
        np.testing.assert_array_equal(a, b)


def test_fp32_gate_8192x2048_bounded():
    print("[gpu] precision gate 8192x2048 Re=200, 20000 steps, window 10000", flush=True)
    res = compare.compare(8192, 2048, 1024.0, 1025.0, 64.0, 200.0, 0.05, 20000, 10000, seed=3)
    _check_gate(res)


def test_u_statistics_of_shedding_run_1600x400():
    """The 1600 x 400 configuration as stated: 200,000 FP64 steps with mean / RMS velocity
    fields over the last 100,000 (a sample every 10th step): the samples are counted right,
    the inlet column's mean is the imposed parabola with zero fluctuation, the walls' mean is
    zero, and the fluctuations peak in the cylinder's wake (shedding) and not upstream."""
    p = compare.make_params(1600, 400, 200.0, 201.0, 20.0, 100.0, 0.06, capi.MRLBM_FP64)
    with capi.Lattice(p) as lat:
        rng = np.random.default_rng(2)
        lat.set_moments(np.ones((1600, 400)), rng.uniform(-1e-4, 1e-4, (1600, 400)),
                        rng.uniform(-1e-4, 1e-4, (1600, 400)))
        print("[gpu] 1600x400 FP64 200000 steps, u statistics over the last 100000", flush=True)
        lat.step(2, 100000)
        o = capi.default_step_opts()
        o.post_process, o.statsbegin, o.statsend, o.stats_interval = 1, 100002, 200001, 10
        lat.step(100002, 100000, o)
        mux, muy, rux, ruy, ns = lat.stats()
        assert lat.check_finite() == 0
    assert ns == 10000
    s = np.arange(400) / 399.0
    np.testing.assert_allclose(mux[0], 4.0 * 0.06 * s * (1.0 - s), rtol=1e-12, atol=1e-17)
    assert np.abs(muy[0]).max() == 0.0 and rux[0].max() < 1e-7 and ruy[0].max() < 1e-7
    assert np.abs(mux[:, 0]).max() == 0.0 and np.abs(mux[:, -1]).max() == 0.0
    wake = ruy[320:600, 150:252].max()      # 3 to 10 diameters behind the cylinder's centre
    upstream = ruy[20:100, 150:252].max()   # 2 to 4.5 diameters ahead of it
    print(f"[gpu] rms uy: wake {wake:.3e}, upstream {upstream:.3e}", flush=True)
    assert wake > 1e-3 and wake > 5 * upstream
