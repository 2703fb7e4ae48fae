"""Four-step and two-step passes against single steps, bit for bit, on shapes the default
cases do not reach: every four-step strip height the library builds (`quad_rows`), and a
cylinder close to a channel wall, where the cylinder's band rectangle (marched row by row
inside the side launch) overlaps the rows next to that wall (marched too; the earlier
rectangle keeps the nodes both hold)."""
import numpy as np
import pytest

from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu

HEIGHTS = {capi.MRLBM_FP32_DEV: (12, 16, 20, 24, 28, 32), capi.MRLBM_FP64: (8, 16, 20, 24, 32, 40)}


def _params(precision, nx, ny, xc, yc, rc, re=100.0, umax=0.05):
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = xc, yc, rc
    p.Re, p.uo = re, umax
    p.inlet_profile = capi.INLET_PARABOLIC
    p.precision = precision
    return p


def _run(p, steps, seed=5, record=False):
    """steps steps from a seeded state; with record, forces over the second half."""
    with capi.Lattice(p) as lat:
        rng = np.random.default_rng(seed)
        lat.set_moments(np.ones((p.nx, p.ny)), rng.uniform(-1e-3, 1e-3, (p.nx, p.ny)),
                        rng.uniform(-1e-3, 1e-3, (p.nx, p.ny)))
        o = capi.default_step_opts()
        if record:
            o.post_process, o.statsbegin, o.statsend = 1, 2 + steps // 2, 10 ** 8
        lat.set_timing(1)
        lat.step(2, steps, o)
        t = lat.last_timing()
        forces = lat.forces() if record else None
        return lat.moments(), forces, t.bulk_steps / max(t.bulk_launches, 1), lat.info.quad_rows


@pytest.mark.parametrize("precision", [capi.MRLBM_FP32_DEV, capi.MRLBM_FP64])
def test_every_four_step_strip_height_matches_single_steps(precision):
    p = _params(precision, 300, 136, 60.0, 68.5, 9.0)
    p.time_blocking = 0
    ref, _, spl, _ = _run(p, 41)
    assert spl == 1.0
    for h in HEIGHTS[precision]:
        q = _params(precision, 300, 136, 60.0, 68.5, 9.0)
        q.quad_rows = h
        print(f"[gpu] 300x136 precision={precision} quad_rows={h}", flush=True)
        got, _, spl, rows = _run(q, 41)
        assert rows == h and spl > 3.0, (rows, spl)  # ten four-step passes and a single step
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"quad_rows {h}")


def test_unbuilt_strip_height_is_refused():
    p = _params(capi.MRLBM_FP64, 300, 136, 60.0, 68.5, 9.0)
    p.quad_rows = 12  # FP64 has no 12-row four-step kernel
    with pytest.raises(capi.MrlbmError) as e:
        capi.Lattice(p)
    assert e.value.code == capi.E_ARG


@pytest.mark.parametrize("precision", [capi.MRLBM_FP32_DEV, capi.MRLBM_FP64])
@pytest.mark.parametrize("yc", [12.0, 125.0])
def test_cylinder_next_to_a_channel_wall_matches_single_steps(precision, yc):
    """Cylinder of radius 5 as close to the bottom (top) channel wall as its interpolation
    stencil allows: its box spans rows 6..16 (119..129 of 136), so the cylinder's rectangle
    of every sub-step (the box dilated by up to 6 + 3) overlaps the band rows at that wall."""
    out = {}
    for spp in (1, 4, 2):
        p = _params(precision, 200, 136, 50.0, yc, 5.0)
        p.time_blocking = 0 if spp == 1 else -1
        p.steps_per_pass = 2 if spp == 2 else 0
        print(f"[gpu] cylinder near the wall y_c={yc} precision={precision} steps per pass {spp}", flush=True)
        out[spp] = _run(p, 60, record=True)
    assert out[1][2] == 1.0 and out[4][2] > 3.0 and out[2][2] == 2.0, [out[k][2] for k in out]
    for spp in (4, 2):
        for a, b in zip(out[spp][0], out[1][0]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(out[spp][1], out[1][1]):
            np.testing.assert_array_equal(a, b)
