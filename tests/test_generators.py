"""Seeded inputs of the adjoint case and of the batched runs (host code, no GPU)."""
import numpy as np

from mrlbm_2d_amd import batch_case as BC
from mrlbm_2d_amd import brinkman as B
from mrlbm_2d_amd import capi


def test_gaussian_bumps_stay_within_two_diameters():
    c = B.ADJ_CASE
    chi = B.adjoint_case_field()
    assert chi.shape == (c["nx"], c["ny"])
    assert np.all(chi >= 0) and 0 < chi.max() <= 0.01 * 8
    x = np.arange(1, c["nx"] + 1)[:, None]
    y = np.arange(1, c["ny"] + 1)[None, :]
    dist = np.sqrt((x - c["xc"]) ** 2 + (y - c["yc"]) ** 2) - c["radius"]
    assert np.all(chi[dist > 4 * c["radius"]] == 0)
    assert chi[dist < 2 * c["radius"]].max() > 1e-3
    np.testing.assert_array_equal(chi, B.adjoint_case_field())


def test_adjoint_case_parameters():
    p = B.adjoint_params(capi.MRLBM_FP32_DEV)
    assert (p.nx, p.ny, p.x_c, p.y_c, p.r_c, p.Re, p.uo) == (512, 128, 100.0, 64.5, 12.0, 100.0, 0.05)
    assert p.inlet_profile == capi.INLET_PARABOLIC and p.label_policy == capi.LABEL_FLUID
    assert p.x_periodic == 0 and p.y_periodic == 0


def test_uneven_runs_follow_their_distribution():
    ps = BC.uneven_runs(256, seed=0)
    nodes = 0
    for p in ps:
        assert 256 <= p.nx <= 2048 and p.nx % 64 == 0 and p.ny == p.nx // 4
        D = 2 * p.r_c
        assert p.ny / 8 <= D <= p.ny / 4
        assert p.x_c == p.nx / 5 and p.y_c == p.ny / 2 + 0.5
        assert 40 <= p.Re <= 300 and p.uo == 0.05
        assert p.precision == capi.MRLBM_FP32_DEV and p.label_policy == capi.LABEL_FLUID
        nodes += p.nx * p.ny
    # E[nx^2/4] over nx uniform in {256..2048 step 64}: ~1.3e8 for 256 runs
    assert 0.8e8 < nodes < 1.6e8
    again = BC.uneven_runs(256, seed=0)
    assert [(p.nx, p.Re, p.r_c) for p in again] == [(p.nx, p.Re, p.r_c) for p in ps]
