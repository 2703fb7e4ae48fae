"""mrlbm_get_populations_noalloc (stages through a fixed per-GPU array, allocates nothing)
against mrlbm_get_populations (allocates a 72 B/node staging buffer per call) through the
C-ABI: bit for bit, on two calls in a row, the second with a smaller grid; the first grid
needs several row bands of the 8 MiB array, the second one band.  Rows beyond the array's
limit are refused with MRLBM_E_ARG."""
import numpy as np
import pytest

from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu


def _lattice(nx, ny, precision, seed, xc, yc, rc):
    p = capi.default_params()
    p.nx, p.ny = nx, ny
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = xc, yc, rc
    p.Re, p.uo = 100.0, 0.05
    p.inlet_profile = capi.INLET_PARABOLIC
    p.label_policy = capi.LABEL_FLUID
    p.precision = precision
    lat = capi.Lattice(p)
    rng = np.random.default_rng(seed)
    lat.set_moments(np.ones((nx, ny)), rng.uniform(-1e-4, 1e-4, (nx, ny)), rng.uniform(-1e-4, 1e-4, (nx, ny)))
    return lat


def test_noalloc_populations_equal_allocating_call_twice():
    # 4096 x 100 FP64: one row of nine populations is 36,864 words, so the array holds 28 rows
    # and the grid takes four bands (28, 28, 28, 16)
    cases = [(4096, 100, capi.MRLBM_FP64, (800.0, 50.5, 10.0), 7), (88, 40, capi.MRLBM_FP32_DEV, (21.0, 20.5, 4.0), 5)]
    for nx, ny, prec, (xc, yc, rc), steps in cases:
        print(f"[gpu] no-alloc populations {nx}x{ny} precision={prec} steps={steps}", flush=True)
        with _lattice(nx, ny, prec, nx + ny, xc, yc, rc) as lat:
            lat.step(2, steps)
            f_ref = lat.populations()
            f_na = lat.populations_noalloc()
            assert np.isfinite(f_ref).all()
            np.testing.assert_array_equal(f_na, f_ref)
            # the state is unchanged by either copy: stepping on gives the same populations again
            lat.step(2 + steps, 1)
            np.testing.assert_array_equal(lat.populations_noalloc(), lat.populations())


def test_noalloc_populations_after_set_populations():
    """Before any step the populations are the ones set (restart.dat read back)."""
    nx, ny = 88, 40
    with _lattice(nx, ny, capi.MRLBM_FP64, 1, 21.0, 20.5, 4.0) as lat:
        lat.step(2, 3)
        f = lat.populations()
    with _lattice(nx, ny, capi.MRLBM_FP64, 2, 21.0, 20.5, 4.0) as lat:
        lat.set_populations(f)
        np.testing.assert_array_equal(lat.populations_noalloc(), f)


def test_noalloc_populations_refuses_rows_beyond_the_array():
    p = capi.default_params()
    p.nx, p.ny = 116509, 4
    p.channel_with_cylinder = 0
    p.y_periodic = 0
    with capi.Lattice(p) as lat:
        with pytest.raises(capi.MrlbmError) as e:
            lat.populations_noalloc()
        assert e.value.code == capi.E_ARG and "116508" in str(e.value)
