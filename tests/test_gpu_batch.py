"""Batched runs (mrlbm_batch_*): runs of uneven size advanced in shared launches and
captured graphs give the bits of the same runs stepped alone, for every bulk variant, grid
mode and one, two or four steps per pass, and the per-run force sums equal the single-run
series summed in order."""
import numpy as np
import pytest

from mrlbm_2d_amd import batch_case as BC
from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu

SEED = 4
SIZES = (128, 192, 256, 320)


def runs(precision):
    return BC.uneven_runs(4, seed=SEED, precision=precision, sizes=SIZES)


def single(p, run, n1, n2):
    """The run alone: n1 steps, then n2 steps with the drag recorded."""
    with capi.Lattice(p) as lat:
        BC.seed_state(lat, SEED, run)
        lat.step(2, n1)
        o = capi.default_step_opts()
        o.post_process = 1
        o.statsbegin = 2 + n1
        o.statsend = 1 + n1 + n2
        lat.step(2 + n1, n2, o)
        _, fd, fl = lat.forces()
        s_fd = 0.0
        s_fl2 = 0.0
        for a, b in zip(fd, fl):  # in step order, as the device sums
            s_fd += a
            s_fl2 += b * b
        return lat.moments(), s_fd, s_fl2, len(fd)


@pytest.mark.parametrize("precision", [capi.MRLBM_FP32_DEV, capi.MRLBM_FP64])
@pytest.mark.parametrize("kernel,persistent,time_blocking,spp", [
    (0, -1, -1, 0), (8, 0, 0, 0), (304, 1, 0, 0), (2, 1, 0, 0), (304, 0, 4, 2), (8, 0, 2, 2), (16, 1, 6, 2),
    (304, 0, 6, 4), (8, 1, -1, 4)])
def test_batch_matches_single_runs(precision, kernel, persistent, time_blocking, spp):
    ps = runs(precision)
    o = capi.default_batch_opts()
    o.kernel, o.persistent, o.graph_steps, o.time_blocking = kernel, persistent, 22, time_blocking
    o.steps_per_pass = spp
    # 22-step graphs: 41 = one graph + direct steps (four-step passes, a pair, a single step),
    # so the second call starts at the other parity; 60 = two graphs + 16 direct
    n1, n2 = 41, 60
    with capi.Batch(ps, o) as b:
        info = b.info
        print(f"[gpu] batch prec {precision} kernel {info.kernel} persistent {info.persistent} time_blocking "
              f"{info.time_blocking} four-step rows {info.quad_rows} grid {info.grid} tiles {info.ntiles} tuned "
              f"{[(info.tune_kernel[i], info.tune_persistent[i], info.tune_tb[i], info.tune_quad[i], round(info.tune_ms[i], 4)) for i in range(info.ntune)]}",
              flush=True)
        if time_blocking > 0 and spp != 4:
            assert info.time_blocking == time_blocking
        if spp == 4:
            assert info.quad_rows in ((20, 24, 12) if precision == capi.MRLBM_FP32_DEV else (20, 16, 8))
        if spp == 2 or time_blocking == 0:
            assert info.quad_rows == 0
        if kernel == 0:
            assert info.ntune >= 2
            best = min(range(info.ntune), key=lambda i: info.tune_ms[i])
            assert info.kernel == info.tune_kernel[best]
        for r in range(len(ps)):
            BC.seed_state(b.run(r), SEED, r)
        b.step(n1)
        b.step(n2, record_forces=True)
        fd, fl, fl2, cnt = b.forces()
        got = [b.run(r).moments() for r in range(len(ps))]
    for r, p in enumerate(ps):
        ref, s_fd, s_fl2, n = single(p, r, n1, n2)
        for a, c in zip(got[r], ref):
            np.testing.assert_array_equal(a, c, err_msg=f"run {r}")
        assert cnt[r] == n == n2
        assert fd[r] == s_fd and fl2[r] == s_fl2, (r, fd[r], s_fd)


def test_batch_rejects_a_run_stepped_alone():
    ps = runs(capi.MRLBM_FP32_DEV)
    o = capi.default_batch_opts()
    o.kernel, o.persistent = 304, 1
    with capi.Batch(ps, o) as b:
        b.run(1).step(2, 1)
        with pytest.raises(capi.MrlbmError) as e:
            b.step(2)
        assert e.value.code == capi.E_STATE
