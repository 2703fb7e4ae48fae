"""The batched configuration at full size: 256 FP32 cylinder runs (nx in 256..2048 step 64,
ny = nx/4, D in [ny/8, ny/4] at (nx/5, ny/2 + 0.5), Re in [40, 300], u_max = 0.05,
parabolic inlet, seed 0) advanced 20,000 steps in one batch (autotuned), the drag summed
over the last 1,000.

* A spread of runs (the smallest and the largest grid, and two between) equals the same
  runs stepped alone, moments bit for bit and drag sums exactly.
* Every run whose relaxation rate omega stays below 1.96 is finite.  Runs above that
  bound may diverge: the regularized scheme goes unstable as omega nears 2.
* One diverging draw (run 66: nx = 256, D = 10.2, Re = 264, omega = 1.983) diverges in the
  FP64 form too, so the divergence is the scheme's, not the FP32 form's.
"""
import numpy as np
import pytest

from mrlbm_2d_amd import batch_case as BC
from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu

SEED, NRUNS, STEPS, WINDOW, OMEGA_STABLE = 0, 256, 20000, 1000, 1.96


def _single(p, r):
    with capi.Lattice(p) as lat:
        BC.seed_state(lat, SEED, r)
        lat.step(2, STEPS - WINDOW)
        o = capi.default_step_opts()
        o.post_process, o.statsbegin, o.statsend = 1, 2 + STEPS - WINDOW, 1 + STEPS
        lat.step(2 + STEPS - WINDOW, WINDOW, o)
        _, fd, _ = lat.forces()
        s = 0.0
        for v in fd:  # in step order, as the device sums
            s += v
        return lat.moments(), s, len(fd)


def test_256_runs_full_size():
    ps = BC.uneven_runs(NRUNS, seed=SEED)
    nodes = sum(p.nx * p.ny for p in ps)
    print(f"[gpu] batch {NRUNS} runs, {nodes} nodes, {STEPS} steps", flush=True)
    with capi.Batch(ps) as b:
        info = b.info
        print(f"[gpu] kernel {info.kernel} persistent {info.persistent} time_blocking {info.time_blocking}", flush=True)
        for r in range(NRUNS):
            BC.seed_state(b.run(r), SEED, r)
        b.step(STEPS - WINDOW)
        b.step(WINDOW, record_forces=True)
        fd, _, _, cnt = b.forces()
        omega = np.array([b.run(r).info.omega for r in range(NRUNS)])
        finite = np.array([b.run(r).check_finite() == 0 for r in range(NRUNS)])
        nx = np.array([p.nx for p in ps])
        spread = sorted({int(np.argmin(nx)), int(np.argmax(nx)), int(np.argsort(nx)[NRUNS // 3]),
                         int(np.argsort(nx)[2 * NRUNS // 3])})
        spread = [r for r in spread if finite[r]]
        got = {r: b.run(r).moments() for r in spread}
    print(f"[gpu] finite {finite.sum()} of {NRUNS}; diverged omegas {sorted(np.round(omega[~finite], 4))}", flush=True)
    assert finite[omega < OMEGA_STABLE].all()
    assert (omega[~finite] > OMEGA_STABLE).all()
    assert len(spread) >= 3
    for r in spread:
        print(f"[gpu] run {r} nx {ps[r].nx} alone", flush=True)
        ref, s_fd, n = _single(ps[r], r)
        for a, c in zip(got[r], ref):
            np.testing.assert_array_equal(a, c, err_msg=f"run {r}")
        assert cnt[r] == n == WINDOW and fd[r] == s_fd, (r, fd[r], s_fd)


def test_diverging_draw_diverges_in_fp64_too():
    ps = BC.uneven_runs(NRUNS, seed=SEED)
    r = 66
    out = {}
    for prec in (capi.MRLBM_FP32_DEV, capi.MRLBM_FP64):
        p = BC.uneven_runs(NRUNS, seed=SEED, precision=prec)[r]
        with capi.Lattice(p) as lat:
            assert lat.info.omega > 1.98 and p.nx == ps[r].nx == 256
            BC.seed_state(lat, SEED, r)
            it, bad_at = 2, None
            while it < 2 + STEPS:
                lat.step(it, 500)
                it += 500
                if lat.check_finite() != 0:
                    bad_at = it - 1
                    break
            out[prec] = bad_at
    print(f"[gpu] run 66 non-finite by step: FP32 {out[capi.MRLBM_FP32_DEV]}, FP64 {out[capi.MRLBM_FP64]}", flush=True)
    assert out[capi.MRLBM_FP32_DEV] is not None and out[capi.MRLBM_FP64] is not None
