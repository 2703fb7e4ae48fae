"""GPU parity: the HIP path through the C-ABI against the reference's own outputs
(tests/golden, written by the reference program) and against the oracle.

FP64 bar: fields and populations bit-identical to the reference after 500-1000 steps
(the kernels keep the reference's operand order and are built without FMA); forces and
surface pressure within 1e-12 relative (the GPU sums the nbct node forces as a tree,
the reference sequentially).
"""
import numpy as np
import pytest

from mrlbm_2d_amd import Lattice, capi, params_from_namelist, step_opts_from_namelist
from tests import golden_cases as gc

pytestmark = pytest.mark.gpu


def run_golden(name, kernel=0, time_blocking=0):
    g = gc.load(name)
    p = g["params"]
    par = params_from_namelist(p)
    par.kernel = kernel
    par.time_blocking = time_blocking
    lat = Lattice(par)
    if "f0" in g:
        lat.set_populations(g["f0"])
    n = int(g["nsteps"])
    print(f"[gpu] golden {name} {lat.nx}x{lat.ny} steps={n} kernel={kernel} time_blocking={time_blocking} "
          f"wall levels={lat.info.levels}", flush=True)
    opts = step_opts_from_namelist(p)
    # stop at iter n, where the reference last appended to rho_mean_inlet.dat (every 100 steps)
    lat.step(2, n - 1, opts)
    g["rho_inf_mean_gpu"] = lat.rho_inf_mean()
    lat.step(n + 1, 1, opts)
    return g, lat


@pytest.mark.parametrize("time_blocking", [0, 4])
@pytest.mark.parametrize("name", gc.names())
def test_fp64_matches_reference(name, time_blocking):
    """Also with two steps per bulk pass: pairs of steps share one pass wherever no
    plot-step zeroing or u sampling falls inside them, single steps elsewhere."""
    g, lat = run_golden(name, time_blocking=time_blocking)
    with lat:
        rho, ux, uy = lat.fields()
        np.testing.assert_array_equal(rho, g["rho"])
        np.testing.assert_array_equal(ux, g["ux"])
        np.testing.assert_array_equal(uy, g["uy"])
        np.testing.assert_array_equal(lat.populations(), g["f"])
        if "forces" in g:
            it, fd, fl = lat.forces()
            np.testing.assert_array_equal(it, g["forces"][:, 0].astype(np.int32))
            # The net force is a sum of nbct node forces of size ~0.1; summing them in a
            # different order moves it by a few 1e-16 of that size, so the absolute floor is
            # 1e-14 (it matters only where the net force cancels to ~1e-13, at rest).
            scale = np.abs(g["forces"][:, 1:]).max()
            atol = max(1e-12 * scale, 1e-14)
            np.testing.assert_allclose(fd, g["forces"][:, 1], rtol=1e-12, atol=atol)
            np.testing.assert_allclose(fl, g["forces"][:, 2], rtol=1e-12, atol=atol)
            _, ps, pf = lat.surface_p()
            np.testing.assert_allclose(ps, g["ps_last"], rtol=1e-12, atol=0)
            np.testing.assert_allclose(pf, g["pfit_last"], rtol=1e-12, atol=0)
            if "rho_mean_inlet" in g:
                assert int(g["rho_mean_inlet"][-1, 0]) == int(g["nsteps"])
                np.testing.assert_allclose(g["rho_inf_mean_gpu"], g["rho_mean_inlet"][-1, 1], rtol=1e-14)


@pytest.mark.parametrize("kernel", [0, 1, 2, 4, 8, 16, 304, 308, 316])
@pytest.mark.parametrize("name", ["p1_walls", "s_periodic", "s_xyper_noise"])
def test_bulk_kernel_variants_agree(name, kernel):
    """Every bulk kernel variant (one node per thread, row loop, unrolled strips, aligned
    64-column waves) gives the reference's bits."""
    g, lat = run_golden(name, kernel=kernel)
    with lat:
        rho, ux, uy = lat.fields()
        np.testing.assert_array_equal(rho, g["rho"])
        np.testing.assert_array_equal(ux, g["ux"])
        np.testing.assert_array_equal(uy, g["uy"])


@pytest.mark.parametrize("time_blocking", [2, 6, 8, 12, 14, 16])
@pytest.mark.parametrize("name", ["p1_walls", "s_periodic", "s_xyper_noise", "chan"])
def test_time_blocking_strips_agree(name, time_blocking):
    g, lat = run_golden(name, time_blocking=time_blocking)
    with lat:
        rho, ux, uy = lat.fields()
        np.testing.assert_array_equal(rho, g["rho"])
        np.testing.assert_array_equal(ux, g["ux"])
        np.testing.assert_array_equal(uy, g["uy"])


@pytest.mark.parametrize("name", ["p1_walls", "s_xyper_noise", "s_noise_re60"])
def test_fp32_variants_identical_and_close_to_reference(name):
    """FP32 deviation form: every bulk variant computes the same per-node arithmetic, so
    all give the same bits; and after 500-1000 steps the fields stay within FP32 accuracy
    of the reference's FP64 run (the form's own error, not a parity bar)."""
    g = gc.load(name)
    p = g["params"]
    results = {}
    for kernel in (0, 1, 2, 4, 8, 16, 304, 308, 316, 1002, 1004, 1006, 1010, 1012, 1016):
        par = params_from_namelist(p, precision=capi.MRLBM_FP32_DEV)
        par.kernel = kernel if kernel < 1000 else 0
        par.time_blocking = kernel - 1000 if kernel >= 1000 else 0
        with Lattice(par) as lat:
            if "f0" in g:
                lat.set_populations(g["f0"])
            print(f"[gpu] fp32 {name} kernel={kernel}", flush=True)
            lat.step(2, int(g["nsteps"]), step_opts_from_namelist(p))
            results[kernel] = lat.fields()
    base = results[1]
    for k, r in results.items():
        for a, b in zip(r, base):
            np.testing.assert_array_equal(a, b, err_msg=f"kernel {k}")
    rho, ux, uy = base
    umax = max(np.abs(g["ux"]).max(), 1e-12)
    drho = np.abs(rho - g["rho"]).max()
    dux = np.abs(ux - g["ux"]).max()
    duy = np.abs(uy - g["uy"]).max()
    print(f"[gpu] fp32 {name}: max |drho| {drho:.3e} |dux| {dux:.3e} |duy| {duy:.3e} (umax {umax:.3e})", flush=True)
    # FP32 rounding of O(1e-7) per step accumulates over 500-1000 steps; the walls' moment
    # solves amplify it by the local condition number (about 10-30 on these grids).  Velocity
    # bound: relative to the flow speed, with an absolute floor of a few FP32 ulps of 1.
    assert drho < 1e-5
    assert dux < max(5e-5 * umax / 0.05, 5e-7)
    assert duy < max(5e-5 * umax / 0.05, 5e-7)


@pytest.mark.parametrize("time_blocking,steps_per_pass,every,launches,steps",
                         [(0, 0, 1, 10, 10), (4, 2, 1, 5, 10), (0, 0, 4, 3, 3), (4, 2, 4, 2, 4), (4, 0, 1, 3, 10),
                          (4, 0, 2, 2, 6)])
def test_bulk_timing_counts_launches(time_blocking, steps_per_pass, every, launches, steps):
    """Timing mode records an event pair around every n-th bulk launch (a two-step launch
    covers two steps, a four-step launch four: 10 steps = 10, 5 or 3 launches, the last of
    them a pair) and leaves the result unchanged; the timed launches' steps are counted."""
    g = gc.load("p1_walls")
    out = []
    for timing in (0, every):
        par = params_from_namelist(g["params"])
        par.time_blocking = time_blocking
        par.steps_per_pass = steps_per_pass
        with Lattice(par) as lat:
            print(f"[gpu] timing={timing} time_blocking={time_blocking} steps_per_pass={steps_per_pass}", flush=True)
            lat.set_timing(timing)
            lat.step(2, 10)
            ms, n = lat.last_bulk_ms()
            if timing:
                t = lat.last_timing()
                assert t.timed_steps == steps and t.bulk_steps == 10
                assert n == launches
                assert 0 < ms < 1e3
                assert lat.last_step_ms() > 0
            else:
                assert n == 0 and ms == 0
            out.append(lat.fields())
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_large_grid_matches_oracle(oracle_mod):
    """A grid of 4.7M nodes with the cylinder wall 64 nodes wide.  FP64 at 4352 x 1088 (the reference's
    L_front/back/top/bot = 8/55/8/7 with ncy = 68) against the oracle, from the rest state
    with seeded +-1e-3 relative noise on every population: a single step from populations,
    one four-step pass, a single step; fields bit for bit and the drag/lift series."""
    p = dict(gc.load("p1_walls")["params"])
    p.update(ncy=68, L_front=8, L_back=55, L_top=8, L_bot=7, Re=200.0, uo=0.05, nx=0, ny=0,
             iplot=100000, statsbegin=2)
    o = oracle_mod.Oracle(gc.oracle_params(oracle_mod, p))
    assert o.nx * o.ny >= 1 << 22
    par = params_from_namelist(p)
    assert (par.nx, par.ny) == (o.nx, o.ny)
    f0 = o.f() * (1.0 + np.random.default_rng(11).uniform(-1e-3, 1e-3, (o.nx, o.ny, 9)))
    f0 = np.asfortranarray(f0)
    o.set_state(f0)
    with Lattice(par) as lat:
        print(f"[gpu] large grid {lat.nx}x{lat.ny} vs oracle, wall levels={lat.info.levels}", flush=True)
        lat.set_populations(f0)
        lat.step(2, 6, step_opts_from_namelist(p))
        fd, fl = [], []
        for it in range(2, 8):
            o.step(it)
            a, b = o.forces()
            fd.append(a)
            fl.append(b)
        rho, ux, uy = lat.fields()
        orho, oux, ouy, *_ = o.moments()
        np.testing.assert_array_equal(rho, orho)
        np.testing.assert_array_equal(ux, oux)
        np.testing.assert_array_equal(uy, ouy)
        it, gfd, gfl = lat.forces()
        np.testing.assert_array_equal(it, np.arange(2, 8, dtype=np.int32))
        # The net force sums nbct node forces of size up to ~1/3 (the wall pressure) that
        # nearly cancel here; another summation order moves it by ~eps * sum|node force| per
        # add, so the absolute bound is 1e-15 per wall node.
        atol = 1e-15 * o.nbct
        np.testing.assert_allclose(gfd, fd, rtol=1e-12, atol=atol)
        np.testing.assert_allclose(gfl, fl, rtol=1e-12, atol=atol)


def test_400x100_channel_20000_steps_matches_oracle(oracle_mod):
    """The CPU-runnable channel at its full length: 400 x 100, cylinder r = 10 (the reference's
    geometry with ncy = 20, L_front/back/top/bot = 3/16/2/2: centre (71, 50)), Re = 20,
    uo = 0.05, from rest, 20000 FP64 steps with the drag/lift series over the last 10000:
    fields bit for bit and every recorded force against the oracle."""
    p = dict(gc.load("p1_walls")["params"])
    p.update(ncy=20, L_front=3, L_back=16, L_top=2, L_bot=2, Re=20.0, uo=0.05, nx=0, ny=0,
             iplot=100000, statsbegin=10002, statsend=200001, post_process=True)
    o = oracle_mod.Oracle(gc.oracle_params(oracle_mod, p))
    assert (o.nx, o.ny) == (400, 100)
    with Lattice(params_from_namelist(p)) as lat:
        print(f"[gpu] 400x100 x 20000 steps vs oracle ({oracle_mod.num_threads()} threads)", flush=True)
        lat.step(2, 20000, step_opts_from_namelist(p))
        rho, ux, uy = lat.fields()
        it, gfd, gfl = lat.forces()
    o.run(2, 10000)
    fd, fl = [], []
    for k in range(10002, 20002):
        o.step(k)
        a, b = o.forces()
        fd.append(a)
        fl.append(b)
    orho, oux, ouy, *_ = o.moments()
    np.testing.assert_array_equal(rho, orho)
    np.testing.assert_array_equal(ux, oux)
    np.testing.assert_array_equal(uy, ouy)
    np.testing.assert_array_equal(it, np.arange(10002, 20002, dtype=np.int32))
    scale = max(np.abs(fd).max(), np.abs(fl).max())
    np.testing.assert_allclose(gfd, fd, rtol=1e-12, atol=1e-12 * scale)
    np.testing.assert_allclose(gfl, fl, rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("precision", [capi.MRLBM_FP64, capi.MRLBM_FP32_DEV])
@pytest.mark.parametrize("name,probes", [
    # the reference's probe points (2d_mrlbm.f90:104-118): in the cylinder box and the bulk
    ("p1_walls", ((29, 21), (33, 21), (37, 21), (41, 21))),
    # row 2 (band at t+1, bulk at t+2), column 1 (band), the outlet column (ring), bulk
    ("p1_walls", ((60, 3), (2, 20), (88, 20), (60, 20))),
    ("s_xyper_noise", ((40, 3), (2, 20), (56, 20), (45, 20))),
    # nodes the four-step bulk writes that its first sub-steps also write (column 4, the
    # band rows near the top, column nx-5) and one that only the bulk writes
    ("p1_walls", ((5, 20), (60, 34), (84, 20), (8, 30))),
])
def test_probes_in_multi_step_passes_match_single_steps(name, probes, precision):
    """Probe records inside two- and four-step passes (nodes the bulk carries between its
    stores recorded by their own kernel, the rest by the side kernels) equal those of single
    steps, bit for bit, with the statistics window opening and closing inside a pass."""
    g = gc.load(name)
    p = dict(g["params"])
    p.update(post_process=True, statsbegin=7, statsend=50)
    out = {}
    for spp in (1, 4, 2):  # steps per pass: single steps, four-step passes, two-step pairs
        par = params_from_namelist(p, precision=precision)
        par.time_blocking = 0 if spp == 1 else -1
        par.steps_per_pass = 2 if spp == 2 else 0
        with Lattice(par) as lat:
            if "f0" in g:
                lat.set_populations(g["f0"])
            print(f"[gpu] probes {name} {probes} precision={precision} steps per pass {spp}", flush=True)
            o = step_opts_from_namelist(p)
            o.nprobe = len(probes)
            for k, (i, j) in enumerate(probes):
                o.probe_i[k], o.probe_j[k] = i, j
            lat.set_timing(1)
            lat.step(2, 60, o)
            _, launches = lat.last_bulk_ms()
            it, v = lat.probes()
            out[spp] = (launches, it, v, lat.fields())
    assert out[1][0] == 60 and out[4][0] < 20 and out[2][0] < 40, [out[k][0] for k in out]  # bulk launches
    np.testing.assert_array_equal(out[1][1], np.arange(7, 51, dtype=np.int32))
    for spp in (4, 2):
        np.testing.assert_array_equal(out[spp][1], out[1][1])
        np.testing.assert_array_equal(out[spp][2], out[1][2])
        for a, b in zip(out[spp][3], out[1][3]):
            np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [capi.MRLBM_FP64, capi.MRLBM_FP32_DEV])
@pytest.mark.parametrize("interval", [1, 2, 3])
def test_u_statistics_in_multi_step_passes_match_single_steps(oracle_mod, precision, interval):
    """Mean / rms velocity accumulators sampled inside four-step passes and two-step pairs
    (the bulk accumulates the sampled levels of the nodes it writes, the side kernels the
    rest) equal those of single steps bit for bit; in FP64 the means also equal the oracle's
    fields averaged over the same steps."""
    g = gc.load("p1_walls")
    p = dict(g["params"])
    p.update(post_process=True, statsbegin=5, statsend=60)
    out = {}
    for spp in (1, 4, 2):  # steps per pass: single steps, four-step passes, two-step pairs
        par = params_from_namelist(p, precision=precision)
        par.time_blocking = 0 if spp == 1 else -1
        par.steps_per_pass = 2 if spp == 2 else 0
        with Lattice(par) as lat:
            print(f"[gpu] u statistics precision={precision} interval={interval} steps per pass {spp}", flush=True)
            o = step_opts_from_namelist(p)
            o.stats_interval = interval
            lat.set_timing(1)
            lat.step(2, 70, o)
            _, launches = lat.last_bulk_ms()
            out[spp] = (launches, lat.stats(), lat.fields())
    assert out[1][0] == 70 and out[4][0] < 20 and out[2][0] < 50, [out[k][0] for k in out]  # bulk launches
    for spp in (4, 2):
        for a, b in zip(out[spp][1], out[1][1]):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(out[spp][2], out[1][2]):
            np.testing.assert_array_equal(a, b)
    if precision == capi.MRLBM_FP64:
        orc = oracle_mod.Oracle(gc.oracle_params(oracle_mod, p))
        su, sv, n = 0.0, 0.0, 0
        for it in range(2, 72):
            orc.step(it)
            if 5 <= it <= 60 and (it - 5) % interval == 0:
                _, ux, uy, *_ = orc.moments()
                su, sv, n = su + ux, sv + uy, n + 1
        mean_ux, mean_uy, _, _, ns = out[4][1]
        assert ns == n
        np.testing.assert_allclose(mean_ux, su / n, rtol=1e-13, atol=1e-17)
        np.testing.assert_allclose(mean_uy, sv / n, rtol=1e-13, atol=1e-17)


def test_full_size_recording_in_passes_matches_single_steps():
    """8192 x 2048 FP64 (the benchmark channel, D = 128) with everything recorded: forces,
    surface pressure, four probes and u statistics every 3rd step inside a window that
    opens and closes mid-pass.  Four (and two) steps per pass against one step per pass:
    fields, force, probe and surface-pressure series and the statistics bit for bit."""
    import bench
    out = {}
    for tb in (0, -1):
        print(f"[gpu] 8192x2048 FP64 recording, time_blocking={tb}", flush=True)
        lat = bench.make_lattice(capi.MRLBM_FP64, time_blocking=tb)
        try:
            o = capi.default_step_opts()
            o.post_process, o.statsbegin, o.statsend = 1, 41, 220
            o.record_surface_p, o.stats_interval = 1, 3
            o.nprobe = 4
            for k, (i, j) in enumerate(((1160, 1026), (1094, 1026), (2000, 3), (8192, 700))):
                o.probe_i[k], o.probe_j[k] = i, j
            lat.step(2, 240, o)
            out[tb] = (lat.fields(), lat.forces(), lat.probes(), lat.surface_p_series(), lat.stats())
        finally:
            lat.close()
    (f0, fo0, p0, s0, st0), (f1, fo1, p1, s1, st1) = out[0], out[-1]
    for a, b in zip(f1, f0):
        np.testing.assert_array_equal(a, b)
    for group1, group0 in ((fo1, fo0), (p1, p0), (s1, s0)):
        for a, b in zip(group1, group0):
            np.testing.assert_array_equal(a, b)
    assert len(fo1[0]) == 180 and st1[4] == st0[4] == 60
    for a, b in zip(st1[:4], st0[:4]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("nx,ny,xper,yper", [(6, 5, False, False), (12, 10, False, False), (12, 10, True, False),
                                             (6, 9, False, True), (17, 16, False, False)])
def test_small_channels_match_oracle(oracle_mod, nx, ny, xper, yper):
    """Plain channels below the sizes of the two-step pass (16 x 16) and of the ring-column
    tiles (8 wide), plus the smallest two-step grid: FP64 fields bit for bit against the
    oracle after 60 steps, with x or y periodic wraps."""
    p = dict(gc.load("chan")["params"])
    p.update(nx=nx, ny=ny, x_periodic=xper, y_periodic=yper)
    o = oracle_mod.Oracle(gc.oracle_params(oracle_mod, p))
    assert (o.nx, o.ny) == (nx, ny)
    o.run(2, 60)
    orho, oux, ouy, *_ = o.moments()
    with Lattice(params_from_namelist(p)) as lat:
        print(f"[gpu] small channel {nx}x{ny} xper={xper} yper={yper}", flush=True)
        lat.step(2, 60, step_opts_from_namelist(p))
        rho, ux, uy = lat.fields()
    np.testing.assert_array_equal(rho, orho)
    np.testing.assert_array_equal(ux, oux)
    np.testing.assert_array_equal(uy, ouy)
