"""Closed-form and brute-force pins of the oracle, for what the reference runs leave
open or pin only in combination.  Each check is chosen so that a plausible slip (a
dropped term, a sign, a transposed operand) fails it."""
import numpy as np
import pytest

from tests import golden_cases as gc

W_TRUE = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


def test_weights_are_real4_rounded(oracle_mod):
    o = oracle_mod.Oracle(oracle_mod.OracleParams(channel_with_cylinder=0, nx=8, ny=8, post_process=0))
    w = o.weights()
    # REAL(4) literals widened (2d_mrlbm.f90:129): c times the true weights, c = 1 + 2^-27
    np.testing.assert_array_equal(w, np.float32(np.array([16, 4, 4, 4, 4, 1, 1, 1, 1], np.float32) / np.float32(36)).astype(np.float64))
    c = 1.0 + 2.0 ** -27
    np.testing.assert_allclose(w / W_TRUE, c, rtol=0, atol=1e-16)
    assert w.sum() == c


def test_rest_state_mass_grows_by_c_per_step(oracle_mod):
    """Fully periodic box at rest: every reconstruction multiplies mass by sum(w)."""
    p = oracle_mod.OracleParams(channel_with_cylinder=0, nx=12, ny=10, x_periodic=1, y_periodic=1,
                                post_process=0, Re=100.0, uo=0.05)
    o = oracle_mod.Oracle(p)
    c = 1.0 + 2.0 ** -27
    for n in range(1, 41):
        o.step(1 + n)
        rho, ux, uy, mxx, myy, mxy = o.moments()
        np.testing.assert_allclose(rho, c ** n, rtol=4e-16 * n, atol=0)
        assert np.all(ux == 0) and np.all(uy == 0) and np.all(mxy == 0)
        np.testing.assert_allclose(mxx, 0.0, atol=1e-16)


def test_uniform_flow_is_preserved(oracle_mod):
    """Periodic box, uniform u: u and the Hermite moments stay uniform (no gradients)."""
    p = oracle_mod.OracleParams(channel_with_cylinder=0, nx=9, ny=7, x_periodic=1, y_periodic=1,
                                post_process=0, Re=50.0, uo=0.05)
    o = oracle_mod.Oracle(p)
    w = o.weights()
    ux0, uy0 = 0.031, -0.017
    cu = EX * ux0 + EY * uy0
    feq = w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux0 ** 2 + uy0 ** 2))
    o.set_state(np.broadcast_to(feq, (9, 7, 9)).copy())
    for it in range(2, 30):
        o.step(it)
    rho, ux, uy, *_ = o.moments()
    assert np.ptp(ux) < 1e-16 and np.ptp(uy) < 1e-16 and np.ptp(rho) < 1e-15
    np.testing.assert_allclose(ux, ux0, rtol=1e-14)
    np.testing.assert_allclose(uy, uy0, rtol=1e-14)


def test_reconstruction_reproduces_moments(oracle_mod):
    """Hermite identities of the regularized populations: sum fout = c rho,
    sum fout c = c rho u, sum fout H = c rho m (brute force on a tiny periodic grid)."""
    p = oracle_mod.OracleParams(channel_with_cylinder=0, nx=6, ny=5, x_periodic=1, y_periodic=1,
                                post_process=0, Re=30.0, uo=0.05)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(0)
    w = o.weights()
    ux, uy = rng.uniform(-0.05, 0.05, (2, 6, 5))
    cu = EX[None, None] * ux[..., None] + EY[None, None] * uy[..., None]
    f = w * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (ux ** 2 + uy ** 2)[..., None])
    o.set_state(f)
    o.step(2)
    fout = o.fout()
    rho, uxm, uym, mxx, myy, mxy = o.moments()
    c = w.sum()
    Hxx = EX ** 2 - 1 / 3
    Hyy = EY ** 2 - 1 / 3
    np.testing.assert_allclose(fout.sum(-1), c * rho, rtol=1e-14)
    np.testing.assert_allclose((fout * EX).sum(-1), c * rho * uxm, atol=1e-16)
    np.testing.assert_allclose((fout * EY).sum(-1), c * rho * uym, atol=1e-16)
    np.testing.assert_allclose((fout * Hxx).sum(-1), c * rho * mxx, atol=1e-16)
    np.testing.assert_allclose((fout * Hyy).sum(-1), c * rho * myy, atol=1e-16)
    np.testing.assert_allclose((fout * EX * EY).sum(-1), c * rho * mxy, atol=1e-16)
    # collision (2d_mrlbm.f90:410-412) on the streamed moments
    om = o.omega
    f2 = o.f()
    r2 = f2.sum(-1)
    u2 = (f2 * EX).sum(-1) / r2
    m2 = (f2 * Hxx).sum(-1) / r2
    o.step(3)
    _, uxn, _, mxxn, _, _ = o.moments()
    np.testing.assert_allclose(uxn, u2, atol=1e-16)
    np.testing.assert_allclose(mxxn, (1 - om) * m2 + om * u2 * u2, atol=1e-16)


def test_streaming_is_a_pure_shift(oracle_mod):
    """Periodic box: after one step, f_k(x) equals fout_k(x - c_k) of the step, with the
    reference's same-row / same-column wraps at the box edges (2d_mrlbm.f90:519-535)."""
    p = oracle_mod.OracleParams(channel_with_cylinder=0, nx=5, ny=4, x_periodic=1, y_periodic=1,
                                post_process=0, Re=30.0, uo=0.05)
    o = oracle_mod.Oracle(p)
    rng = np.random.default_rng(1)
    w = o.weights()
    f = w * (1 + rng.uniform(-0.01, 0.01, (5, 4, 9)))
    o.set_state(f)
    o.step(2)
    fout, fs = o.fout(), o.f()
    for k in range(9):
        for i in range(5):
            for j in range(4):
                si, sj = i - EX[k], j - EY[k]
                if 0 <= sj < 4 and 0 <= si < 5:
                    src = (si, sj)
                elif not (0 <= sj < 4):
                    src = (i, sj % 4)      # y-wrap, same column, applied last
                else:
                    src = (si % 5, j)      # x-wrap, same row
                assert fs[i, j, k] == fout[src[0], src[1], k], (i, j, k)


def test_solve_system_is_cramer(oracle_mod):
    rng = np.random.default_rng(2)
    for _ in range(20):
        A = rng.normal(size=(3, 3))
        b = rng.normal(size=3)
        np.testing.assert_allclose(oracle_mod.solve_system(A, b), np.linalg.solve(A, b), rtol=1e-10)


def test_quadratic_fit_is_exact_for_quadratics(oracle_mod):
    rng = np.random.default_rng(3)
    for _ in range(20):
        a0, a1, a2 = rng.normal(size=3)
        th = rng.uniform(0, 2 * np.pi)
        rs = rng.uniform(5, 30)
        rr = rs + np.sqrt(2) * np.array([1, 2, 3])
        pts = [(r * np.cos(th), r * np.sin(th), a0 + a1 * r + a2 * r * r) for r in rr]
        xs, ys = rs * np.cos(th), rs * np.sin(th)
        args = [v for pnt in pts for v in pnt] + [xs, ys]
        np.testing.assert_allclose(oracle_mod.quadratic_interpolation(*args), a0 + a1 * rs + a2 * rs * rs,
                                   rtol=1e-9, atol=1e-9)


def test_gaussian_smoothing_brute_force(oracle_mod):
    rng = np.random.default_rng(4)
    th = np.sort(rng.uniform(0, 2 * np.pi, 37))
    p = rng.normal(size=37)
    W = np.exp(-((th[:, None] - th[None, :]) ** 2) / (2 * 0.1 ** 2))
    np.testing.assert_allclose(oracle_mod.gaussian_smoothing(th, p), (W @ p) / W.sum(1), rtol=1e-13)


def test_geometry_matches_input_stat(oracle_mod):
    """input_stat.dat:2-3 holds nbct = 160 and r_cy = 20.39610 for the ncy = 40 cylinder."""
    o = oracle_mod.Oracle(oracle_mod.OracleParams(ncy=40, L_front=2, L_back=3, L_top=2, L_bot=2))
    assert o.nbct == 160
    # r_cyl = sqrt(416) = 20.396078...; the file carries it to about 5 significant decimals
    assert o.r_cyl == np.sqrt(416.0)
    assert abs(o.r_cyl - 20.39610) < 3e-5


def test_labels_follow_the_closed_rule(oracle_mod):
    """Every boundary label of integer-centred cylinders is one the reference sets
    (2d_mrlbm.f90:2011-2057); nbct = 4*ncy for even ncy."""
    for ncy in (6, 8, 10, 16, 20, 24):
        o = oracle_mod.Oracle(oracle_mod.OracleParams(ncy=ncy, L_front=2, L_back=3, L_top=2, L_bot=2))
        lab = set(o.boundary()["label"].tolist())
        assert lab <= {1, 2, 3, 4, 5, 7, 8, 10, 11, 12, 13, 14}
        assert o.nbct == 4 * ncy


@pytest.mark.parametrize("name", gc.names("statistics"))
def test_statistics_matches_reference_postprocessor(oracle_mod, name):
    g = gc.load(name)
    r = oracle_mod.statistics(g["fd"], g["fl"], float(g["uo"]), float(g["rho_infty"]), float(g["r_cy"]))
    # plots/drag_lift.dat is list-directed text (17 significant digits)
    np.testing.assert_allclose(r["st"], g["st"], rtol=1e-15)
    np.testing.assert_allclose(r["cd"], g["cd"], rtol=1e-15)
    np.testing.assert_allclose(r["cl_rms"], g["cl_rms"], rtol=1e-15)
    t = np.arange(len(g["fl"]))
    peaks = np.where((g["fl"][1:-1] > g["fl"][:-2]) & (g["fl"][1:-1] > g["fl"][2:]))[0] + 2
    assert r["cycle_start"] == peaks[0] and r["period"] == peaks[1] - peaks[0] == 137
    assert t.size == 1500


def test_statistics_no_two_peaks(oracle_mod):
    fl = np.linspace(0, 1, 50)
    assert oracle_mod.statistics(np.ones(50), fl, 0.05, 1.0, 10.0) is None
