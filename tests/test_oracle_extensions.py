"""Pins for the oracle's extensions beyond the reference (explicit grid and cylinder centre,
parabolic inlet with Re on the mean velocity, the label-fluid policy, the Brinkman field),
by properties the mathematics fixes:

* an explicit integer centre reproduces the reference's derived geometry and its golden
  outputs bit for bit;
* the parabolic inlet column equals 4 uo s (1 - s) and nu follows Re on 2/3 uo;
* a cylinder-free channel with the parabolic inlet holds the Poiseuille profile;
* the label-fluid policy removes exactly the nodes of labels 0/6/9/15, whose reference
  masks are empty (NaN in the reference), and keeps r_cyl;
* a zero Brinkman field changes nothing; a uniform one makes a uniform flow decay as
  (1 - chi)^n wherever the domain's outer ring (where chi does not act) is more than n
  nodes away; on a cylinder it leaves the wall nodes of the first step untouched and scales
  every other interior velocity by (1 - chi).
"""
import numpy as np
import pytest

from tests import golden_cases as gc


def _params(po, **kw):
    p = po.OracleParams(post_process=0, statsbegin=10 ** 8, statsend=10 ** 8 + 1, iplot=10 ** 8)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _equilibrium(nx, ny, ux, uy, w):
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], float)
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], float)
    ux = np.broadcast_to(ux, (nx, ny))
    uy = np.broadcast_to(uy, (nx, ny))
    cu = ex * ux[..., None] + ey * uy[..., None]
    return w * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * (ux * ux + uy * uy)[..., None])


@pytest.mark.parametrize("name", ["p1_walls", "s_periodic", "s_noise_re60"])
def test_explicit_integer_centre_reproduces_reference_geometry_and_run(oracle_mod, name):
    po = oracle_mod
    g = gc.load(name)
    ref = gc.oracle_params(po, g["params"])
    o_ref = po.Oracle(ref)
    ncy = ref.ncy
    xc = ref.L_front * ncy + ncy // 2 + 1           # 2d_mrlbm.f90:946-948
    yc = ref.L_bot * ncy + ncy // 2
    ex = gc.oracle_params(po, g["params"])
    ex.geometry_explicit, ex.nx, ex.ny = 1, o_ref.nx, o_ref.ny
    ex.x_c, ex.y_c, ex.r_c = float(xc), float(yc), float(ncy // 2)
    o = po.Oracle(ex)
    assert (o.nx, o.ny, o.nbct, o.r_cyl, o.nu) == (o_ref.nx, o_ref.ny, o_ref.nbct, o_ref.r_cyl, o_ref.nu)
    for a, b in zip(o.boundary().values(), o_ref.boundary().values()):
        np.testing.assert_array_equal(a, b)
    if "f0" in g:
        o.set_state(g["f0"])
    o.run(2, int(g["nsteps"]))
    np.testing.assert_array_equal(o.f(), g["f"])


def test_parabolic_inlet_column_and_viscosity(oracle_mod):
    po = oracle_mod
    uo, re, ny = 0.05, 20.0, 100
    p = _params(po, geometry_explicit=1, nx=400, ny=ny, x_c=80.0, y_c=50.5, r_c=10.0, Re=re, uo=uo,
                inlet_parabolic=1, label_fluid=1, y_periodic=0)
    o = po.Oracle(p)
    # Re on the mean of the parabola, 2/3 uo, and the diameter 2 r_cyl (2d_mrlbm.f90:134)
    assert o.nu == pytest.approx((2.0 / 3.0) * uo * 2.0 * o.r_cyl / re, rel=1e-15)
    assert o.omega == pytest.approx(1.0 / (3.0 * o.nu + 0.5), rel=1e-15)
    o.run(2, 3)
    _, ux, uy, *_ = o.moments()
    s = np.arange(ny) / (ny - 1.0)
    np.testing.assert_allclose(ux[0], 4.0 * uo * s * (1.0 - s), rtol=1e-15, atol=1e-19)
    assert ux[0, 0] == 0.0 and ux[0, -1] == 0.0 and np.all(uy[0] == 0.0)
    # the column's sum in closed form: with s_j = j / m (j = 0..m, m = ny - 1),
    # sum_j 4 s_j (1 - s_j) = 4 (sum j / m - sum j^2 / m^2)
    n1 = ny - 1.0
    assert ux[0].sum() == pytest.approx(4.0 * uo * ((n1 * (n1 + 1) / 2) / n1 - (n1 * (n1 + 1) * (2 * n1 + 1) / 6) / n1 ** 2),
                                        rel=1e-13)


def test_parabolic_inlet_holds_poiseuille_profile(oracle_mod):
    po = oracle_mod
    nx, ny, uo = 64, 21, 0.05
    p = _params(po, channel_with_cylinder=0, nx=nx, ny=ny, Re=5.0, uo=uo, inlet_parabolic=1, y_periodic=0)
    o = po.Oracle(p)
    o.run(2, 6000)
    _, ux, uy, *_ = o.moments()
    s = np.arange(ny) / (ny - 1.0)
    want = 4.0 * uo * s * (1.0 - s)
    for i in (nx // 4, nx // 2, 3 * nx // 4):
        err = np.abs(ux[i] - want).max() / uo
        assert err < (5e-3 if i == nx // 2 else 1e-2), (i, err)
        assert np.abs(uy[i]).max() < 1e-3 * uo
    # symmetric about the mid-channel row (an inlet profile or wall rule off by a row is not)
    np.testing.assert_allclose(ux[1:-1], ux[1:-1, ::-1], rtol=0, atol=1e-12)
    # (the mass flux differs between columns by ~0.7% at steady state: the regularized wall
    # rule, 2d_mrlbm.f90:1327, does not conserve mass, so flux is not a pin here)


def test_label_fluid_policy_removes_exactly_the_unmasked_labels(oracle_mod):
    po = oracle_mod
    kw = dict(geometry_explicit=1, nx=400, ny=100, x_c=80.0, y_c=50.5, r_c=10.0, Re=20.0, uo=0.05,
              inlet_parabolic=1, y_periodic=0)
    o_ref = po.Oracle(_params(po, **kw))
    o_fl = po.Oracle(_params(po, label_fluid=1, **kw))
    b_ref, b_fl = o_ref.boundary(), o_fl.boundary()
    unset = np.isin(b_ref["label"], [0, 6, 9, 15])
    assert unset.sum() == 2                       # the half-integer centre's two label-15 nodes
    assert o_fl.nbct == o_ref.nbct - unset.sum() and o_fl.r_cyl == o_ref.r_cyl
    np.testing.assert_array_equal(b_fl["bou_i"], b_ref["bou_i"][~unset])
    np.testing.assert_array_equal(b_fl["bou_j"], b_ref["bou_j"][~unset])
    np.testing.assert_array_equal(b_fl["label"], b_ref["label"][~unset])
    fluid_ref, _, bound_ref = o_ref.geometry()
    fluid_fl, _, bound_fl = o_fl.geometry()
    for i, j in zip(b_ref["bou_i"][unset] - 1, b_ref["bou_j"][unset] - 1):
        assert bound_ref[i, j] == 1 and bound_fl[i, j] == 0 and fluid_fl[i, j] == 1
    # with the reference's empty masks those nodes divide by zero (2d_mrlbm.f90:1713)
    o_ref.run(2, 2)
    o_fl.run(2, 200)
    rho_ref = o_ref.moments()[0]
    assert np.isnan(rho_ref).any()
    assert np.isfinite(np.stack(o_fl.moments())).all()


def test_zero_brinkman_field_changes_nothing(oracle_mod):
    po = oracle_mod
    g = gc.load("p1_walls")
    a = po.Oracle(gc.oracle_params(po, g["params"]))
    b = po.Oracle(gc.oracle_params(po, g["params"]))
    b.set_brinkman(np.zeros((a.nx, a.ny)))
    a.run(2, 300)
    b.run(2, 300)
    np.testing.assert_array_equal(a.f(), b.f())
    b.set_brinkman(None)
    a.run(302, 10)
    b.run(302, 10)
    np.testing.assert_array_equal(a.f(), b.f())


def test_uniform_brinkman_field_decays_uniform_flow(oracle_mod):
    po = oracle_mod
    nx = ny = 48
    chi, u0, v0, n = 0.01, 0.02, -0.01, 16
    p = _params(po, channel_with_cylinder=0, nx=nx, ny=ny, Re=10.0, uo=0.0, x_periodic=1, y_periodic=1)
    o = po.Oracle(p)
    w = o.weights()
    o.set_state(_equilibrium(nx, ny, u0, v0, w))
    o.set_brinkman(np.full((nx, ny), chi))
    c = w.sum()                                   # the REAL(4) weights' sum (2d_mrlbm.f90:129)
    for t in range(1, n + 1):
        o.step(1 + t)
        rho, ux, uy, *_ = o.moments()
        core = (slice(t + 1, nx - t - 1), slice(t + 1, ny - t - 1))   # more than t nodes from the ring
        np.testing.assert_allclose(ux[core], u0 * (1 - chi) ** t, rtol=1e-13)
        np.testing.assert_allclose(uy[core], v0 * (1 - chi) ** t, rtol=1e-13)
        np.testing.assert_allclose(rho[core], c ** t, rtol=1e-14)   # the start's populations sum to c
        if t == 1:                                # the ring is not penalized
            np.testing.assert_allclose(ux[0], u0, rtol=1e-14)
            np.testing.assert_allclose(ux[:, -1], u0, rtol=1e-14)


def test_brinkman_field_skips_ring_and_wall_nodes(oracle_mod):
    po = oracle_mod
    kw = dict(geometry_explicit=1, nx=128, ny=48, x_c=30.0, y_c=24.5, r_c=6.0, Re=50.0, uo=0.05,
              inlet_parabolic=1, label_fluid=1, y_periodic=0)
    a = po.Oracle(_params(po, **kw))
    b = po.Oracle(_params(po, **kw))
    rng = np.random.default_rng(4)
    ux0 = rng.uniform(-1e-3, 1e-3, (a.nx, a.ny))
    uy0 = rng.uniform(-1e-3, 1e-3, (a.nx, a.ny))
    f0 = _equilibrium(a.nx, a.ny, ux0, uy0, a.weights())
    a.set_state(f0)
    b.set_state(f0)
    chi = rng.uniform(0.0, 0.3, (a.nx, a.ny))
    b.set_brinkman(chi)
    a.step(2)
    b.step(2)
    ra, uxa, uya, *ma = a.moments()
    rb, uxb, uyb, *mb = b.moments()
    _, _, bound = a.geometry()
    interior = np.zeros((a.nx, a.ny), bool)
    interior[1:-1, 1:-1] = True
    pen = interior & (bound == 0)
    np.testing.assert_array_equal(rb, ra)                         # the density is untouched
    np.testing.assert_array_equal(uxb[~pen], uxa[~pen])           # ring and wall nodes
    np.testing.assert_array_equal(uyb[~pen], uya[~pen])
    for x, y in zip(ma, mb):
        np.testing.assert_array_equal(y[bound == 1], x[bound == 1])
    np.testing.assert_allclose(uxb[pen], uxa[pen] * (1 - chi[pen]), rtol=1e-15, atol=1e-20)
    np.testing.assert_allclose(uyb[pen], uya[pen] * (1 - chi[pen]), rtol=1e-15, atol=1e-20)
