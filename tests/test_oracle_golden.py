"""The oracle against the reference's own outputs (tests/golden/*.npz).

Fixtures come from the reference program compiled from its own sources
(2d_mrlbm.f90) and run single-threaded; see tests/golden/make_fixtures.py.
"""
import numpy as np
import pytest

from tests import golden_cases as gc


@pytest.mark.parametrize("name", gc.names())
def test_oracle_matches_reference(oracle_mod, name):
    g = gc.load(name)
    o = oracle_mod.Oracle(gc.oracle_params(oracle_mod, g["params"]))
    assert (o.nx, o.ny) == (int(g["nx"]), int(g["ny"]))
    if "nbct" in g:
        assert o.nbct == int(g["nbct"])
        assert abs(o.r_cyl - float(g["r_cyl"])) < 1e-13
        np.testing.assert_allclose(np.sort(o.boundary()["theta"]), g["theta"], rtol=0, atol=1e-15)
    assert abs(1.0 / o.omega - float(g["tau"])) < 1e-14
    if "f0" in g:
        o.set_state(g["f0"])
    nsteps = int(g["nsteps"])
    fd, fl = [], []
    for it in range(2, 2 + nsteps):
        o.step(it)
        if "forces" in g:
            a, b = o.forces()
            fd.append(a)
            fl.append(b)
    # The restatement is bit-exact against the flang build of the reference.
    np.testing.assert_array_equal(o.f(), g["f"])
    rho, ux, uy, *_ = o.moments()
    np.testing.assert_array_equal(rho, g["rho"])
    np.testing.assert_array_equal(ux, g["ux"])
    np.testing.assert_array_equal(uy, g["uy"])
    if "forces" in g:
        # forces_time.dat is list-directed text with 16-17 significant digits
        np.testing.assert_allclose(np.array(fd), g["forces"][:, 1], rtol=1e-15, atol=1e-17)
        np.testing.assert_allclose(np.array(fl), g["forces"][:, 2], rtol=1e-15, atol=1e-17)
        ps, pf = o.surface_p()
        np.testing.assert_array_equal(ps, g["ps_last"])
        np.testing.assert_array_equal(pf, g["pfit_last"])
