"""Discrete adjoint of the time-averaged drag (mrlbm_adjoint_drag): gradients with respect
to the Brinkman solid fraction field and the inlet velocity, checked against central
finite differences of the forward run (FP64), against the FP64 adjoint (FP32 form), and
through a torch.autograd.Function wrapper (the only place torch meets the adjoint)."""
import numpy as np
import pytest

from mrlbm_2d_amd import capi

pytestmark = pytest.mark.gpu

NX, NY = 96, 40
XC, YC, RC = 24.0, 20.5, 4.5          # half-integer centre, as the 512x128 adjoint case
RE, UO = 40.0, 0.05
NSTEPS, WINDOW, K = 300, 150, 64        # 300 = 4*64 + 44: a partial last segment


def params(precision=capi.MRLBM_FP64, uo=UO, omega=0.0):
    p = capi.default_params()
    p.nx, p.ny = NX, NY
    p.x_periodic, p.y_periodic = 0, 0
    p.channel_with_cylinder = 1
    p.x_c, p.y_c, p.r_c = XC, YC, RC
    p.Re, p.uo = RE, uo
    p.omega = omega
    p.inlet_profile = capi.INLET_PARABOLIC
    p.precision = precision
    return p


def bumps(seed, amp, n=6):
    """Smooth Gaussian bumps within 2D of the cylinder centre."""
    rng = np.random.default_rng(seed)
    x = np.arange(1, NX + 1)[:, None]
    y = np.arange(1, NY + 1)[None, :]
    chi = np.zeros((NX, NY))
    for _ in range(n):
        r = rng.uniform(RC + 1.5, 4 * RC)
        th = rng.uniform(0, 2 * np.pi)
        cx, cy = XC + r * np.cos(th), YC + r * np.sin(th)
        chi += amp * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 2.0 ** 2))
    return chi


def drag_objective(p, chi):
    """J = mean Cd over the window from a plain forward run."""
    with capi.Lattice(p) as lat:
        lat.set_brinkman(chi)
        o = capi.default_step_opts()
        o.post_process = 1
        o.statsbegin = 2 + NSTEPS - WINDOW
        o.statsend = 1 + NSTEPS
        lat.step(2, NSTEPS, o)
        _, fd, _ = lat.forces()
        assert len(fd) == WINDOW
        uref = (2.0 / 3.0) * p.uo
        return float(np.sum(fd) / WINDOW / (0.5 * uref * uref * 2 * lat.info.r_cyl))


def adjoint(p, chi):
    with capi.Lattice(p) as lat:
        lat.set_brinkman(chi)
        res, g = lat.adjoint_drag(2, NSTEPS, WINDOW, K)
        omega = lat.info.omega
        return res, g, omega


@pytest.fixture(scope="module")
def fp64():
    chi = bumps(7, 0.01)
    res, g, omega = adjoint(params(), chi)
    print(f"[gpu] adjoint fp64: J {res.J:.10f} dJ/duo {res.dJ_duo:.6e} fwd {res.forward_ms:.2f} ms "
          f"adj {res.adjoint_ms:.2f} ms", flush=True)
    return chi, res, g, omega


def test_serial_recompute_gives_the_same_gradient(fp64):
    """One segment buffer on one stream, and the recompute of the next segment overlapped
    on a second stream, run the same kernels on the same states: identical outputs."""
    chi, res, g, _ = fp64
    with capi.Lattice(params()) as lat:
        lat.set_brinkman(chi)
        res1, g1 = lat.adjoint_drag(2, NSTEPS, WINDOW, K, serial_recompute=True)
    assert res1.J == res.J and res1.dJ_duo == res.dJ_duo
    np.testing.assert_array_equal(g1, g)


def test_objective_matches_forward_run(fp64):
    chi, res, _, _ = fp64
    # same forces; only the order of the window sum differs
    assert abs(res.J - drag_objective(params(), chi)) <= 1e-14 * abs(res.J)


def test_gradient_zero_where_chi_has_no_effect(fp64):
    _, _, g, _ = fp64
    with capi.Lattice(params()) as lat:
        cls = lat.node_class()
    assert np.all(g[0, :] == 0) and np.all(g[-1, :] == 0) and np.all(g[:, 0] == 0) and np.all(g[:, -1] == 0)
    assert np.all(g[cls == 3] == 0)
    assert np.abs(g).max() > 0


def test_directional_derivative_matches_central_difference(fp64):
    chi, res, g, _ = fp64
    v = bumps(11, 1.0)
    ad = float(np.sum(g * v))
    err = {}
    for eps in (1e-4, 1e-5, 1e-6):
        fd = (drag_objective(params(), chi + eps * v) - drag_objective(params(), chi - eps * v)) / (2 * eps)
        err[eps] = abs(ad - fd) / abs(fd)
        print(f"[gpu] directional eps {eps:.0e}: adjoint {ad:.10e} fd {fd:.10e} rel {err[eps]:.2e}", flush=True)
    # central differences converge to the adjoint as eps^2 until round-off takes over
    assert min(err.values()) <= 1e-6


def test_pointwise_derivatives_match_central_difference(fp64):
    chi, _, g, _ = fp64
    flat = np.argsort(np.abs(g).ravel())[::-1]
    picks = [np.unravel_index(flat[r], g.shape) for r in (0, 7, 40)]
    # the largest gradient next to each wall, the inlet and the outlet column
    picks.append((int(np.argmax(np.abs(g[:, 1]))), 1))
    picks.append((int(np.argmax(np.abs(g[:, NY - 2]))), NY - 2))
    picks.append((1, int(np.argmax(np.abs(g[1, :])))))
    picks.append((NX - 2, int(np.argmax(np.abs(g[NX - 2, :])))))
    eps = 1e-5
    for (i, j) in picks:
        e = np.zeros_like(chi)
        e[i, j] = eps
        fd = (drag_objective(params(), chi + e) - drag_objective(params(), chi - e)) / (2 * eps)
        print(f"[gpu] node ({i},{j}): adjoint {g[i, j]:.10e} fd {fd:.10e}", flush=True)
        assert abs(g[i, j] - fd) <= 1e-5 * abs(fd) + 1e-9


def test_inlet_velocity_derivative_matches_central_difference(fp64):
    chi, res, _, omega = fp64
    h = 1e-6
    jp = drag_objective(params(uo=UO + h, omega=omega), chi)
    jm = drag_objective(params(uo=UO - h, omega=omega), chi)
    fd = (jp - jm) / (2 * h)
    print(f"[gpu] dJ/duo: adjoint {res.dJ_duo:.10e} fd {fd:.10e}", flush=True)
    assert abs(res.dJ_duo - fd) <= 1e-5 * abs(fd)


def test_fp32_adjoint_close_to_fp64(fp64):
    chi, r64, g64, _ = fp64
    r32, g32, _ = adjoint(params(capi.MRLBM_FP32_DEV), chi)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"[gpu] fp32 adjoint: J {r32.J:.8f} (fp64 {r64.J:.8f}) dJ/duo {r32.dJ_duo:.6e} "
          f"(fp64 {r64.dJ_duo:.6e}) grad rel L2 {rel:.3e}", flush=True)
    assert abs(r32.J - r64.J) <= 1e-4 * abs(r64.J)
    assert abs(r32.dJ_duo - r64.dJ_duo) <= 1e-3 * abs(r64.dJ_duo)
    assert rel < 1e-2


def test_torch_autograd_function_wraps_the_adjoint(fp64):
    torch = pytest.importorskip("torch")
    chi0, _, _, omega = fp64

    class DragCd(torch.autograd.Function):
        @staticmethod
        def forward(ctx, chi, uo):
            p = params(uo=float(uo), omega=omega)
            res, g, _ = adjoint(p, chi.detach().cpu().numpy())
            ctx.save_for_backward(torch.from_numpy(g), torch.tensor(res.dJ_duo, dtype=torch.float64))
            return torch.tensor(res.J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            g, duo = ctx.saved_tensors
            return gout * g, gout * duo

    chi = torch.tensor(chi0, dtype=torch.float64, requires_grad=True)
    uo = torch.tensor(UO, dtype=torch.float64, requires_grad=True)
    J = 3.0 * DragCd.apply(chi, uo)
    J.backward()
    v = torch.tensor(bumps(13, 1.0))
    eps = 1e-4
    with torch.no_grad():
        fd = 3.0 * (DragCd.apply(chi + eps * v, uo) - DragCd.apply(chi - eps * v, uo)) / (2 * eps)
    ad = float(torch.sum(chi.grad * v))
    assert abs(ad - float(fd)) <= 1e-6 * abs(float(fd))
    assert uo.grad is not None and np.isfinite(float(uo.grad))


def test_adjoint_512x128_case():
    """The 512 x 128 adjoint case (D = 24 at (100, 64.5), Re = 100, 4000 steps, mean Cd over
    the last 2000, checkpoints every 200): FP64 adjoint against a central difference along a
    smooth direction, FP32 adjoint against FP64."""
    from mrlbm_2d_amd import brinkman as B
    c = B.ADJ_CASE
    chi = B.adjoint_case_field()
    r64, g64, _ = B.run_adjoint_case(capi.MRLBM_FP64, chi)
    v = B.gaussian_bumps(c["nx"], c["ny"], c["xc"], c["yc"], c["radius"], 1.0, seed=9)
    ad = float(np.sum(g64 * v))
    eps = 1e-6
    fd = (B.drag_objective(capi.MRLBM_FP64, chi + eps * v) - B.drag_objective(capi.MRLBM_FP64, chi - eps * v)) / (2 * eps)
    r32, g32, _ = B.run_adjoint_case(capi.MRLBM_FP32_DEV, chi)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"[gpu] 512x128 fp64: J {r64.J:.10f} directional adjoint {ad:.10e} fd {fd:.10e}; "
          f"adjoint/forward {r64.adjoint_ms / r64.forward_ms:.2f}; fp32: J {r32.J:.8f} grad rel L2 {rel:.3e} "
          f"adjoint/forward {r32.adjoint_ms / r32.forward_ms:.2f}", flush=True)
    assert abs(ad - fd) <= 1e-4 * abs(fd)
    assert abs(r32.J - r64.J) <= 1e-4 * abs(r64.J)
    assert rel < 1e-2


def test_cylinder_next_to_the_outlet_column():
    """A wall whose interpolation stencil reaches column nx-2, where the outlet copies its
    values (the two share special rows, so the transposed interpolation runs in one
    workgroup): the directional derivative along smooth chi bumps and dJ/duo against
    central differences of the forward run."""
    nx, ny, nsteps, window, K = 48, 24, 200, 100, 48

    def par(uo=UO, omega=0.0):
        p = capi.default_params()
        p.nx, p.ny = nx, ny
        p.x_periodic, p.y_periodic = 0, 0
        p.channel_with_cylinder = 1
        p.x_c, p.y_c, p.r_c = 39.0, 12.5, 3.0
        p.label_policy = capi.LABEL_FLUID  # this small wall has labels the reference leaves unset
        p.Re, p.uo, p.omega = 20.0, uo, omega
        p.inlet_profile = capi.INLET_PARABOLIC
        p.precision = capi.MRLBM_FP64
        return p

    def objective(p, chi):
        with capi.Lattice(p) as lat:
            lat.set_brinkman(chi)
            o = capi.default_step_opts()
            o.post_process = 1
            o.statsbegin = 2 + nsteps - window
            o.statsend = 1 + nsteps
            lat.step(2, nsteps, o)
            _, fd, _ = lat.forces()
            uref = (2.0 / 3.0) * p.uo
            return float(np.sum(fd) / window / (0.5 * uref * uref * 2 * lat.info.r_cyl))

    x = np.arange(1, nx + 1)[:, None]
    y = np.arange(1, ny + 1)[None, :]
    chi = 0.01 * np.exp(-((x - 33.0) ** 2 + (y - 15.0) ** 2) / 8.0)
    with capi.Lattice(par()) as lat:
        assert lat.info.n_stencil > 0
        lat.set_brinkman(chi)
        res, g = lat.adjoint_drag(2, nsteps, window, K)
        omega = lat.info.omega
    print(f"[gpu] near-outlet adjoint: J {res.J:.10f} dJ/duo {res.dJ_duo:.6e}", flush=True)
    direction = np.exp(-((x - 36.0) ** 2 + (y - 9.0) ** 2) / 6.0)
    eps = 1e-6
    fd = (objective(par(omega=omega), chi + eps * direction)
          - objective(par(omega=omega), chi - eps * direction)) / (2 * eps)
    ad = float(np.sum(g * direction))
    assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12
    h = 1e-6
    fdu = (objective(par(uo=UO + h, omega=omega), chi) - objective(par(uo=UO - h, omega=omega), chi)) / (2 * h)
    assert abs(res.dJ_duo - fdu) <= 1e-5 * abs(fdu)
