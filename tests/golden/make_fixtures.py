"""Writes the golden fixtures under tests/golden/ by running the reference itself.

The reference program (2d_mrlbm.f90) is compiled from its own sources with flang by
oracle/Makefile into oracle/_ref/mrlbm_ref.  Each case below writes an input.dat in
the reference's namelist format (input.dat:1-45, with the closing '/' that flang's
namelist reader needs), optionally a restart/restart.dat carrying a seeded noisy
initial state (read back by the reference at 2d_mrlbm.f90:208-212, 2204-2220), runs
the reference single-threaded (&Parallel num_threads=1) and stores:

* the final restart record: f(nx,ny,0:8), ux, uy, rho (2d_mrlbm.f90:2186-2202);
* the force series data_mean/forces_time.dat (2d_mrlbm.f90:492-494);
* the last surface-pressure records (2d_mrlbm.f90:609-616);
* rho_mean_inlet.dat (2d_mrlbm.f90:362-366), theta.dat (2d_mrlbm.f90:1158-1162);
* r_cyl, nbct, nu, tau from simulation_parameters.dat (2d_mrlbm.f90:157-168);
* the size and SHA-256 digest of every file the run wrote (plot3d grid and function
  files, master.p3d, probe series, ...; files_json).

Usage: python tests/golden/make_fixtures.py   (needs /root/reference and flang)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_BIN = os.path.join(REPO, "oracle", "_ref", "mrlbm_ref")

BASE = dict(nx=320, ny=32, ncy=8, L_front=2, L_back=4, L_top=2, L_bot=2, Re=20.0, uo=0.05,
            iplot=100000, statsbegin=2, statsend=200001, post_process=True, x_periodic=False,
            y_periodic=False, channel_with_cylinder=True, channel_with_square=False,
            incomp_bc=False, incomp_bulk=False, vel_interp=True, mom_interp=True,
            rotated_coordinate=True, strong_mass_bc=True, eq_mass_bc=False, weak_mass_bc=False)

# name -> (overrides, nsteps, noise seed or None)
CASES = {
    "p1_walls": (dict(L_back=8, iplot=250), 1000, None),
    "s_periodic": (dict(y_periodic=True), 1000, None),
    "s_eq": (dict(strong_mass_bc=False, eq_mass_bc=True), 500, None),
    "s_weak": (dict(strong_mass_bc=False, weak_mass_bc=True), 500, None),
    "s_unrot": (dict(rotated_coordinate=False), 500, None),
    "s_incomp_bc": (dict(incomp_bc=True), 500, None),
    "s_incomp_bulk": (dict(incomp_bulk=True), 500, None),
    "s_novel": (dict(vel_interp=False), 500, None),
    "s_xyper_noise": (dict(x_periodic=True, y_periodic=True), 500, 11),
    "s_noise_re60": (dict(Re=60.0, iplot=400), 1000, 12),
    "chan": (dict(channel_with_cylinder=False, nx=48, ny=24, post_process=False), 500, None),
}


def fbool(v):
    return ".true." if v else ".false."


def write_input(path, p, max_iter, irestart):
    txt = f"""&DomainSize
\tnx = {p['nx']}
\tny = {p['ny']}
\tlattice_type = 9
/
&Cylinder
\tncy = {p['ncy']}
\tL_front = {p['L_front']}
\tL_back = {p['L_back']}
\tL_top = {p['L_top']}
\tL_bot = {p['L_bot']}
/
&Numbers
\tRe = {p['Re']!r}
/
&Parallel
\tnum_threads=1
/
&Controls
\tuo = {p['uo']!r}
\trho_infty = 1.0
\tiplot = {p['iplot']}
\tmax_iter = {max_iter}
\tisave = {max_iter}
\tirestart = {irestart}
\tstatsbegin = {p['statsbegin']}
\tstatsend = {p['statsend']}
\tiplotstats = 1000
/
&LogicalControls
"""
    for k in ("post_process", "x_periodic", "y_periodic", "channel_with_cylinder",
              "channel_with_square", "incomp_bc", "incomp_bulk", "vel_interp", "mom_interp",
              "rotated_coordinate", "strong_mass_bc", "eq_mass_bc", "weak_mass_bc"):
        txt += f"\t{k} = {fbool(p[k])}\n"
    txt += "/\n"
    with open(path, "w") as fh:
        fh.write(txt)


def grid_size(p):
    if p["channel_with_cylinder"]:
        ncy = p["ncy"]
        return ncy * (p["L_front"] + 1 + p["L_back"]), ncy * (p["L_top"] + 1 + p["L_bot"])
    return p["nx"], p["ny"]


def noisy_state(nx, ny, seed):
    """Equilibrium populations at rho = 1 with u = uniform noise in [-1e-4, 1e-4]."""
    w = np.array([16, 4, 4, 4, 4, 1, 1, 1, 1], np.float32) / np.float32(36)
    w = w.astype(np.float64)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1], np.float64)
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1], np.float64)
    rng = np.random.default_rng(seed)
    ux = rng.uniform(-1e-4, 1e-4, (nx, ny))
    uy = rng.uniform(-1e-4, 1e-4, (nx, ny))
    cu = ex[None, None, :] * ux[..., None] + ey[None, None, :] * uy[..., None]
    usq = (ux * ux + uy * uy)[..., None]
    f = w[None, None, :] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
    return f, ux, uy


def write_restart(path, it, f, ux, uy, rho):
    body = (np.array([it], np.int32).tobytes() + np.asfortranarray(f).tobytes(order="F")
            + np.zeros(f.size).tobytes() + np.asfortranarray(ux).tobytes(order="F")
            + np.asfortranarray(uy).tobytes(order="F") + np.asfortranarray(rho).tobytes(order="F"))
    marker = np.array([len(body)], np.int32).tobytes()
    with open(path, "wb") as fh:
        fh.write(marker + body + marker)


def read_restart(path, nx, ny):
    b = open(path, "rb").read()
    n = nx * ny
    it = int(np.frombuffer(b[4:8], np.int32)[0])
    d = np.frombuffer(b[8:-4], np.float64)
    shp = dict(order="F")
    return (it, d[:9 * n].reshape(nx, ny, 9, **shp), d[18 * n:19 * n].reshape(nx, ny, **shp),
            d[19 * n:20 * n].reshape(nx, ny, **shp), d[20 * n:21 * n].reshape(nx, ny, **shp))


def read_unformatted(path):
    b = open(path, "rb").read()
    out, pos = [], 0
    while pos < len(b):
        m = int(np.frombuffer(b[pos:pos + 4], np.int32)[0])
        rec = b[pos + 4:pos + 4 + m]
        pos += 8 + m
        out.append((int(np.frombuffer(rec[:4], np.int32)[0]), np.frombuffer(rec[4:], np.float64)))
    return out


def run_case(name, overrides, nsteps, seed):
    p = dict(BASE)
    p.update(overrides)
    nx, ny = grid_size(p)
    max_iter = 1 + nsteps
    with tempfile.TemporaryDirectory() as d:
        for sub in ("output", "restart", "data_probe", "data_geo", "data_mean"):
            os.makedirs(os.path.join(d, sub))
        f0 = None
        if seed is not None:
            f0, ux0, uy0 = noisy_state(nx, ny, seed)
            write_restart(os.path.join(d, "restart", "restart.dat"), 1, f0, ux0, uy0, np.ones((nx, ny)))
        write_input(os.path.join(d, "input.dat"), p, max_iter, 1 if seed is not None else 0)
        r = subprocess.run([REF_BIN], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout}\n{r.stderr}")
        it, f, ux, uy, rho = read_restart(os.path.join(d, "restart", "restart.dat"), nx, ny)
        assert it == max_iter, (it, max_iter)
        out = dict(f=f, ux=ux, uy=uy, rho=rho, iter_last=it, iter_first=2, nsteps=nsteps,
                   nx=nx, ny=ny)
        if f0 is not None:
            out["f0"] = f0
        params = {}
        for line in open(os.path.join(d, "simulation_parameters.dat")):
            key, _, val = line.strip().rpartition(" ")
            params[key.strip()] = val
        if p["channel_with_cylinder"]:
            out["r_cyl"] = float(params["Maximum_radius"])
            out["nbct"] = int(params["No. of points on cylinder:"])
            out["theta"] = np.loadtxt(os.path.join(d, "data_mean", "theta.dat"))
            out["bound_xy"] = np.loadtxt(os.path.join(d, "data_geo", "bound.dat"))
        out["tau"] = float(params["tau:"])
        if p["post_process"]:
            fr = np.atleast_2d(np.loadtxt(os.path.join(d, "data_mean", "forces_time.dat")))
            out["forces"] = fr
            ps = read_unformatted(os.path.join(d, "data_mean", "p_surface_time.dat"))
            pf = read_unformatted(os.path.join(d, "data_mean", "p_surface_smooth_time.dat"))
            out["ps_iter"] = np.array([ps[-1][0]])
            out["ps_last"] = ps[-1][1]
            out["pfit_last"] = pf[-1][1]
            rm = os.path.join(d, "rho_mean_inlet.dat")
            if os.path.getsize(rm):
                out["rho_mean_inlet"] = np.atleast_2d(np.loadtxt(rm))
        out["files_json"] = np.array(json.dumps(file_digests(d)))
        out["simulation_parameters"] = np.array(open(os.path.join(d, "simulation_parameters.dat")).read())
        out["params_json"] = np.array(json.dumps(p))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(f"{name}: {nx}x{ny} steps={nsteps} nbct={out.get('nbct')} ok", flush=True)


def file_digests(d):
    """{relative path: [size, SHA-256]} of every file a run wrote under its directory (the
    run's input and its restart state included)."""
    import hashlib
    out = {}
    for root, _, names in os.walk(d):
        for n in names:
            path = os.path.join(root, n)
            with open(path, "rb") as fh:
                data = fh.read()
            out[os.path.relpath(path, d)] = [len(data), hashlib.sha256(data).hexdigest()]
    return out


STATS_BIN = os.path.join(REPO, "oracle", "_ref", "statistics_ref")


def write_unformatted(path, iters, rows):
    with open(path, "wb") as fh:
        for it, row in zip(iters, rows):
            body = np.array([it], np.int32).tobytes() + np.asarray(row, np.float64).tobytes()
            m = np.array([len(body)], np.int32).tobytes()
            fh.write(m + body + m)


def run_statistics_case(name="stats_synth", nstep=1500, nbct=24, period=137, seed=5):
    """The reference post-processor (2d_statistics.f90) on a synthetic shedding signal:
    lift = A sin(2 pi t/period + phase) with a slow drift, drag = C0 + B cos(4 pi t/period),
    surface pressure = smooth function of theta plus a periodic ripple."""
    rng = np.random.default_rng(seed)
    t = np.arange(nstep, dtype=np.float64)
    ph = rng.uniform(0, 2 * np.pi)
    fl = 0.02 * np.sin(2 * np.pi * t / period + ph) + 1e-6 * t / nstep
    fd = 0.11 + 0.004 * np.cos(4 * np.pi * t / period + ph)
    theta = np.sort(rng.uniform(0, 2 * np.pi, nbct))
    ps = (1.0 / 3.0 + 0.01 * np.cos(theta)[None, :]
          + 0.001 * np.sin(2 * np.pi * t / period)[:, None] * np.sin(theta)[None, :])
    pfit = ps + 1e-4 * np.cos(3 * theta)[None, :]
    uo, rho_infty, r_cy = 0.05, 1.0123, 20.39610
    iters = np.arange(2, 2 + nstep, dtype=np.int32)
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "data_mean"))
        os.makedirs(os.path.join(d, "plots"))
        with open(os.path.join(d, "input_stat.dat"), "w") as fh:
            fh.write(f"&parameters\n\tnbct = {nbct}\n\tr_cy = {r_cy!r}\n\tuo = {uo!r}\n"
                     f"\trho_infty = {rho_infty!r}\n/\n")
        with open(os.path.join(d, "data_mean", "forces_time.dat"), "w") as fh:
            for it, a, b in zip(iters, fd, fl):
                fh.write(f" {int(it)} {float(a)!r} {float(b)!r}\n")
        with open(os.path.join(d, "data_mean", "theta.dat"), "w") as fh:
            for x in theta:
                fh.write(f" {float(x)!r}\n")
        write_unformatted(os.path.join(d, "data_mean", "p_surface_time.dat"), iters, ps)
        write_unformatted(os.path.join(d, "data_mean", "p_surface_smooth_time.dat"), iters, pfit)
        r = subprocess.run([STATS_BIN], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)
        vals = {}
        for line in open(os.path.join(d, "plots", "drag_lift.dat")):
            key, _, v = line.rpartition(":")
            vals[key.strip()] = float(v)
        cp = np.loadtxt(os.path.join(d, "plots", "cp.dat"))
        cpf = np.loadtxt(os.path.join(d, "plots", "cp_fit.dat"))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), fd=fd, fl=fl, theta=theta, ps=ps, pfit=pfit,
                            uo=uo, rho_infty=rho_infty, r_cy=r_cy, iters=iters,
                            st=vals["Strouhal number (St)"], cd=vals["Drag coefficinet (C_D)"],
                            cl_rms=vals["Lift coefficinet (C_L"], cp=cp, cp_fit=cpf,
                            params_json=np.array(json.dumps({"kind": "statistics"})))
        print(f"{name}: St={vals['Strouhal number (St)']} Cd={vals['Drag coefficinet (C_D)']} ok", flush=True)


def _sha256(*arrays):
    import hashlib
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.asfortranarray(a, np.float64).tobytes(order="F"))
    return h.hexdigest()


# Vortex shedding that the reference can express: ncy = 20 with L = 3/16/2/2 (400 x 100,
# centre (71, 50), r_cyl = 10.44), channel walls, Re = 100, uniform inlet uo = 0.05, from
# a seeded noisy restart state.  Shedding sets in after ~20k steps and has saturated by
# ~30k (lift amplitude 2.41e-2, period ~1.7k steps).
SHED = dict(ncy=20, L_front=3, L_back=16, L_top=2, L_bot=2, Re=100.0, uo=0.05, iplot=10 ** 7)
SHED_STEPS, SHED_WINDOW, SHED_SEED, SHED_STRIDE = 50000, 20000, 21, 5


def run_shedding_case(name="shed_re100"):
    """The reference over SHED_STEPS steps with the statistics window on the last
    SHED_WINDOW steps, then the reference post-processor (2d_statistics.f90) on the files
    that run wrote, with input_stat.dat filled from the run (nbct, r_cyl, uo and the
    measured inlet-density mean, as a user of the reference does).  Stored: the post-
    processor's St, Cd, Cl_rms, Cp profiles; every SHED_STRIDE-th force record; SHA-256
    digests of the final rho, ux, uy and f (the whole fields, bit for bit) and a few
    slices of them; the inlet-density means."""
    p = dict(BASE)
    p.update(SHED)
    statsbegin = 2 + SHED_STEPS - SHED_WINDOW
    p.update(statsbegin=statsbegin, statsend=10 ** 7)
    nx, ny = grid_size(p)
    max_iter = 1 + SHED_STEPS
    with tempfile.TemporaryDirectory() as d:
        for sub in ("output", "restart", "data_probe", "data_geo", "data_mean", "plots"):
            os.makedirs(os.path.join(d, sub))
        f0, ux0, uy0 = noisy_state(nx, ny, SHED_SEED)
        write_restart(os.path.join(d, "restart", "restart.dat"), 1, f0, ux0, uy0, np.ones((nx, ny)))
        write_input(os.path.join(d, "input.dat"), p, max_iter, 1)
        r = subprocess.run([REF_BIN], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference failed\n{r.stdout}\n{r.stderr}")
        it, f, ux, uy, rho = read_restart(os.path.join(d, "restart", "restart.dat"), nx, ny)
        assert it == max_iter, (it, max_iter)
        params = {}
        for line in open(os.path.join(d, "simulation_parameters.dat")):
            key, _, val = line.strip().rpartition(" ")
            params[key.strip()] = val
        r_cyl = float(params["Maximum_radius"])
        nbct = int(params["No. of points on cylinder:"])
        fr = np.atleast_2d(np.loadtxt(os.path.join(d, "data_mean", "forces_time.dat")))
        assert fr.shape[0] == SHED_WINDOW and int(fr[0, 0]) == statsbegin, fr[:2]
        rm = np.atleast_2d(np.loadtxt(os.path.join(d, "rho_mean_inlet.dat")))
        rho_infty = float(rm[-1, 1])
        ps = read_unformatted(os.path.join(d, "data_mean", "p_surface_time.dat"))
        pf = read_unformatted(os.path.join(d, "data_mean", "p_surface_smooth_time.dat"))
        with open(os.path.join(d, "input_stat.dat"), "w") as fh:
            fh.write(f"&parameters\n\tnbct = {nbct}\n\tr_cy = {r_cyl!r}\n\tuo = {p['uo']!r}\n"
                     f"\trho_infty = {rho_infty!r}\n/\n")
        r = subprocess.run([STATS_BIN], cwd=d, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(r.stdout + r.stderr)
        vals = {}
        for line in open(os.path.join(d, "plots", "drag_lift.dat")):
            key, _, v = line.rpartition(":")
            vals[key.strip()] = float(v)
        ncycle = [l for l in r.stdout.splitlines() if "cycle" in l]
        sel = np.arange(0, SHED_WINDOW, SHED_STRIDE)
        out = dict(
            nx=nx, ny=ny, nsteps=SHED_STEPS, iter_first=2, seed=SHED_SEED, statsbegin=statsbegin,
            f0_sha256=np.array(_sha256(f0)),
            fields_sha256=np.array(_sha256(rho, ux, uy)), f_sha256=np.array(_sha256(f)),
            rho_row=rho[:, ny // 2].copy(), ux_row=ux[:, ny // 2].copy(), uy_col=uy[nx // 4, :].copy(),
            r_cyl=r_cyl, nbct=nbct, rho_infty=rho_infty, rho_mean_inlet=rm,
            force_stride=SHED_STRIDE, forces=fr[sel], theta=np.loadtxt(os.path.join(d, "data_mean", "theta.dat")),
            ps_last=ps[-1][1], pfit_last=pf[-1][1], ps_iter=np.array([ps[-1][0]]),
            st=vals["Strouhal number (St)"], cd=vals["Drag coefficinet (C_D)"], cl_rms=vals["Lift coefficinet (C_L"],
            cp=np.loadtxt(os.path.join(d, "plots", "cp.dat")), cp_fit=np.loadtxt(os.path.join(d, "plots", "cp_fit.dat")),
            stats_stdout=np.array("\n".join(ncycle)),
            params_json=np.array(json.dumps(dict(p, kind="shedding"))))
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), **out)
        print(f"{name}: {nx}x{ny} steps={SHED_STEPS} nbct={nbct} St={out['st']} Cd={out['cd']} "
              f"Cl_rms={out['cl_rms']} ({ncycle}) ok", flush=True)


def main(names=None):
    if not os.path.exists(REF_BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle")], check=True)
    for name, (ov, n, seed) in CASES.items():
        if names and name not in names:
            continue
        run_case(name, ov, n, seed)
    if not names or "stats_synth" in names:
        run_statistics_case()
    if not names or "shed_re100" in names:
        run_shedding_case()


if __name__ == "__main__":
    main(sys.argv[1:])
