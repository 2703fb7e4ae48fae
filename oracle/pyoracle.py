"""ctypes binding of the CPU oracle (oracle/mrlbm_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP path; only tests/,
``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg import this module.
It restates the reference time loop of sakthi-lbm/mrlbm_2d (2d_mrlbm.f90:229-535)
with the reference's own data layout f(nx,ny,0:8), column-major.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, fields

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmrlbm_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build(force: bool = False) -> str:
    """Compile the oracle (and the reference into _ref/ when its sources exist)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


@dataclass
class OracleParams:
    """The reference's namelist values (2d_mrlbm.f90:2227-2233, input.dat)."""
    nx: int = 320
    ny: int = 32
    ncy: int = 8
    L_front: int = 2
    L_back: int = 4
    L_top: int = 2
    L_bot: int = 2
    Re: float = 20.0
    uo: float = 0.05
    iplot: int = 100000
    statsbegin: int = 2
    statsend: int = 200001
    post_process: int = 1
    x_periodic: int = 0
    y_periodic: int = 0
    channel_with_cylinder: int = 1
    channel_with_square: int = 0
    incomp_bc: int = 0
    incomp_bulk: int = 0
    vel_interp: int = 1
    mom_interp: int = 1
    rotated_coordinate: int = 1
    strong_mass_bc: int = 1
    eq_mass_bc: int = 0
    weak_mass_bc: int = 0
    # extensions beyond the reference (all 0 for the reference's own cases): grid and
    # cylinder centre/radius given directly, parabolic inlet with Re on 2/3 uo, and the
    # nodes of labels 0/6/9/15 as fluid (oracle/mrlbm_oracle.h)
    geometry_explicit: int = 0
    x_c: float = 0.0
    y_c: float = 0.0
    r_c: float = 0.0
    inlet_parabolic: int = 0
    label_fluid: int = 0


_FLOAT_FIELDS = {"Re", "uo", "x_c", "y_c", "r_c"}


class _CParams(C.Structure):
    _fields_ = [(f.name, C.c_double if f.name in _FLOAT_FIELDS else C.c_int)
                for f in fields(OracleParams)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(_CParams)]
        L.orc_destroy.argtypes = [vp]
        for name in ("orc_nx", "orc_ny", "orc_nbct"):
            getattr(L, name).restype = C.c_int
            getattr(L, name).argtypes = [vp]
        for name in ("orc_r_cyl", "orc_omega", "orc_nu", "orc_rho_inf_mean"):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = [vp]
        L.orc_weights.argtypes = [vp, dp]
        L.orc_get_geometry.argtypes = [vp, ip, ip, ip]
        L.orc_get_boundary.argtypes = [vp, ip, ip, ip, dp, ip, ip, ip, ip, ip, ip, dp, ip]
        L.orc_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_f.argtypes = [vp, dp]
        L.orc_get_fout.argtypes = [vp, dp]
        L.orc_get_moments.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.orc_step.argtypes = [vp, C.c_int]
        L.orc_set_brinkman.argtypes = [vp, dp]
        L.orc_run.argtypes = [vp, C.c_int, C.c_int]
        L.orc_get_forces.argtypes = [vp, dp, dp]
        L.orc_get_surface_p.argtypes = [vp, dp, dp]
        L.orc_statistics.restype = C.c_int
        L.orc_statistics.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_double,
                                     ip, ip, ip, dp, dp, dp]
        L.orc_gaussian_smoothing.argtypes = [C.c_int, dp, dp, dp]
        L.orc_quadratic_interpolation.restype = C.c_double
        L.orc_quadratic_interpolation.argtypes = [C.c_double] * 11
        L.orc_solve_system.argtypes = [dp, dp, dp]
        L.orc_num_threads.restype = C.c_int
        L.orc_num_threads.argtypes = []
        _lib = L
    return _lib


def num_threads() -> int:
    """OpenMP threads of the oracle's per-node sweeps (OMP_NUM_THREADS, else all cores)."""
    return int(lib().orc_num_threads())


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


class Oracle:
    """One reference run held on the CPU.  Arrays use the reference's column-major
    layout: a field is (nx, ny) Fortran order, f is (nx, ny, 9) Fortran order."""

    def __init__(self, params: OracleParams):
        self.params = params
        cp = _CParams(**{f.name: getattr(params, f.name) for f in fields(OracleParams)})
        self._h = lib().orc_create(C.byref(cp))
        L = lib()
        self.nx, self.ny, self.nbct = L.orc_nx(self._h), L.orc_ny(self._h), L.orc_nbct(self._h)
        self.N = self.nx * self.ny
        self.r_cyl = L.orc_r_cyl(self._h)
        self.omega = L.orc_omega(self._h)
        self.nu = L.orc_nu(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    def weights(self):
        w = np.zeros(9)
        lib().orc_weights(self._h, _dp(w))
        return w

    def geometry(self):
        a = [np.zeros(self.N, np.int32) for _ in range(3)]
        lib().orc_get_geometry(self._h, *[_ip(x) for x in a])
        return [x.reshape(self.nx, self.ny, order="F") for x in a]

    def boundary(self):
        b = self.nbct
        out = dict(bou_i=np.zeros(b, np.int32), bou_j=np.zeros(b, np.int32),
                   label=np.zeros(b, np.int32), delta=np.zeros(b),
                   ip2=np.zeros(4 * b, np.int32), jp2=np.zeros(4 * b, np.int32),
                   ip3=np.zeros(4 * b, np.int32), jp3=np.zeros(4 * b, np.int32),
                   ip4=np.zeros(4 * b, np.int32), jp4=np.zeros(4 * b, np.int32),
                   theta=np.zeros(b), idx=np.zeros(b, np.int32))
        L = lib()
        L.orc_get_boundary(self._h, _ip(out["bou_i"]), _ip(out["bou_j"]), _ip(out["label"]),
                           _dp(out["delta"]), _ip(out["ip2"]), _ip(out["jp2"]), _ip(out["ip3"]),
                           _ip(out["jp3"]), _ip(out["ip4"]), _ip(out["jp4"]), _dp(out["theta"]),
                           _ip(out["idx"]))
        return out

    def set_state(self, f, ux=None, uy=None, rho=None):
        f = np.ascontiguousarray(np.asarray(f, np.float64).reshape(-1, order="F"))
        assert f.size == 9 * self.N
        def flat(a):
            return None if a is None else np.ascontiguousarray(np.asarray(a, np.float64).reshape(-1, order="F"))
        ux, uy, rho = flat(ux), flat(uy), flat(rho)
        nul = C.POINTER(C.c_double)()
        lib().orc_set_state(self._h, _dp(f), _dp(ux) if ux is not None else nul,
                            _dp(uy) if uy is not None else nul, _dp(rho) if rho is not None else nul)

    def f(self):
        a = np.zeros(9 * self.N)
        lib().orc_get_f(self._h, _dp(a))
        return a.reshape(self.nx, self.ny, 9, order="F")

    def fout(self):
        a = np.zeros(9 * self.N)
        lib().orc_get_fout(self._h, _dp(a))
        return a.reshape(self.nx, self.ny, 9, order="F")

    def moments(self):
        a = [np.zeros(self.N) for _ in range(6)]
        lib().orc_get_moments(self._h, *[_dp(x) for x in a])
        return [x.reshape(self.nx, self.ny, order="F") for x in a]

    def set_brinkman(self, chi=None):
        """Brinkman solid fraction chi (nx, ny) column-major, or None to remove it."""
        if chi is None:
            lib().orc_set_brinkman(self._h, C.POINTER(C.c_double)())
            return
        a = np.ascontiguousarray(np.asarray(chi, np.float64).reshape(-1, order="F"))
        assert a.size == self.N
        lib().orc_set_brinkman(self._h, _dp(a))

    def step(self, it: int):
        lib().orc_step(self._h, it)

    def run(self, iter_first: int, nsteps: int):
        lib().orc_run(self._h, iter_first, nsteps)

    def forces(self):
        fd, fl = C.c_double(), C.c_double()
        lib().orc_get_forces(self._h, C.byref(fd), C.byref(fl))
        return fd.value, fl.value

    def surface_p(self):
        ps, pf = np.zeros(self.nbct), np.zeros(self.nbct)
        lib().orc_get_surface_p(self._h, _dp(ps), _dp(pf))
        return ps, pf

    def rho_inf_mean(self):
        return lib().orc_rho_inf_mean(self._h)


def statistics(fd, fl, uo, rho_infty, r_cy):
    """2d_statistics.f90 restated; returns a dict or None when two peaks are missing."""
    fd = np.ascontiguousarray(fd, np.float64)
    fl = np.ascontiguousarray(fl, np.float64)
    cs, per, nc = C.c_int(), C.c_int(), C.c_int()
    st, cd, cl = C.c_double(), C.c_double(), C.c_double()
    rc = lib().orc_statistics(len(fd), _dp(fd), _dp(fl), uo, rho_infty, r_cy, C.byref(cs),
                              C.byref(per), C.byref(nc), C.byref(st), C.byref(cd), C.byref(cl))
    if rc != 0:
        return None
    return dict(cycle_start=cs.value, period=per.value, n_cycle=nc.value, st=st.value,
                cd=cd.value, cl_rms=cl.value)


def gaussian_smoothing(theta, p):
    theta = np.ascontiguousarray(theta, np.float64)
    p = np.ascontiguousarray(p, np.float64)
    out = np.zeros_like(p)
    lib().orc_gaussian_smoothing(len(p), _dp(theta), _dp(p), _dp(out))
    return out


def quadratic_interpolation(*args):
    return lib().orc_quadratic_interpolation(*[float(a) for a in args])


def solve_system(A, b):
    A = np.ascontiguousarray(A, np.float64).reshape(9)
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros(3)
    lib().orc_solve_system(_dp(A), _dp(b), _dp(x))
    return x
