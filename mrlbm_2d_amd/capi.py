"""ctypes binding of the C-ABI in include/mrlbm.h (libmrlbm.so, gfx950 HIP).

The library is the product; this module only marshals numpy arrays across the C-ABI.
Loading fails loudly when the built library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_lib", "libmrlbm.so")

MRLBM_FP64, MRLBM_FP32_DEV = 0, 1
MASS_STRONG, MASS_EQ, MASS_WEAK = 0, 1, 2
INLET_UNIFORM, INLET_PARABOLIC = 0, 1
LABEL_REFERENCE, LABEL_FLUID = 0, 1
E_ARG, E_HIP, E_GEOMETRY, E_NONFINITE, E_NOMEM, E_STATE = -1, -2, -3, -4, -5, -6

# Every symbol include/mrlbm.h declares.
EXPORTS = (
    "mrlbm_create", "mrlbm_destroy", "mrlbm_last_error", "mrlbm_abi_version", "mrlbm_default_params",
    "mrlbm_default_step_opts", "mrlbm_get_info", "mrlbm_get_boundary", "mrlbm_get_node_class",
    "mrlbm_set_populations", "mrlbm_set_moments", "mrlbm_step", "mrlbm_synchronize",
    "mrlbm_get_fields", "mrlbm_get_moments", "mrlbm_get_populations", "mrlbm_get_populations_noalloc",
    "mrlbm_get_forces",
    "mrlbm_get_surface_p", "mrlbm_get_surface_p_series", "mrlbm_get_rho_inf_mean",
    "mrlbm_get_probes", "mrlbm_get_stats", "mrlbm_reset_stats", "mrlbm_max_ux",
    "mrlbm_check_finite", "mrlbm_set_timing", "mrlbm_last_step_ms", "mrlbm_last_bulk_ms",
    "mrlbm_last_timing",
    "mrlbm_set_brinkman", "mrlbm_default_adjoint_opts", "mrlbm_adjoint_drag",
    "mrlbm_default_batch_opts", "mrlbm_batch_create", "mrlbm_batch_destroy", "mrlbm_batch_last_error",
    "mrlbm_batch_run", "mrlbm_batch_autotune", "mrlbm_batch_step", "mrlbm_batch_forces",
    "mrlbm_batch_reset_forces", "mrlbm_batch_get_info",
)


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32),
        ("x_periodic", C.c_int32), ("y_periodic", C.c_int32),
        ("channel_with_cylinder", C.c_int32),
        ("x_c", C.c_double), ("y_c", C.c_double), ("r_c", C.c_double),
        ("Re", C.c_double), ("uo", C.c_double), ("omega", C.c_double),
        ("inlet_profile", C.c_int32), ("mass_bc", C.c_int32), ("rotated_coordinate", C.c_int32),
        ("vel_interp", C.c_int32), ("mom_interp", C.c_int32),
        ("incomp_bc", C.c_int32), ("incomp_bulk", C.c_int32),
        ("precision", C.c_int32), ("label_policy", C.c_int32), ("strip_rows", C.c_int32),
        ("kernel", C.c_int32), ("block_order", C.c_int32), ("time_blocking", C.c_int32),
        ("steps_per_pass", C.c_int32), ("quad_rows", C.c_int32), ("reserved", C.c_int32 * 2),
    ]


class StepOpts(C.Structure):
    _fields_ = [
        ("post_process", C.c_int32), ("statsbegin", C.c_int32), ("statsend", C.c_int32),
        ("iplot", C.c_int32), ("record_surface_p", C.c_int32), ("stats_interval", C.c_int32),
        ("nprobe", C.c_int32), ("probe_i", C.c_int32 * 4), ("probe_j", C.c_int32 * 4),
        ("reserved", C.c_int32 * 6),
    ]


class Timing(C.Structure):
    _fields_ = [("step_ms", C.c_double), ("timed_ms", C.c_double), ("bulk_mean_ms", C.c_double),
                ("bulk_median_ms", C.c_double), ("bulk_min_ms", C.c_double), ("bulk_max_ms", C.c_double),
                ("bulk_launches", C.c_int32), ("bulk_steps", C.c_int32), ("timed_launches", C.c_int32),
                ("timed_steps", C.c_int32), ("reserved", C.c_int32 * 4)]


class Info(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nbct", C.c_int32), ("precision", C.c_int32),
        ("r_cyl", C.c_double), ("nu", C.c_double), ("omega", C.c_double),
        ("x_c", C.c_double), ("y_c", C.c_double), ("r_c", C.c_double),
        ("bytes_per_step", C.c_int64), ("box", C.c_int32 * 4), ("n_stencil", C.c_int32),
        ("levels", C.c_int32), ("pair_rows", C.c_int32), ("quad_rows", C.c_int32),
    ]


class AdjointOpts(C.Structure):
    _fields_ = [
        ("iter_first", C.c_int32), ("nsteps", C.c_int32), ("window", C.c_int32),
        ("checkpoint_interval", C.c_int32), ("serial_recompute", C.c_int32), ("reserved", C.c_int32 * 3),
    ]


class AdjointResult(C.Structure):
    _fields_ = [
        ("J", C.c_double), ("dJ_duo", C.c_double), ("forward_ms", C.c_double),
        ("adjoint_ms", C.c_double), ("reserved", C.c_double * 4),
    ]


class BatchOpts(C.Structure):
    _fields_ = [
        ("kernel", C.c_int32), ("persistent", C.c_int32), ("graph_steps", C.c_int32),
        ("tune_steps", C.c_int32), ("time_blocking", C.c_int32), ("steps_per_pass", C.c_int32),
        ("reserved", C.c_int32 * 2),
    ]


class BatchInfo(C.Structure):
    _fields_ = [
        ("nruns", C.c_int32), ("kernel", C.c_int32), ("persistent", C.c_int32), ("grid", C.c_int32),
        ("ntiles", C.c_int32), ("graph_steps", C.c_int32), ("nodes", C.c_int64),
        ("time_blocking", C.c_int32), ("ntune", C.c_int32),
        ("tune_kernel", C.c_int32 * 16), ("tune_persistent", C.c_int32 * 16), ("tune_tb", C.c_int32 * 16),
        ("tune_ms", C.c_double * 16), ("quad_rows", C.c_int32), ("tune_quad", C.c_int32 * 16),
    ]


class MrlbmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mrlbm error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH):
    """Loads libmrlbm.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(path)
    vp, dp, ip, i8p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_int8)
    L.mrlbm_create.argtypes = [C.POINTER(Params), C.POINTER(vp)]
    L.mrlbm_destroy.argtypes = [vp]
    L.mrlbm_last_error.argtypes = [vp]
    L.mrlbm_last_error.restype = C.c_char_p
    L.mrlbm_default_params.argtypes = [C.POINTER(Params)]
    L.mrlbm_default_params.restype = None
    L.mrlbm_default_step_opts.argtypes = [C.POINTER(StepOpts)]
    L.mrlbm_default_step_opts.restype = None
    L.mrlbm_get_info.argtypes = [vp, C.POINTER(Info)]
    L.mrlbm_get_boundary.argtypes = [vp, ip, ip, ip, dp, ip]
    L.mrlbm_get_node_class.argtypes = [vp, i8p]
    L.mrlbm_set_populations.argtypes = [vp, dp]
    L.mrlbm_set_moments.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.mrlbm_step.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(StepOpts)]
    L.mrlbm_synchronize.argtypes = [vp]
    L.mrlbm_get_fields.argtypes = [vp, dp, dp, dp]
    L.mrlbm_get_moments.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.mrlbm_get_populations.argtypes = [vp, dp]
    L.mrlbm_get_populations_noalloc.argtypes = [vp, dp]
    L.mrlbm_get_forces.argtypes = [vp, ip, dp, dp, C.c_int32, ip]
    L.mrlbm_get_surface_p.argtypes = [vp, ip, dp, dp]
    L.mrlbm_get_surface_p_series.argtypes = [vp, ip, dp, dp, C.c_int32, ip]
    L.mrlbm_get_rho_inf_mean.argtypes = [vp, dp]
    L.mrlbm_get_probes.argtypes = [vp, ip, dp, C.c_int32, ip]
    L.mrlbm_get_stats.argtypes = [vp, dp, dp, dp, dp, C.POINTER(C.c_int64)]
    L.mrlbm_reset_stats.argtypes = [vp]
    L.mrlbm_max_ux.argtypes = [vp, dp]
    L.mrlbm_check_finite.argtypes = [vp]
    L.mrlbm_set_timing.argtypes = [vp, C.c_int32]
    L.mrlbm_last_step_ms.argtypes = [vp, dp]
    L.mrlbm_last_bulk_ms.argtypes = [vp, dp, C.POINTER(C.c_int32)]
    L.mrlbm_last_timing.argtypes = [vp, C.POINTER(Timing)]
    L.mrlbm_set_brinkman.argtypes = [vp, dp]
    L.mrlbm_default_adjoint_opts.argtypes = [C.POINTER(AdjointOpts)]
    L.mrlbm_default_adjoint_opts.restype = None
    L.mrlbm_adjoint_drag.argtypes = [vp, C.POINTER(AdjointOpts), dp, C.POINTER(AdjointResult)]
    L.mrlbm_default_batch_opts.argtypes = [C.POINTER(BatchOpts)]
    L.mrlbm_default_batch_opts.restype = None
    L.mrlbm_batch_create.argtypes = [C.POINTER(Params), C.c_int32, C.POINTER(BatchOpts), C.POINTER(vp)]
    L.mrlbm_batch_destroy.argtypes = [vp]
    L.mrlbm_batch_last_error.argtypes = [vp]
    L.mrlbm_batch_last_error.restype = C.c_char_p
    L.mrlbm_batch_run.argtypes = [vp, C.c_int32]
    L.mrlbm_batch_run.restype = vp
    L.mrlbm_batch_autotune.argtypes = [vp]
    L.mrlbm_batch_step.argtypes = [vp, C.c_int32, C.c_int32]
    L.mrlbm_batch_forces.argtypes = [vp, dp, dp, dp, C.POINTER(C.c_int64)]
    L.mrlbm_batch_reset_forces.argtypes = [vp]
    L.mrlbm_batch_get_info.argtypes = [vp, C.POINTER(BatchInfo)]
    _lib = L
    return L


def default_params() -> Params:
    p = Params()
    load().mrlbm_default_params(C.byref(p))
    return p


def default_step_opts() -> StepOpts:
    o = StepOpts()
    load().mrlbm_default_step_opts(C.byref(o))
    return o


def default_adjoint_opts() -> AdjointOpts:
    o = AdjointOpts()
    load().mrlbm_default_adjoint_opts(C.byref(o))
    return o


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else C.POINTER(C.c_double)()


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else C.POINTER(C.c_int32)()


def _fort(a, shape):
    """Host array in the reference's column-major layout, flattened."""
    a = np.asarray(a, dtype=np.float64)
    assert a.size == int(np.prod(shape)), (a.shape, shape)
    return np.ascontiguousarray(a.reshape(shape, order="F").reshape(-1, order="F"))


class Lattice:
    """One run on the GPU behind the C-ABI (handle + device state)."""

    def __init__(self, params: Params, _handle=None):
        self._L = load()
        self._owned = _handle is None
        if _handle is None:
            h = C.c_void_p()
            rc = self._L.mrlbm_create(C.byref(params), C.byref(h))
            if rc != 0:
                raise MrlbmError(rc, self._L.mrlbm_last_error(None).decode())
        else:
            h = C.c_void_p(_handle)
        self._h = h
        self.params = params
        info = Info()
        self._check(self._L.mrlbm_get_info(self._h, C.byref(info)))
        self.info = info
        self.nx, self.ny, self.nbct = info.nx, info.ny, info.nbct
        self.N = self.nx * self.ny

    def _check(self, rc):
        if rc != 0:
            raise MrlbmError(rc, self._L.mrlbm_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                self._L.mrlbm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- geometry
    def boundary(self):
        b = self.nbct
        bi, bj, lab, idx = (np.zeros(b, np.int32) for _ in range(4))
        th = np.zeros(b)
        self._check(self._L.mrlbm_get_boundary(self._h, _ip(bi), _ip(bj), _ip(lab), _dp(th), _ip(idx)))
        return dict(bou_i=bi, bou_j=bj, label=lab, theta=th, idx=idx)

    def node_class(self):
        c = np.zeros(self.N, np.int8)
        self._check(self._L.mrlbm_get_node_class(self._h, c.ctypes.data_as(C.POINTER(C.c_int8))))
        return c.reshape(self.nx, self.ny, order="F")

    # -- state
    def set_populations(self, f):
        a = _fort(f, (self.nx, self.ny, 9))
        self._check(self._L.mrlbm_set_populations(self._h, _dp(a)))

    def set_moments(self, rho, ux, uy, mxx=None, myy=None, mxy=None):
        sh = (self.nx, self.ny)
        arrs = [_fort(x, sh) if x is not None else None for x in (rho, ux, uy, mxx, myy, mxy)]
        self._check(self._L.mrlbm_set_moments(self._h, *[_dp(x) for x in arrs]))

    def set_brinkman(self, chi):
        """Brinkman solid fraction (nx, ny); None removes it."""
        arr = _fort(chi, (self.nx, self.ny)) if chi is not None else None
        self._check(self._L.mrlbm_set_brinkman(self._h, _dp(arr)))

    def adjoint_drag(self, iter_first: int, nsteps: int, window: int, checkpoint_interval: int,
                     serial_recompute: bool = False):
        """Forward run with checkpoints, then the adjoint sweep.  Returns (result, dJ/dchi)
        with dJ/dchi as an (nx, ny) array."""
        o = default_adjoint_opts()
        o.iter_first, o.nsteps, o.window, o.checkpoint_interval = iter_first, nsteps, window, checkpoint_interval
        o.serial_recompute = 1 if serial_recompute else 0
        g = np.zeros(self.N)
        r = AdjointResult()
        self._check(self._L.mrlbm_adjoint_drag(self._h, C.byref(o), _dp(g), C.byref(r)))
        return r, g.reshape(self.nx, self.ny, order="F")

    def step(self, iter_first: int, nsteps: int, opts: StepOpts | None = None):
        self._check(self._L.mrlbm_step(self._h, int(iter_first), int(nsteps),
                                       C.byref(opts) if opts is not None else None))

    def synchronize(self):
        self._check(self._L.mrlbm_synchronize(self._h))

    def fields(self):
        a = [np.zeros(self.N) for _ in range(3)]
        self._check(self._L.mrlbm_get_fields(self._h, *[_dp(x) for x in a]))
        return [x.reshape(self.nx, self.ny, order="F") for x in a]

    def moments(self):
        a = [np.zeros(self.N) for _ in range(6)]
        self._check(self._L.mrlbm_get_moments(self._h, *[_dp(x) for x in a]))
        return [x.reshape(self.nx, self.ny, order="F") for x in a]

    def populations(self):
        a = np.zeros(9 * self.N)
        self._check(self._L.mrlbm_get_populations(self._h, _dp(a)))
        return a.reshape(self.nx, self.ny, 9, order="F")

    def populations_noalloc(self):
        """As populations(), through the entry point that allocates nothing on the device."""
        a = np.zeros(9 * self.N)
        self._check(self._L.mrlbm_get_populations_noalloc(self._h, _dp(a)))
        return a.reshape(self.nx, self.ny, 9, order="F")

    # -- records
    def forces(self, max_n: int = 1 << 20):
        it = np.zeros(max_n, np.int32)
        fd, fl = np.zeros(max_n), np.zeros(max_n)
        n = C.c_int32()
        self._check(self._L.mrlbm_get_forces(self._h, _ip(it), _dp(fd), _dp(fl), max_n, C.byref(n)))
        k = n.value
        return it[:k], fd[:k], fl[:k]

    def surface_p(self):
        it = C.c_int32()
        ps, pf = np.zeros(self.nbct), np.zeros(self.nbct)
        self._check(self._L.mrlbm_get_surface_p(self._h, C.byref(it), _dp(ps), _dp(pf)))
        return it.value, ps, pf

    def surface_p_series(self, max_n: int = 4096):
        it = np.zeros(max_n, np.int32)
        ps, pf = np.zeros(max_n * self.nbct), np.zeros(max_n * self.nbct)
        n = C.c_int32()
        self._check(self._L.mrlbm_get_surface_p_series(self._h, _ip(it), _dp(ps), _dp(pf), max_n, C.byref(n)))
        k = n.value
        return it[:k], ps[:k * self.nbct].reshape(k, self.nbct), pf[:k * self.nbct].reshape(k, self.nbct)

    def probes(self, max_n: int = 1 << 16):
        it = np.zeros(max_n, np.int32)
        v = np.zeros(12 * max_n)
        n = C.c_int32()
        self._check(self._L.mrlbm_get_probes(self._h, _ip(it), _dp(v), max_n, C.byref(n)))
        k = n.value
        return it[:k], v[:12 * k].reshape(k, 4, 3)

    def rho_inf_mean(self):
        v = C.c_double()
        self._check(self._L.mrlbm_get_rho_inf_mean(self._h, C.byref(v)))
        return v.value

    def stats(self):
        a = [np.zeros(self.N) for _ in range(4)]
        ns = C.c_int64()
        self._check(self._L.mrlbm_get_stats(self._h, *[_dp(x) for x in a], C.byref(ns)))
        return [x.reshape(self.nx, self.ny, order="F") for x in a] + [ns.value]

    def reset_stats(self):
        self._check(self._L.mrlbm_reset_stats(self._h))

    def max_ux(self):
        v = C.c_double()
        self._check(self._L.mrlbm_max_ux(self._h, C.byref(v)))
        return v.value

    def check_finite(self):
        return self._L.mrlbm_check_finite(self._h)

    def set_timing(self, on=True):
        """False/0 off; True/1 time every bulk launch; n > 1 every n-th launch."""
        self._check(self._L.mrlbm_set_timing(self._h, int(on)))

    def last_step_ms(self):
        v = C.c_double()
        self._check(self._L.mrlbm_last_step_ms(self._h, C.byref(v)))
        return v.value

    def last_bulk_ms(self):
        """(mean ms of one bulk launch, launches timed) over the last timed step call."""
        v, n = C.c_double(), C.c_int32()
        self._check(self._L.mrlbm_last_bulk_ms(self._h, C.byref(v), C.byref(n)))
        return v.value, n.value

    def last_timing(self) -> Timing:
        """Bulk launches of the last step call: all, timed, and the steps each set advanced."""
        t = Timing()
        self._check(self._L.mrlbm_last_timing(self._h, C.byref(t)))
        return t


def default_batch_opts() -> BatchOpts:
    o = BatchOpts()
    load().mrlbm_default_batch_opts(C.byref(o))
    return o


class Batch:
    """Independent runs advanced together (mrlbm_batch_*).  run(r) is a Lattice view of run r
    for setting and reading its state; step the runs only through the batch."""

    def __init__(self, params_list, opts: BatchOpts | None = None):
        self._L = load()
        arr = (Params * len(params_list))(*params_list)
        o = opts if opts is not None else default_batch_opts()
        b = C.c_void_p()
        rc = self._L.mrlbm_batch_create(arr, len(params_list), C.byref(o), C.byref(b))
        if rc != 0:
            raise MrlbmError(rc, self._L.mrlbm_batch_last_error(None).decode())
        self._b = b
        self.params = list(params_list)
        self.nruns = len(params_list)
        self._views = [Lattice(p, _handle=self._L.mrlbm_batch_run(b, r)) for r, p in enumerate(params_list)]

    def _check(self, rc):
        if rc != 0:
            raise MrlbmError(rc, self._L.mrlbm_batch_last_error(self._b).decode())

    def run(self, r: int) -> Lattice:
        return self._views[r]

    def step(self, nsteps: int, record_forces: bool = False):
        self._check(self._L.mrlbm_batch_step(self._b, int(nsteps), 1 if record_forces else 0))

    def autotune(self):
        self._check(self._L.mrlbm_batch_autotune(self._b))

    def forces(self):
        fd, fl, fl2 = (np.zeros(self.nruns) for _ in range(3))
        n = np.zeros(self.nruns, np.int64)
        self._check(self._L.mrlbm_batch_forces(self._b, _dp(fd), _dp(fl), _dp(fl2),
                                                n.ctypes.data_as(C.POINTER(C.c_int64))))
        return fd, fl, fl2, n

    def reset_forces(self):
        self._check(self._L.mrlbm_batch_reset_forces(self._b))

    @property
    def info(self) -> BatchInfo:
        i = BatchInfo()
        self._check(self._L.mrlbm_batch_get_info(self._b, C.byref(i)))
        return i

    def close(self):
        if getattr(self, "_b", None):
            for v in self._views:
                v.close()
            self._L.mrlbm_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
