"""ndspmhd_amd: the ndspmhd hot path (neighbour finding, h-rho iteration,
MHD/dust force sums, leapfrog) as HIP kernels for MI355X (gfx950).

The compute lives in libndsp_hip.so behind the C-ABI of include/ndsp.h; this
module is a thin ctypes host that mirrors the reference's call interface:
`Simulation.derivs()` (src/derivs.f90:27), `.step()`
(src/stepND_leapfrog_mhd.f90:31), and the pieces `.set_linklist()`
(src/linkND.f90:45), `.iterate_density()` (src/iterate_density.f90:43),
`.get_rates()` (src/ratesND_mhd.f90:29).  There is no CPU fallback: if the
library is missing or no GPU is present, constructing a Simulation raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .options import Options, STATE_FIELDS, empty_state, check_state  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libndsp_hip.so")
_lib = None

PREC_MIXED = 0
PREC_FP64 = 1

FIELDS3 = ("vel", "Bevol", "alpha", "Bfield", "force", "dBevoldt", "daldt", "curlB",
           "gradpsi", "deltav", "graddivv")
CONSERVED = ("ekin", "etherm", "emag", "etot", "momx", "momy", "momz", "momtot",
             "angx", "angy", "angz", "angtot", "rhomin", "rhomax", "rhomean",
             "totmassgas", "totmassdust")
# symbols include/ndsp.h declares
SYMBOLS = ("ndsp_create", "ndsp_destroy", "ndsp_upload", "ndsp_download",
           "ndsp_set_linklist", "ndsp_iterate_density", "ndsp_get_rates",
           "ndsp_derivs", "ndsp_step", "ndsp_set_dt", "ndsp_conserved", "ndsp_ev_columns",
           "ndsp_neighbours", "ndsp_set_timing", "ndsp_kernel_times", "ndsp_set_graph",
           "ndsp_mlp_config", "ndsp_mlp_set_params", "ndsp_mlp_get_params",
           "ndsp_mlp_forward", "ndsp_mlp_infer", "ndsp_mlp_forward_noalloc", "ndsp_mlp_grad", "ndsp_mlp_train",
           "ndsp_batch_create", "ndsp_batch_destroy", "ndsp_batch_step", "ndsp_batch_autotune", "ndsp_autotune",
           "ndsp_set_reorder", "ndsp_set_ireal", "ndsp_set_tile_shapes",
           "ndsp_copy_probe",
           "ndsp_last_error")
MLP_NPARAM = 1313


class NdspError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("ndsp error %d: %s" % (code, msg))
        self.code = code


class _Stats(ctypes.Structure):
    _fields_ = [("itsdensity", ctypes.c_int32), ("nmin", ctypes.c_int32),
                ("nmax", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("ncalc_mean", ctypes.c_double), ("pairs_density", ctypes.c_int64),
                ("pairs_force", ctypes.c_int64)]


class _Dt(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in ("dt", "dtcourant", "dtforce", "dtdrag",
                                                "vsig2max", "hhmax", "stressmax", "time")]


def load_library(path: str | None = None):
    """Load libndsp_hip.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIBPATH
    if not os.path.exists(path):
        raise OSError("libndsp_hip.so is not built (run python -m ndspmhd_amd.build)")
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int32)
    L.ndsp_create.argtypes = [ip, dp, ctypes.c_int64, ctypes.c_int32, ctypes.POINTER(vp)]
    L.ndsp_destroy.argtypes = [vp]
    L.ndsp_upload.argtypes = [vp] + [dp] * 9 + [ip]
    L.ndsp_download.argtypes = [vp, ctypes.c_char_p, dp, ip]
    L.ndsp_set_linklist.argtypes = [vp]
    L.ndsp_iterate_density.argtypes = [vp, ctypes.POINTER(_Stats)]
    L.ndsp_get_rates.argtypes = [vp, ctypes.POINTER(_Dt)]
    L.ndsp_derivs.argtypes = [vp, ctypes.POINTER(_Stats), ctypes.POINTER(_Dt)]
    L.ndsp_step.argtypes = [vp, ctypes.POINTER(_Stats), ctypes.POINTER(_Dt)]
    L.ndsp_set_dt.argtypes = [vp, ctypes.c_double, ctypes.c_double]
    L.ndsp_conserved.argtypes = [vp, dp]
    L.ndsp_ev_columns.argtypes = [vp, ctypes.c_double, ctypes.c_double, dp, ip]
    L.ndsp_neighbours.argtypes = [vp, ctypes.c_int32, ip, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64)]
    L.ndsp_set_timing.argtypes = [vp, ctypes.c_int32]
    L.ndsp_set_graph.argtypes = [vp, ctypes.c_int32]
    L.ndsp_set_reorder.argtypes = [vp, ctypes.c_int32]
    L.ndsp_set_ireal.argtypes = [vp, ip]
    L.ndsp_set_tile_shapes.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]
    L.ndsp_autotune.argtypes = [vp, ctypes.c_int32, ip]
    L.ndsp_copy_probe.argtypes = [vp, ctypes.c_int64]
    L.ndsp_batch_create.argtypes = [ctypes.POINTER(vp), ctypes.c_int32, ctypes.POINTER(vp)]
    L.ndsp_batch_destroy.argtypes = [vp]
    L.ndsp_batch_step.argtypes = [vp, ctypes.POINTER(_Stats), ctypes.POINTER(_Dt)]
    L.ndsp_batch_autotune.argtypes = [vp, ctypes.c_int32, ip]
    L.ndsp_kernel_times.argtypes = [vp, dp, ctypes.c_int32, ip, ctypes.c_char_p, ctypes.c_int32]
    fp = ctypes.POINTER(ctypes.c_float)
    L.ndsp_mlp_config.argtypes = [vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_int32,
                                  ctypes.c_int32, ctypes.c_int32]
    L.ndsp_mlp_set_params.argtypes = [vp, fp, ctypes.c_int32]
    L.ndsp_mlp_get_params.argtypes = [vp, fp, ctypes.c_int32, ip]
    L.ndsp_mlp_forward.argtypes = [vp, fp, ctypes.c_int32, fp]
    L.ndsp_mlp_forward_noalloc.argtypes = [vp, fp, ctypes.c_int32, fp]
    L.ndsp_mlp_infer.argtypes = [vp, fp, ctypes.c_int32, fp]
    L.ndsp_mlp_grad.argtypes = [vp, fp, fp, ctypes.c_int32, fp, dp]
    L.ndsp_mlp_train.argtypes = [vp, fp, fp, ctypes.c_int32, ctypes.c_int32, ctypes.c_double, dp]
    L.ndsp_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class Simulation:
    """One simulation resident in HBM (an ndsp_ctx)."""

    def __init__(self, opts: Options, state: dict, precision: int = PREC_MIXED):
        L = load_library()
        self.opts = opts
        self.n = check_state(state, opts.ndim)
        self._iopt = np.ascontiguousarray(opts.iopt(), dtype=np.int32)
        self._dopt = np.ascontiguousarray(opts.dopt(), dtype=np.float64)
        h = ctypes.c_void_p()
        self._check(L.ndsp_create(_ip(self._iopt), _dp(self._dopt), self.n, precision, ctypes.byref(h)))
        self.h = h
        self.upload(state)
        self.stats = None
        self.dt_info = None

    def _check(self, rc):
        if rc != 0:
            raise NdspError(rc, load_library().ndsp_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            load_library().ndsp_destroy(self.h)
            self.h = None

    __del__ = close

    def upload(self, st: dict):
        a = {k: np.ascontiguousarray(st[k], dtype=np.float64)
             for k in ("x", "vel", "pmass", "hh", "en", "Bevol", "alpha", "psi", "dustevol")}
        it = np.ascontiguousarray(st["itype"], dtype=np.int32)
        self.itype = it.copy()
        self._check(load_library().ndsp_upload(self.h, *[_dp(a[k]) for k in (
            "x", "vel", "pmass", "hh", "en", "Bevol", "alpha", "psi", "dustevol")], _ip(it)))
        if "ireal" in st:
            ir = np.ascontiguousarray(st["ireal"], dtype=np.int32)
            self._check(load_library().ndsp_set_ireal(self.h, _ip(ir)))

    def get(self, name: str) -> np.ndarray:
        w = 3 if name in FIELDS3 else (self.opts.ndim if name == "x" else 1)
        out = np.empty(self.n * w)
        wid = ctypes.c_int32()
        self._check(load_library().ndsp_download(self.h, name.encode(), _dp(out), ctypes.byref(wid)))
        return out.reshape(self.n, w) if w > 1 else out

    def _stats(self, s, d):
        if s is not None:
            self.stats = {k: getattr(s, k) for k, _ in _Stats._fields_ if k != "pad"}
        if d is not None:
            self.dt_info = {k: getattr(d, k) for k, _ in _Dt._fields_}

    # --- the reference's hot-path entry points ---
    def set_linklist(self):
        self._check(load_library().ndsp_set_linklist(self.h))

    def iterate_density(self):
        s = _Stats()
        self._check(load_library().ndsp_iterate_density(self.h, ctypes.byref(s)))
        self._stats(s, None)
        return self.stats

    def get_rates(self):
        d = _Dt()
        self._check(load_library().ndsp_get_rates(self.h, ctypes.byref(d)))
        self._stats(None, d)
        return self.dt_info

    def derivs(self):
        s, d = _Stats(), _Dt()
        self._check(load_library().ndsp_derivs(self.h, ctypes.byref(s), ctypes.byref(d)))
        self._stats(s, d)
        return self.dt_info

    def step(self):
        s, d = _Stats(), _Dt()
        self._check(load_library().ndsp_step(self.h, ctypes.byref(s), ctypes.byref(d)))
        self._stats(s, d)
        return self.dt_info

    def set_dt(self, dt: float, time: float = 0.0):
        self._check(load_library().ndsp_set_dt(self.h, float(dt), float(time)))

    def conserved(self) -> dict:
        out = np.empty(len(CONSERVED))
        self._check(load_library().ndsp_conserved(self.h, _dp(out)))
        return dict(zip(CONSERVED, out))

    def ev_columns(self, t: float, dt: float) -> np.ndarray:
        """the numbers of one .ev line at time t with timestep dt, in the
        order evwrite writes them (ndsp_ev_columns; ndspmhd_amd.evfile names
        and formats them)"""
        out = np.zeros(25)
        nc = ctypes.c_int32()
        self._check(load_library().ndsp_ev_columns(self.h, float(t), float(dt), _dp(out), ctypes.byref(nc)))
        return out[: nc.value].copy()

    def neighbours(self, sym: bool = False) -> np.ndarray:
        L = load_library()
        npairs = ctypes.c_int64()
        rc = L.ndsp_neighbours(self.h, int(sym), None, 0, ctypes.byref(npairs))
        buf = np.empty(max(1, npairs.value) * 3, dtype=np.int32)
        self._check(L.ndsp_neighbours(self.h, int(sym), _ip(buf), npairs.value, ctypes.byref(npairs)))
        return buf[: 3 * npairs.value].reshape(-1, 3)

    # --- h-predictor MLP ---
    def mlp_config(self, enable=True, train=True, lr=1e-3, batch=16384, iters=4, seed_after=2):
        self._check(load_library().ndsp_mlp_config(self.h, int(enable), int(train), float(lr),
                                                   int(batch), int(iters), int(seed_after)))

    def mlp_set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._check(load_library().ndsp_mlp_set_params(self.h, _fp(p), p.size))

    def mlp_get_params(self):
        p = np.empty(MLP_NPARAM, dtype=np.float32)
        k = ctypes.c_int32()
        self._check(load_library().ndsp_mlp_get_params(self.h, _fp(p), p.size, ctypes.byref(k)))
        return p, k.value

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        self._check(load_library().ndsp_mlp_forward(self.h, _fp(X), X.shape[0], _fp(y)))
        return y

    def mlp_infer(self, X):
        """the seed's inference arithmetic (matrix cores, BF16 activations)"""
        X = np.ascontiguousarray(X, dtype=np.float32)
        y = np.empty(X.shape[0], dtype=np.float32)
        self._check(load_library().ndsp_mlp_infer(self.h, _fp(X), X.shape[0], _fp(y)))
        return y

    def mlp_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(MLP_NPARAM, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(load_library().ndsp_mlp_grad(self.h, _fp(X), _fp(T), X.shape[0], _fp(g), ctypes.byref(loss)))
        return loss.value, g

    def mlp_train(self, X, T, iters, lr):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(load_library().ndsp_mlp_train(self.h, _fp(X), _fp(T), X.shape[0], int(iters), float(lr),
                                                  ctypes.byref(loss)))
        return loss.value

    def set_timing(self, on: bool):
        self._check(load_library().ndsp_set_timing(self.h, int(on)))

    def autotune(self, reps: int = 2) -> dict:
        """choose the tile shapes by timing derivs (ndsp_autotune); the state
        is left unchanged"""
        ch = np.zeros(3, dtype=np.int32)
        self._check(load_library().ndsp_autotune(self.h, int(reps), _ip(ch)))
        return {"force_workgroup": int(ch[0]), "density_workgroup": int(ch[1]), "cells_per_radius": int(ch[2])}

    def set_tile_shapes(self, force_workgroup: int, density_workgroup: int, cells_per_radius: int):
        """tile shapes as autotune returns them (ndsp_set_tile_shapes)"""
        self._check(load_library().ndsp_set_tile_shapes(self.h, int(force_workgroup), int(density_workgroup),
                                                         int(cells_per_radius)))

    def copy_probe(self, nbytes: int):
        """plain device copy of nbytes, then a scattered read of as many
        64-byte lines (counter calibration, ndsp_copy_probe)"""
        self._check(load_library().ndsp_copy_probe(self.h, int(nbytes)))

    def set_reorder(self, every: int):
        """permute the state into cell order every `every` steps (0 = never;
        ndsp_set_reorder)"""
        self._check(load_library().ndsp_set_reorder(self.h, int(every)))

    def set_graph(self, on: bool):
        """replay each step as two recorded HIP graphs (ndsp_set_graph)"""
        self._check(load_library().ndsp_set_graph(self.h, int(on)))

    def kernel_times(self) -> dict:
        L = load_library()
        n = ctypes.c_int32()
        ms = np.zeros(64)
        names = ctypes.create_string_buffer(4096)
        self._check(L.ndsp_kernel_times(self.h, _dp(ms), 64, ctypes.byref(n), names, 4096))
        labels = names.value.decode().split(",") if n.value else []
        return dict(zip(labels, ms[: n.value]))


class Batch:
    """Independent simulations stepped together (ndsp_batch_*): both halves of
    every member's leapfrog step replayed as two recorded HIP graphs in which
    each phase is one launch over all members.  The batch keeps references to the
    Simulation objects; each member's state, stats and dt_info are updated
    as by Simulation.step."""

    def __init__(self, sims):
        L = load_library()
        self.sims = list(sims)
        arr = (ctypes.c_void_p * len(self.sims))(*[s.h.value for s in self.sims])
        h = ctypes.c_void_p()
        self._check(L.ndsp_batch_create(arr, len(self.sims), ctypes.byref(h)))
        self.h = h
        self._st = (_Stats * len(self.sims))()
        self._dt = (_Dt * len(self.sims))()

    def _check(self, rc):
        if rc != 0:
            raise NdspError(rc, load_library().ndsp_last_error().decode())

    def step(self):
        self._check(load_library().ndsp_batch_step(self.h, self._st, self._dt))
        for k, sim in enumerate(self.sims):
            sim._stats(self._st[k], self._dt[k])

    def autotune(self, reps: int = 2) -> dict:
        """tile shapes of the batch's merged launches, timed over batch steps
        (ndsp_batch_autotune); every member's state is left unchanged"""
        ch = np.zeros(2, dtype=np.int32)
        self._check(load_library().ndsp_batch_autotune(self.h, int(reps), _ip(ch)))
        return {"workgroup": int(ch[0]), "cells_per_radius": int(ch[1])}

    def close(self):
        if getattr(self, "h", None):
            load_library().ndsp_batch_destroy(self.h)
            self.h = None

    __del__ = close
