"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product package
(ndspmhd_amd) never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

FIELDS3 = ("vel", "Bevol", "alpha", "Bfield", "force", "dBevoldt", "daldt",
           "curlB", "gradpsi", "deltav", "graddivv")


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        L.orc_new.restype = ctypes.c_void_p
        L.orc_new.argtypes = [ip, dp, ctypes.c_int32, ctypes.c_char_p]
        L.orc_free.argtypes = [ctypes.c_void_p]
        L.orc_set_state.argtypes = [ctypes.c_void_p] + [dp] * 9 + [ip]
        L.orc_derivs.argtypes = [ctypes.c_void_p]
        L.orc_step.argtypes = [ctypes.c_void_p]
        L.orc_predict.argtypes = [ctypes.c_void_p]
        L.orc_set.argtypes = [ctypes.c_void_p, ctypes.c_char_p, dp]
        L.orc_get.argtypes = [ctypes.c_void_p, ctypes.c_char_p, dp]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.orc_set_scalar.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
        L.orc_set_ireal.argtypes = [ctypes.c_void_p, ip]
        L.orc_neighbours.restype = ctypes.c_int64
        L.orc_neighbours.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ctypes.c_int64]
        L.orc_conserved.argtypes = [ctypes.c_void_p, dp]
        L.orc_kernel_table.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.orc_interpolate.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp, dp]
        L.orc_message.restype = ctypes.c_char_p
        L.orc_message.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


CONSERVED = ("ekin", "etherm", "emag", "etot", "momx", "momy", "momz", "momtot",
             "angx", "angy", "angz", "angtot", "rhomin", "rhomax", "rhomean",
             "totmassgas", "totmassdust")


class OracleError(RuntimeError):
    pass


class Oracle:
    """One simulation held by the CPU oracle."""

    def __init__(self, opts, state):
        L = lib()
        self.opts = opts
        self.n = int(state["pmass"].shape[0])
        self._iopt = np.ascontiguousarray(opts.iopt(), dtype=np.int32)
        self._dopt = np.ascontiguousarray(opts.dopt(), dtype=np.float64)
        err = ctypes.create_string_buffer(256)
        h = L.orc_new(_ip(self._iopt), _dp(self._dopt), self.n, err)
        if not h:
            raise OracleError(err.value.decode())
        self.h = ctypes.c_void_p(h)
        a = {k: np.ascontiguousarray(state[k], dtype=np.float64)
             for k in ("x", "vel", "pmass", "hh", "en", "Bevol", "alpha", "psi", "dustevol")}
        it = np.ascontiguousarray(state["itype"], dtype=np.int32)
        L.orc_set_state(self.h, *[_dp(a[k]) for k in ("x", "vel", "pmass", "hh", "en",
                                                       "Bevol", "alpha", "psi", "dustevol")], _ip(it))
        if "ireal" in state:
            ir = np.ascontiguousarray(state["ireal"], dtype=np.int32)
            self._check(L.orc_set_ireal(self.h, _ip(ir)), "set_ireal")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_free(h)
            self.h = None

    def _check(self, rc, what):
        if rc != 0:
            raise OracleError("%s failed (%d): %s" % (what, rc, lib().orc_message(self.h).decode()))

    def derivs(self):
        self._check(lib().orc_derivs(self.h), "derivs")

    def step(self):
        self._check(lib().orc_step(self.h), "step")

    def predict(self):
        """the predictor half of a step alone (no evaluation, no corrector)"""
        self._check(lib().orc_predict(self.h), "predict")

    def get(self, name):
        w = 3 if name in FIELDS3 else (self.opts.ndim if name == "x" else 1)
        out = np.empty(self.n * w)
        r = lib().orc_get(self.h, name.encode(), _dp(out))
        if r < 0:
            raise KeyError(name)
        return out.reshape(self.n, w) if w > 1 else out

    def set(self, name, values):
        """overwrite a per-particle field (state fields and the rates a step
        starts from)"""
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        w = 3 if name in FIELDS3 else (self.opts.ndim if name == "x" else 1)
        if a.size != self.n * w or lib().orc_set(self.h, name.encode(), _dp(a)) < 0:
            raise KeyError(name)

    def scalar(self, name):
        return lib().orc_scalar(self.h, name.encode())

    def set_scalar(self, name, v):
        if lib().orc_set_scalar(self.h, name.encode(), float(v)) != 0:
            raise KeyError(name)

    def neighbours(self, sym=False):
        cap = self.n * 128
        while True:
            buf = np.empty(cap * 3, dtype=np.int32)
            n = lib().orc_neighbours(self.h, int(sym), _ip(buf), cap)
            if n >= 0:
                return buf[: 3 * n].reshape(n, 3).copy()
            cap = -n

    def conserved(self):
        out = np.empty(len(CONSERVED))
        lib().orc_conserved(self.h, _dp(out))
        return dict(zip(CONSERVED, out))


def kernel_table(ikernel, ndim):
    w = np.empty(4001); g = np.empty(4001); gg = np.empty(4001); r = np.empty(1)
    lib().orc_kernel_table(ikernel, ndim, _dp(w), _dp(g), _dp(gg), _dp(r))
    return w, g, gg, float(r[0])


def interpolate(ikernel, ndim, q2):
    w = np.empty(1); g = np.empty(1); gg = np.empty(1)
    lib().orc_interpolate(ikernel, ndim, float(q2), _dp(w), _dp(g), _dp(gg))
    return float(w[0]), float(g[0]), float(gg[0])
