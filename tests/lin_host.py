"""ctypes binding of libatmos_lin_host.so: the linearised step's program and
per-point bodies run by host loops (csrc/lin_host.cpp), for CPU checks of
the tangent-linear and adjoint models.  Arrays are halo'd (nk, ny+14, nx+14)."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ATM_LIN_HOST_LIB names another build (the sanitizer build)
LIB = os.environ.get("ATM_LIN_HOST_LIB") or os.path.join(ROOT, "atmosphere-model_amd", "lib", "libatmos_lin_host.so")
H = 7
STATE = ("dp", "pt", "ud", "vd")


def _lib():
    L = C.CDLL(LIB)
    dp = C.POINTER(C.c_double)
    L.lin_host_create.restype = C.c_void_p
    L.lin_host_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, dp, dp, dp]
    L.lin_host_destroy.argtypes = [C.c_void_p]
    L.lin_host_put.argtypes = [C.c_void_p, C.c_int, dp]
    L.lin_host_get.argtypes = [C.c_void_p, C.c_int, dp]
    L.lin_host_tl_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
    L.lin_host_ad_step.argtypes = [C.c_void_p, C.c_double, C.c_int]
    L.lin_host_tape.argtypes = [C.c_int]
    L.lin_host_tape_stat.argtypes = [C.c_int]
    L.lin_host_tape_stat.restype = C.c_longlong
    return L


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def halo(a):
    """periodic halo of interior arrays (..., ny, nx) -> (..., ny+14, nx+14)"""
    return np.pad(a, [(0, 0)] * (a.ndim - 2) + [(H, H), (H, H)], mode="wrap")


class HostLin:
    def __init__(self, ic, gravity=9.807, coriolis=1e-5, divergence_damping=0.15, vorticity_damping=0.005):
        self.L = _lib()
        self.nlay, self.ny, self.nx = ic["dp"].shape
        prm = np.array([gravity, coriolis, divergence_damping, vorticity_damping])
        self.h = self.L.lin_host_create(self.nx, self.ny, self.nlay, ic["dx"], ic["dy"], _p(prm),
                                        _p(np.ascontiguousarray(ic["ak"], dtype=float)),
                                        _p(np.ascontiguousarray(ic["bk"], dtype=float)))
        assert self.h
        self.put(12, halo(np.asarray(ic["gzs"], dtype=float)[None])[0])
        self.put(13, np.zeros((self.nlay + 1, self.ny + 2 * H, self.nx + 2 * H)))
        self.set_state({"dp": ic["dp"], "pt": ic["pt"], "ud": ic["u"], "vd": ic["v"]})

    def __del__(self):
        if getattr(self, "h", None):
            self.L.lin_host_destroy(self.h)
            self.h = None

    def put(self, which, a):
        a = np.ascontiguousarray(a, dtype=float)
        assert self.L.lin_host_put(self.h, which, _p(a)) == 0

    def get(self, which, nk=None):
        nk = self.nlay if nk is None else nk
        a = np.zeros((nk, self.ny + 2 * H, self.nx + 2 * H))
        assert self.L.lin_host_get(self.h, which, _p(a)) == 0
        return a

    def _set(self, base, d):
        for i, n in enumerate(STATE):
            self.put(base + i, halo(np.asarray(d[n], dtype=float)))

    def _get(self, base):
        return {n: self.get(base + i)[:, H:-H, H:-H].copy() for i, n in enumerate(STATE)}

    def set_state(self, d):
        self._set(0, d)

    def state(self):
        return self._get(0)

    def set_tangent(self, d):
        self._set(4, d)

    def tangent(self):
        return self._get(4)

    def set_adjoint(self, d):
        self._set(8, d)

    def adjoint(self):
        return self._get(8)

    def tl_step(self, dt, remap=False):
        assert self.L.lin_host_tl_step(self.h, dt, int(remap)) == 0

    def ad_step(self, dt, remap=False):
        assert self.L.lin_host_ad_step(self.h, dt, int(remap)) == 0


def tape(mode):
    """Branch-decision tape of the host program: "off", "record" (clears),
    "replay" (from the start)."""
    _lib().lin_host_tape({"off": 0, "record": 1, "replay": 2}[mode])


def tape_stats():
    L = _lib()
    return {k: int(L.lin_host_tape_stat(i)) for i, k in enumerate(("recorded", "replayed", "flips", "overrun"))}
