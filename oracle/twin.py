"""ctypes wrapper of the C twin (oracle/ed_twin.c).  TEST INFRASTRUCTURE ONLY."""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "libedtwin.so"
_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE)], check=True, capture_output=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        l = C.CDLL(str(LIB))
        l.twin_create.restype = C.c_void_p
        l.twin_create.argtypes = [C.c_int] * 4 + [C.POINTER(C.c_int), C.POINTER(C.c_double),
                                                  C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int]
        l.twin_destroy.argtypes = [C.c_void_p]
        l.twin_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        l.twin_matvec.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        l.twin_lanczos.restype = C.c_int
        l.twin_lanczos.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_uint64,
                                   C.POINTER(C.c_double), C.POINTER(C.c_double)]
        l.twin_threads.restype = C.c_int
        _lib = l
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class TwinOperator:
    def __init__(self, num_sites, n_up, n_dn, bonds, hops, eps, u, width):
        self.bonds = np.ascontiguousarray(np.asarray(bonds, dtype=np.int32).reshape(-1, 2))
        self.hops = np.ascontiguousarray(np.broadcast_to(np.asarray(hops, np.float64), (len(self.bonds),)))
        self.eps = np.ascontiguousarray(np.broadcast_to(np.asarray(eps, np.float64), (num_sites,)))
        self.u = np.ascontiguousarray(np.broadcast_to(np.asarray(u, np.float64), (num_sites,)))
        b = self.bonds.ravel() if len(self.bonds) else np.zeros(2, np.int32)
        self._b = np.ascontiguousarray(b, dtype=np.int32)
        h = self.hops if len(self.hops) else np.zeros(1)
        self._h = np.ascontiguousarray(h)
        self.ptr = lib().twin_create(num_sites, n_up, n_dn, len(self.bonds),
                                     self._b.ctypes.data_as(C.POINTER(C.c_int)), _dp(self._h), _dp(self.eps),
                                     _dp(self.u), int(width))
        if not self.ptr:
            raise ValueError("twin_create rejected the model")
        nu, nd = C.c_int64(), C.c_int64()
        lib().twin_dims(self.ptr, C.byref(nu), C.byref(nd))
        self.num_up, self.num_dn = nu.value, nd.value
        self.size = self.num_up * self.num_dn

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().twin_destroy(self.ptr)
            self.ptr = None

    def matvec(self, x, out=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = out if out is not None else np.empty_like(x)
        lib().twin_matvec(self.ptr, _dp(x), _dp(y))
        return y

    def lanczos(self, iters, v0=None, seed=0):
        a = np.zeros(iters)
        b = np.zeros(iters)
        v = None if v0 is None else np.ascontiguousarray(v0, np.float64)
        m = lib().twin_lanczos(self.ptr, iters, _dp(v) if v is not None else None, seed, _dp(a), _dp(b))
        if m < 0:
            raise MemoryError("twin_lanczos allocation failed")
        return a[:m], b[:m]


def threads() -> int:
    return lib().twin_threads()
