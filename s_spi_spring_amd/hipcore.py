"""ctypes binding of libspring_hip.so (the C ABI declared in include/spring_hip.h).

The library is built in-tree (s_spi_spring_amd/lib/libspring_hip.so) by
``s_spi_spring_amd.build.build_all()``. Loading fails loudly when it is missing:
there is no fallback implementation of any operation.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libspring_hip.so")

ER, HIO, RAAR = 0, 1, 2
ALG = {"ER": ER, "HIO": HIO, "RAAR": RAAR}

_lib = None

# name -> (restype, argtypes); mirrors include/spring_hip.h
_VP = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
SIGNATURES = {
    "spr_supported_size": (_I, [_I]),
    "spr_create": (_I, [ctypes.POINTER(_VP), _I, _I, _VP, _I]),
    "spr_destroy": (_I, [_VP]),
    "spr_last_error": (ctypes.c_char_p, [_VP]),
    "spr_size": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "spr_set_pattern": (_I, [_VP, _VP]),
    "spr_get_pattern": (_I, [_VP, _VP]),
    "spr_set_population": (_I, [_VP, _I, _VP, _VP]),
    "spr_get_population": (_I, [_VP, _I, _VP, _VP]),
    "spr_ft": (_I, [_VP]),
    "spr_ift": (_I, [_VP]),
    "spr_iterate": (_I, [_VP, _I, _I, _F, _F]),
    "spr_iterate_async": (_I, [_VP, _I, _I, _F, _F]),
    "spr_sync": (_I, [_VP]),
    "spr_device_buffers": (_I, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "spr_stream": (_I, [_VP, ctypes.POINTER(_VP)]),
    "spr_timing": (_I, [_VP, _I]),
    "spr_timing_read": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]),
}

KIND_COL_PROJ, KIND_ROW_UPDATE, KIND_ROW_OTHER, KIND_COL_OTHER = 0, 1, 2, 3


class SpringHipError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libspring_hip.so and declare its signatures. Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SpringHipError(
            f"{path} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Backend:
    """One population-resident backend (spr_ctx) for an n x n pattern."""

    def __init__(self, amp: np.ndarray, capacity: int, device: int = 0):
        self.lib = load()
        amp = np.ascontiguousarray(amp, dtype=np.float32)
        n = amp.shape[0]
        if amp.shape != (n, n):
            raise ValueError("amplitude must be square")
        h = ctypes.c_void_p()
        rc = self.lib.spr_create(ctypes.byref(h), n, int(capacity), _ptr(amp), int(device))
        if rc != 0:
            raise SpringHipError(self.lib.spr_last_error(None).decode())
        self.h = h
        self.n = n
        self.capacity = int(capacity)
        self.P = 0

    def _check(self, rc):
        if rc != 0:
            raise SpringHipError(self.lib.spr_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.spr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pattern(self, amp):
        a = np.ascontiguousarray(amp, dtype=np.float32)
        self._check(self.lib.spr_set_pattern(self.h, _ptr(a)))

    def get_pattern(self):
        a = np.empty((self.n, self.n), dtype=np.float32)
        self._check(self.lib.spr_get_pattern(self.h, _ptr(a)))
        return a

    def set_population(self, rho, support=None):
        rho = np.ascontiguousarray(rho, dtype=np.complex64)
        if rho.ndim == 2:
            rho = rho[None]
        P = rho.shape[0]
        s = None
        if support is not None:
            s = np.ascontiguousarray(np.broadcast_to(support, rho.shape), dtype=np.int32)
        self._check(self.lib.spr_set_population(self.h, P, _ptr(rho), _ptr(s)))
        self.P = P

    def get_population(self, P=None, support=True):
        P = self.P if P is None else P
        rho = np.empty((P, self.n, self.n), dtype=np.complex64)
        s = np.empty((P, self.n, self.n), dtype=np.int32) if support else None
        self._check(self.lib.spr_get_population(self.h, P, _ptr(rho), _ptr(s)))
        return (rho, s) if support else rho

    def ft(self):
        self._check(self.lib.spr_ft(self.h))

    def ift(self):
        self._check(self.lib.spr_ift(self.h))

    def iterate(self, alg, niter, beta=0.9, reality=0.0):
        alg = ALG[alg] if isinstance(alg, str) else int(alg)
        self._check(self.lib.spr_iterate(self.h, alg, int(niter), float(beta), float(reality)))

    def iterate_async(self, alg, niter, beta=0.9, reality=0.0):
        alg = ALG[alg] if isinstance(alg, str) else int(alg)
        self._check(self.lib.spr_iterate_async(self.h, alg, int(niter), float(beta), float(reality)))

    def sync(self):
        self._check(self.lib.spr_sync(self.h))

    def timing(self, enable: bool):
        self._check(self.lib.spr_timing(self.h, int(bool(enable))))

    def timing_read(self, kind):
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        self._check(self.lib.spr_timing_read(self.h, int(kind), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def stream(self):
        s = ctypes.c_void_p()
        self._check(self.lib.spr_stream(self.h, ctypes.byref(s)))
        return s.value
