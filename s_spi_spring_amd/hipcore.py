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
# SPRING_HIP_LIB selects another build of the same library (e.g. a tuning variant).
LIB_PATH = os.environ.get("SPRING_HIP_LIB") or os.path.join(_HERE, "lib", "libspring_hip.so")

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
    "spr_set_active": (_I, [_VP, _I]),
    "spr_gather": (_I, [_VP, _I, _VP]),
    "spr_put_individuals": (_I, [_VP, _I, _VP, _VP, _VP]),
    "spr_fetch_individuals": (_I, [_VP, _I, _VP, _VP, _VP]),
    "spr_ft": (_I, [_VP]),
    "spr_ift": (_I, [_VP]),
    "spr_iterate": (_I, [_VP, _I, _I, _F, _F]),
    "spr_iterate_async": (_I, [_VP, _I, _I, _F, _F]),
    "spr_sync": (_I, [_VP]),
    "spr_device_buffers": (_I, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    "spr_stream": (_I, [_VP, ctypes.POINTER(_VP)]),
    "spr_timing": (_I, [_VP, _I]),
    "spr_set_fused": (_I, [_VP, _I]),
    "spr_get_fused": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "spr_timing_read": (_I, [_VP, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]),
    "spr_sw": (_I, [_VP, _F, _F, _F]),
    "spr_swa": (_I, [_VP, _F, _I, _F]),
    "spr_evaluate": (_I, [_VP, _F, _VP, _VP, _VP, _VP, _VP]),
    "spr_symmetry": (_I, [_VP, _F, _F, _VP]),
    "spr_run": (_I, [_VP, _VP, _I, _I, _F]),
    "spr_align": (_I, [_VP, _VP, _F, _F, _VP, _VP, _VP]),
    "spr_average": (_I, [_VP, _VP, _I, _VP, _VP]),
    "spr_average_module": (_I, [_VP, _VP, _I, _I, _VP]),
    "spr_crossover": (_I, [_VP, _I, _VP, _VP, _I, _VP, _VP, _F, _F, _F, _F, _F, _I, _VP, _VP]),
    "spr_device_count": (_I, [ctypes.POINTER(_I)]),
    "spr_device_name": (_I, [_I, ctypes.c_char_p, _I]),
    "spr_loss_grad": (_I, [_VP, _I, _VP, _VP, _VP, _I, _F, _VP]),
    "spr_loss_grad_noalloc": (_I, [_I, _I, _VP, _VP, _VP, _VP, _I, _F, _I, _VP]),
    "spr_refine": (_I, [_VP, _I, _F, _I, _F, _VP]),
    "spr_refine_support": (_I, [_VP, _I, _F, _I, _F, _VP]),
    "spr_batch_create": (_I, [ctypes.POINTER(_VP), _I, _VP, _VP, _VP, _I]),
    "spr_batch_destroy": (_I, [_VP]),
    "spr_batch_last_error": (ctypes.c_char_p, [_VP]),
    "spr_batch_graphs": (_I, [_VP, _I]),
    "spr_batch_concurrent": (_I, [_VP, _I]),
    "spr_batch_set_population": (_I, [_VP, _I, _VP, _VP]),
    "spr_batch_get_population": (_I, [_VP, _I, _VP, _VP]),
    "spr_batch_run": (_I, [_VP, _VP, _I, _I, _F]),
    "spr_batch_evaluate": (_I, [_VP, _F, _VP, _VP, _VP, _VP]),
    "spr_batch_autotune": (_I, [_VP, _I]),
    "spr_batch_set_chunk": (_I, [_VP, ctypes.c_longlong]),
    "spr_variants": (_I, [_I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "spr_set_chunk": (_I, [_VP, ctypes.c_longlong]),
    "spr_get_chunk": (_I, [_VP, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(_I)]),
    "spr_autotune": (_I, [_VP, _I, _VP, _VP]),
    "spr_row_variant_info": (_I, [_I, _I, ctypes.POINTER(_I)]),
    "spr_set_variants": (_I, [_VP, _I, _I]),
    "spr_get_variants": (_I, [_VP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "spr_default_variants": (_I, [_I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "spr_set_default_variants": (_I, [_I, _I, _I]),
}


class SprCrossDesc(ctypes.Structure):
    """struct spr_cross_desc of include/spring_hip.h."""
    _fields_ = [(k, ctypes.c_int) for k in ("p0", "p1", "p2", "p3", "crossft", "crosssym", "crosssizex",
                                             "crosssizey", "squaresizex", "squaresizey", "offsetx", "offsety")] + \
               [("rotangle", ctypes.c_float), ("crossmap_offset", ctypes.c_int), ("phase", ctypes.c_float * 8)]


class SprOp(ctypes.Structure):
    """struct spr_op of include/spring_hip.h."""
    _fields_ = [("kind", ctypes.c_int), ("iterations", ctypes.c_int), ("beta", ctypes.c_float),
                ("sigma", ctypes.c_float), ("threshold", ctypes.c_float), ("pixels", ctypes.c_int)]


OP_KIND = {"ER": 0, "HIO": 1, "RAAR": 2, "SW": 3, "SWA": 4, "GRAD": 5}


def make_ops(sequence):
    """[(name, [params])] in Solver.set_sequence form -> ctypes array of SprOp.
    ER [n]; HIO/RAAR [n, beta]; SW [sigma, threshold]; SWA [sigma, pixels]."""
    ops = (SprOp * max(1, len(sequence)))()
    for i, (name, p) in enumerate(sequence):
        o = ops[i]
        o.kind = OP_KIND[name]
        if name == "ER":
            o.iterations = int(p[0])
        elif name in ("HIO", "RAAR"):
            o.iterations, o.beta = int(p[0]), float(p[1])
        elif name == "SW":
            o.sigma, o.threshold = float(p[0]), float(p[1])
        elif name == "GRAD":
            o.iterations, o.beta, o.threshold = int(p[0]), float(p[1]), 1e-3
            o.pixels = int(p[2]) if len(p) > 2 else 0
        else:
            o.sigma, o.pixels = float(p[0]), int(p[1])
    return ops

KIND_COL_PROJ, KIND_ROW_UPDATE, KIND_ROW_OTHER, KIND_COL_OTHER, KIND_BLOCK = 0, 1, 2, 3, 4


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


def loss_grad_noalloc(n, P, amp_ptr, rho_ptr, grad_ptr, loss_ptr=None, kind=0, eps=1e-3, want_grad=True,
                      stream=None):
    """spr_loss_grad_noalloc on device addresses: losses (FP64, P entries at
    loss_ptr, optional) and gradients of P n x n fields, queued on `stream`
    without any allocation or synchronisation."""
    lib = load()
    rc = lib.spr_loss_grad_noalloc(int(n), int(P), ctypes.c_void_p(amp_ptr), ctypes.c_void_p(rho_ptr),
                                   ctypes.c_void_p(grad_ptr), ctypes.c_void_p(loss_ptr or 0), int(kind), float(eps),
                                   int(bool(want_grad)), ctypes.c_void_p(stream or 0))
    if rc != 0:
        raise SpringHipError(lib.spr_last_error(None).decode())


def variant_counts(n):
    """(column variants, row variants) compiled for size n."""
    lib = load()
    c, r = ctypes.c_int(), ctypes.c_int()
    if lib.spr_variants(int(n), ctypes.byref(c), ctypes.byref(r)) != 0:
        raise SpringHipError(lib.spr_last_error(None).decode())
    return c.value, r.value


def row_variant_recomputes_state(n, variant):
    """Whether row variant `variant` of size n recomputes rho_n from its row
    transform instead of storing it (spr_row_variant_info)."""
    lib = load()
    r = ctypes.c_int()
    if lib.spr_row_variant_info(int(n), int(variant), ctypes.byref(r)) != 0:
        raise SpringHipError(lib.spr_last_error(None).decode())
    return bool(r.value)


def default_variants(n):
    """(column, row) variants that new handles of size n start with."""
    lib = load()
    c, r = ctypes.c_int(), ctypes.c_int()
    if lib.spr_default_variants(int(n), ctypes.byref(c), ctypes.byref(r)) != 0:
        raise SpringHipError(lib.spr_last_error(None).decode())
    return c.value, r.value


def set_default_variants(n, col, row):
    lib = load()
    if lib.spr_set_default_variants(int(n), int(col), int(row)) != 0:
        raise SpringHipError(lib.spr_last_error(None).decode())


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
        self.device = int(device)
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

    def set_active(self, P):
        self._check(self.lib.spr_set_active(self.h, int(P)))
        self.P = int(P)

    def gather(self, src):
        """Slot i <- slot src[i] (-1: leave slot i)."""
        src = np.ascontiguousarray(src, dtype=np.int32)
        self._check(self.lib.spr_gather(self.h, len(src), _ptr(src)))

    def put_individuals(self, slots, rho, support=None):
        """Upload rho[k] (and support[k]) into slot slots[k]."""
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        rho = [np.ascontiguousarray(r, dtype=np.complex64) for r in rho]
        rp = (ctypes.c_void_p * len(rho))(*[r.ctypes.data for r in rho])
        sp = None
        if support is not None:
            support = [np.ascontiguousarray(x, dtype=np.int32) for x in support]
            sp = (ctypes.c_void_p * len(support))(*[x.ctypes.data for x in support])
        self._check(self.lib.spr_put_individuals(self.h, len(slots), _ptr(slots), rp, sp))

    def fetch_individuals(self, slots, support=True):
        slots = np.ascontiguousarray(slots, dtype=np.int32)
        rho = [np.empty((self.n, self.n), np.complex64) for _ in slots]
        rp = (ctypes.c_void_p * len(rho))(*[r.ctypes.data for r in rho])
        sup, sp = None, None
        if support:
            sup = [np.empty((self.n, self.n), np.int32) for _ in slots]
            sp = (ctypes.c_void_p * len(sup))(*[x.ctypes.data for x in sup])
        self._check(self.lib.spr_fetch_individuals(self.h, len(slots), _ptr(slots), rp, sp))
        return (rho, sup) if support else rho

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

    def sw(self, sigma, threshold, reality=0.0):
        self._check(self.lib.spr_sw(self.h, float(sigma), float(threshold), float(reality)))

    def swa(self, sigma, pixels, reality=0.0):
        self._check(self.lib.spr_swa(self.h, float(sigma), int(pixels), float(reality)))

    def evaluate(self, reality=0.0, symmetry=True):
        """Per individual: dict of error_fourier, error_real, complexsum, modulesum[, symmetry]."""
        P = self.P
        out = {k: np.zeros(P, np.float32) for k in ("error_fourier", "error_real", "complexsum", "modulesum")}
        sym = np.zeros(P, np.float32) if symmetry else None
        self._check(self.lib.spr_evaluate(self.h, float(reality), _ptr(out["error_fourier"]), _ptr(out["error_real"]),
                                          _ptr(out["complexsum"]), _ptr(out["modulesum"]), _ptr(sym)))
        if symmetry:
            out["symmetry"] = sym
        return out

    def symmetry(self, symmin=0.0, symmax=1.0):
        s = np.zeros(self.P, np.float32)
        self._check(self.lib.spr_symmetry(self.h, float(symmin), float(symmax), _ptr(s)))
        return s

    def run(self, sequence, nrep=1, reality=0.0):
        ops = make_ops(sequence)
        self._check(self.lib.spr_run(self.h, ctypes.cast(ops, ctypes.c_void_p), len(sequence), int(nrep),
                                     float(reality)))

    def align(self, ref_rho, phase_direct, phase_flip, minr=0.0, maxr=2.0):
        """Align every resident individual to ref_rho; returns (P, 3) shifts (sx, sy, flip)."""
        ref = np.ascontiguousarray(ref_rho, dtype=np.complex64)
        pd = np.ascontiguousarray(phase_direct, dtype=np.float32)
        pf = np.ascontiguousarray(phase_flip, dtype=np.float32)
        sh = np.zeros((self.P, 3), np.int32)
        self._check(self.lib.spr_align(self.h, _ptr(ref), float(minr), float(maxr), _ptr(pd), _ptr(pf), _ptr(sh)))
        return sh

    def average(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        rho = np.zeros((self.n, self.n), np.complex64)
        s = np.zeros((self.n, self.n), np.int32)
        self._check(self.lib.spr_average(self.h, _ptr(idx), len(idx), _ptr(rho), _ptr(s)))
        return rho, s

    def average_module(self, idx, ft=True):
        idx = np.ascontiguousarray(idx, dtype=np.int32)
        out = np.zeros((self.n, self.n), np.float32)
        self._check(self.lib.spr_average_module(self.h, _ptr(idx), len(idx), int(bool(ft)), _ptr(out)))
        return out

    def loss_grad(self, P, rho_ptr, grad_ptr, kind=0, eps=1e-3, stream=None, want_loss=True):
        """Loss and gradient of P fields at device addresses rho_ptr -> grad_ptr
        (grad_ptr None: the losses only)."""
        loss = np.zeros(P, np.float64) if want_loss else None
        self._check(self.lib.spr_loss_grad(self.h, int(P), ctypes.c_void_p(rho_ptr), ctypes.c_void_p(grad_ptr or 0),
                                           _ptr(loss), int(kind), float(eps), ctypes.c_void_p(stream or 0)))
        return loss

    def device_buffers(self):
        """Device addresses (rho, uint8 supports, amplitude) of the resident population."""
        r, s, a = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        self._check(self.lib.spr_device_buffers(self.h, ctypes.byref(r), ctypes.byref(s), ctypes.byref(a)))
        return r.value, s.value, a.value

    def refine(self, steps, lr=1e-2, kind=0, eps=1e-3):
        loss = np.zeros(self.P, np.float64)
        self._check(self.lib.spr_refine(self.h, int(steps), float(lr), int(kind), float(eps), _ptr(loss)))
        return loss

    def refine_support(self, steps, lr_rel=1e-2, kind=0, eps=1e-3):
        loss = np.zeros(self.P, np.float64)
        self._check(self.lib.spr_refine_support(self.h, int(steps), float(lr_rel), int(kind), float(eps), _ptr(loss)))
        return loss

    def timing(self, enable: bool, stride: int = 1):
        """Bracket every `stride`-th launch of each kind with HIP events (0/False: off)."""
        self._check(self.lib.spr_timing(self.h, int(stride) if enable else 0))

    def timing_read(self, kind):
        ms = ctypes.c_double()
        cnt = ctypes.c_long()
        self._check(self.lib.spr_timing_read(self.h, int(kind), ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value

    def set_chunk(self, nbytes):
        """Working set per chunk of an iteration block (0: the whole population)."""
        self._check(self.lib.spr_set_chunk(self.h, int(nbytes)))

    def chunk(self):
        """(bytes, individuals) per chunk of an iteration block."""
        b, k = ctypes.c_longlong(), ctypes.c_int()
        self._check(self.lib.spr_get_chunk(self.h, ctypes.byref(b), ctypes.byref(k)))
        return b.value, k.value

    def set_fused(self, mode):
        """On-chip iteration blocks for 64^2 and 128^2 fields: -1 from a
        population of 8 (default), 0 never, 1 always."""
        self._check(self.lib.spr_set_fused(self.h, int(mode)))

    def fused(self):
        """(mode, whether the next iteration block runs on chip)."""
        m, a = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.spr_get_fused(self.h, ctypes.byref(m), ctypes.byref(a)))
        return m.value, bool(a.value)

    def variants(self):
        """(column variant, row variant) of this handle."""
        c, r = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.spr_get_variants(self.h, ctypes.byref(c), ctypes.byref(r)))
        return c.value, r.value

    def set_variants(self, col, row):
        self._check(self.lib.spr_set_variants(self.h, int(col), int(row)))

    def autotune(self, reps=5):
        """Time every variant on the resident population and keep the fastest;
        returns (column ms per variant, row ms per variant)."""
        nc, nr = variant_counts(self.n)
        cm, rm = np.zeros(nc, np.float32), np.zeros(nr, np.float32)
        self._check(self.lib.spr_autotune(self.h, int(reps), _ptr(cm), _ptr(rm)))
        return cm, rm

    def stream(self):
        s = ctypes.c_void_p()
        self._check(self.lib.spr_stream(self.h, ctypes.byref(s)))
        return s.value


class Batch:
    """Many patterns of several sizes on one device (spr_batch): every ER/HIO/RAAR
    block covers all individuals of all patterns, one launch set per size.

    amps: corner-origin amplitudes (n_i x n_i, any supported sizes); populations:
    individuals per pattern."""

    def __init__(self, amps, populations, device: int = 0):
        self.lib = load()
        self.amps = [np.ascontiguousarray(a, dtype=np.float32) for a in amps]
        for a in self.amps:
            if a.ndim != 2 or a.shape[0] != a.shape[1]:
                raise ValueError("every amplitude must be square")
        self.sizes = np.array([a.shape[0] for a in self.amps], np.int32)
        self.populations = np.array([int(p) for p in populations], np.int32)
        if len(self.populations) != len(self.amps):
            raise ValueError("one population size per pattern")
        ptrs = (ctypes.c_void_p * len(self.amps))(*[a.ctypes.data for a in self.amps])
        h = ctypes.c_void_p()
        rc = self.lib.spr_batch_create(ctypes.byref(h), len(self.amps), _ptr(self.sizes), _ptr(self.populations),
                                       ctypes.cast(ptrs, ctypes.c_void_p), int(device))
        if rc != 0:
            raise SpringHipError(self.lib.spr_batch_last_error(None).decode())
        self.h = h

    @property
    def total(self):
        return int(self.populations.sum())

    def _check(self, rc):
        if rc != 0:
            raise SpringHipError(self.lib.spr_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.spr_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graphs(self, enable: bool):
        self._check(self.lib.spr_batch_graphs(self.h, 1 if enable else 0))

    def concurrent(self, enable: bool):
        """Size classes of an iteration block side by side (True) or in turn (False, the default)."""
        self._check(self.lib.spr_batch_concurrent(self.h, 1 if enable else 0))

    def set_population(self, k, rho, support=None):
        n, P = int(self.sizes[k]), int(self.populations[k])
        rho = np.ascontiguousarray(np.broadcast_to(rho, (P, n, n)), dtype=np.complex64)
        s = None if support is None else np.ascontiguousarray(np.broadcast_to(support, (P, n, n)), dtype=np.int32)
        self._check(self.lib.spr_batch_set_population(self.h, int(k), _ptr(rho), _ptr(s)))

    def get_population(self, k, support=True):
        n, P = int(self.sizes[k]), int(self.populations[k])
        rho = np.empty((P, n, n), np.complex64)
        s = np.empty((P, n, n), np.int32) if support else None
        self._check(self.lib.spr_batch_get_population(self.h, int(k), _ptr(rho), _ptr(s)))
        return (rho, s) if support else rho

    def run(self, sequence, nrep=1, reality=0.0):
        ops = make_ops(sequence)
        self._check(self.lib.spr_batch_run(self.h, ctypes.cast(ops, ctypes.c_void_p), len(sequence), int(nrep),
                                           float(reality)))

    def set_chunk(self, nbytes):
        """spr_set_chunk on every size class (see Backend.set_chunk)."""
        self._check(self.lib.spr_batch_set_chunk(self.h, int(nbytes)))

    def autotune(self, reps=5):
        """spr_autotune on every size class (see Backend.autotune)."""
        self._check(self.lib.spr_batch_autotune(self.h, int(reps)))

    def evaluate(self, reality=0.0):
        """Dict of per-pattern arrays: error_fourier, error_real, complexsum, modulesum."""
        out = {k: np.zeros(self.total, np.float32) for k in ("error_fourier", "error_real", "complexsum", "modulesum")}
        self._check(self.lib.spr_batch_evaluate(self.h, float(reality), _ptr(out["error_fourier"]),
                                                _ptr(out["error_real"]), _ptr(out["complexsum"]),
                                                _ptr(out["modulesum"])))
        bounds = np.concatenate([[0], np.cumsum(self.populations)])
        return {k: [v[bounds[i]:bounds[i + 1]] for i in range(len(self.populations))] for k, v in out.items()}
