"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package. Each function wraps
the restatement in oracle/spring_oracle.hpp (which cites the reference lines it
follows). Complex fields are numpy complex64 (FP32 oracle) or complex128 (FP64
oracle) arrays of shape (n, n) or (P, n, n); supports are int32.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _rt(dtype):
    if dtype in (np.complex64, np.float32):
        return np.float32, "f32", np.complex64
    return np.float64, "f64", np.complex128


def _cfield(x, ctype):
    return np.ascontiguousarray(x, dtype=ctype)


def fft2(f, sign):
    """Unnormalised 2-D DFT with kernel exp(sign*2*pi*i*k.x/N)."""
    rt, sfx, ct = _rt(f.dtype.type)
    out = _cfield(f, ct).copy()
    getattr(lib(), "orc_fft2_" + sfx)(ctypes.c_int(out.shape[-1]), _p(out), ctypes.c_int(sign))
    return out


def ft(f):
    rt, sfx, ct = _rt(f.dtype.type)
    out = _cfield(f, ct).copy()
    getattr(lib(), "orc_ft_" + sfx)(ctypes.c_int(out.shape[-1]), _p(out))
    return out


def ift(f):
    rt, sfx, ct = _rt(f.dtype.type)
    out = _cfield(f, ct).copy()
    getattr(lib(), "orc_ift_" + sfx)(ctypes.c_int(out.shape[-1]), _p(out))
    return out


def project(F, amp):
    rt, sfx, ct = _rt(F.dtype.type)
    out = _cfield(F, ct).copy()
    a = np.ascontiguousarray(amp, dtype=rt)
    getattr(lib(), "orc_project_" + sfx)(ctypes.c_int(out.shape[-1]), _p(out), _p(a))
    return out


ALG = {"ER": 0, "HIO": 1, "RAAR": 2}


def apply_support(alg, rho, S, buf=None, beta=0.9, reality=0.0):
    """One real-space update; returns (rho, buf)."""
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct).copy()
    b = _cfield(buf if buf is not None else rho, ct).copy()
    s = np.ascontiguousarray(S, dtype=np.int32)
    fn = getattr(lib(), "orc_apply_support_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_float if rt == np.float32 else ctypes.c_double, ctypes.c_float]
    fn(r.shape[-1], ALG[alg], _p(r), _p(s), _p(b), beta, reality)
    return r, b


def iterate(alg, rho, S, amp, niter, beta=0.9, reality=0.0, nthreads=0):
    """GPU_context::HIO/RAAR/ER on one field (n,n) or a population (P,n,n)."""
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct).copy()
    n = r.shape[-1]
    P = 1 if r.ndim == 2 else r.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(S, r.shape), dtype=np.int32)
    a = np.ascontiguousarray(amp, dtype=rt)
    fn = getattr(lib(), "orc_iterate_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int, ctypes.c_int,
                   ctypes.c_float if rt == np.float32 else ctypes.c_double, ctypes.c_float, ctypes.c_int]
    fn(n, P, _p(r), _p(s), _p(a), ALG[alg], int(niter), beta, reality, int(nthreads))
    return r


def fft2_fast(f, sign):
    """The throughput port's unnormalised 2-D DFT (fast_port.hpp), FP32."""
    out = _cfield(f, np.complex64).copy()
    if lib().orc_fft2_fast(ctypes.c_int(out.shape[-1]), _p(out), ctypes.c_int(sign)) != 0:
        raise ValueError("fast port: lengths 2^a 3^b only")
    return out


def apply_support_fast(alg, rho, S, buf=None, beta=0.9, reality=0.0):
    """The throughput port's real-space update, FP32; returns (rho, buf)."""
    r = _cfield(rho, np.complex64).copy()
    b = _cfield(buf if buf is not None else rho, np.complex64).copy()
    s = np.ascontiguousarray(S, dtype=np.int32)
    fn = lib().orc_apply_support_fast
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_float, ctypes.c_float]
    fn(r.shape[-1], ALG[alg], _p(r), _p(s), _p(b), beta, reality)
    return r, b


def iterate_fast(alg, rho, S, amp, niter, beta=0.9, reality=0.0, nthreads=0):
    """The throughput port of iterate() in FP32 (fast_port.hpp), for the CPU
    baseline: same driver and per-pixel operators, a vectorised transform."""
    r = _cfield(rho, np.complex64).copy()
    n = r.shape[-1]
    P = 1 if r.ndim == 2 else r.shape[0]
    s = np.ascontiguousarray(np.broadcast_to(S, r.shape), dtype=np.int32)
    a = np.ascontiguousarray(amp, dtype=np.float32)
    fn = lib().orc_iterate_fast
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                   ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    if fn(n, P, _p(r), _p(s), _p(a), ALG[alg], int(niter), beta, reality, int(nthreads)) != 0:
        raise ValueError("fast port: lengths 2^a 3^b only")
    return r


def blur(rho, amp, sigma):
    """Shrink-wrap blurred modulus field b (complex) of SW/SWA."""
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    out = np.empty_like(r)
    a = np.ascontiguousarray(amp, dtype=rt)
    fn = getattr(lib(), "orc_blur_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
    fn(r.shape[-1], _p(r), _p(a), sigma, _p(out))
    return out


def sw(rho, S, amp, sigma, threshold):
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    s = np.ascontiguousarray(S, dtype=np.int32).copy()
    a = np.ascontiguousarray(amp, dtype=rt)
    fn = getattr(lib(), "orc_sw_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float]
    fn(r.shape[-1], _p(r), _p(s), _p(a), sigma, threshold)
    return s


def swa(rho, S, amp, sigma, pixels):
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    s = np.ascontiguousarray(S, dtype=np.int32).copy()
    a = np.ascontiguousarray(amp, dtype=rt)
    fn = getattr(lib(), "orc_swa_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_int]
    fn(r.shape[-1], _p(r), _p(s), _p(a), sigma, int(pixels))
    return s


def get_error(rho, S, amp, reality=0.0):
    """Returns (error_fourier, error_real, complexsum, modulesum)."""
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    s = np.ascontiguousarray(S, dtype=np.int32)
    a = np.ascontiguousarray(amp, dtype=rt)
    out = np.zeros(4, dtype=np.float64)
    fn = getattr(lib(), "orc_get_error_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p]
    fn(r.shape[-1], _p(r), _p(s), _p(a), reality, _p(out))
    return tuple(float(v) for v in out)


def get_symmetry(rho, amp, symmin=0.0, symmax=1.0):
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    a = np.ascontiguousarray(amp, dtype=rt)
    out = np.zeros(1, dtype=np.float64)
    fn = getattr(lib(), "orc_get_symmetry_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float, ctypes.c_void_p]
    fn(r.shape[-1], _p(r), _p(a), symmin, symmax, _p(out))
    return float(out[0])


def get_shift(rho, ref, amp, minr=0.0, maxr=2.0, unit_scale_ft=True):
    """Returns (sx, sy, phase, flip)."""
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct)
    f = _cfield(ref, ct)
    a = np.ascontiguousarray(amp, dtype=rt)
    oi = np.zeros(3, dtype=np.int32)
    of = np.zeros(1, dtype=np.float32)
    fn = getattr(lib(), "orc_get_shift_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float, ctypes.c_float,
                   ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    fn(r.shape[-1], _p(r), _p(f), _p(a), minr, maxr, int(bool(unit_scale_ft)), _p(oi), _p(of))
    return int(oi[0]), int(oi[1]), float(of[0]), bool(oi[2])


def reshift(rho, S, sx, sy, phase, flip):
    rt, sfx, ct = _rt(rho.dtype.type)
    r = _cfield(rho, ct).copy()
    s = np.ascontiguousarray(S, dtype=np.int32).copy()
    fn = getattr(lib(), "orc_reshift_" + sfx)
    fn.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_float,
                   ctypes.c_int]
    fn(r.shape[-1], _p(r), _p(s), int(sx), int(sy), float(phase), int(bool(flip)))
    return r, s


def cross_desc_arrays(desc):
    """desc: dict with crosssizex, crosssizey, squaresizex, squaresizey, offsetx, offsety,
    crossft, crosssym, rotangle, crossdiff, use_average."""
    di = np.array([desc["crosssizex"], desc["crosssizey"], desc["squaresizex"], desc["squaresizey"],
                   desc["offsetx"], desc["offsety"], int(desc["crossft"]), int(desc["crosssym"])], dtype=np.int32)
    df = np.array([desc["rotangle"], desc["crossdiff"], desc["use_average"]], dtype=np.float32)
    return di, df


def cross_cells(n, desc):
    di, df = cross_desc_arrays(desc)
    cells = np.zeros((n, n), dtype=np.int32)
    lib().orc_cross_cells(ctypes.c_int(n), _p(di), _p(df), _p(cells))
    return cells


def cross_child(amp, amp_avg, parents, avg, crossmap, desc, minr=0.0, maxr=2.0, reality=0.0, update_support=True):
    """parents: [(rho0,S0),(rho1,S1),(rho2,S2),(rho3,S3)]; returns (child_rho, child_S)."""
    rt, sfx, ct = _rt(parents[0][0].dtype.type)
    r0 = _cfield(parents[0][0], ct).copy()
    s0 = np.ascontiguousarray(parents[0][1], dtype=np.int32).copy()
    others = []
    for rho, S in parents[1:]:
        others += [_cfield(rho, ct), np.ascontiguousarray(S, dtype=np.int32)]
    av = _cfield(avg, ct)
    a = np.ascontiguousarray(amp, dtype=rt)
    aa = np.ascontiguousarray(amp_avg, dtype=rt)
    cm = np.ascontiguousarray(crossmap, dtype=np.int32)
    di, df = cross_desc_arrays(desc)
    fn = getattr(lib(), "orc_cross_child_" + sfx)
    fn.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 14 + [ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                             ctypes.c_int]
    fn(r0.shape[-1], _p(a), _p(aa), _p(r0), _p(s0), *[_p(o) for o in others], _p(av), _p(cm), _p(di), _p(df),
       minr, maxr, reality, int(bool(update_support)))
    return r0, s0


def prepare_pattern(intensity, center=(-1, -1), use_bounds=False):
    """Solver::set_boundaries + prepare: centred intensity -> corner-origin amplitude (float32)."""
    I = np.ascontiguousarray(intensity, dtype=np.float32)
    n = I.shape[0]
    cx, cy = center
    if cx < 0:
        cx = n // 2
    if cy < 0:
        cy = n // 2
    out = np.zeros_like(I)
    lib().orc_prepare_pattern(_p(I), ctypes.c_int(n), ctypes.c_int(cx), ctypes.c_int(cy),
                              ctypes.c_int(int(bool(use_bounds))), _p(out))
    return out


def rng_uniform(seed, count):
    out = np.zeros(count, dtype=np.float32)
    lib().orc_rng_uniform(ctypes.c_int(seed), ctypes.c_int(count), _p(out))
    return out


def crossover_draws(seed, nchild, P, sorted_idx, suppdim, n, crossexp=1.0, crossftin=1.0, crossbalance=0.6,
                    stream=True):
    """Random draws of Solver::Crossover (src/solver.cpp:742-801). stream=True:
    nchild consecutive children from one context seeded `seed`; False: child i
    from its own context seeded (i+1)*seed. Returns a list of dicts (p1, p2, p3,
    crossft, crosssizex/y, squaresizex/y, offsetx/y, rotangle, crossmap)."""
    si = np.ascontiguousarray(sorted_idx, dtype=np.int32)
    di = np.zeros(11 * nchild, np.int32)
    df = np.zeros(nchild, np.float32)
    cap = 4 * (n + 64) * (n + 64)
    cm = np.zeros(cap, np.int32)
    cl = np.zeros(nchild, np.int32)
    args = [int(P), _p(si), ctypes.c_float(suppdim), int(n), ctypes.c_float(crossexp), ctypes.c_float(crossftin),
            ctypes.c_float(crossbalance), _p(di), _p(df), _p(cm), _p(cl), int(cap)]
    if stream:
        rc = lib().orc_crossover_draws_stream(int(nchild), int(seed), *args)
    else:
        seeds = np.array([(i + 1) * seed for i in range(nchild)], np.int32)
        rc = lib().orc_crossover_draws(int(nchild), _p(seeds), *args)
    if rc != 0:
        raise RuntimeError("crossmap buffer too small")
    out, off = [], 0
    for c in range(nchild):
        o = di[11 * c:11 * c + 11]
        out.append(dict(p1=int(o[0]), p2=int(o[1]), p3=int(o[2]), crosssizex=int(o[3]), crosssizey=int(o[4]),
                        squaresizex=int(o[5]), squaresizey=int(o[6]), offsetx=int(o[7]), offsety=int(o[8]),
                        crossft=int(o[9]), rotangle=float(df[c]), crossmap=cm[off:off + cl[c]].copy()))
        off += int(cl[c])
    return out


def _init_out(n, count):
    return np.zeros((count, n, n), np.complex64), np.zeros((count, n, n), np.int32)


def init_disk(amp, seed, count, support_size, itemsize, nitems, exponent, scaling, phase):
    """Solver::Initialize_rec with a disk support (src/solver.cpp:870-957): `count`
    individuals drawn in turn from one context RNG seeded `seed`."""
    a = np.ascontiguousarray(amp, dtype=np.float32)
    n = a.shape[0]
    fp = np.array([itemsize[0], itemsize[1], exponent[0], exponent[1], scaling[0], scaling[1], phase], np.float32)
    ip = np.array(nitems, np.int32)
    rho, S = _init_out(n, count)
    lib().orc_init_disk(int(n), _p(a), int(seed), int(count), int(support_size), _p(fp), _p(ip), _p(rho), _p(S))
    return rho, S


def init_ref(ref_rho, ref_S, seed, count, itemsize, nitems, exponent, scaling, phase):
    """Initialize_rec from a reference reconstruction (src/solver.cpp:961-1058)."""
    r = np.ascontiguousarray(ref_rho, dtype=np.complex64)
    s = np.ascontiguousarray(ref_S, dtype=np.int32)
    n = r.shape[0]
    fp = np.array([itemsize[0], itemsize[1], exponent[0], exponent[1], scaling[0], scaling[1], phase], np.float32)
    ip = np.array(nitems, np.int32)
    rho, S = _init_out(n, count)
    lib().orc_init_ref(int(n), _p(r), _p(s), int(seed), int(count), _p(fp), _p(ip), _p(rho), _p(S))
    return rho, S


def init_supp(amp, support, seed, count, initfrac, scaling, phase):
    """Initialize_rec from a support with random Fourier cells (src/solver.cpp:1061-1174)."""
    a = np.ascontiguousarray(amp, dtype=np.float32)
    s = np.ascontiguousarray(support, dtype=np.int32)
    n = a.shape[0]
    fp = np.array([initfrac, scaling[0], scaling[1], phase], np.float32)
    rho, S = _init_out(n, count)
    lib().orc_init_supp(int(n), _p(a), _p(s), int(seed), int(count), _p(fp), _p(rho), _p(S))
    return rho, S
