"""Reconstruction quality against a known object (FP64, host).

real_space_error(rho, truth): relative L2 distance min over the ambiguities of
phase retrieval from a Fourier modulus — cyclic translation, global phase and
the twin (complex conjugate point reflection) — of ||c rho' - truth|| / ||truth||
with rho' the shifted (and possibly twinned) reconstruction (the alignment
spring/result.py:748-790 performs for saved reconstructions). The translation
is the peak of the FFT cross-correlation of the complex fields; both twin
choices are tried. c is the least-squares complex factor <rho', truth> /
<rho', rho'>: its phase is the global phase, and its modulus absorbs the
overall scale, which a photon-count pattern fixes only up to the unknown
total (scale=False keeps |c| = 1).
"""
from __future__ import annotations

import numpy as np


def _best_shift(a, b):
    """Integer cyclic shift s maximising |<roll(a, s), b>|."""
    c = np.fft.ifft2(np.fft.fft2(b) * np.conj(np.fft.fft2(a)))
    i = int(np.argmax(np.abs(c)))
    return np.unravel_index(i, a.shape)


def _aligned_error(a, truth, scale):
    s = _best_shift(a, truth)
    a = np.roll(a, s, axis=(0, 1))
    ip = np.vdot(a, truth)
    if abs(ip) == 0:
        c = 1.0
    else:
        c = ip / np.vdot(a, a).real if scale else ip / abs(ip)
    return float(np.linalg.norm(c * a - truth) / np.linalg.norm(truth))


def real_space_error(rho, truth, scale=True):
    a = np.asarray(rho, dtype=np.complex128)
    t = np.asarray(truth, dtype=np.complex128)
    twin = np.conj(np.roll(a[::-1, ::-1], (1, 1), axis=(0, 1)))
    return min(_aligned_error(a, t, scale), _aligned_error(twin, t, scale))


def fourier_error(rho, amp):
    """E_F of get_error for a density already projected on its support:
    sqrt(sum_{A>=0} (A - |FT rho|)^2 / sum_{A>=0} A^2), FT = exp(+2 pi i k.x/N)."""
    n = rho.shape[-1]
    F = np.abs(np.fft.ifft2(np.asarray(rho, np.complex128)) * n * n)
    A = np.asarray(amp, np.float64)
    k = A >= 0
    return float(np.sqrt(np.sum((A[k] - F[k]) ** 2) / np.sum(A[k] ** 2)))
