"""Pins the oracle's FFT (oracle/spring_oracle.hpp: fft_line/fft2_raw/ft/ift).

Checks: numpy FP64 transforms (a third-party implementation), closed forms
(delta -> constant, shifted delta -> plane wave, plane wave -> delta), Parseval,
linearity and the round trip, for power-of-two, radix-3/5 and prime sizes.
Conventions (src/gpu/gpu_context.cu:748-774): FT = N^2 * numpy.ifft2, IFT = numpy.fft2 / N^2.
"""
import numpy as np
import pytest

from conftest import rel_l2

SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 24, 30, 32, 48, 64, 96, 128, 11 * 4, 384]


def rnd(n, seed=0, dtype=np.complex128):
    r = np.random.default_rng(seed)
    return (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))).astype(dtype)


@pytest.mark.parametrize("n", SIZES)
def test_ft_matches_numpy_fp64(orc, n):
    x = rnd(n)
    assert rel_l2(orc.ft(x), np.fft.ifft2(x) * n * n) < 1e-13
    assert rel_l2(orc.ift(x), np.fft.fft2(x) / (n * n)) < 1e-13


@pytest.mark.parametrize("n", [8, 64, 128, 384, 512])
def test_ft_fp32_accuracy(orc, n):
    x = rnd(n, 1)
    got = orc.ft(x.astype(np.complex64))
    assert got.dtype == np.complex64
    assert rel_l2(got, np.fft.ifft2(x) * n * n) < 2e-6 * np.log2(n)


@pytest.mark.parametrize("n", [16, 24, 64])
def test_closed_forms(orc, n):
    k = np.arange(n)
    # delta at the origin -> all ones
    d = np.zeros((n, n), np.complex128)
    d[0, 0] = 1
    assert np.allclose(orc.ft(d), 1.0, atol=1e-13)
    # delta at (x0, y0) -> exp(+2 pi i (kx x0 + ky y0) / N)
    x0, y0 = 3, 5
    d = np.zeros((n, n), np.complex128)
    d[y0, x0] = 1
    want = np.exp(2j * np.pi * (k[None, :] * x0 + k[:, None] * y0) / n)
    assert np.allclose(orc.ft(d), want, atol=1e-12)
    # plane wave exp(-2 pi i (k0 . x)/N) -> N^2 delta(k0) under FT
    kx0, ky0 = 2, n - 3
    w = np.exp(-2j * np.pi * (kx0 * k[None, :] + ky0 * k[:, None]) / n)
    F = orc.ft(w)
    want = np.zeros((n, n))
    want[ky0, kx0] = n * n
    assert np.allclose(F, want, atol=1e-9 * n * n)
    # IFT of the constant one field is a delta of height 1 at the origin
    I = orc.ift(np.ones((n, n), np.complex128))
    want = np.zeros((n, n))
    want[0, 0] = 1
    assert np.allclose(I, want, atol=1e-13)


@pytest.mark.parametrize("n", [12, 32, 48])
def test_parseval_linearity_roundtrip(orc, n):
    x, y = rnd(n, 2), rnd(n, 3)
    X = orc.ft(x)
    assert np.isclose(np.sum(np.abs(X) ** 2), n * n * np.sum(np.abs(x) ** 2), rtol=1e-12)
    a, b = 0.3 - 1.1j, 2.0 + 0.5j
    assert rel_l2(orc.ft(a * x + b * y), a * X + b * orc.ft(y)) < 1e-13
    assert rel_l2(orc.ift(orc.ft(x)), x) < 1e-13
    # raw transforms of both signs
    assert rel_l2(orc.fft2(x, -1), np.fft.fft2(x)) < 1e-13
    assert rel_l2(orc.fft2(x, +1), np.fft.ifft2(x) * n * n) < 1e-13
