"""The CPU baseline's throughput port of the FP32 oracle (oracle/fast_port.hpp).

bench.py times this port, not the product, as its CPU baseline, so it has to
compute what the oracle computes. Pinned here:
  * its 2-D transform against numpy FP64 (both signs, every factor it supports:
    radix 8, 4, 2 and 3 stages, and lengths not a multiple of its 16-line blocks);
  * its support update bit for bit against the oracle's apply_support_* (ER, HIO,
    RAAR; reality 0 and 0.3), on fields seeded with the values where the reality
    window needs atan2: negative real parts with zero, tiny or signed-zero
    imaginary parts;
  * its driver against the FP32 oracle's iterate: one iteration within relative
    L2 1e-5, and twenty (where HIO and RAAR amplify rounding differences) as
    close to the FP64 oracle as the FP32 oracle is, within a factor of 4, as the
    GPU path is held in tests/test_gpu_iterate.py.
"""
import numpy as np
import pytest

from helpers import case


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n", [8, 24, 64, 96, 128, 384, 512])
def test_transform_against_numpy(orc, n):
    rng = np.random.default_rng(n)
    f = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    f64 = f.astype(np.complex128)
    for sign, want in ((+1, np.fft.ifft2(f64) * n * n), (-1, np.fft.fft2(f64))):
        got = orc.fft2_fast(f, sign)
        e = rel(got, want)
        print(f"n={n} sign={sign:+d}: rel-L2 vs numpy FP64 {e:.2e}")
        assert e < 1e-6


def test_transform_rejects_other_factors(orc):
    with pytest.raises(ValueError):
        orc.fft2_fast(np.zeros((10, 10), np.complex64), 1)


def _edge_field(rng, n):
    rho = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    flat = rho.reshape(-1)
    tiny = np.array([0.0, -0.0, 1e-9, -1e-9, 3e-8, -3e-8, 1e-7, 1e-6, 2e-6], np.float32)
    k = 9 * 8
    re = -np.abs(flat[:k].real).astype(np.float32)
    flat[:k] = re + 1j * (np.tile(tiny, 8) * np.abs(re))
    flat[k:k + 4] = [complex(-0.0, 0.0), complex(-0.0, -0.0), complex(0.0, -0.0), complex(-1.0, -0.0)]
    return rho


@pytest.mark.parametrize("alg", ["ER", "HIO", "RAAR"])
@pytest.mark.parametrize("reality", [0.0, 0.3])
def test_update_is_bit_identical(orc, alg, reality):
    rng = np.random.default_rng(5)
    n = 64
    rho = _edge_field(rng, n)
    buf = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    if alg == "RAAR":   # put the edge values into q = 2 rho - buf
        buf.reshape(-1)[:80] = (2 * rho.reshape(-1)[:80] - _edge_field(rng, n).reshape(-1)[:80])
    S = (rng.random((n, n)) < 0.6).astype(np.int32)
    want_r, want_b = orc.apply_support(alg, rho, S, buf, 0.87, reality)
    got_r, got_b = orc.apply_support_fast(alg, rho, S, buf, 0.87, reality)
    assert np.array_equal(got_r.view(np.uint32), want_r.view(np.uint32))
    if alg != "ER":
        assert np.array_equal(got_b.view(np.uint32), want_b.view(np.uint32))


@pytest.mark.parametrize("n", [128, 384])
@pytest.mark.parametrize("alg,beta", [("ER", 0.0), ("HIO", 0.9), ("RAAR", 0.87)])
def test_one_iteration_matches_the_oracle(orc, n, alg, beta):
    amp, rho, supp = case(n, 2, seed=3)
    want = orc.iterate(alg, rho, supp, amp, 1, beta, 0.0)
    got = orc.iterate_fast(alg, rho, supp, amp, 1, beta, 0.0)
    e = rel(got, want)
    print(f"n={n} {alg} x1: rel-L2 vs FP32 oracle {e:.2e}")
    assert e < 1e-5


@pytest.mark.parametrize("alg,beta", [("ER", 0.0), ("HIO", 0.9), ("RAAR", 0.87)])
def test_twenty_iterations_track_fp64_like_the_oracle(orc, alg, beta):
    amp, rho, supp = case(128, 2, seed=7)
    want = orc.iterate(alg, rho.astype(np.complex128), supp, amp.astype(np.float64), 20, beta, 0.0)
    ref32 = orc.iterate(alg, rho, supp, amp, 20, beta, 0.0)
    got = orc.iterate_fast(alg, rho, supp, amp, 20, beta, 0.0)
    e, e32 = rel(got, want), rel(ref32, want)
    print(f"{alg} x20 n=128: rel-L2 vs FP64 oracle: port {e:.2e}, FP32 oracle {e32:.2e}")
    assert e < 4 * max(e32, 1e-6)
    if alg == "ER":
        assert e < 1e-4


def test_population_and_threads(orc):
    """Several individuals on several threads equal one at a time."""
    amp, rho, supp = case(128, 4, seed=9)
    many = orc.iterate_fast("RAAR", rho, supp, amp, 3, 0.87, 0.0, nthreads=3)
    for i in range(4):
        one = orc.iterate_fast("RAAR", rho[i], supp[i], amp, 3, 0.87, 0.0, nthreads=1)
        assert np.array_equal(many[i], one)
