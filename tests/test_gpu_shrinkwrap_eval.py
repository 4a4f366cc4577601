"""Parity of shrink-wrap (SW/SWA), fitness (get_error), symmetry and the op-list
runner with the CPU oracle, through the C ABI.

Supports are integer masks decided by FP32 comparisons on a blurred modulus.
The bar is bit-exact masks on these fixed seeds. The GPU and the oracle compute
that modulus with different FFTs, so a pixel whose blurred modulus lies within
float rounding of the threshold (relative 1e-5, the reported "rounding band")
could in principle be decided differently; on these inputs no pixel is (the
counts are printed). Scalar metrics: relative error <= 1e-5 against the FP64
oracle.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import case

pytestmark = pytest.mark.gpu


def blurred_modulus64(orc, rho, amp, sigma):
    return np.abs(orc.blur(rho.astype(np.complex128), amp.astype(np.float64), sigma))


def check_mask(got, want, b, thr, tag):
    ambiguous = np.abs(b - thr) <= 1e-5 * thr
    diff = (got != want)
    bad = diff & ~ambiguous
    print(f"{tag}: {diff.sum()} differing pixels, {ambiguous.sum()} in the rounding band, {bad.sum()} outside it")
    assert diff.sum() == 0


@pytest.mark.parametrize("n,P", [(128, 6), (384, 2), (512, 3)])
@pytest.mark.parametrize("sigma,thr", [(3.0, 0.2), (1.5, 0.07)])
def test_sw(hip, orc, n, P, sigma, thr):
    amp, rho, supp = case(n, P, seed=11)
    be = hip.Backend(amp, capacity=P)
    be.set_population(rho, supp)
    be.sw(sigma, thr)
    got_rho, got_S = be.get_population()
    assert np.array_equal(got_rho, rho)          # SW restores the density
    for i in range(P):
        b = blurred_modulus64(orc, rho[i], amp, sigma)
        want = orc.sw(rho[i].astype(np.complex128), supp[i], amp.astype(np.float64), sigma, thr)
        check_mask(got_S[i], want, b, thr * b.max(), f"SW n={n} i={i}")


@pytest.mark.parametrize("n,P", [(128, 4), (512, 2), (768, 1)])
@pytest.mark.parametrize("pixels", [0, 1, 500, 4000, 10 ** 7])
def test_swa(hip, orc, n, P, pixels):
    amp, rho, supp = case(n, P, seed=12)
    be = hip.Backend(amp, capacity=P)
    be.set_population(rho, supp)
    be.swa(2.0, pixels)
    got_rho, got_S = be.get_population()
    assert np.array_equal(got_rho, rho)
    for i in range(P):
        if pixels <= 0:
            assert got_S[i].sum() == 0
            continue
        k = min(pixels, n * n)
        assert got_S[i].sum() >= k
        b = blurred_modulus64(orc, rho[i], amp, 2.0)
        thr = np.sort(b.ravel())[n * n - k]
        want = orc.swa(rho[i].astype(np.complex128), supp[i], amp.astype(np.float64), 2.0, pixels)
        check_mask(got_S[i], want, b, thr, f"SWA n={n} k={k} i={i}")


@pytest.mark.parametrize("n,P", [(128, 5), (384, 2), (512, 3), (768, 2)])
@pytest.mark.parametrize("reality", [0.0, 0.5])
def test_evaluate(hip, orc, n, P, reality):
    amp, rho, supp = case(n, P, seed=13)
    # move towards a consistent state so the errors are not trivial
    rho = orc.iterate("ER", rho, supp, amp, 3, 0.0, reality)
    be = hip.Backend(amp, capacity=P)
    be.set_population(rho, supp)
    ev = be.evaluate(reality)
    for i in range(P):
        ef, er_, cs, ms = orc.get_error(rho[i].astype(np.complex128), supp[i], amp.astype(np.float64), reality)
        sym = orc.get_symmetry(rho[i].astype(np.complex128), amp.astype(np.float64), 0.0, 1.0)
        for name, want in (("error_fourier", ef), ("error_real", er_), ("complexsum", cs), ("modulesum", ms),
                           ("symmetry", sym)):
            got = float(ev[name][i])
            assert abs(got - want) <= 1e-5 * abs(want) + 1e-12, (name, i, got, want)


def test_symmetry_window(hip, orc):
    amp, rho, supp = case(128, 3, seed=14)
    be = hip.Backend(amp, capacity=3)
    be.set_population(rho, supp)
    got = be.symmetry(0.1, 0.8)
    for i in range(3):
        want = orc.get_symmetry(rho[i].astype(np.complex128), amp.astype(np.float64), 0.1, 0.8)
        assert abs(got[i] - want) <= 1e-5 * want


def test_run_sequence_matches_individual_calls(hip, orc):
    amp, rho, supp = case(128, 4, seed=15)
    seq = [("HIO", [5, 0.9]), ("SW", [2.0, 0.1]), ("ER", [3]), ("SWA", [2.0, 900]), ("RAAR", [2, 0.87])]
    a = hip.Backend(amp, capacity=4)
    a.set_population(rho, supp)
    a.run(seq, nrep=2, reality=0.0)
    b = hip.Backend(amp, capacity=4)
    b.set_population(rho, supp)
    for _ in range(2):
        b.iterate("HIO", 5, 0.9)
        b.sw(2.0, 0.1)
        b.iterate("ER", 3)
        b.swa(2.0, 900)
        b.iterate("RAAR", 2, 0.87)
    ra, sa = a.get_population()
    rb, sb = b.get_population()
    assert np.array_equal(ra, rb) and np.array_equal(sa, sb)
