"""Batches of patterns of several sizes (spr_batch) through the C ABI.

Every individual's arithmetic in a batch is the arithmetic of the single-pattern
handle (the same kernels, reading the individual's own amplitude), and every
reduction has a fixed order, so the bar is bit-identity with one single-pattern
handle per pattern, whose parity with the oracle the other GPU tests establish.
A block replayed from its HIP graph must equal the same launches issued
directly, on every replay, with ordinary work and synchronisations between
replays.
"""
import numpy as np
import pytest

from helpers import case
from s_spi_spring_amd import hipcore

pytestmark = pytest.mark.gpu

SIZES = [128, 384, 128, 256, 768, 384]
POPS = [3, 2, 2, 2, 1, 2]
SEQ = [("HIO", [5, 0.9]), ("SW", [2.0, 0.1]), ("RAAR", [3, 0.87]), ("GRAD", [3, 1e-2]), ("SWA", [2.0, 900]),
       ("ER", [2])]


def cases():
    return [case(n, P, seed=20 + i) for i, (n, P) in enumerate(zip(SIZES, POPS))]


@pytest.mark.parametrize("chunk", [None, 1])
def test_batch_equals_single_pattern_handles(hip, chunk):
    """chunk = 1: iteration blocks in chunks of one individual, so that every
    chunk after the first reads its amplitude through an offset pattern map."""
    cs = cases()
    print("batch: sizes", SIZES, "populations", POPS, "chunk", chunk, flush=True)
    b = hipcore.Batch([c[0] for c in cs], POPS)
    if chunk is not None:
        b.set_chunk(1)
    for k, (amp, rho, supp) in enumerate(cs):
        b.set_population(k, rho, supp)
    b.run(SEQ, nrep=2, reality=0.0)
    ev = b.evaluate(0.0)
    for k, (amp, rho, supp) in enumerate(cs):
        one = hip.Backend(amp, capacity=len(rho))
        one.set_population(rho, supp)
        one.run(SEQ, nrep=2, reality=0.0)
        want_rho, want_s = one.get_population()
        got_rho, got_s = b.get_population(k)
        assert np.array_equal(got_s, want_s), k
        assert np.array_equal(got_rho, want_rho), k
        e1 = one.evaluate(0.0, symmetry=False)
        for name in ("error_fourier", "error_real", "complexsum", "modulesum"):
            assert np.array_equal(ev[name][k], e1[name]), (k, name)
        one.close()
    b.close()


@pytest.mark.parametrize("concurrent", [False, True])
def test_graph_replays_equal_direct_launches(hip, concurrent):
    cs = cases()
    a = hipcore.Batch([c[0] for c in cs], POPS)
    d = hipcore.Batch([c[0] for c in cs], POPS)
    a.concurrent(concurrent)
    d.concurrent(not concurrent)
    d.graphs(False)
    for k, (amp, rho, supp) in enumerate(cs):
        a.set_population(k, rho, supp)
        d.set_population(k, rho, supp)
    for rep in range(3):      # the first call captures, the others replay
        a.run([("RAAR", [4, 0.87])])
        d.run([("RAAR", [4, 0.87])])
        ea, ed = a.evaluate(0.0), d.evaluate(0.0)
        for k in range(len(SIZES)):
            assert np.array_equal(a.get_population(k, support=False), d.get_population(k, support=False)), (rep, k)
            assert np.array_equal(ea["error_fourier"][k], ed["error_fourier"][k]), (rep, k)
        a.run([("SW", [2.5, 0.12])])      # eager work between replays
        d.run([("SW", [2.5, 0.12])])
    a.close()
    d.close()


def test_batch_argument_errors(hip):
    amp128 = case(128, 1)[0]
    with pytest.raises(hipcore.SpringHipError, match="unsupported grid size"):
        hipcore.Batch([amp128, np.zeros((100, 100), np.float32)], [2, 2])
    with pytest.raises(hipcore.SpringHipError, match="empty population"):
        hipcore.Batch([amp128], [0])
    b = hipcore.Batch([amp128], [2])
    assert b.lib.spr_batch_get_population(b.h, 1, None, None) != 0
    assert b"out of range" in b.lib.spr_batch_last_error(b.h)
    with pytest.raises(hipcore.SpringHipError, match="negative iteration"):
        b.run([("HIO", [-1, 0.9])])
    b.close()


def test_batch_of_all_config_sizes_against_the_fp64_oracle(hip, orc):
    """A batch with every size of the 256-pattern workload (256^2, 384^2, 512^2,
    768^2, 1024^2; two individuals each) against the FP64 oracle directly: one
    RAAR iteration with its trailing projection per individual (relative L2
    <= 1e-5) and the Fourier error of the result (relative <= 1e-5)."""
    from conftest import rel_l2
    sizes, pops = [256, 384, 512, 768, 1024], [2, 2, 2, 2, 2]
    cs = [case(n, P, seed=40 + i) for i, (n, P) in enumerate(zip(sizes, pops))]
    print("batch: sizes", sizes, "populations", pops, flush=True)
    b = hipcore.Batch([c[0] for c in cs], pops)
    for k, (amp, rho, supp) in enumerate(cs):
        b.set_population(k, rho, supp)
    b.run([("RAAR", [1, 0.87])], reality=0.0)
    ev = b.evaluate(0.0)
    for k, (amp, rho, supp) in enumerate(cs):
        got = b.get_population(k, support=False)
        amp64 = amp.astype(np.float64)
        for i in range(pops[k]):
            want = orc.iterate("RAAR", rho[i].astype(np.complex128), supp[i], amp64, 1, 0.87, 0.0)
            e = rel_l2(got[i], want)
            ef = orc.get_error(want, supp[i], amp64, 0.0)[0]
            print(f"n={sizes[k]} individual {i}: RAAR step rel-L2 vs FP64 {e:.2e}; E_F {ev['error_fourier'][k][i]:.6f} "
                  f"vs {ef:.6f}", flush=True)
            assert e < 1e-5
            assert abs(ev["error_fourier"][k][i] - ef) <= 1e-5 * ef
    b.close()
