"""Device residency of individuals between Solver calls.

Backend level: slots moved with spr_gather and copied with
spr_put_individuals / spr_fetch_individuals keep every value (bit-exact).
Solver level: a Solver that keeps individuals resident across calls gives
bit-identical results to fresh Solvers (nothing resident) on copies of the
same individuals, also after the Python side changes an individual.
"""
import numpy as np
import pytest

from helpers import case
from s_spi_spring_amd import synth

pytestmark = pytest.mark.gpu


def test_slot_gather_put_fetch(hip):
    n, cap = 128, 7
    amp, rho, supp = case(n, cap, seed=21)
    supp = supp * np.arange(1, cap + 1)[:, None, None]          # any nonzero int32 counts as inside
    be = hip.Backend(amp, capacity=cap)
    slots = [3, 0, 6, 1, 5, 2, 4]
    be.put_individuals(slots, list(rho), list(supp))
    got, gs = be.fetch_individuals(slots)
    for k in range(cap):
        assert np.array_equal(got[k], rho[k]) and np.array_equal(gs[k], (supp[k] != 0).astype(np.int32))
    # slot i <- slot src[i]; -1 keeps slot i
    src = [6, 6, -1, 0, 1, 5, 3]
    before = be.fetch_individuals(list(range(cap)))
    be.gather(src)
    after = be.fetch_individuals(list(range(cap)))
    for i, s in enumerate(src):
        j = i if s < 0 else s
        assert np.array_equal(after[0][i], before[0][j]) and np.array_equal(after[1][i], before[1][j])
    # only the density of a slot
    be.put_individuals([2], [rho[0]])
    assert np.array_equal(be.fetch_individuals([2], support=False)[0], rho[0])
    assert np.array_equal(be.fetch_individuals([2])[1][0], after[1][2])


@pytest.fixture(scope="module")
def pr():
    from s_spi_spring_amd import spring_core
    return spring_core.phase_retrieval


def _pop(pr, S, P, n):
    return S.InitializePop(popsize=P, suppsize=n // 3, itemsize=[n // 3 * 0.2, n // 3 * 0.9], nitems=[2, 8],
                           exponent=[0.5, 0.5], scaling=[0.5, 1.5], phaserange=0.1, suppcut=0)


def _clone(pr, recs):
    return [pr.Reconstruction(r.get_density(), r.get_support()) for r in recs]


def _state(recs):
    return [(r.get_density(), r.get_support(), r.get_error()) for r in recs]


def _same(a, b):
    return all(np.array_equal(x[0], y[0]) and np.array_equal(x[1], y[1]) and x[2] == y[2] for x, y in zip(a, b))


def test_resident_solver_matches_fresh_solvers(pr):
    n, P = 128, 5
    I, _ = synth.polygon_pattern(seed=2, n=n)

    def solver():
        return pr.Solver(I, reality=0.0, realpen=0.0, nthreads=P, use_bounds=False, gpus=[0], seed=17)

    S = solver()
    A = _pop(pr, S, P, n)
    B = _pop(pr, S, P, n)
    A0, B0 = _clone(pr, A), _clone(pr, B)
    seq = [("HIO", [5.0, 0.9]), ("SW", [2.0, 0.1]), ("ER", [3.0])]

    # the resident Solver: interleaved populations, a single individual, reordering
    S.set_sequence(seq)
    S.Sequence(A)
    S.Sequence(B)
    S.SW(A, 2.0, 0.15)
    avg = S.GetAverage(A, 1.0)
    S.Evaluate([avg])
    A2 = [A[3], A[1], A[4], A[0], A[2]]
    S.Sequence(A2)
    A[1].set_density(A[1].get_density() * np.complex64(1.5))   # a change made on the host
    S.Sequence(A)

    # the same steps, each on a fresh Solver (nothing resident)
    def fresh(op):
        s = solver()
        s.set_sequence(seq)
        op(s)

    fresh(lambda s: s.Sequence(A0))
    fresh(lambda s: s.Sequence(B0))
    fresh(lambda s: s.SW(A0, 2.0, 0.15))
    s = solver()
    avg0 = s.GetAverage(A0, 1.0)
    s.Evaluate([avg0])
    fresh(lambda s: s.Sequence([A0[3], A0[1], A0[4], A0[0], A0[2]]))
    A0[1].set_density(A0[1].get_density() * np.complex64(1.5))
    fresh(lambda s: s.Sequence(A0))

    assert _same(_state(A), _state(A0))
    assert _same(_state(B), _state(B0))
    assert np.array_equal(avg.get_density(), avg0.get_density()) and avg.get_error() == avg0.get_error()


@pytest.mark.parametrize("k", [1, 2, 3])
def test_groups_of_k_match_whole_population(pr, k):
    """Solver.set_max_batch(k): sequences and single operations run k
    individuals per launch, each group copied in from the host and back; every
    individual's arithmetic is the same, so the results are bit-identical to
    the whole population in one launch (256^2: the two-pass kernels at every
    group size)."""
    n, P = 256, 5
    I, _ = synth.ellipsoids_pattern(seed=4, n=n)

    def solver():
        return pr.Solver(I, reality=0.0, realpen=0.0, nthreads=P, use_bounds=False, gpus=[0], seed=17)

    S = solver()
    A = _pop(pr, S, P, n)
    B = _clone(pr, A)
    seq = [("RAAR", [4.0, 0.87]), ("SW", [2.0, 0.1]), ("HIO", [3.0, 0.9]), ("ER", [2.0])]
    whole, grouped = solver(), solver()
    grouped.set_max_batch(k)
    assert grouped.get_max_batch() == k and whole.get_max_batch() == 0
    for s, pop in ((whole, A), (grouped, B)):
        s.set_sequence(seq)
        s.Sequence(pop, 2)
        s.SWA(pop, 2.0, 900)
        s.ER(pop, 2)
    assert _same(_state(A), _state(B))
