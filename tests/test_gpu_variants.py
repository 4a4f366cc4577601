"""Kernel variants of the iteration passes and the autotuner, through the C ABI.

Every variant must meet the single-step bar against the FP64 oracle (relative
L2 <= 1e-5, supports unchanged). Variants that differ only in the tile width,
the row-pass threads or the prefetch order run the same operations; the
compiler may contract multiply-adds differently in each, so one step of two
such variants must agree within relative L2 2e-6 (a few float roundings). A row
variant that recomputes the HIO/RAAR state from its row transform instead of
storing it must agree with its stored-state twin to the same bar. The
autotuner must leave the resident population unchanged and keep the fastest
pair it measured.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import case
from s_spi_spring_amd import hipcore

pytestmark = pytest.mark.gpu

SIZES = [256, 384, 512, 768, 1024, 2048]
# (size, kind, variant, variant with the same radix plan)
SAME_PLAN = [(256, "col", 1, 0), (256, "col", 2, 0), (384, "col", 1, 0), (384, "col", 2, 0), (384, "col", 3, 0),
             (512, "col", 1, 0), (512, "col", 3, 0), (768, "col", 1, 0), (768, "col", 2, 0), (1024, "col", 2, 1),
             (1024, "col", 3, 1), (1024, "col", 4, 1), (256, "row", 1, 0), (384, "row", 1, 0), (512, "row", 1, 0),
             (768, "row", 1, 0), (768, "row", 2, 0), (1024, "row", 2, 1), (1024, "row", 3, 1),
             (384, "col", 4, 0), (512, "col", 4, 0), (512, "col", 5, 0), (768, "col", 3, 0), (768, "col", 4, 0),
             (768, "col", 5, 0), (1024, "col", 5, 1), (1024, "col", 6, 1), (1024, "col", 7, 0), (2048, "col", 1, 0),
             (1024, "row", 4, 1), (2048, "row", 2, 1), (512, "col", 7, 0), (768, "col", 7, 0)]


# (size, row variant that recomputes rho_n, its twin that stores it: same plan,
# threads and padding)
RECOMPUTED = [(256, 3, 2), (384, 3, 2), (512, 3, 0), (512, 4, 2), (768, 4, 3), (1024, 6, 5), (1024, 7, 1),
              (2048, 4, 3), (2048, 5, 2)]


def pairs(n):
    nc, nr = hipcore.variant_counts(n)
    return [(c, 0) for c in range(nc)] + [(0, r) for r in range(1, nr)]


@pytest.mark.parametrize("n", SIZES)
def test_every_variant_meets_the_oracle_bar(hip, orc, n):
    P = 2 if n <= 1024 else 1
    amp, rho, supp = case(n, P, seed=31)
    want = orc.iterate("RAAR", rho.astype(np.complex128), supp, amp.astype(np.float64), 1, 0.87, 0.0)
    for cv, rv in pairs(n):
        print(f"n={n} variant col={cv} row={rv}", flush=True)
        be = hip.Backend(amp, capacity=P)
        be.set_variants(cv, rv)
        be.set_population(rho, supp)
        be.iterate("RAAR", 1, 0.87, 0.0)
        got, s = be.get_population()
        assert np.array_equal(s, supp)
        e = rel_l2(got, want)
        print(f"   rel-L2 vs fp64 oracle {e:.2e}")
        assert e < 1e-5
        be.close()


@pytest.mark.parametrize("n,kind,v,ref", SAME_PLAN)
def test_same_plan_variants_agree(hip, n, kind, v, ref):
    amp, rho, supp = case(n, 2, seed=32)
    out = []
    for var in (v, ref):
        be = hip.Backend(amp, capacity=2)
        be.set_variants(var, 0) if kind == "col" else be.set_variants(0, var)
        be.set_population(rho, supp)
        be.iterate("HIO", 1, 0.9, 0.0)
        out.append(be.get_population(support=False))
        be.close()
    e = rel_l2(out[0], out[1])
    print(f"n={n} {kind} variant {v} vs {ref}: rel-L2 {e:.1e}")
    assert e < 2e-6


def test_row_variant_info():
    for n, rs, stored in RECOMPUTED:
        assert hipcore.row_variant_recomputes_state(n, rs)
        assert not hipcore.row_variant_recomputes_state(n, stored)
    with pytest.raises(hipcore.SpringHipError):
        hipcore.row_variant_recomputes_state(512, 99)
    with pytest.raises(hipcore.SpringHipError):
        hipcore.row_variant_recomputes_state(500, 0)


@pytest.mark.parametrize("n,rs,stored", RECOMPUTED)
@pytest.mark.parametrize("alg,beta", [("HIO", 0.9), ("RAAR", 0.87)])
def test_recomputed_state_agrees_with_the_stored_state(hip, n, rs, stored, alg, beta):
    """A row variant that recomputes rho_n from its row transform (rowIFFT of
    the rowFFT the iteration keeps, / N) instead of reading it runs one more
    transform; one iteration agrees with its stored-state twin within relative
    L2 2e-6 (float rounding), with the supports unchanged."""
    P = 2 if n <= 1024 else 1
    amp, rho, supp = case(n, P, seed=34)
    out = []
    for rv in (rs, stored):
        be = hip.Backend(amp, capacity=P)
        be.set_variants(0, rv)
        be.set_population(rho, supp)
        be.iterate(alg, 1, beta, 0.0)
        out.append(be.get_population())
        be.close()
    assert np.array_equal(out[0][1], out[1][1])
    e = rel_l2(out[0][0], out[1][0])
    print(f"n={n} {alg}: row variant {rs} (recomputed state) vs {stored}: rel-L2 {e:.1e}")
    assert e < 2e-6


def test_autotune_keeps_state_and_picks_the_fastest(hip):
    n = 384
    before = hipcore.default_variants(n)
    amp, rho, supp = case(n, 8, seed=33)
    be = hip.Backend(amp, capacity=8)
    be.set_population(rho, supp)
    cm, rm = be.autotune(reps=3)
    print("column ms", cm, "row ms", rm)
    assert np.all(cm > 0) and np.all(rm > 0)
    got, s = be.get_population()
    assert np.array_equal(got, rho) and np.array_equal(s, supp)
    pick = lambda t: int(np.nonzero(t <= t.min() * 1.015)[0][0])
    assert be.variants() == (pick(cm), pick(rm))
    assert hipcore.default_variants(n) == be.variants()
    fresh = hip.Backend(amp, capacity=2)
    assert fresh.variants() == be.variants()
    fresh.close()
    be.close()
    hipcore.set_default_variants(n, *before)     # later tests see the compiled defaults


def test_variant_argument_errors(hip):
    amp, rho, supp = case(128, 1)
    be = hip.Backend(amp, capacity=1)
    nc, nr = hipcore.variant_counts(128)
    with pytest.raises(hipcore.SpringHipError, match="out of range"):
        be.set_variants(nc, 0)
    with pytest.raises(hipcore.SpringHipError, match="resident"):
        be.autotune()          # no resident individuals yet
    be.close()
    with pytest.raises(hipcore.SpringHipError, match="out of range"):
        hipcore.set_default_variants(128, 0, nr)
