"""On-chip iteration blocks (spr_set_fused) for 64^2 and 128^2 fields.

One workgroup per individual runs a whole ER/HIO/RAAR block with the field in
LDS. Bar: one iteration from an identical state within relative L2 1e-5 of the
FP32 and FP64 oracles, supports unchanged; over many iterations as close to
the FP64 oracle as the FP32 oracle is (factor 4), as for the two-pass form.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import case
from s_spi_spring_amd import hipcore

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.mark.parametrize("n,P", [(64, 3), (128, 4)])
@pytest.mark.parametrize("alg,beta", [("HIO", 0.9), ("RAAR", 0.87), ("ER", 0.0)])
@pytest.mark.parametrize("reality", [0.0, 0.5])
def test_fused_one_iteration(hip, orc, n, P, alg, beta, reality):
    print(f"fused one iteration: n={n} P={P} {alg} reality={reality}", flush=True)
    amp, rho, supp = case(n, P, seed=31)
    be = hip.Backend(amp, capacity=P)
    be.set_fused(1)
    be.set_population(rho, supp)
    assert be.fused() == (1, True)
    be.iterate(alg, 1, beta, reality)
    got, s = be.get_population()
    want32 = orc.iterate(alg, rho, supp, amp, 1, beta, reality)
    want64 = orc.iterate(alg, rho.astype(np.complex128), supp, amp.astype(np.float64), 1, beta, reality)
    assert np.array_equal(s, supp)
    e32, e64 = rel_l2(got, want32), rel_l2(got, want64)
    print(f"  rel-L2 vs fp32 oracle {e32:.2e}, vs fp64 {e64:.2e}")
    assert e32 < TOL and e64 < TOL
    be.close()


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("alg,beta", [("HIO", 0.9), ("RAAR", 0.87), ("ER", 0.0)])
def test_fused_zero_iterations_is_one_projection(hip, orc, n, alg, beta):
    amp, rho, supp = case(n, 2, seed=32)
    be = hip.Backend(amp, capacity=2)
    be.set_fused(1)
    be.set_population(rho, supp)
    be.iterate(alg, 0, beta, 0.0)
    got = be.get_population(support=False)
    want = np.stack([orc.ift(orc.project(orc.ft(r.astype(np.complex128)), amp)) for r in rho])
    assert rel_l2(got, want) < TOL
    be.close()


@pytest.mark.parametrize("alg,beta", [("HIO", 0.9), ("RAAR", 0.87), ("ER", 0.0)])
def test_fused_twenty_iterations(hip, orc, alg, beta):
    n, P = 128, 3
    amp, rho, supp = case(n, P, seed=33)
    be = hip.Backend(amp, capacity=P)
    be.set_fused(1)
    be.set_population(rho, supp)
    be.iterate(alg, 20, beta, 0.0)
    got = be.get_population(support=False)
    want = orc.iterate(alg, rho.astype(np.complex128), supp, amp.astype(np.float64), 20, beta, 0.0)
    ref32 = orc.iterate(alg, rho, supp, amp, 20, beta, 0.0)
    e, e32 = rel_l2(got, want), rel_l2(ref32, want)
    print(f"fused {alg} x20 n={n}: rel-L2 vs fp64 oracle: HIP {e:.2e}, fp32 oracle {e32:.2e}")
    assert e < 4 * max(e32, 1e-6)
    if alg == "ER":
        assert e < 1e-4
    be.close()


def test_fused_agrees_with_two_pass(hip):
    """30 ER iterations (contractive, so rounding differences do not grow)
    through both forms agree within rounding."""
    n, P = 128, 5
    amp, rho, supp = case(n, P, seed=34)
    out = []
    for mode in (1, 0):
        be = hip.Backend(amp, capacity=P)
        be.set_fused(mode)
        be.set_population(rho, supp)
        assert be.fused()[1] == (mode == 1)
        be.iterate("ER", 30, 0.0, 0.0)
        out.append(be.get_population(support=False))
        be.close()
    e = rel_l2(out[0], out[1])
    print(f"ER x30 on chip vs two-pass: rel-L2 {e:.2e}")
    assert e < 1e-5


def test_fused_mode_selection_and_errors(hip):
    amp, rho, supp = case(128, 8, seed=35)
    be = hip.Backend(amp, capacity=8)
    be.set_population(rho[:7], supp[:7])
    assert be.fused() == (-1, False)       # 7 individuals: two-pass form
    be.set_population(rho, supp)
    assert be.fused() == (-1, True)        # 8 individuals: on chip
    be.set_fused(0)
    assert be.fused() == (0, False)
    with pytest.raises(hipcore.SpringHipError):
        be.set_fused(2)
    be.close()
    amp2, rho2, supp2 = case(256, 1, seed=35)
    be2 = hip.Backend(amp2, capacity=1)
    with pytest.raises(hipcore.SpringHipError):
        be2.set_fused(1)
    assert be2.fused() == (-1, False)
    be2.close()


def test_fused_timing_kind(hip):
    amp, rho, supp = case(128, 2, seed=36)
    be = hip.Backend(amp, capacity=2)
    be.set_fused(1)
    be.set_population(rho, supp)
    be.timing(True)
    be.iterate("RAAR", 10, 0.87, 0.0)
    be.iterate("HIO", 5, 0.9, 0.0)
    ms, cnt = be.timing_read(hip.KIND_BLOCK)
    assert cnt == 2 and ms > 0
    assert be.timing_read(hip.KIND_ROW_UPDATE)[1] == 0
    be.timing(False)
    be.close()


def test_fused_in_a_batch_reads_each_pattern(hip, orc):
    """A batch of 128^2 patterns whose size class holds 96 individuals runs its
    blocks on chip; every individual reads its own pattern's amplitude (through
    the slot -> pattern map)."""
    cs = [case(128, 32, seed=40 + k) for k in range(3)]
    b = hipcore.Batch([c[0] for c in cs], [32, 32, 32])
    for k, (amp, rho, supp) in enumerate(cs):
        b.set_population(k, rho, supp)
    b.run([("RAAR", [1, 0.87])])
    for k, (amp, rho, supp) in enumerate(cs):
        got = b.get_population(k, support=False)
        sel = slice(0, 32, 8)
        want = orc.iterate("RAAR", rho[sel].astype(np.complex128), supp[sel], amp.astype(np.float64), 1, 0.87, 0.0)
        e = rel_l2(got[sel], want)
        print(f"batch pattern {k}: rel-L2 vs fp64 oracle {e:.2e}")
        assert e < TOL
    b.close()
