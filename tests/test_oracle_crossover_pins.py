"""Independent pins of the oracle's crossover child and of the per-child draws.

Child (src/gpu/gpu_context.cu:1661-1694 prepare_parents, :2044-2117
get_new_individual, :1699-1796 cross_density, :1799-1853 cross_density_real,
:1856-1883 cross_support). The parents differ: equal moduli with different
phase fields whose sums have phase 0, so the oracle's alignment of every role
to parent 1 is the identity (asserted first: shift (0, 0), no flip, phase
~1e-17). The child then has the closed form, evaluated here in numpy FP64:
  rho_k' = ER(rho_k, S_k)                      k = 0..3
  crossft = 1:  R_k = FT(rho_k'), A = FT(avg),
                R_1 <- R_1 * newmod/|R_1| at A < 0 pixels (use_average > 0:
                newmod = ua (amp_avg + |A|)/2 + (1-ua)|R_1|; use_average < 0:
                |R_1| pulled by |ua| towards |A| below it or amp_avg above it),
                child = IFT(where(cell, R_1 + crossdiff (R_2 - R_3), R_0))
  crossft = 0:  child = where(cell, rho_1' + crossdiff (rho_2' - rho_3'), rho_0')
  S = [S_1 + S_2 S_3 S_0 > 0.5],  child <- ER(child, S)
with cell = crossmap[cell index] == 1 from the rotated grid of
:1717-1739 (evaluated here in numpy, double). A wrong sign of the difference
term, swapped map polarity, a dropped modulus rule or a transposed role fails.

Draws (src/solver.cpp:742-801): two consecutive children of one context seeded
17, derived here from numpy's MT19937 (init_genrand seeding) and libstdc++'s
float uniform (raw / 2^32, float32), in the reference's order: u^crossexp for
parent 1, parents 2 and 3, crossft, the two square sizes, the two offsets, the
rotation, then one draw per crossmap cell. The result is also written out as a
golden list.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import NumpyContextRng

F32 = np.float32


def _base_modulus(n, seed):
    r = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n]
    c = n / 2
    m = (((x - c + 1) / (n / 4)) ** 2 + ((y - c + 1) / (n / 5.8)) ** 2 <= 1)
    m |= (x - c > n / 8) & (x - c < n / 4.5) & (np.abs(y - c + n / 4.5) < n / 10)   # off-centre tab
    return m * (r.random((n, n)) + 0.5), m.astype(np.int32)


def _phase_variant(M, seed):
    """Same modulus, its own smooth random phase, sum rotated to phase 0."""
    g = np.random.default_rng(seed)
    z = M * np.exp(1j * g.uniform(-0.5, 0.5, M.shape))
    return z * np.exp(-1j * np.angle(z.sum()))


def _cells_numpy(n, d):
    k = np.arange(n)
    kk = np.where(k >= n // 2, k - n, k)
    X, Y = np.meshgrid(kk, kk)
    R = np.sqrt((X * X + Y * Y).astype(np.float64))
    R = np.where(R >= n / 2, n // 2 - 1, R)
    a = np.arctan2(Y.astype(np.float64), X.astype(np.float64))
    if d["crosssym"]:
        a = np.where(a < 0, a + np.pi, a)
    a = a + np.float64(F32(d["rotangle"]))
    xx = np.trunc(np.cos(a) * R).astype(int) + n // 2
    yy = np.trunc(np.sin(a) * R).astype(int) + n // 2
    return (xx + d["offsetx"]) // d["squaresizex"] + ((yy + d["offsety"]) // d["squaresizey"]) * d["crosssizex"]


def _child_numpy(n, parents, avg, amp, amp_avg, crossmap, d):
    FT = lambda f: np.fft.ifft2(f) * n * n      # noqa: E731  exp(+2 pi i k.x/N), unnormalised
    IFT = lambda f: np.fft.fft2(f) / (n * n)    # noqa: E731
    r = [rho * S for rho, S in parents]          # ER at reality 0 (no negative-real pixel occurs)
    S0, S1, S2, S3 = [S for _, S in parents]
    cell = crossmap[_cells_numpy(n, d)] == 1
    ua, cd = d["use_average"], d["crossdiff"]
    if d["crossft"]:
        R = [FT(x) for x in r]
        A = FT(avg)
        R1 = R[1].copy()
        m1 = np.abs(R1)
        neg = (amp < 0) & (m1 > 0)
        if ua > 0:
            newmod = ua * (amp_avg + np.abs(A)) / 2 + (1 - ua) * m1
            R1 = np.where(neg, R1 * newmod / np.where(m1 > 0, m1, 1), R1)
        elif ua < 0:
            loc, lo, hi = abs(ua), np.abs(A), amp_avg
            newmod = np.where(m1 < lo, lo * loc + m1 * (1 - loc), np.where(m1 > hi, hi * loc + m1 * (1 - loc), m1))
            R1 = np.where(neg, R1 * newmod / np.where(m1 > 0, m1, 1), R1)
        child = IFT(np.where(cell, R1 + cd * (R[2] - R[3]), R[0]))
    else:
        child = np.where(cell, r[1] + cd * (r[2] - r[3]), r[0])
    S = ((S1 + S2 * S3 * S0) > 0.5).astype(np.int32)
    return child * S, S, cell


@pytest.mark.parametrize("n,crossft,crosssym,use_average", [
    (32, 1, 0, 0.0), (32, 1, 0, 0.5), (32, 1, 1, -0.5), (48, 1, 1, 0.3), (32, 0, 0, 0.0), (48, 0, 1, -0.5)])
def test_child_with_distinct_parents_matches_closed_form(orc, n, crossft, crosssym, use_average):
    M, base = _base_modulus(n, 3)
    rho = [_phase_variant(M, s) for s in (11, 12, 13, 14)]
    avg = _phase_variant(M, 15)
    g = np.random.default_rng(16)
    # four different supports, each covering the object (the alignment reads moduli only)
    supp = [np.maximum(base, (g.random((n, n)) < 0.3)).astype(np.int32) for _ in range(4)]
    amp = np.abs(np.fft.ifft2(rho[0]) * n * n)
    amp[0:3, 0:3] = -1
    amp[:, n // 2 + 3] = -1                     # a masked column off the centre: A < 0 pixels
    amp_avg = np.abs(np.fft.ifft2(rho[1]) * n * n) * g.uniform(0.5, 2.0, (n, n))
    for k in (0, 2, 3):
        assert orc.get_shift(rho[k], rho[1], amp, 0.0, 2.0, unit_scale_ft=False)[:2] == (0, 0)
    sx, sy, ph, flip = orc.get_shift(avg, rho[1], amp, 0.0, 2.0, unit_scale_ft=False)
    assert (sx, sy, flip) == (0, 0, False) and abs(ph) < 1e-12
    sq = (7, 6) if crossft else (5, 4)
    d = dict(crosssizex=(n + sq[0] - 1) // sq[0] + 1, crosssizey=(n + sq[1] - 1) // sq[1] + 1, squaresizex=sq[0],
             squaresizey=sq[1], offsetx=2, offsety=3, crossft=crossft, crosssym=crosssym,
             rotangle=float(F32(0.7)), crossdiff=float(F32(0.4)), use_average=float(F32(use_average)))
    cm = (g.random(d["crosssizex"] * d["crosssizey"]) < 0.5).astype(np.int32)
    parents = list(zip(rho, supp))
    want, wS, cell = _child_numpy(n, parents, avg, amp, amp_avg, cm, d)
    assert 0.1 < cell.mean() < 0.9                               # both kinds of cells are present
    if crossft and use_average != 0:
        assert np.count_nonzero(cell & (amp < 0)) > 0            # the modulus rule is exercised
    got, gS = orc.cross_child(amp, amp_avg, parents, avg, cm, d, 0.0, 2.0, 0.0, True)
    assert np.array_equal(gS, wS)
    assert rel_l2(got, want) < 1e-12
    # the FP32 oracle within float rounding of the same closed form
    p32 = [(r.astype(np.complex64), S) for r, S in parents]
    got32, gS32 = orc.cross_child(amp.astype(F32), amp_avg.astype(F32), p32, avg.astype(np.complex64), cm, d,
                                  0.0, 2.0, 0.0, True)
    assert np.array_equal(gS32, wS)
    assert rel_l2(got32, want) < 1e-5
    # the pins are sensitive: a flipped difference sign or swapped map polarity is far off
    d2 = dict(d, crossdiff=-d["crossdiff"])
    assert rel_l2(_child_numpy(n, parents, avg, amp, amp_avg, cm, d2)[0], want) > 1e-2
    assert rel_l2(_child_numpy(n, parents, avg, amp, amp_avg, 1 - cm, d)[0], want) > 1e-2


def test_use_average_negative_branches_all_occur(orc):
    """use_average < 0 has three outcomes per pixel (pulled up to |A|, pulled
    down to amp_avg, kept); the closed-form case above covers all three."""
    n = 32
    M, base = _base_modulus(n, 3)
    rho1 = _phase_variant(M, 12)
    avg = _phase_variant(M, 15)
    g = np.random.default_rng(16)
    for _ in range(4):
        g.random((n, n))
    amp_avg = np.abs(np.fft.ifft2(rho1) * n * n) * g.uniform(0.5, 2.0, (n, n))
    m1 = np.abs(np.fft.ifft2(rho1 * base) * n * n)
    lo = np.abs(np.fft.ifft2(avg) * n * n)
    up, down = m1 < lo, (m1 >= lo) & (m1 > amp_avg)
    kept = ~up & ~down
    assert up.sum() > 0 and down.sum() > 0 and kept.sum() > 0


# ---------------------------------------------------------------------------
# Draws
# ---------------------------------------------------------------------------
def draws_numpy(g, P, sorted_idx, suppdim, n, crossexp, crossftin, crossbalance):
    """One child's draws in the order of src/solver.cpp:750-787."""
    suppdim = F32(suppdim)
    e = int(F32(F32(g.rand(0, 1)) ** F32(crossexp)) * F32(P))
    out = dict(p1=int(sorted_idx[e]))
    out["p2"] = int(g.rand(0, P - 0.0001))
    out["p3"] = int(g.rand(0, P - 0.0001))
    out["crossft"] = int(g.rand(0, 1) < F32(crossftin))
    if out["crossft"]:
        lo, hi = np.ceil(F32(F32(n) / suppdim)), np.ceil(n / 4.0)
    else:
        lo, hi = suppdim / 10.0 + 1.0, suppdim / 2.0
    out["squaresizex"] = int(g.rand(lo, hi))
    out["squaresizey"] = int(g.rand(lo, hi))
    out["crosssizex"] = (n + out["squaresizex"] - 1) // out["squaresizex"] + 1
    out["crosssizey"] = (n + out["squaresizey"] - 1) // out["squaresizey"] + 1
    out["offsetx"] = int(g.rand(0, out["squaresizex"]))
    out["offsety"] = int(g.rand(0, out["squaresizey"]))
    out["rotangle"] = float(g.rand(0, np.pi))
    out["crossmap"] = np.array([int(g.rand(0, 1) < F32(crossbalance))
                                for _ in range(out["crosssizex"] * out["crosssizey"])], np.int32)
    return out


# Two consecutive children of the context seeded 17: P = 6, population sorted
# as SORTED, suppdim 40, N = 128, crossbalance 0.6; crossexp 1 and 2. Crossmaps
# by length, count of 1 cells and first 12 entries.
SORTED = [3, 1, 4, 0, 5, 2]
GOLDEN = {
    1.0: [dict(p1=1, p2=1, p3=3, crossft=1, squaresizex=9, squaresizey=28, crosssizex=16, crosssizey=6, offsetx=0,
               offsety=22, rotangle=2.4723877906799316, cm=(96, 61, [1, 0, 1, 0, 1, 1, 0, 1, 0, 1, 1, 0])),
          dict(p1=1, p2=2, p3=5, crossft=1, squaresizex=13, squaresizey=17, crosssizex=11, crosssizey=9, offsetx=6,
               offsety=2, rotangle=2.8987843990325928, cm=(99, 59, [0, 1, 1, 1, 1, 0, 1, 1, 0, 1, 1, 1]))],
}


@pytest.mark.parametrize("crossexp", [1.0, 2.0])
@pytest.mark.parametrize("crossftin", [1.0, 0.5])
def test_two_children_draws_match_numpy_mt19937(orc, crossexp, crossftin):
    g = NumpyContextRng(17)
    want = [draws_numpy(g, 6, SORTED, 40.0, 128, crossexp, crossftin, 0.6) for _ in range(2)]
    got = orc.crossover_draws(17, 2, 6, SORTED, 40.0, 128, crossexp, crossftin, 0.6, stream=True)
    for w, o in zip(want, got):
        for k in ("p1", "p2", "p3", "crossft", "squaresizex", "squaresizey", "crosssizex", "crosssizey",
                  "offsetx", "offsety"):
            assert w[k] == o[k], (k, w[k], o[k])
        assert w["rotangle"] == o["rotangle"]
        assert np.array_equal(w["crossmap"], o["crossmap"])
    if crossexp == 1.0 and crossftin == 1.0:
        for w, gold in zip(want, GOLDEN[1.0]):
            for k, v in gold.items():
                if k == "cm":
                    cm = w["crossmap"]
                    assert (len(cm), int(cm.sum()), cm[:12].tolist()) == v
                else:
                    assert w[k] == v, (k, w[k], v)


def test_float_cell_map_differs_in_few_pixels(orc):
    """The oracle and the GPU evaluate the rotated cell grid in double; the
    reference evaluates it in float with fast-math intrinsics. A float32
    evaluation (numpy) of 64 children's maps at 512^2 (the draws of a
    population of 64, suppdim 120) puts 30 of 16.8 M pixels (1.8e-6) in another
    cell, 14 of them with another crossover choice (0.2 pixels per child): those
    whose rotated coordinate lies within float rounding of a cell edge. The
    count is printed and bounded here."""
    n = 512
    draws = orc.crossover_draws(17, 64, 64, list(range(64)), 120.0, n, 1.0, 1.0, 0.6, stream=False)
    k = np.arange(n)
    kk = np.where(k >= n // 2, k - n, k)
    X, Y = np.meshgrid(kk, kk)
    R = np.sqrt((X * X + Y * Y).astype(F32))
    R = np.where(R >= F32(n / 2), F32(n // 2 - 1), R).astype(F32)
    diff_cells = diff_sel = 0
    for d in draws:
        d.update(crosssym=0, crossdiff=0.0, use_average=0.0)
        cells64 = orc.cross_cells(n, d)
        a = (np.arctan2(Y.astype(F32), X.astype(F32)) + F32(d["rotangle"])).astype(F32)
        xx = np.trunc((np.cos(a) * R).astype(F32)).astype(int) + n // 2
        yy = np.trunc((np.sin(a) * R).astype(F32)).astype(int) + n // 2
        cells32 = (xx + d["offsetx"]) // d["squaresizex"] + ((yy + d["offsety"]) // d["squaresizey"]) * d["crosssizex"]
        diff_cells += int(np.count_nonzero(cells32 != cells64))
        cm = d["crossmap"]
        diff_sel += int(np.count_nonzero(cm[cells32] != cm[cells64]))
    total = len(draws) * n * n
    print(f"float32 cell maps, {len(draws)} children at {n}^2: {diff_cells} of {total} pixels in another cell, "
          f"{diff_sel} with another crossover choice")
    assert diff_cells < 1e-5 * total
