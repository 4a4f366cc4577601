"""Pins the oracle's alignment, symmetry, crossover and host-side restatements.

Alignment (src/gpu/gpu_context.cu:1226-1371): known cyclic shifts, point flips
and global phases are recovered exactly. Symmetry (:1378-1463): exactly 1 for a
real centrosymmetric object, below 1 otherwise. Crossover
(src/gpu/gpu_context.cu:1699-1883, src/solver.cpp:716-817): cell map against an
independent numpy evaluation, cross_support truth table, the child against a
numpy composition. RNG (src/context.h:13-18): std::mt19937's raw stream against
numpy's MT19937 with the reference (init_genrand) seeding, and the float
uniform derived from it. Pattern preparation (src/solver.cpp:173-289): closed
forms of the radial-profile bound on a ring of constant intensity.
"""
import numpy as np
import pytest

from conftest import rel_l2


def blob(n, seed, cx=None, cy=None):
    r = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n]
    cx = n / 2 if cx is None else cx
    cy = n / 2 if cy is None else cy
    m = ((x - cx) / 9.0) ** 2 + ((y - cy) / 6.0) ** 2 <= 1
    m |= (x - cx - 6 > 0) & (x - cx - 6 < 5) & (np.abs(y - cy + 7) < 3)   # asymmetric tab
    rho = m * (r.random((n, n)) + 0.5) * np.exp(1j * r.uniform(-0.2, 0.2, (n, n)))
    return rho, m.astype(np.int32)


@pytest.mark.parametrize("dx,dy", [(0, 0), (5, -3), (-7, 11), (13, 2)])
def test_shift_recovered(orc, dx, dy):
    n = 64
    amp = np.ones((n, n))
    ref, Sref = blob(n, 1)
    phi = 0.7
    moved = np.roll(ref, (dy, dx), axis=(0, 1)) * np.exp(1j * phi)
    Smoved = np.roll(Sref, (dy, dx), axis=(0, 1))
    sx, sy, ph, flip = orc.get_shift(moved, ref, amp, 0.0, 2.0, unit_scale_ft=True)
    assert not flip
    assert (sx % n, sy % n) == ((-dx) % n, (-dy) % n)
    assert abs(ph + phi) < 1e-5
    back, Sback = orc.reshift(moved, Smoved, sx, sy, ph, flip)
    assert np.array_equal(Sback, Sref)
    assert rel_l2(back, ref) < 1e-6


def test_flip_recovered(orc):
    n = 64
    amp = np.ones((n, n))
    ref, Sref = blob(n, 2)
    # twin: point flip p -> N^2-1-p with conjugation, then a cyclic shift
    twin = np.conj(ref.ravel()[::-1].reshape(n, n))
    Stwin = Sref.ravel()[::-1].reshape(n, n)
    twin = np.roll(twin, (4, -6), axis=(0, 1))
    Stwin = np.roll(Stwin, (4, -6), axis=(0, 1))
    sx, sy, ph, flip = orc.get_shift(twin, ref, amp, 0.0, 2.0, unit_scale_ft=True)
    assert flip
    back, Sback = orc.reshift(twin, Stwin, sx, sy, ph, flip)
    # The twin correlation peak is the shift for a reflection through the origin
    # (x -> -x), while the shift kernel flips with p -> N^2-1-p (x -> N-1-x), so a
    # twin aligned the reference's way ends one pixel off on both axes
    # (src/gpu/gpu_context.cu:1316-1320, 1344).
    assert np.array_equal(Sback, np.roll(Sref, (-1, -1), axis=(0, 1)))
    assert rel_l2(back, np.roll(ref, (-1, -1), axis=(0, 1))) < 1e-6


def test_reshift_matches_numpy(orc):
    n = 16
    r = np.random.default_rng(3)
    rho = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    S = (r.random((n, n)) < 0.5).astype(np.int32)
    for sx, sy, ph, flip in [(3, -2, 0.4, False), (-5, 7, -1.1, True), (15, -15, 0.0, True)]:
        got, gS = orc.reshift(rho, S, sx, sy, ph, flip)
        want = np.zeros_like(rho)
        wS = np.zeros_like(S)
        for y in range(n):
            for x in range(n):
                nx, ny = (x + sx) % n, (y + sy) % n
                i = nx + n * ny
                if flip:
                    i = n * n - 1 - i
                z = rho[y, x]
                # the shift phase is a float parameter (FLOAT shiftphase)
                want.flat[i] = abs(z) * np.exp(1j * ((-1 if flip else 1) * np.angle(z) + np.float32(ph)))
                wS.flat[i] = S[y, x]
        assert np.array_equal(gS, wS)
        assert rel_l2(got, want) < 1e-12


def test_symmetry(orc):
    n = 64
    amp = np.ones((n, n))
    y, x = np.mgrid[0:n, 0:n]
    sym = (((x - 32) / 10.0) ** 2 + ((y - 32) / 5.0) ** 2 <= 1).astype(np.complex128)
    assert abs(orc.get_symmetry(sym, amp) - 1.0) < 1e-9
    asym, _ = blob(n, 4)
    s = orc.get_symmetry(asym, amp)
    assert 0 < s < 0.99


def test_cross_cells_numpy(orc):
    n = 64
    for crosssym in (0, 1):
        desc = dict(crosssizex=9, crosssizey=11, squaresizex=8, squaresizey=7, offsetx=3, offsety=5,
                    crossft=1, crosssym=crosssym, rotangle=np.float32(1.2345), crossdiff=0.4, use_average=0.0)
        cells = orc.cross_cells(n, desc)
        k = np.arange(n)
        kk = np.where(k >= n // 2, k - n, k)
        X, Y = np.meshgrid(kk, kk)
        R = np.sqrt((X * X + Y * Y).astype(np.float64))
        R = np.where(R >= n / 2, n // 2 - 1, R)
        a = np.arctan2(Y.astype(np.float64), X.astype(np.float64))
        if crosssym:
            a = np.where(a < 0, a + np.pi, a)
        a = a + np.float64(desc["rotangle"])
        xx = np.trunc(np.cos(a) * R).astype(int) + n // 2
        yy = np.trunc(np.sin(a) * R).astype(int) + n // 2
        want = (xx + 3) // 8 + ((yy + 5) // 7) * 9
        assert np.array_equal(cells, want)
        assert cells.min() >= 0 and cells.max() < 9 * 11


def test_cross_support_truth_table(orc):
    # child support S = [S1 + S2*S3*S0 > 0.5]
    n = 4
    combos = np.array([[a, b, c, d] for a in (0, 1) for b in (0, 1) for c in (0, 1) for d in (0, 1)])
    S0 = combos[:, 0].reshape(n, n).astype(np.int32)
    S1 = combos[:, 1].reshape(n, n).astype(np.int32)
    S2 = combos[:, 2].reshape(n, n).astype(np.int32)
    S3 = combos[:, 3].reshape(n, n).astype(np.int32)
    want = (S1 | (S2 & S3 & S0)).astype(np.int32)
    rho = np.ones((n, n), np.complex128)
    amp = np.ones((n, n))
    desc = dict(crosssizex=2, crosssizey=2, squaresizex=n, squaresizey=n, offsetx=0, offsety=0,
                crossft=0, crosssym=0, rotangle=0.0, crossdiff=0.0, use_average=0.0)
    cm = np.zeros(4, np.int32)
    # identical zero-shift parents: alignment is the identity, so the support is the truth table
    _, S = orc.cross_child(amp, amp, [(rho, S0), (rho, S1), (rho, S2), (rho, S3)], rho, cm, desc)
    assert np.array_equal(S, want)


def test_cross_child_matches_numpy_composition(orc):
    """With parents already aligned (copies of one density), the child is the
    per-cell mix rho1' + F (rho2 - rho3) of the supported spectra, inverse
    transformed and projected on the child support."""
    n = 32
    rho, S = blob(n, 5)
    amp = np.abs(np.fft.ifft2(rho) * n * n)
    amp[0:2, 0:2] = -1
    desc = dict(crosssizex=6, crosssizey=6, squaresizex=7, squaresizey=6, offsetx=2, offsety=1,
                crossft=1, crosssym=0, rotangle=np.float32(0.3), crossdiff=0.4, use_average=0.0)
    cm = (np.random.default_rng(6).random(36) < 0.5).astype(np.int32)
    r2 = rho * 1.0
    r3 = rho * 1.0
    child, cS = orc.cross_child(amp, amp, [(rho, S), (rho, S), (r2, S), (r3, S)], rho, cm, desc)
    # all four parents are identical: rho2 - rho3 = 0 and rho1 = rho0, so the child is ER(rho0)
    ts = S * (np.abs(np.angle(rho)) <= np.pi)
    assert np.array_equal(cS, S)
    assert rel_l2(child, rho * ts) < 1e-12


def test_mt19937_stream(orc):
    bg = np.random.MT19937(0)
    for seed in (17, 42, 5489, 34):
        bg._legacy_seeding(seed)
        raw = bg.random_raw(64).astype(np.uint64)
        want = (raw.astype(np.float32) / np.float32(4294967296.0)).astype(np.float32)
        want = np.where(want >= 1, np.nextafter(np.float32(1), np.float32(0)), want)
        got = orc.rng_uniform(seed, 64)
        assert np.array_equal(got, want)


def test_prepare_pattern_closed_forms(orc):
    n = 64
    r = np.random.default_rng(7)
    I = r.integers(0, 100, (n, n)).astype(np.float32)
    I[30:34, 30:34] = -1
    I[5, 5] = -2
    A = orc.prepare_pattern(I, use_bounds=False)
    want = np.fft.ifftshift(np.where(I >= 0, np.sqrt(np.maximum(I, 0)), -1)).astype(np.float32)
    assert np.array_equal(A, want)
    # with bounds: a masked pixel on a ring of constant intensity v gets A = -2 - sqrt(v)
    y, x = np.mgrid[0:n, 0:n]
    d = np.sqrt((x - 32.0) ** 2 + (y - 32.0) ** 2)
    I = np.full((n, n), 9.0, np.float32)
    I[(d > 14) & (d < 18)] = 25.0
    I[32, 48] = -1     # on the ring, radius 16
    I[50, 32] = -2     # saturated -> free (-1)
    A = orc.prepare_pattern(I, use_bounds=True)
    Ac = np.fft.fftshift(A)
    assert np.isclose(Ac[32, 48], -2 - 5.0)
    assert Ac[50, 32] == -1.0
    assert np.isclose(Ac[10, 10], 3.0)
