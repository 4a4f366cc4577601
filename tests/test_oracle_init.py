"""Pins of the oracle's population initialisers (Solver::Initialize_rec,
src/solver.cpp:870-1174) against closed forms evaluated in numpy FP64 from draws
made with numpy's MT19937 (the reference's seeding and float uniform, pinned in
test_oracle_align_cross.py::test_mt19937_stream).

Disk variant (:870-957): the support is the disk of diameter s centred at
(N/2, N/2); each item is a projected sphere 2 sqrt(sigma^2 - r^2) e^{i phase};
|rho| <- |rho|^gamma; the density is scaled so that the measured Fourier energy
equals sum A^2 times the drawn scaling.
Reference variant (:961-1058): spheres centred on support pixels picked by
rank, confined to the reference support, |sum rho| = |sum rho_ref| x scaling.
Support variant (:1061-1174): FT(S), each selected cell of the rotated grid
scaled and phased, IFT, then the energy scaling; with initfrac = 0 this is S
times the closed-form factor.
"""
import numpy as np
import pytest

from conftest import rel_l2
from helpers import F32, NumpyContextRng


def _amp(n, seed=0):
    r = np.random.default_rng(seed)
    a = r.random((n, n)).astype(np.float32) * 50
    a[0:3, 0:3] = -1
    a[:, 7] = -1
    return a


def _energy_factor(rho, amp):
    n = rho.shape[0]
    F = np.fft.ifft2(rho.astype(np.complex128)) * n * n
    k = amp >= 0
    return np.sqrt(np.sum(amp[k].astype(np.float64) ** 2) / np.sum(np.abs(F[k]) ** 2))


def _spheres(g, nitems_in, itemsize, exponent_in, scaling_in, phase, centre):
    """Draws of one individual in the reference's order; centre(g, sigma) -> (cx, cy)."""
    nitems = int(g.rand(nitems_in[0], nitems_in[1] + 0.99))
    expo = g.rand(*exponent_in)
    scaling = g.rand(*scaling_in)
    items = []
    for _ in range(nitems):
        s2 = g.rand(F32(itemsize[0]) ** 2, F32(itemsize[1]) ** 2)
        sigma = F32(np.sqrt(s2) / 2.0)
        cx, cy = centre(g, sigma)
        ph = g.rand(-phase * np.pi, phase * np.pi)
        items.append((float(sigma), float(cx), float(cy), float(ph)))
    return float(expo), float(scaling), items


def _sum_spheres(X, Y, items, expo, mask):
    d = np.zeros(X.shape, np.complex128)
    for sigma, cx, cy, ph in items:
        r = np.sqrt((X - cx) ** 2 + (Y - cy) ** 2)
        inside = mask & (r < sigma)
        d[inside] += 2 * np.sqrt(sigma ** 2 - r[inside] ** 2) * np.exp(1j * ph)
    m = np.abs(d)
    return np.where(m > 0, d / np.where(m > 0, m, 1) * m ** expo, 0)


@pytest.mark.parametrize("n,s,nitems,expo", [(64, 30, (1, 1), (1.0, 1.0)), (64, 31, (2, 6), (0.5, 0.7)),
                                             (128, 44, (3, 8), (0.4, 1.2))])
def test_init_disk_closed_form(orc, n, s, nitems, expo):
    amp = _amp(n)
    itemsize, scaling, phase = (s * 0.2, s * 0.9), (0.5, 1.5), 0.3
    rho, S = orc.init_disk(amp, 17, 3, s, itemsize, nitems, expo, scaling, phase)
    g = NumpyContextRng(17)
    iy, ix = np.mgrid[0:s, 0:s]
    disk = np.sqrt((ix - s / 2.0) ** 2 + (iy - s / 2.0) ** 2) <= s / 2.0
    o = n // 2 - s // 2
    for k in range(3):
        e, sc, items = _spheres(g, nitems, itemsize, expo, scaling, phase,
                                lambda g_, sg: (g_.rand(sg, F32(s) - sg), g_.rand(sg, F32(s) - sg)))
        want_S = np.zeros((n, n), np.int32)
        want_S[o:o + s, o:o + s] = disk
        assert np.array_equal(S[k], want_S)
        local = _sum_spheres(ix.astype(float), iy.astype(float), items, e, disk)
        want = np.zeros((n, n), np.complex128)
        want[o:o + s, o:o + s] = local
        want *= _energy_factor(want, amp) * sc
        # float32 sin(acos(r/sigma)) near the sphere edges and |rho|^gamma of
        # small moduli leave ~1e-5; any misplaced, missing or mis-scaled sphere is O(1)
        err = rel_l2(rho[k], want)
        print(f"n={n} s={s} individual {k}: {len(items)} spheres, rel-L2 {err:.2e}")
        assert err < 1e-4, k


def test_init_ref_closed_form(orc):
    n = 64
    r = np.random.default_rng(3)
    ref_S = (r.random((n, n)) < 0.25).astype(np.int32)
    ref_S[:, :5] = 0
    ref = (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))).astype(np.complex64) * ref_S
    itemsize, nitems, expo, scaling, phase = (4.0, 20.0), (2, 7), (0.5, 1.0), (0.5, 1.5), 0.25
    rho, S = orc.init_ref(ref, ref_S, 17, 3, itemsize, nitems, expo, scaling, phase)
    flat = ref_S.ravel()
    # the pixel of rank k: the first index with k support pixels before it
    before = np.concatenate([[0], np.cumsum(flat > 0)[:-1]])

    def centre(g_, sg):
        target = int(g_.rand(0, int(flat.sum())))
        idx = int(np.argmax(before == target))
        return idx % n, idx // n

    g = NumpyContextRng(17)
    Y, X = np.mgrid[0:n, 0:n]
    newnorm = abs(ref.astype(np.complex128).sum())
    for k in range(3):
        e, sc, items = _spheres(g, nitems, itemsize, expo, scaling, phase, centre)
        want = _sum_spheres(X.astype(float), Y.astype(float), items, e, ref_S > 0)
        want *= newnorm / abs(want.sum()) * sc
        assert np.array_equal(S[k], ref_S)
        err = rel_l2(rho[k], want)
        print(f"reference variant, individual {k}: {len(items)} spheres, rel-L2 {err:.2e}")
        assert err < 1e-4, k
        assert abs(abs(rho[k].astype(np.complex128).sum()) - newnorm * sc) <= 1e-5 * newnorm * sc


def test_init_supp_closed_forms(orc):
    n = 64
    amp = _amp(n, 1)
    y, x = np.mgrid[0:n, 0:n]
    supp = (((x - 30) / 9.0) ** 2 + ((y - 33) / 6.0) ** 2 <= 1).astype(np.int32)
    # initfrac 0: no cell is modified, the density is the support times the energy factor
    rho, S = orc.init_supp(amp, supp, 17, 2, 0.0, (0.5, 1.5), 0.4)
    g = NumpyContextRng(17)
    for k in range(2):
        nf = g.rand(0.5, 1.5)
        sqx, sqy = int(g.rand(1, F32(n) / 16.0)), int(g.rand(1, F32(n) / 16.0))
        cells = ((n + sqx - 1) // sqx + 1) * ((n + sqy - 1) // sqy + 1)
        for _ in range(3 + cells):      # offsets, rotation, one draw per cell (none selected)
            g.rand(0, 1)
        want = supp * _energy_factor(supp.astype(np.complex128), amp) * nf
        assert np.array_equal(S[k], supp)
        assert rel_l2(rho[k], want) < 2e-6
    # initfrac 1: every cell scaled by U(0.5, 2) and phased; the composition in
    # numpy (cell index from a double evaluation of the rotated grid)
    rho, S = orc.init_supp(amp, supp, 23, 1, 1.0, (0.5, 1.5), 0.4)
    g = NumpyContextRng(23)
    nf = g.rand(0.5, 1.5)
    sqx, sqy = int(g.rand(1, F32(n) / 16.0)), int(g.rand(1, F32(n) / 16.0))
    csx, csy = (n + sqx - 1) // sqx + 1, (n + sqy - 1) // sqy + 1
    offx, offy = int(g.rand(0, sqx)), int(g.rand(0, sqy))
    rot = g.rand(0, np.pi)
    sc = np.zeros(csx * csy)
    ph = np.zeros(csx * csy)
    for i in range(csx * csy):
        assert g.rand(0, 1) < 1.0
        sc[i], ph[i] = g.rand(0.5, 2), g.rand(-np.pi * 0.4, np.pi * 0.4)
    k = np.arange(n)
    kk = np.where(k >= n // 2, k - n, k)
    KX, KY = np.meshgrid(kk, kk)
    R = np.sqrt((KX * KX + KY * KY).astype(float))
    R = np.where(R >= n / 2, n // 2 - 1, R)
    a = np.arctan2(KY.astype(float), KX.astype(float)) + float(rot)
    cx = np.trunc(np.cos(a) * R).astype(int) + n // 2
    cy = np.trunc(np.sin(a) * R).astype(int) + n // 2
    ci = (cx + offx) // sqx + ((cy + offy) // sqy) * csx
    F = np.fft.ifft2(supp.astype(np.complex128)) * n * n
    F = F * sc[ci] * np.exp(1j * ph[ci])
    d = np.fft.fft2(F) / (n * n)
    want = d * _energy_factor(d, amp) * nf
    assert np.array_equal(S[0], supp)
    assert rel_l2(rho[0], want) < 1e-5
