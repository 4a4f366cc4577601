"""Seeded synthetic diffraction patterns for the benchmark and test workloads.

Each generator returns (intensity, rho_true): the centred intensity in the
reference's input encoding (>= 0 measured photon counts, -1 masked;
spring/pattern_manager.py:94-96, src/solver.cpp:194-201) and the complex
ground-truth object it was made from. Intensities follow the reference's "FT"
convention (exp(+2 pi i k.x/N), src/gpu/gpu_context.cu:748-754):
I = Poisson(photons * |FT rho|^2 / sum |FT rho|^2), centred with fftshift.

Workloads:
  polygon_pattern      128^2 convex polygon, weak random phase, 1e7 photons, 5 px beamstop
  ellipsoids_pattern   512^2 overlapping projected ellipsoids, 1e9 photons, 12 px beamstop,
                       two 4 px detector-gap strips, 0.5 % random saturated pixels masked
  octahedron_pattern   1024^2 truncated-octahedron silhouette with rough boundary
  batch_patterns       many polygon patterns of sizes drawn from {256, 384, 512, 768, 1024},
                       objects scaled by n/128, photons log-uniform in [1e7, 1e10]
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage


def ft_plus(rho):
    """Unnormalised DFT with kernel exp(+2 pi i k.x / N)."""
    n0, n1 = rho.shape[-2:]
    return np.fft.ifft2(rho) * (n0 * n1)


def _intensity(rho, photons, rng):
    F = ft_plus(rho.astype(np.complex128))
    p = np.abs(F) ** 2
    lam = photons * p / p.sum()
    I = rng.poisson(lam).astype(np.float32)
    return np.fft.fftshift(I)


def _beamstop(n, radius):
    y, x = np.mgrid[0:n, 0:n]
    c = n // 2
    return (x - c) ** 2 + (y - c) ** 2 <= radius ** 2


def polygon_pattern(seed=0, n=128, photons=1e7, beamstop=5, scale=1.0):
    """Uniform-density convex polygon (6-12 vertices, radius 15-25 px times
    `scale`) with a weak random phase (uniform +-0.3 rad), centred in an n^2
    grid; the beamstop radius is scaled likewise."""
    rng = np.random.default_rng(seed)
    nv = int(rng.integers(6, 13))
    ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
    rad = rng.uniform(15, 25, nv) * scale
    beamstop = int(round(beamstop * scale))
    vx = n / 2 + rad * np.cos(ang)
    vy = n / 2 + rad * np.sin(ang)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    inside = np.ones((n, n), bool)
    # convex hull of the vertices (sorted by angle around the centre) -> half-plane tests
    from scipy.spatial import ConvexHull
    pts = np.stack([vx, vy], 1)
    hull = ConvexHull(pts)
    for a, b, c in hull.equations:
        inside &= (a * x + b * y + c) <= 0
    phase = rng.uniform(-0.3, 0.3, (n, n))
    rho = inside * np.exp(1j * phase)
    I = _intensity(rho, photons, rng)
    I[_beamstop(n, beamstop)] = -1
    return I, rho.astype(np.complex64)


def ellipsoids_pattern(seed=0, n=512, photons=1e9, beamstop=12, gap=4, sat_frac=0.005):
    """1-5 overlapping projected ellipsoids (semi-axes 20-60 px) with amplitude in
    [0.5, 1] on their union and a smooth phase within +-0.5 rad."""
    rng = np.random.default_rng(seed)
    k = int(rng.integers(1, 6))
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    thick = np.zeros((n, n))
    phase = np.zeros((n, n))
    for _ in range(k):
        a, b = rng.uniform(20, 60, 2)
        rc, tc = 80 * np.sqrt(rng.random()), rng.uniform(0, 2 * np.pi)
        cx, cy = n / 2 + rc * np.cos(tc), n / 2 + rc * np.sin(tc)
        th = rng.uniform(0, np.pi)
        xr = (x - cx) * np.cos(th) + (y - cy) * np.sin(th)
        yr = -(x - cx) * np.sin(th) + (y - cy) * np.cos(th)
        q = 1 - (xr / a) ** 2 - (yr / b) ** 2
        t = np.sqrt(np.clip(q, 0, None))
        thick += t
        gx, gy, g0 = rng.uniform(-1, 1, 3)
        ramp = (gx * xr / a + gy * yr / b) * 0.25 + g0 * 0.25
        phase += np.where(q > 0, ramp, 0)
    union = thick > 0
    amp = np.zeros((n, n))
    if union.any():
        tv = thick[union]
        amp[union] = 0.5 + 0.5 * (tv - tv.min()) / max(tv.max() - tv.min(), 1e-12)
    phase = np.clip(phase, -0.5, 0.5)
    rho = amp * np.exp(1j * phase)
    I = _intensity(rho, photons, rng)
    mask = _beamstop(n, beamstop)
    c = n // 2
    mask[c - gap // 2:c + (gap - gap // 2), :] = True
    mask[:, c - gap // 2:c + (gap - gap // 2)] = True
    sat = rng.random((n, n)) < sat_frac
    I[mask | sat] = -1
    return I, rho.astype(np.complex64)


def octahedron_pattern(seed=0, n=1024, photons=1e10, beamstop=12, roughness=2.0):
    """Truncated-octahedron silhouette (octagon-like projection, radius 80-150 px)
    with its boundary roughened by Gaussian-smoothed noise of sigma 2 px."""
    rng = np.random.default_rng(seed)
    R = rng.uniform(80, 150)
    rot = rng.uniform(0, np.pi / 4)
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    xc, yc = x - n / 2, y - n / 2
    ang = np.arctan2(yc, xc) - rot
    r = np.hypot(xc, yc)
    # regular octagon with truncated corners: distance to the nearest edge normal
    sector = np.pi / 4
    a = np.mod(ang + sector / 2, sector) - sector / 2
    edge = R * np.cos(sector / 2) / np.cos(a)
    noise = ndimage.gaussian_filter(rng.standard_normal((n, n)), roughness) * roughness * 4
    inside = r <= edge + noise
    thick = np.where(inside, np.sqrt(np.clip(1 - (r / (R * 1.2)) ** 2, 0.05, None)), 0)
    rho = thick * np.exp(1j * rng.uniform(-0.1, 0.1) * inside)
    I = _intensity(rho, photons, rng)
    I[_beamstop(n, beamstop)] = -1
    return I, rho.astype(np.complex64)


def corner_amplitude(intensity):
    """Reference preparation without bounds (src/solver.cpp:173-206 with
    use_bounds = 0): sqrt of the counts, -1 for masked pixels, origin moved
    from the centre (N/2, N/2) to the corner."""
    I = np.asarray(intensity, dtype=np.float32)
    A = np.where(I >= 0, np.sqrt(np.maximum(I, 0)), -1.0).astype(np.float32)
    return np.fft.ifftshift(A)


def disk_support(n, diameter):
    """Initial support of Solver::Initialize_rec: a disk of `diameter` px centred at (N/2, N/2)."""
    y, x = np.mgrid[0:n, 0:n]
    s = diameter
    c = n // 2 - s // 2
    ix, iy = x - c, y - c
    r = np.sqrt((ix - s / 2.0) ** 2 + (iy - s / 2.0) ** 2)
    return ((ix >= 0) & (ix < s) & (iy >= 0) & (iy < s) & (r <= s / 2.0)).astype(np.int32)


def random_start(n, P, support, seed=0, scale=1.0):
    """P random complex starting densities confined to `support`."""
    rng = np.random.default_rng(seed)
    rho = (rng.random((P, n, n)) * np.exp(1j * rng.uniform(-0.2, 0.2, (P, n, n)))) * support
    return (rho * scale).astype(np.complex64)


BATCH_SIZES = (256, 384, 512, 768, 1024)


def batch_sizes(count, seed=0):
    """Sizes of `count` batch patterns, drawn uniformly from BATCH_SIZES."""
    rng = np.random.default_rng(seed)
    return [int(BATCH_SIZES[i]) for i in rng.integers(0, len(BATCH_SIZES), count)]


def batch_patterns(count, seed=0):
    """`count` independent polygon patterns for a mixed-size batch: sizes from
    batch_sizes(count, seed), objects as polygon_pattern scaled by n/128,
    photons log-uniform in [1e7, 1e10]. Yields (intensity, rho_true) lazily."""
    rng = np.random.default_rng(seed + 1)
    photons = 10.0 ** rng.uniform(7, 10, count)
    for i, n in enumerate(batch_sizes(count, seed)):
        yield polygon_pattern(seed=seed * 100_003 + i, n=n, photons=float(photons[i]), scale=n / 128)
