"""Shared inputs of the parity tests: realistic patterns with a starting population."""
import numpy as np

from s_spi_spring_amd import synth


def case(n, P, seed=0):
    """(amp, rho, supp): corner-origin amplitude of a synthetic pattern of size n,
    P random starting densities and their (disk) supports."""
    if n <= 128:
        I, truth = synth.polygon_pattern(seed=seed, n=n)
    elif n <= 512:
        I, truth = synth.ellipsoids_pattern(seed=seed, n=n)
    else:
        I, truth = synth.octahedron_pattern(seed=seed, n=n)
    amp = synth.corner_amplitude(I)
    supp = synth.disk_support(n, max(8, n // 3))
    rho = synth.random_start(n, P, supp, seed=seed + 1, scale=float(np.sqrt(np.maximum(I, 0).sum()) / supp.sum()))
    return amp, rho, np.broadcast_to(supp, rho.shape).astype(np.int32).copy()


F32 = np.float32


class NumpyContextRng:
    """std::mt19937(seed) + uniform_real_distribution<float>(0, 1) as libstdc++
    computes it: float(raw) / 2^32 in float32, 1 mapped to the float below 1;
    rand(a, b) = u * (b - a) + a in float32 (src/context.h:13-18)."""

    def __init__(self, seed):
        self.bg = np.random.MT19937(0)
        self.bg._legacy_seeding(seed)

    def rand(self, a=0.0, b=1.0):
        u = F32(self.bg.random_raw()) / F32(4294967296.0)
        if u >= 1:
            u = np.nextafter(F32(1), F32(0))
        return F32(F32(u * F32(F32(b) - F32(a))) + F32(a))
