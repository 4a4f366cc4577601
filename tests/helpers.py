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
