"""Algorithm schedules of the benchmark workloads, as [(name, [params])] lists
for Solver.set_sequence / MPR.run(schedule=...).

plain_small   128^2 polygon: 10 blocks of 20 iterations (7 HIO(beta 0.9) blocks,
              one HIO(10)+ER(10) block, two ER blocks), shrink-wrap SW(sigma_k, 0.2)
              after each block with sigma_k = 3 - 1.5 k/9.
plain_512     512^2 ellipsoids: RAAR 400 (beta 0.87) + HIO 100 (beta 0.95) + ER 50 in
              25-iteration blocks, SW after each block (22 per generation,
              sigma 3 -> 1.5, threshold SW_THRESHOLD_512 = 0.03, the reference's
              default shrink-wrap threshold, spring/settings_manager.py:119;
              at 0.2 the support of this object collapses within a generation).
grad_512      the same blocks with a support-projected Adam refinement of the masked
              Poisson likelihood after each RAAR block (GRAD: 10 steps, step 3e-2
              of the RMS density on the support; of 20/10 steps, 1e-2/3e-2 and the
              Poisson/amplitude losses this reached E_F < 0.05 soonest on 4 seeds).

time_to_rf runs one memetic reconstruction of the 512^2 ellipsoid pattern with
one of these schedules and reports the wall-clock time from population
initialisation to the first generation whose best Fourier error is below the
target, with the real-space error of the best individual against the object.
"""
from __future__ import annotations

import time

import numpy as np


def _sigma(k, nblocks, start=3.0, end=1.5):
    return start - (start - end) * k / max(nblocks - 1, 1)


def plain_small(generation=0, area=None):
    blocks = [[("HIO", [20, 0.9])]] * 7 + [[("HIO", [10, 0.9]), ("ER", [10])], [("ER", [20])], [("ER", [20])]]
    seq = []
    for k, b in enumerate(blocks):
        seq += b + [("SW", [_sigma(k, len(blocks)), 0.2])]
    return seq


SW_THRESHOLD_512 = 0.03


def _blocks_512(grad, steps=10, lr=3e-2, kind=0, raar=16, hio=4, er=2, grad_all=False):
    g = [("GRAD", [steps, lr, kind])] if grad else []
    blocks = []
    for _ in range(raar):
        blocks.append([("RAAR", [25, 0.87])] + g)
    blocks += [[("HIO", [25, 0.95])] + (g if grad_all else [])] * hio
    blocks += [[("ER", [25])] + (g if grad_all else [])] * er
    return blocks


def blocks_512(raar=16, hio=4, er=2, grad_steps=0, lr=3e-2, kind=0, polish=None, grad_all=False):
    """A 512^2 schedule with `raar` RAAR, `hio` HIO and `er` ER blocks of 25
    iterations (shrink-wrap after each block, sigma 3 -> 1.5), with
    grad_steps > 0 that many support-projected Adam steps after each RAAR block,
    (grad_all: after every block), and with polish = (steps, lr, kind)
    support-projected Adam steps after the last shrink-wrap (on the support the
    generation ends with): a callable (generation, area, threshold=None) ->
    sequence."""
    def sched(generation=0, area=None, threshold=None):
        blocks = _blocks_512(grad_steps > 0, grad_steps, lr, kind, raar, hio, er, grad_all)
        th = SW_THRESHOLD_512 if threshold is None else threshold
        seq = []
        for k, b in enumerate(blocks):
            seq += b + [("SW", [_sigma(k, len(blocks)), th])]
        if polish:
            seq.append(("GRAD", [polish[0], polish[1], polish[2]]))
        return seq
    return sched


def plain_512(generation=0, area=None, threshold=None):
    blocks = _blocks_512(False)
    th = SW_THRESHOLD_512 if threshold is None else threshold
    seq = []
    for k, b in enumerate(blocks):
        seq += b + [("SW", [_sigma(k, len(blocks)), th])]
    return seq


def grad_512(generation=0, area=None, threshold=None, steps=10, lr=3e-2, kind=0):
    blocks = _blocks_512(True, steps, lr, kind)
    th = SW_THRESHOLD_512 if threshold is None else threshold
    seq = []
    for k, b in enumerate(blocks):
        seq += b + [("SW", [_sigma(k, len(blocks)), th])]
    return seq


def iterations(seq):
    """Individual-iterations of ER/HIO/RAAR in a sequence (per individual)."""
    return sum(int(p[0]) for name, p in seq if name in ("ER", "HIO", "RAAR"))


def differentiation_1024(generation=0, area=None, threshold=None, blocks=8, steps=20, lr=1e-2):
    """The low-count differentiation workload's generation: `blocks` x [RAAR 25,
    `steps` steps of support-projected Adam on the masked Poisson likelihood
    (step lr x the RMS density on the support), shrink-wrap], then ER 25 x 2."""
    th = SW_THRESHOLD_512 if threshold is None else threshold
    body = [[("RAAR", [25, 0.87]), ("GRAD", [steps, lr, 0])] for _ in range(blocks)] + [[("ER", [25])]] * 2
    seq = []
    for k, b in enumerate(body):
        seq += b + [("SW", [_sigma(k, len(body)), th])]
    return seq


SCHEDULES = {"plain": plain_512, "grad": grad_512}


def support_size(truth):
    """2 * (largest distance of an object pixel from the grid centre) + 4."""
    n = truth.shape[0]
    y, x = np.nonzero(np.abs(truth) > 0)
    return int(np.ceil(2 * np.sqrt(((x - n / 2) ** 2 + (y - n / 2) ** 2).max()) + 4))


def noise_floor(I, truth):
    """E_F of the object itself against its own noisy pattern, with the
    least-squares scale: the lowest Fourier error a reconstruction can be
    expected to reach."""
    from . import synth
    A = synth.corner_amplitude(I).astype(np.float64)
    F = np.abs(synth.ft_plus(np.asarray(truth, np.complex128)))
    k = A >= 0
    s = (A[k] * F[k]).sum() / (F[k] ** 2).sum()
    return float(np.sqrt(((A[k] - s * F[k]) ** 2).sum() / (A[k] ** 2).sum()))


def time_to_rf(seed, form="plain", target=0.05, generations=10, popsize=64, n=512, photons=1e9, threshold=None,
               schedule=None, max_batch=0):
    """One memetic run on the ellipsoid pattern of `seed`; returns the time to
    best E_F < target (None when not reached within `generations`), the
    generation reached, the final Fourier and real-space errors of the best
    individual, the best E_F after every generation and the pattern's noise
    floor. max_batch > 0 runs the sequences in groups of that many individuals
    (Solver.set_max_batch)."""
    from . import Pattern, Settings, metrics, synth
    from .mpr import MPR
    I, truth = synth.ellipsoids_pattern(seed=seed, n=n, photons=photons)
    s = Settings()
    s.set("global", "popsize", popsize).set("global", "generations", generations)
    s.set("init", "supportsize", support_size(truth)).set("IA", "reality", 0.0)
    s.set_advanced("seed", 1000 + seed)
    m = MPR(Pattern.from_intensity(I), s)
    m.max_batch = max_batch
    t0 = time.perf_counter()
    sched = SCHEDULES[form] if schedule is None else schedule
    res = m.run(target_error=target, log=False,
                schedule=lambda g, area: sched(g, area, threshold=threshold))
    wall = time.perf_counter() - t0
    h = m.history[-1]
    reached = h["best_fourier_error"] < target
    return {"seed": seed, "form": form, "max_batch": max_batch, "time_to_target": h["time"] if reached else None,
            "wall": wall, "generation": h["generation"], "fourier_error": h["best_fourier_error"],
            "real_space_error": metrics.real_space_error(res.get_density_best(crop=False), truth),
            "history": [round(x["best_fourier_error"], 5) for x in m.history],
            "noise_floor": noise_floor(I, truth),
            "iterations": iterations(sched()) * (h["generation"] + 1)}


def differentiation_run(seed, generations=2, popsize=32, n=1024, photons=1e8):
    """The memetic loop on a low-count truncated-octahedron pattern with the
    differentiation_1024 schedule; returns the best Fourier error and the
    wall time after every generation and the real-space error of the best
    individual at the end."""
    from . import Pattern, Settings, metrics, synth
    from .mpr import MPR
    I, truth = synth.octahedron_pattern(seed=seed, n=n, photons=photons)
    s = Settings()
    s.set("global", "popsize", popsize).set("global", "generations", generations)
    s.set("init", "supportsize", support_size(truth)).set("IA", "reality", 0.0)
    s.set_advanced("seed", 2000 + seed)
    m = MPR(Pattern.from_intensity(I), s)
    res = m.run(log=False, schedule=lambda g, area: differentiation_1024(g, area))
    seq = differentiation_1024()
    return {"seed": seed, "history": [round(h["best_fourier_error"], 5) for h in m.history],
            "times": [round(h["time"], 3) for h in m.history],
            "real_space_error": metrics.real_space_error(res.get_density_best(crop=False), truth),
            "noise_floor": noise_floor(I, truth),
            "iterations_per_generation": iterations(seq),
            "adam_steps_per_generation": sum(int(p[0]) for name, p in seq if name == "GRAD")}
