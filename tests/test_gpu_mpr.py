"""End-to-end memetic loop on the GPU (small workload: 128^2 polygon pattern,
population 8, the plain_small schedule, one crossover generation), and
seed reproducibility of the whole loop."""
import numpy as np
import pytest

from s_spi_spring_amd import Pattern, Settings, synth, workloads

pytestmark = pytest.mark.gpu


def run_small(seed, generations=1):
    from s_spi_spring_amd.mpr import MPR
    I, truth = synth.polygon_pattern(seed=seed, n=128)
    s = Settings()
    s.set("global", "popsize", 8).set("global", "generations", generations).set("global", "threads", 8)
    s.set("init", "supportsize", 60).set("IA", "reality", 0.0)
    s.set_advanced("seed", 42)
    m = MPR(Pattern.from_intensity(I), s)
    res = m.run(schedule=workloads.plain_small, log=False)
    return m, res, truth


def test_small_memetic_run():
    m, res, truth = run_small(0)
    h = m.history
    assert len(h) == 2 and all(np.isfinite(x["best_fourier_error"]) for x in h)
    assert h[-1]["best_fourier_error"] <= h[0]["best_fourier_error"] * 1.0001
    assert h[-1]["best_fourier_error"] < 0.5
    assert res.get_support_best(crop=False).sum() > 0
    assert res.get_density().shape[0] <= 128 and res.step == 1 and res.total == 1
    st = res.get_stats()
    assert len(st["best"]["error"]) == 2 and len(st["all"]["error"][0]) == 8


def test_memetic_run_is_reproducible():
    m1, r1, _ = run_small(1)
    m2, r2, _ = run_small(1)
    e1 = [x["best_fourier_error"] for x in m1.history]
    e2 = [x["best_fourier_error"] for x in m2.history]
    assert e1 == e2
    assert np.array_equal(r1.get_support_best(crop=False), r2.get_support_best(crop=False))
