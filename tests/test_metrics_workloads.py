"""Quality metric (FP64) and workload schedules."""
import numpy as np

from s_spi_spring_amd import metrics, synth, workloads


def test_real_space_error_invariances():
    I, t = synth.polygon_pattern(seed=3, n=64)
    t = t.astype(np.complex128)
    assert metrics.real_space_error(t, t) == 0.0
    assert metrics.real_space_error(np.roll(t, (5, -3), (0, 1)) * np.exp(0.7j), t) < 1e-12
    twin = np.conj(t[::-1, ::-1])
    assert metrics.real_space_error(np.roll(twin, (2, 7), (0, 1)), t) < 1e-12
    noisy = t + 0.05 * np.random.default_rng(0).standard_normal(t.shape)
    e = metrics.real_space_error(noisy, t)
    assert 0 < e < 0.5


def test_fourier_error_closed_form():
    I, t = synth.polygon_pattern(seed=4, n=64)
    F = np.abs(np.fft.ifft2(t.astype(np.complex128)) * 64 * 64)
    assert metrics.fourier_error(t, F) < 1e-12
    assert abs(metrics.fourier_error(t, 2 * F) - 0.5) < 1e-12


def test_schedules():
    s = workloads.plain_small()
    assert workloads.iterations(s) == 200 and [x[0] for x in s].count("SW") == 10
    assert s[1] == ("SW", [3.0, 0.2]) and s[-1] == ("SW", [1.5, 0.2])
    p = workloads.plain_512()
    assert workloads.iterations(p) == 550 and [x[0] for x in p].count("SW") == 22
    g = workloads.grad_512()
    assert workloads.iterations(g) == 550 and [x[0] for x in g].count("GRAD") == 16


def test_differentiation_schedule():
    """The low-count differentiation workload's generation: 8 x [RAAR 25, GRAD 20
    (step 1e-2, Poisson), SW] + 2 x [ER 25, SW]."""
    from s_spi_spring_amd import workloads
    seq = workloads.differentiation_1024()
    names = [name for name, _ in seq]
    assert workloads.iterations(seq) == 250
    assert names.count("GRAD") == 8 and names.count("SW") == 10
    assert all(p == [20, 1e-2, 0] for name, p in seq if name == "GRAD")
    assert names[-4:] == ["ER", "SW", "ER", "SW"]
    sig = [p[0] for name, p in seq if name == "SW"]
    assert sig[0] == 3.0 and sig[-1] == 1.5 and all(a >= b for a, b in zip(sig, sig[1:]))
