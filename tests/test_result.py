"""Result (the object MPR.run returns) and Stats.print, on synthetic data:
cropping, the reconstructed pattern and residuals against the closed form, the
PRTF of an exact reconstruction, alignment of shifted / twinned copies."""
import numpy as np

from s_spi_spring_amd import Result, Settings, Stats, synth


def _result(n=64, seed=0):
    I, truth = synth.polygon_pattern(seed=seed, n=n, photons=1e9)
    S = (np.abs(truth) > 0).astype(np.int32)
    # the pattern this object gives in the orientation of Result (|FFT(conj rho)|^2, centred)
    pat = np.abs(np.fft.fftshift(np.fft.fft2(np.conj(truth.astype(np.complex128))))) ** 2
    r = Result()
    st = Stats()
    r.set(pat, truth, S, truth, S, [3, 7], st, Settings(), {"run": 1}, step=2, total=5)
    return r, truth, S, pat


def test_crop_and_accessors():
    r, truth, S, pat = _result()
    c = r.get_density()
    ys, xs = np.nonzero(S)
    assert c.shape[0] >= ys.max() - ys.min() + 1 and c.shape[1] >= xs.max() - xs.min() + 1
    assert c.shape[0] < S.shape[0] and c.shape[1] < S.shape[1]
    assert np.isclose(np.abs(c).sum(), np.abs(truth).sum())          # the crop keeps the whole object
    assert np.array_equal(r.get_support_best(crop=False), S)
    assert r.get_filename([3, 7], "x") == "3_7_x_rec.h5"


def test_rec_pattern_residuals_and_prtf():
    r, truth, S, pat = _result()
    assert np.allclose(r.get_rec_pattern(), pat)
    assert np.abs(r.get_residuals()).max() <= 1e-6 * pat.max()
    q, prtf = r.get_prtf()
    assert len(q) > 10 and np.allclose(prtf, 1.0)                      # an exact reconstruction


def test_reshift_recovers_translation_phase_and_twin():
    r, truth, S, pat = _result(seed=1)
    t = truth.astype(np.complex128)
    moved = np.roll(t, (9, -4), (0, 1)) * np.exp(0.8j)
    assert np.abs(r.reshift(moved, crop=False) - t).max() < 1e-12 * np.abs(t).max() * t.size
    twin = np.conj(np.roll(t[::-1, ::-1], (2, 5), (0, 1)))
    assert np.abs(r.reshift(twin, crop=False) - t).max() < 1e-12 * np.abs(t).max() * t.size


def test_stats_print_and_save_without_h5py(capsys):
    st = Stats()

    class Rec:
        def __init__(self, e):
            self.e = e

        def get_error(self):
            return self.e

        def get_reality(self):
            return 0.1

        def get_symmetry(self):
            return 0.2

        def get_module_sum(self):
            return 3.0

        def get_complex_sum(self):
            return 1.0

        def get_support(self):
            return np.ones((4, 4))

    recs = [Rec(0.5), Rec(0.25)]
    st.add(recs, recs[1], recs[0], 1, 0, 0.0)
    st.print()
    out = capsys.readouterr().out
    assert "OS: 1.00" in out and "Best" in out and "Err. 2.500e-01" in out
    r, *_ = _result()
    try:
        import h5py  # noqa: F401
    except ImportError:
        r.save("/nonexistent-dir")
        assert "Failed to save file" in capsys.readouterr().out
