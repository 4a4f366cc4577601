"""spring_core.pattern (detector-image preprocessing, host C++) against closed
forms and the plain-Python oracle (oracle/pattern_oracle.py) on small images."""
import math

import numpy as np
import pytest

from oracle import pattern_oracle as po


@pytest.fixture(scope="module")
def pm():
    from s_spi_spring_amd import spring_core
    return spring_core.pattern


def test_oracle_closed_forms():
    # constant -> constant; a linear ramp is reproduced at the scaled positions
    c = np.full((16, 16), 2.5)
    assert np.allclose(po.downscale(c, [8, 8]), 2.5)
    y, x = np.mgrid[0:16, 0:16].astype(float)
    ramp = 3.0 * x + 0.5 * y
    o = po.downscale(ramp, [8, 8])
    Y, X = np.mgrid[0:8, 0:8].astype(float)
    inner = (slice(1, 7), slice(1, 7))
    assert np.allclose(o[inner], (3.0 * 2 * X + 0.5 * 2 * Y)[inner])
    # exact identity crop
    m = np.ones((16, 16))
    p, mm = po.cut_and_downscale(ramp, m, [8.0, 8.0], [16, 16], [16, 16])
    assert np.allclose(p, ramp) and np.allclose(mm, 1)
    # Ewald map: a far detector (R -> inf) leaves positions and intensities
    q = po.downscale_and_correct(ramp, m, [8.0, 8.0], 1e12, 1.0, [16, 16])
    assert np.allclose(q, ramp, atol=1e-6)
    # radial profile: a constant image gives the constant; a radial ramp gives
    # increasing values within one bracket of each node
    assert np.allclose(po.radial_profile(c, 1.0, [8.0, 8.0], [0, 2, 4, 6], [-180, 180]), 2.5)
    img = np.hypot(x - 8, y - 8)
    prof = po.radial_profile(img, 1.0, [8.0, 8.0], [0, 2, 4, 6], [-180, 180])
    assert np.all(np.diff(prof) > 0) and np.all(np.abs(prof - [0, 2, 4, 6]) < 2)


def test_ewald_radius_map(pm):
    """A single pixel at distance r lands at 2R sin(theta/2)/dx_new, theta = atan(r dx / R),
    with intensity 1/cos^3(theta) (the splat spreads it over <= 4 pixels)."""
    n, R, dx = 64, 40.0, 1.0
    c = [32.0, 32.0]
    r = 20
    d = np.zeros((n, n))
    d[32, 32 + r] = 1.0
    m = np.ones((n, n))
    o = pm.downscale_and_correct(d, m, c, R, dx, [n, n])
    far = 32.0
    dx_new = R * 2 * math.sin(math.atan2(dx * far, R) / 2) / far
    theta = math.atan2(r * dx, R)
    want_x = 32 + R * 2 * math.sin(theta / 2) / dx_new
    row = np.where(o[32] > 0, o[32], 0)
    got_x = np.sum(np.arange(n) * row) / np.sum(row)
    assert abs(got_x - want_x) < 0.5
    # with only that pixel valid, every output pixel it reaches holds value x 1/cos^3(theta)
    m1 = np.zeros((n, n))
    m1[32, 32 + r] = 1
    o1 = pm.downscale_and_correct(d, m1, c, R, dx, [n, n])
    hit = o1 >= 0
    assert 1 <= hit.sum() <= 4 and np.allclose(o1[hit], 1 / math.cos(theta) ** 3, rtol=1e-12)


def test_quirk_downscale_uses_x_scale_for_y(pm):
    d = np.ones((32, 64))                       # 32 rows, 64 columns
    o = pm.downscale(d, [32, 32])                # s = 64/32 = 2 for both axes
    assert np.all(o[:16] == 1) and np.all(o[17:] == -1)


@pytest.mark.parametrize("seed", [0, 1])
def test_against_oracle(pm, seed):
    rng = np.random.default_rng(seed)
    ny, nx = 24, 28
    d = rng.random((ny, nx)) * 100
    d[rng.random((ny, nx)) < 0.1] = -1
    m = (rng.random((ny, nx)) > 0.15).astype(float)
    c = [13.3, 11.7]
    tol = dict(rtol=1e-12, atol=1e-12)
    assert np.allclose(pm.downscale(d, [10, 9]), po.downscale(d, [10, 9]), **tol)
    assert np.allclose(pm.downscale_and_correct(d, m, c, 50.0, 1.2, [12, 11]),
                       po.downscale_and_correct(d, m, c, 50.0, 1.2, [12, 11]), **tol)
    for got, want in zip(pm.cut_downscale_and_correct(d, m, c, 50.0, 1.2, [20, 18], [10, 9]),
                         po.cut_downscale_and_correct(d, m, c, 50.0, 1.2, [20, 18], [10, 9])):
        assert np.allclose(got, want, **tol)
    for got, want in zip(pm.cut_and_downscale(d, m, c, [20, 18], [10, 9]),
                         po.cut_and_downscale(d, m, c, [20, 18], [10, 9])):
        assert np.allclose(got, want, **tol)
    ang = [0.0, 1.5, 3.0, 6.0, 9.0, 14.0]
    for phi in ([-180, 180], [10, 100], [200, 300]):
        assert np.allclose(pm.get_radial_profile(d, 1.1, c, ang, phi), po.radial_profile(d, 1.1, c, ang, phi), **tol)


def test_shapes_and_errors(pm):
    d = np.ones((20, 30))
    assert pm.downscale(d, [15, 10]).shape == (10, 15)
    p, m = pm.cut_and_downscale(d, np.ones((20, 30)), [15.0, 10.0], [20, 20], [8, 6])
    assert p.shape == (6, 8) and m.shape == (6, 8)
    with pytest.raises(ValueError):
        pm.cut_and_downscale(d, np.ones((30, 20)), [15.0, 10.0], [20, 20], [8, 6])
    assert list(pm.get_radial_profile(d, 1.0, [15.0, 10.0], [1.0], [-180, 180])) == [-1.0]


def test_pattern_encoding_crop_mask_saturation():
    from s_spi_spring_amd import Pattern
    rng = np.random.default_rng(3)
    img = rng.random((40, 50)) * 10
    img[5, 12] = np.nan
    img[20, 30] = 1e6
    mask = np.zeros((40, 50))
    mask[0:3, :] = 1
    p = Pattern(img, mask, center=[10, 25], cropsize=32, satvalue=1e5).get()
    assert p.shape == (32, 32)
    # crop rows [10 - 16, 10 + 16) = [-6, 26), columns [9, 41)
    assert np.allclose(p[20, 10], np.float32(img[14, 19]))
    assert p[20 + 6, 30 - 9] == -2                          # saturated
    assert p[5 + 6, 12 - 9] == -1                           # NaN -> masked
    assert np.all(p[0:6, :] == 0)                           # zero padding above the image
    assert np.all(p[6:9, :] == -1)                          # image rows 0-2 are masked


def test_pattern_downscale_path():
    from s_spi_spring_amd import Pattern
    img = np.full((64, 64), 7.0)
    mask = np.zeros((64, 64))
    mask[:, :16] = 1                                        # left quarter masked
    p = Pattern(img, mask, center=[32, 32], cropsize=64, size=32).get()
    assert p.shape == (32, 32)
    assert np.allclose(p[:, 10:], 7.0)
    assert np.all(p[:, :7] == -1)
