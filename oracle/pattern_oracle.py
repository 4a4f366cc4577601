"""Test infrastructure only (not part of the product): a plain-Python
restatement of the reference's detector-image preprocessing, used by the tests
as an independent check of spring_core.pattern on small images.

Follows src/pattern/pattern.cpp (downscale :16-81, downscale_and_correct
:85-183, cut_downscale_and_correct :190-318, cut_and_downscale :323-438,
get_radial_profile :455-516); images are [row][column], points (x, y).
Pinned by tests/test_pattern_prep.py against closed forms (constants, linear
ramps, the Ewald radius map, exact identity crops).
"""
import math

import numpy as np


def _splat(acc, wsum, fx, fy, val):
    ny, nx = acc.shape
    ix, iy = math.floor(fx), math.floor(fy)
    ax, ay = fx - ix, fy - iy
    for (dx, dy, w) in ((0, 0, (1 - ax) * (1 - ay)), (1, 0, ax * (1 - ay)), (0, 1, (1 - ax) * ay), (1, 1, ax * ay)):
        X, Y = ix + dx, iy + dy
        if 0 <= X < nx and 0 <= Y < ny:
            acc[Y, X] += val * w
            wsum[Y, X] += w


def _finish(acc, w, ref):
    out = np.full(acc.shape, -1.0)
    keep = w > 0.1 * ref
    out[keep] = acc[keep] / w[keep]
    return out


def downscale(data, newsize):
    ny, nx = data.shape
    acc = np.zeros((newsize[1], newsize[0]))
    w = np.zeros_like(acc)
    s = nx / newsize[0]
    for y in range(ny):
        for x in range(nx):
            if data[y, x] >= 0:
                _splat(acc, w, x / s, y / s, data[y, x])
    return _finish(acc, w, w)


def _ewald(nx, ny, c, R, dx):
    far = max(c[0], c[1], nx - c[0], ny - c[1])
    dx_new = R * 2 * math.sin(math.atan2(dx * far, R) / 2) / far

    def f(x, y):
        rx, ry = x - c[0], y - c[1]
        phi = math.atan2(ry, rx)
        theta = math.atan2(math.hypot(rx, ry) * dx, R)
        rp = R * 2 * math.sin(theta / 2) / dx_new
        return rp * math.cos(phi), rp * math.sin(phi), 1 / math.cos(theta) ** 3
    return f


def downscale_and_correct(data, mask, center, R, dx, newsize):
    ny, nx = data.shape
    acc = np.zeros((newsize[1], newsize[0]))
    w = np.zeros_like(acc)
    s0, s1 = nx / newsize[0], ny / newsize[1]
    f = _ewald(nx, ny, center, R, dx)
    for y in range(ny):
        for x in range(nx):
            if mask[y, x] > 0.5:
                ox, oy, g = f(x, y)
                _splat(acc, w, (ox + center[0]) / s0, (oy + center[1]) / s1, data[y, x] * g)
    return _finish(acc, w, w.max())


def cut_downscale_and_correct(data, mask, center, R, dx, cutsize, newsize):
    ny, nx = data.shape
    acc, w = np.zeros((newsize[1], newsize[0])), np.zeros((newsize[1], newsize[0]))
    macc, mw = np.zeros_like(acc), np.zeros_like(acc)
    s0, s1 = cutsize[0] / newsize[0], cutsize[1] / newsize[1]
    f = _ewald(nx, ny, center, R, dx)
    for y in range(ny):
        for x in range(nx):
            ox, oy, g = f(x, y)
            fx, fy = ox / s0 + newsize[0] // 2, oy / s1 + newsize[1] // 2
            if mask[y, x] > 0.5:
                _splat(acc, w, fx, fy, data[y, x] * g)
            _splat(macc, mw, fx, fy, mask[y, x])
    return _finish(acc, w, w.max()), _finish(macc, mw, mw.max())


def cut_and_downscale(data, mask, center, cutsize, newsize):
    ny, nx = data.shape
    acc, w = np.zeros((newsize[1], newsize[0])), np.zeros((newsize[1], newsize[0]))
    macc, mw = np.zeros_like(acc), np.zeros_like(acc)
    s0, s1 = cutsize[0] / newsize[0], cutsize[1] / newsize[1]
    for y in range(ny):
        for x in range(nx):
            fx, fy = (x - center[0]) / s0 + newsize[0] // 2, (y - center[1]) / s1 + newsize[1] // 2
            if mask[y, x] > 0.5:
                _splat(acc, w, fx, fy, data[y, x])
            _splat(macc, mw, fx, fy, mask[y, x])
    return _finish(acc, w, mw), _finish(macc, mw, mw.max())


def radial_profile(data, dtheta, center, angles, phi):
    ny, nx = data.shape
    na = len(angles)
    prof, norm = np.zeros(na), np.zeros(na)
    for y in range(ny):
        for x in range(nx):
            rx, ry = x - center[0], y - center[1]
            a = math.atan2(ry, rx) / math.pi * 180
            while a + 360 <= phi[1]:
                a += 360
            if not (phi[0] <= a <= phi[1]):
                continue
            r = math.hypot(rx, ry) * dtheta
            k = next((i for i in range(na - 1) if angles[i] <= r < angles[i + 1]), None)
            if k is None or data[y, x] < 0:
                continue
            f = (angles[k + 1] - r) / (angles[k + 1] - angles[k])
            prof[k] += data[y, x] * f
            norm[k] += f
            prof[k + 1] += data[y, x] * (1 - f)
            norm[k + 1] += 1 - f
    return np.where(norm > 4, prof / np.where(norm > 4, norm, 1), -1.0)
