"""Diffraction pattern container with the reference's input encoding.

Pattern(pattern, mask, center, cropsize, pid, size, satvalue, metadata) follows
spring/pattern_manager.py:6-98: the pattern and the mask are cropped to
[cropsize, cropsize] around center = (row, column) with zero padding, pixels
above `satvalue` are flagged saturated, and with size < cropsize both are
rescaled with spring_core.pattern.downscale (the native preprocessing module).
In the result masked pixels are -1, saturated pixels -2 and the remaining
negative values 0.
"""
from __future__ import annotations

import numpy as np


def _downscale(data, size):
    from . import spring_core      # host C++ module (no device needed)
    return spring_core.pattern.downscale(np.ascontiguousarray(data, dtype=np.float64), [size, size])


def resize_pattern(data, mask, newsize):
    """pattern_manager.py:6-18: shift the counts by 1e3 so that masked pixels
    (-1) are the only negative ones, downscale, and mask what the downscaling
    could not fill."""
    tmp = np.array(data, dtype=np.float64) + 1e3
    tmp[mask > 0.5] = -1
    out = _downscale(tmp, newsize)
    newmask = np.zeros(out.shape)
    newmask[out < 0] = 1
    return out - 1e3, newmask


def cut(data, center, size):
    """pattern_manager.py:21-30: rows and columns [c - size//2, c + size//2) of
    the image zero-padded by twice its height on every side."""
    data = np.asarray(data, dtype=np.float64)
    frame = int(data.shape[0] * 2)
    t = int(data.shape[0] + 2 * frame)
    tmp = np.zeros([t, t])
    tmp[frame:data.shape[0] + frame, frame:data.shape[1] + frame] = data
    c0, c1 = int(center[0]) + frame, int(center[1]) + frame
    return np.copy(tmp[c0 - size // 2:c0 + size // 2, c1 - size // 2:c1 + size // 2])


class Pattern:
    def __init__(self, pattern, mask, center, cropsize: int = None, pid: list = (0,), size: int = None,
                 satvalue: float = None, metadata=None):
        self.origpattern = np.copy(pattern)
        self.origmask = np.copy(mask)
        self.pid = list(pid)
        self.metadata = {} if metadata is None else metadata
        data = np.array(pattern, dtype=np.float64)
        pixmask = np.array(mask, dtype=np.float64)
        pixmask[np.isnan(data)] = 1
        if cropsize is None:
            cropsize = int(np.amax(data.shape))
        if size is None:
            size = cropsize
        data = cut(data, center, cropsize)
        pixmask = cut(pixmask, center, cropsize)
        satmask = np.zeros(pixmask.shape)
        if satvalue is not None:
            satmask[data > satvalue] = 1
            pixmask[satmask == 1] = 1
        data[pixmask == 1] = -1
        data[satmask == 1] = -1
        resizedmask = np.copy(pixmask)
        if size < cropsize:
            data, resizedmask = resize_pattern(data, pixmask, size)
            satmask = _downscale(satmask, size)
            pixmask = _downscale(pixmask, size)
        pixmask[pixmask > 0.25] = 1
        pixmask[pixmask <= 0.25] = 0
        pixmask[np.logical_and(pixmask < 0.5, resizedmask > 0.25)] = 1
        satmask[satmask > 0.05] = 1
        satmask[satmask <= 0.05] = 0
        data[data < 0] = 0
        data[pixmask == 1] = -1
        data[satmask == 1] = -2
        self.pattern = data.astype(np.float32)

    def get(self):
        return self.pattern

    def get_pid(self):
        return self.pid

    def get_metadata(self):
        return self.metadata

    @classmethod
    def from_intensity(cls, intensity, pid=(0,), metadata=None):
        """Wrap an already centred intensity in the reference encoding (-1 masked, -2 saturated)."""
        obj = cls.__new__(cls)
        obj.origpattern = np.copy(intensity)
        obj.origmask = (np.asarray(intensity) < 0).astype(np.float64)
        obj.pid = list(pid)
        obj.metadata = {} if metadata is None else metadata
        obj.pattern = np.asarray(intensity, dtype=np.float32)
        return obj
