"""Diffraction pattern container with the reference's input encoding.

Pattern(pattern, mask, center, cropsize, pid, size, satvalue, metadata) follows
spring/pattern_manager.py:48-98 for cropping and masking: NaN and masked
pixels become -1, pixels above `satvalue` become -2 (saturated), other negative
values are clipped to 0. Rescaling (size < cropsize) uses the reference's
native downscaling, which is outside the accelerated path and not provided here.
"""
from __future__ import annotations

import numpy as np


def _cut(data, center, size):
    """Square crop of side `size` centred on (y, x) = center, zero padded."""
    out = np.zeros((size, size), dtype=np.float64)
    cy, cx = int(round(center[0])), int(round(center[1]))
    y0, x0 = cy - size // 2, cx - size // 2
    ys, xs = max(0, y0), max(0, x0)
    ye, xe = min(data.shape[0], y0 + size), min(data.shape[1], x0 + size)
    if ye > ys and xe > xs:
        out[ys - y0:ye - y0, xs - x0:xe - x0] = data[ys:ye, xs:xe]
    return out


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
        if size != cropsize:
            raise NotImplementedError("rescaling of the pattern is not provided; crop to the final size instead")
        data = _cut(np.nan_to_num(data), center, cropsize)
        pixmask = _cut(pixmask, center, cropsize)
        sat = np.zeros_like(pixmask)
        if satvalue is not None:
            sat[data > satvalue] = 1
        data[data < 0] = 0
        data[pixmask > 0.25] = -1
        data[sat > 0.05] = -2
        self.pattern = data.astype(np.float32)

    def get(self):
        return self.pattern

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
