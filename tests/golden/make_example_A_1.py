"""Writes example_A_1_256.npz: the reference's recorded example pattern
(examples/data/example_A_1.npz of the reference, loaded with
np.load(allow_pickle=False)) preprocessed as examples/Reconstruction.ipynb
cell 3 does:

    Pattern(pattern=data['pattern'], mask=data['mask'], center=data['center'],
            pid=data['pid'], cropsize=1024, size=256, satvalue=3e4)

The 256^2 result (float32, -1 masked, -2 saturated) and the pid are stored.
The downscaling of this image by spring_core.pattern is checked here against
the plain-Python restatement oracle/pattern_oracle.py before writing.

Usage: python tests/golden/make_example_A_1.py [path to the reference checkout]
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import pattern_oracle  # noqa: E402
from s_spi_spring_amd import Pattern  # noqa: E402
from s_spi_spring_amd.pattern import cut  # noqa: E402


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = np.load(os.path.join(ref, "examples", "data", "example_A_1.npz"), allow_pickle=False)
    patt = Pattern(pattern=data["pattern"], mask=data["mask"], center=data["center"], pid=data["pid"],
                   cropsize=1024, size=256, satvalue=3e4)
    # the native downscaling of this image against the oracle's restatement
    img = cut(np.array(data["pattern"], np.float64), data["center"], 1024) + 1e3
    m = cut(np.array(data["mask"], np.float64), data["center"], 1024)
    img[m > 0.5] = -1
    from s_spi_spring_amd import spring_core
    got = spring_core.pattern.downscale(np.ascontiguousarray(img), [256, 256])
    want = pattern_oracle.downscale(img, [256, 256])
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"downscale 1024 -> 256 against the oracle: max relative difference {err:.2e}")
    assert err < 1e-12
    out = os.path.join(HERE, "example_A_1_256.npz")
    np.savez_compressed(out, pattern=patt.get(), pid=np.asarray(data["pid"], np.int64))
    p = patt.get()
    print(f"wrote {out}: {p.shape} {p.dtype}, {np.count_nonzero(p == -1)} masked, "
          f"{np.count_nonzero(p == -2)} saturated, sum {p[p > 0].sum():.6g}")


if __name__ == "__main__":
    main()
