"""MI355X-native backend of the SPRING memetic phase-retrieval engine.

The hot path (per-individual HIO/RAAR/ER iterations, shrink-wrap, fitness,
alignment, averaging and crossover of a population) runs as hand-written gfx950
HIP kernels in lib/libspring_hip.so, reached through the C ABI of
include/spring_hip.h. Above it:

  spring_core          pybind11 module with the reference's Python surface
                       (phase_retrieval.Solver / Reconstruction, GPUtils)
  hipcore              ctypes binding of the C ABI
  MPR, Settings, Pattern, AlgManager, Stats
                       the memetic driver and its configuration
  grad                 differentiable masked Fourier loss (adjoint FFT) for PyTorch

install_aliases() registers `spring_core` and `spring` as module names so that
code written against the reference imports this package unchanged.
"""
from __future__ import annotations

import sys

try:
    # libspring_hip.so binds to the HIP runtime loaded under the soname
    # `libamdhip64.so`; importing torch first makes that torch's runtime, so that
    # torch tensors and this library share one HIP/HSA runtime and one device
    # context in the process (two runtimes cannot both open the GPU).
    import torch  # noqa: F401
except ImportError:
    pass

from .alg_manager import ER, HIO, RAAR, SW, AlgManager  # noqa: F401
from .pattern import Pattern  # noqa: F401
from .result import Result  # noqa: F401
from .settings import Settings  # noqa: F401
from .stats import Stats  # noqa: F401


def __getattr__(name):
    if name == "MPR":
        from .mpr import MPR
        return MPR
    raise AttributeError(name)


def install_aliases():
    """Make `import spring_core` / `from spring_core.phase_retrieval import Solver`
    and `from spring import MPR, Settings, Pattern` resolve to this package."""
    import types

    from . import spring_core
    from .mpr import MPR
    sys.modules.setdefault("spring_core", spring_core)
    sys.modules.setdefault("spring_core.phase_retrieval", spring_core.phase_retrieval)
    sys.modules.setdefault("spring_core.GPUtils", spring_core.GPUtils)
    spring = types.ModuleType("spring")
    spring.MPR, spring.Settings, spring.Pattern, spring.Result = MPR, Settings, Pattern, Result
    spring.Reconstruction = spring_core.phase_retrieval.Reconstruction
    sys.modules.setdefault("spring", spring)
    return spring
