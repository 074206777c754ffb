"""The device boundary as the solvers see it, and the three precisions.

Mirrors the reference's backend.py:1-99: `to_numpy`, `array_module`,
`free_gpu`, `precision(name)` and the `Precision` objects passed to the
solvers.  CuPy is replaced by libdyfox.so (hand-written gfx950 HIP behind
a C ABI); `device()` returns that binding and raises if it is missing.

    waves = solve_dislocation(segs, s_dev, grid, xtal, precision=MIXED)

FP64    complex128 exit waves, FP64 state and e^{iH}.
FP32    complex64 exit waves, FP32 state and e^{iH} (H reduced mod 2 pi
        in FP64 first): the reference's "draft" mode.
MIXED   complex64 exit waves, FP32 state carried with an FP32
        compensation term (two-float), FP32 e^{iH} from FP64-reduced H.
"""

import numpy as np

WAVE = 64            # lanes per wavefront on gfx950


def device():
    """The native binding with its device initialised, or an error naming
    what is missing (no CPU fallback exists)."""
    from . import _native
    _native.lib_ready()
    return _native


def to_numpy(a):
    """Host copy of `a`, device array or not."""
    return a.get() if hasattr(a, "get") else np.asarray(a)


def array_module(a):
    """numpy for host arrays, the device binding for device arrays."""
    from . import _native
    return _native if isinstance(a, _native.DeviceArray) else np


def free_gpu():
    """Release the solvers' cached workspaces; arrays free themselves when
    collected."""
    import gc
    import sys
    mod = sys.modules.get(__package__ + ".laue.solver")
    if mod is not None:
        mod.release_workspaces()
    gc.collect()


class Precision:
    """One precision: dtypes, the library code, and the element size of
    the H grid it builds."""

    _CODES = {"fp64": 0, "fp32": 1, "mixed": 2}

    def __init__(self, name):
        if name not in self._CODES:
            raise ValueError(
                f"precision is 'fp64', 'fp32' or 'mixed', not {name!r}")
        self.name = name
        self.code = self._CODES[name]
        self.is_fp32 = name != "fp64"
        # H grids follow the reference: float64 in FP64, float32 otherwise
        self.real = np.float32 if name == "fp32" else np.float64
        self.complex = np.complex128 if name == "fp64" else np.complex64
        self.itemsize = np.dtype(self.real).itemsize

    def scalar(self, x):
        return self.real(x)

    def __repr__(self):
        return f"Precision({self.name!r})"


FP64 = Precision("fp64")
FP32 = Precision("fp32")
MIXED = Precision("mixed")


def precision(name):
    """FP64, FP32 or MIXED by name (a Precision passes through)."""
    if isinstance(name, Precision):
        return name
    s = str(name).lower()
    if s in ("mixed", "fp32c", "compensated"):
        return MIXED
    return FP32 if s in ("fp32", "32", "single") else FP64
