"""dyf_laue_march_noalloc (the march with its coefficient table in a
library-held device array) against dyf_laue_march (the same march with a
caller workspace), through the C ABI: two calls in a row, the second
smaller, each bit-identical to the existing entry point; a call above the
array's maximum is rejected before anything is launched."""

import ctypes as C

import numpy as np
import pytest

from dyfox_amd import _native as N
from dyfox_amd.backend import FP64, MIXED
from dyfox_amd.laue import LaueGrid
from dyfox_amd.laue import solver as LS
from dyfox_amd.units import deviation_from_urad
from tests.helpers import geometry_n, oracle_exit

pytestmark = pytest.mark.gpu


def _both(grid, xtal, segs, s, prec):
    """(existing entry point's output, no-allocation entry point's)."""
    src = LS.prepare_phase(segs, grid, xtal, precision=prec)
    a = LS.laue_args(grid, xtal)
    sd = N.DeviceArray.from_host(np.asarray(s, np.float64))
    n_e = len(s)
    shape = (n_e,) + grid.shape
    nb = C.c_size_t()
    N.call("dyf_laue_coef_bytes", C.byref(a), n_e, C.byref(nb))
    ws = N.DeviceArray.empty((nb.value,), np.uint8)
    want = N.DeviceArray.empty(shape, prec.complex)
    got = N.DeviceArray.empty(shape, prec.complex)
    line = None if src.line is None else C.byref(src.line)
    N.call("dyf_laue_march", C.byref(a), src.mode, prec.code, src.ptr,
           src.h_prec, line,
           C.c_void_p(sd.ptr), n_e, C.c_void_p(ws.ptr),
           C.c_void_p(want.ptr), None)
    N.call("dyf_laue_march_noalloc", C.byref(a), src.mode, prec.code,
           src.ptr, src.h_prec, line,
           C.c_void_p(sd.ptr), n_e, C.c_void_p(got.ptr), None)
    return want.get(), got.get()


def test_noalloc_march_equals_workspace_march(dev, diamond400):
    big = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    seg = geometry_n(diamond400, big)
    s = deviation_from_urad(np.linspace(-150.0, 150.0, 6), diamond400)
    print("no-alloc call 1: t 40 um, 6 angles, FP64", flush=True)
    want, got = _both(big, diamond400, seg, s, FP64)
    assert np.array_equal(got, want)
    ref = oracle_exit(seg, s, big, diamond400)
    assert np.abs(got - ref).max() <= 1e-10 * np.abs(ref).max()
    # second call: a thinner grid (fewer steps) and fewer angles, so the
    # array still holds the first call's longer table
    small = LaueGrid(diamond400, 20e-6, 0.25e-6, 6e-6)
    seg2 = geometry_n(diamond400, small, sign=-1)
    s2 = deviation_from_urad(np.array([30.0, -75.0]), diamond400)
    for prec in (FP64, MIXED):
        print(f"no-alloc call 2: t 20 um, 2 angles, {prec.name}", flush=True)
        want, got = _both(small, diamond400, seg2, s2, prec)
        assert np.array_equal(got, want)


def test_noalloc_rejects_calls_above_its_maximum(dev, diamond400):
    lib = N.load_library()
    cap = lib.dyf_laue_march_noalloc_max_bytes()
    assert cap == 4 << 20
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    a = LS.laue_args(grid, diamond400)
    per = 48 * (grid.Nz - 1) + 16          # coefficient bytes per angle
    n_e = cap // per + 1
    assert n_e * per > cap
    # buffers of the full size, so even a call that were not rejected
    # would stay inside them
    sd = N.DeviceArray.from_host(np.zeros(n_e))
    out = N.DeviceArray.empty((n_e,) + grid.shape, np.complex128)
    with pytest.raises(ValueError, match="scratch"):
        N.call("dyf_laue_march_noalloc", C.byref(a), N.MODE_PRISTINE,
               N.FP64, None, N.FP64, None, C.c_void_p(sd.ptr), n_e,
               C.c_void_p(out.ptr), None)
