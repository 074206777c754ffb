"""The reference's error contract, raised by the device path.

* MemoryError when the H grid would not fit (laue/solver.py:49-54: more
  than 0.7 of free device memory) and when a workgroup would need more
  LDS than the device has (laue/kernels.py:378-386; 160 KB on gfx950);
* KeyError for an unknown march mode or precision through the C ABI
  (laue/kernels.py:365-367);
* no launch is left enqueued by a call that fails: the next solve is
  still exact.
"""

import ctypes

import numpy as np
import pytest

from dyfox_amd import _native as N
from dyfox_amd.backend import to_numpy
from dyfox_amd.laue import LaueGrid, solve_dislocation
from dyfox_amd.laue import solver as LS
from dyfox_amd.units import deviation_from_urad
from tests.helpers import geometry_n, oracle_exit

pytestmark = pytest.mark.gpu


def test_grid_larger_than_free_memory_is_a_memory_error(dev, diamond400,
                                                        monkeypatch):
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    seg = geometry_n(diamond400, grid)
    s = deviation_from_urad(np.array([0.0]), diamond400)
    real = N.mem_info
    # 1 MB "free": the e^{iH} table (and an H grid) need more than 0.7 of it
    monkeypatch.setattr(N, "mem_info", lambda: (1 << 20, real()[1]))
    with pytest.raises(MemoryError):
        solve_dislocation(seg, s, grid, diamond400)
    with pytest.raises(MemoryError):
        solve_dislocation(seg, s, grid, diamond400,
                          H_gpu=np.zeros((grid.shape[0], grid.Nz,
                                          grid.shape[1])))
    monkeypatch.setattr(N, "mem_info", real)
    got = to_numpy(solve_dislocation(seg, s, grid, diamond400))
    want = oracle_exit(seg, s, grid, diamond400)
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_lds_over_the_device_limit_is_a_memory_error(dev, diamond400):
    """A non-exact grid (forced Nz: the interpolating general path) whose
    x-extent needs 8 (Nx + 1) doubles of LDS above 160 KB."""
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 4e-6, beam_thickness=600e-6,
                    Nz=61)
    Nx = grid.geometry()[0]
    assert 8 * (Nx + 1) * 8 > 160 * 1024
    assert LS.plan(grid, diamond400)[0] == N.PATH_GENERAL
    s = deviation_from_urad(np.array([0.0, 10.0]), diamond400)
    with pytest.raises(MemoryError, match="LDS"):
        to_numpy(solve_dislocation([], s, grid, diamond400))
    # a narrower beam on the same depth grid fits and runs
    ok = LaueGrid(diamond400, 40e-6, 0.25e-6, 4e-6, Nz=61)
    out = to_numpy(solve_dislocation([], s, ok, diamond400))
    want = oracle_exit([], s, ok, diamond400)
    assert np.abs(out - want).max() <= 1e-10 * np.abs(want).max()


def test_unknown_mode_and_precision_raise_key_error(dev, diamond400):
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    a = LS.laue_args(grid, diamond400)
    nb = ctypes.c_size_t()
    with pytest.raises(KeyError):
        N.call("dyf_laue_phase_bytes", ctypes.byref(a), 5, N.FP64,
               ctypes.byref(nb))
    with pytest.raises(KeyError):
        N.call("dyf_laue_phase_bytes", ctypes.byref(a), N.MODE_GRID, 3,
               ctypes.byref(nb))
    sd = N.DeviceArray.from_host(np.zeros(1))
    out = N.DeviceArray.empty((1,) + grid.shape, np.complex128)
    with pytest.raises(KeyError):
        N.call("dyf_laue_march", ctypes.byref(a), 4, N.FP64, None, N.FP64,
               None, ctypes.c_void_p(sd.ptr), 1, None,
               ctypes.c_void_p(out.ptr), None)
