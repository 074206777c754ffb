"""The C ABI: the library loads and exports every symbol include/dyfox.h
declares (no device calls, runs on the CPU)."""

import ctypes
import os
import re

import pytest

from dyfox_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(dyf_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail("libdyfox.so is not built (python -m dyfox_amd.build)")
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in header_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(header_symbols()) >= 20


def test_binding_covers_the_header():
    declared = set(header_symbols())
    bound = set(_native.declared_symbols())
    assert declared <= bound | {"dyf_last_error"}
    assert bound <= declared


def test_structs_match_the_header_layout():
    assert ctypes.sizeof(_native.LaueArgs) == 12 * 8 + 4 * 4
    assert ctypes.sizeof(_native.LineParams) == 12 * 8
    assert ctypes.sizeof(_native.VolumeDesc) == 24
    assert ctypes.sizeof(_native.BraggArgs) == 7 * 8 + 6 * 4
    assert ctypes.sizeof(_native.PupilArgs) == 2 * 4 + 6 * 8 + 4 * 4
    assert ctypes.sizeof(_native.FixedOpts) == 16
    assert ctypes.sizeof(_native.FixedInfo) == 16


def _host_call(name, *args):
    """Call a host-only entry point (argument checks and sizing: no device
    call is made) through the bare library; (code, message)."""
    lib = _native.load_library()
    code = getattr(lib, name)(*args)
    return code, lib.dyf_last_error().decode()


def _laue_args(diamond400):
    from dyfox_amd.laue import LaueGrid
    from dyfox_amd.laue.solver import laue_args
    return laue_args(LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6), diamond400)


def test_unknown_mode_and_precision_are_key_errors(diamond400):
    """The reference raises KeyError for an unknown kernel name
    (laue/kernels.py:365-367); the C ABI returns DYF_EKEY, which the
    binding maps to KeyError."""
    a = _laue_args(diamond400)
    nb = ctypes.c_size_t()
    code, msg = _host_call("dyf_laue_phase_bytes", ctypes.byref(a), 7,
                           _native.FP64, ctypes.byref(nb))
    assert code == _native.DYF_EKEY and "mode" in msg
    code, msg = _host_call("dyf_laue_phase_bytes", ctypes.byref(a),
                           _native.MODE_GRID, 9, ctypes.byref(nb))
    assert code == _native.DYF_EKEY and "precision" in msg
    with pytest.raises(KeyError):
        _native._raise(code, "dyf_laue_phase_bytes")
    code, _ = _host_call("dyf_laue_phase_bytes", ctypes.byref(a),
                         _native.MODE_GRID, _native.FP64, ctypes.byref(nb))
    assert code == _native.DYF_OK and nb.value > 0


def test_bad_arguments_are_value_errors(diamond400):
    a = _laue_args(diamond400)
    a.Nx = 0
    code, msg = _host_call("dyf_laue_coef_bytes", ctypes.byref(a), 4,
                           ctypes.byref(ctypes.c_size_t()))
    assert code == _native.DYF_EINVAL and "grid" in msg
    with pytest.raises(ValueError):
        _native._raise(code, "dyf_laue_coef_bytes")
    with pytest.raises(MemoryError):
        _native._raise(_native.DYF_ENOMEM, "x")


def test_bad_precision_name_is_a_value_error():
    """backend.py:59-60: only the named precisions exist."""
    from dyfox_amd.backend import FP32, FP64, Precision, precision
    with pytest.raises(ValueError):
        Precision("fp16")
    # the lookup by flag keeps the reference's forgiving default
    assert precision("single") is FP32 and precision("whatever") is FP64
