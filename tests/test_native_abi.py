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
