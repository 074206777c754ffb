"""Test configuration.

Markers: `gpu` (needs an MI355X and libdyfox.so; never skipped silently:
on a GPU run a missing library is an error), `slow`.  The CPU suite
(-m "not gpu") covers the oracle against closed forms and the reference's
known-answer values, the host logic, and the C ABI's exports.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X and libdyfox.so")
    config.addinivalue_line("markers", "slow: more than about ten seconds")


@pytest.fixture
def rng():
    return np.random.default_rng(20260909)


@pytest.fixture(scope="session")
def diamond400():
    """Diamond (400) at 17 keV, f' = 0.003, f'' = 0.002."""
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    return CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002,
                                         label="Diamond(400)")


@pytest.fixture(scope="session")
def small_grid(diamond400):
    """t = 40 um, dx = 0.25 um, beam 8 um: Nz 73, (Ny, Nx) = (96, 224)."""
    from dyfox_amd.laue import LaueGrid
    return LaueGrid(diamond400, t_crystal=40e-6, dx=0.25e-6,
                    beam_width=8e-6, beam_thickness=1e-6)


@pytest.fixture(scope="session")
def dev():
    """The native library with its device initialised (errors, never
    skips, when the library or the device is missing)."""
    from dyfox_amd.backend import device
    return device()


@pytest.fixture(autouse=True)
def _collect():
    yield
    import gc
    gc.collect()
