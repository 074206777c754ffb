"""Takagi-Taupin diffraction in transmission (Laue geometry), on MI355X.

    from dyfox_amd.laue import LaueGrid, solve_dislocation, WaveCache

    grid = LaueGrid(xtal, t_crystal=273e-6, dx=0.25e-6, beam_width=120e-6)
    stack = solve_dislocation(segs, s_dev, grid, xtal)

Importing this package touches no device; the first solve does.
"""

from .cache import WaveCache
from .grid import LaueGrid
from .solver import solve_dislocation, solve_pristine

__all__ = ["LaueGrid", "WaveCache", "solve_dislocation", "solve_pristine"]
