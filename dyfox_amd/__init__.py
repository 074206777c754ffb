"""dyfox_amd: a Takagi-Taupin dynamical-diffraction engine for MI355X.

Drop-in for the DYFOX hot path:

    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.laue import LaueGrid, solve_dislocation, WaveCache
    from dyfox_amd.bragg import BraggConfig, solve_surface_batch

Importing touches no device; the first solve loads libdyfox.so.
"""

__all__ = ["backend", "crystal", "units", "dislocations", "laue", "bragg"]
