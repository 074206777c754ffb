"""Takagi-Taupin diffraction in reflection (Bragg geometry), on MI355X.

    from dyfox_amd.bragg import BraggConfig, bragg_g_vector, solve_surface_batch

    g = bragg_g_vector(xtal)
    cfg = BraggConfig(xtal, t_crystal=30e-6, dx=0.05e-6)
    surface, info = solve_surface_batch(segs, s_dev, cfg, ys, xtal, g)

Boundary conditions are two-point: D_0(z = 0) = incident at the front,
D_g(z = t) = 0 at the back.  Conventions match the Laue side:
dD/ds = sigma D, couplings sigma_h e^{-iH} and sigma_hbar e^{+iH},
alpha = -2 pi s_dev, and the exit wave restores the factored phase,
D_g e^{-iH}.  Importing this package touches no device.
"""

from .geometry import (BraggConfig, bragg_g_vector, edge_segment,
                       threading_segment)
from .reference import (darwin_width_urad, reflectivity_riccati,
                        reflectivity_semi_infinite)
from .solver import (perfect_symbol, solve_bragg, solve_bragg_stable,
                     solve_surface_batch, solve_surface_fields)

__all__ = ["BraggConfig", "bragg_g_vector", "edge_segment",
           "threading_segment", "darwin_width_urad", "reflectivity_riccati",
           "reflectivity_semi_infinite", "perfect_symbol", "solve_bragg",
           "solve_bragg_stable", "solve_surface_batch",
           "solve_surface_fields"]
