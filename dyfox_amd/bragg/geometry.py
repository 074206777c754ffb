"""Symmetric-Bragg geometry: the grid and dislocations to put in it.

g lies along the surface normal; the two characteristics share a lateral
direction and run opposite ways in depth, so dz = dx tan(theta_B) walks
exactly one pixel per step (the mirror of the Laue rule).  The lateral
roll of the march is periodic, so the grid is padded by the round-trip
walk (reference bragg/geometry.py:18-109, whose sizing rules are kept so
the same inputs give the same grid):

    Nz = round(t / dz) + 1,  t snapped to (Nz - 1) dz
    Nx = 2 (round(half_x / dx) + 2 Nz + 16),  ds = dz / sin(theta_B)
"""

import numpy as np


def bragg_g_vector(xtal):
    """g along the surface normal with the reflection's magnitude."""
    return np.array([0.0, 0.0, float(np.linalg.norm(xtal.g_vec))])


class BraggConfig:
    """Grid for a symmetric-Bragg solve."""

    def __init__(self, xtal, t_crystal=30e-6, dx=0.05e-6, half_x=12e-6,
                 beam_thickness=None, beam_centre=0.0):
        self.xtal = xtal
        self.dx = float(dx)
        self.dz = float(dx) * xtal.tan_tB
        self.Nz = int(round(float(t_crystal) / self.dz)) + 1
        self.t_crystal = (self.Nz - 1) * self.dz
        self.half_x = float(half_x)
        self.Nx = int(2 * (round(half_x / self.dx) + 2 * self.Nz + 16))
        self.ds = self.dz / xtal.sin_tB
        # None: a plane wave over the whole entrance face; a number: a
        # top hat of that full width centred on beam_centre
        self.beam_thickness = beam_thickness
        self.beam_centre = float(beam_centre)

    def incident(self, xp=np, n=1):
        """D_0 on the entrance face, shape (n, Nx), complex128 (host)."""
        if self.beam_thickness is None:
            return np.ones((n, self.Nx), dtype=np.complex128)
        x = (np.arange(self.Nx) - self.Nx / 2) * self.dx
        top = (np.abs(x - self.beam_centre)
               < 0.5 * self.beam_thickness).astype(np.complex128)
        return np.broadcast_to(top, (n, self.Nx)).copy()

    def x_axis(self):
        return (np.arange(self.Nx) - self.Nx / 2) * self.dx

    def z_axis(self):
        return np.arange(self.Nz) * self.dz

    def __repr__(self):
        return (f"BraggConfig(t={self.t_crystal*1e6:.2f}um, "
                f"dx={self.dx*1e9:.0f}nm, dz={self.dz*1e9:.0f}nm, "
                f"Nz={self.Nz}, Nx={self.Nx})")


def edge_segment(sign, cfg, b_mag, depth=None):
    """Edge line along y, b in the (x, z) plane at 45 degrees to the
    normal (g.b = 2 for diamond (400)), at `depth` (default t/4)."""
    depth = cfg.t_crystal * 0.25 if depth is None else depth
    xi = np.array([0.0, 1.0, 0.0])
    bv = sign * b_mag * np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    core = np.array([0.0, 0.0, depth])
    return [(core - 400e-6 * xi, core + 400e-6 * xi, bv)]


def threading_segment(kind, sign, b_mag, g_mag, psi=np.pi / 4,
                      y0=0.0, x0=0.0, azimuth_deg=0.0):
    """A line threading the surface at angle psi from the normal, with
    g.b = 2 sign for g along the normal (pure edge or pure screw), its
    in-plane projection optionally turned by `azimuth_deg`."""
    xi = np.array([0.0, np.sin(psi), np.cos(psi)])
    if azimuth_deg:
        c, s = np.cos(np.radians(azimuth_deg)), np.sin(np.radians(azimuth_deg))
        xi = np.array([c * xi[0] - s * xi[1], s * xi[0] + c * xi[1], xi[2]])
    bz = 2.0 / g_mag
    if kind == "Screw":
        bv = sign * b_mag * xi
    else:
        by = -bz / np.tan(psi)
        bx = np.sqrt(max(b_mag ** 2 - bz ** 2 - by ** 2, 0.0))
        bv = sign * np.array([bx, by, bz])
    core = np.array([x0, y0, 0.0])
    return [(core - 300e-6 * xi, core + 300e-6 * xi, bv)]
