"""The Laue sampling grid and the depth step it implies.

Unless `Nz` is forced, dz = dx / tan(theta_B): one depth step carries a
characteristic of the Borrmann fan exactly one pixel sideways, so the
march transports by an index shift, bit-exactly (laue/grid.py:1-89 of the
reference, whose sizing rules this follows so that grids, and therefore
exit waves and cache keys, come out identical):

    Nz = round(t / dz) + 1, t snapped to (Nz - 1) dz
    Nx = ceil32(max(ceil((beam_eff + 2 t tan tB + 2 pad dx) / dx), 128))
    Ny = ceil32(max(ceil((beam_width + 2 pad dx) / dx), 64))
    ds = dz / cos tB,  shift_px = tan tB dz / dx (snapped within 1e-9)

`Ny=` is an addition: it fixes the number of rows (e.g. 256 rows all
inside the field of view) instead of deriving it from the beam width.
"""

import numpy as np


class LaueGrid:
    """Grid and beam settings for one solve (see module docstring)."""

    def __init__(self, xtal, t_crystal, dx, beam_width, *,
                 beam_thickness=1.0e-6, beam_sigma=0.0, Nz=None,
                 grid_pad=30, Ny=None):
        self.xtal = xtal
        self.beam_thickness = float(beam_thickness)
        self.beam_sigma = float(beam_sigma)
        self.dx = float(dx)
        self.beam_width = float(beam_width)
        self.grid_pad = int(grid_pad)
        self.Ny_override = None if Ny is None else int(Ny)
        if Nz is None:
            dz = self.dx / xtal.tan_tB
            self.Nz = int(round(float(t_crystal) / dz)) + 1
            self.t_crystal = (self.Nz - 1) * dz
            self.exact = True
        else:
            self.Nz = int(Nz)
            self.t_crystal = float(t_crystal)
            dz = self.t_crystal / (self.Nz - 1)
            self.exact = abs(xtal.tan_tB * dz / self.dx - 1.0) < 1e-9

    def replace(self, **kw):
        """A copy with some settings changed."""
        base = dict(xtal=self.xtal, t_crystal=self.t_crystal, dx=self.dx,
                    beam_width=self.beam_width,
                    beam_thickness=self.beam_thickness,
                    beam_sigma=self.beam_sigma, grid_pad=self.grid_pad,
                    Ny=self.Ny_override)
        if "Nz" not in kw and not self.exact:
            base["Nz"] = self.Nz
        base.update(kw)
        return LaueGrid(**base)

    def sheet_arg(self):
        """The march's beam parameter: half-width, or minus the sigma."""
        if self.beam_sigma > 0:
            return -self.beam_sigma
        return self.beam_thickness / 2.0

    def geometry(self):
        """(Nx, Ny, dz, ds, shift_px) of this grid."""
        xtal = self.xtal
        pad = self.grid_pad * self.dx
        eff = (6.0 * self.beam_sigma if self.beam_sigma > 0
               else self.beam_thickness)
        extent_x = eff + 2 * (self.t_crystal * xtal.tan_tB) + 2 * pad
        Nx = max(int(np.ceil(extent_x / self.dx)), 128)
        Nx = int(np.ceil(Nx / 32)) * 32
        if self.Ny_override is not None:
            Ny = self.Ny_override
        else:
            Ny = max(int(np.ceil((self.beam_width + 2 * pad) / self.dx)), 64)
            Ny = int(np.ceil(Ny / 32)) * 32
        dz = self.t_crystal / (self.Nz - 1)
        ds = dz / xtal.cos_tB
        shift_px = xtal.tan_tB * dz / self.dx
        if abs(shift_px - round(shift_px)) < 1e-9:
            shift_px = float(round(shift_px))
        return Nx, Ny, dz, ds, shift_px

    @property
    def shape(self):
        """(Ny, Nx), the shape of one exit wave."""
        Nx, Ny = self.geometry()[:2]
        return Ny, Nx

    def marched_rows(self):
        """Rows inside the field of view, |(iy - Ny/2) dx| < beam_width/2,
        evaluated in FP64 (the same test the march applies)."""
        Ny = self.geometry()[1]
        y = (np.arange(Ny) - Ny // 2) * self.dx
        return np.nonzero(np.abs(y) < self.beam_width / 2.0)[0]

    def pixel_steps(self, n_e):
        """Work of one solve: marched rows x Nx x (Nz - 1) x n_e."""
        Nx = self.geometry()[0]
        return len(self.marched_rows()) * Nx * (self.Nz - 1) * int(n_e)

    def key(self):
        """Identity of this grid for a cache key (an Ny override is part
        of it only when set, so default grids hash as the reference's)."""
        k = dict(beam_thickness=self.beam_thickness,
                 beam_sigma=self.beam_sigma, t_crystal=self.t_crystal,
                 Nz=self.Nz, dx=self.dx, beam_width=self.beam_width,
                 grid_pad=self.grid_pad)
        if self.Ny_override is not None:
            k["Ny"] = self.Ny_override
        return k

    def __repr__(self):
        Nx, Ny = self.geometry()[:2]
        return (f"LaueGrid({Ny}x{Nx}, Nz={self.Nz}, "
                f"t={self.t_crystal * 1e6:.1f}um, "
                f"dx={self.dx * 1e9:.0f}nm"
                + ("" if self.exact else ", interpolating") + ")")
