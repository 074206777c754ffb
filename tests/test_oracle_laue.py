"""Pins the Laue oracle (oracle/laue.py, oracle/volterra.py).

Each check is chosen so a plausible mistake in the oracle fails it:

* discrete Pendelloesung: with H = 0 and a beam wider than the fan, the
  interior of the march equals (M^(Nz-1))_21 computed by matrix powers
  (a wrong sign of alpha, a swapped sigma_h/sigma_hbar, a missing ep or
  a transposed M all change it);
* continuous Pendelloesung (the reference's closed-form check,
  tests/test_laue_gpu.py:59-70) on an interpolating grid;
* a uniform shear H = q y gives an exit phase slope of -q (pins the
  e^{-iH} coupling sign, tests/test_laue_gpu.py:73-90);
* the winding of H about the core equals g.b (tests/test_laue_gpu.py:93-113),
  zero b gives exactly zero (141-145), superposition over segments;
* the x-directed line's closed-form H equals the grid H;
* the interpolating march with shift exactly 1 reproduces the pure
  index shift.
"""

import numpy as np
import pytest

from dyfox_amd.dislocations import g_dot_b
from dyfox_amd.laue import LaueGrid
from dyfox_amd.units import deviation_from_urad
from oracle import closed_forms as CF, laue as OL, volterra as OV
from tests.helpers import geometry_n, grid_args, mixed_on_x


def test_discrete_pendellosung(diamond400):
    grid = LaueGrid(diamond400, 30e-6, 0.25e-6, 2e-6, beam_thickness=200e-6)
    g = grid_args(grid, diamond400)
    s = deviation_from_urad(np.array([-40.0, 0.0, 25.0, 117.5]), diamond400)
    out = OL.march_grid(g, s)
    Nx, Ny = g["Nx"], g["Ny"]
    for j, sd in enumerate(s):
        want = CF.pendellosung_discrete(sd, g["sig0"], g["sig_h"],
                                        g["sig_hbar"], g["ds"], grid.Nz)
        got = out[j, Ny // 2, Nx // 2]
        assert abs(got - want) <= 1e-12 * abs(want) + 1e-15


def test_continuous_pendellosung(diamond400):
    grid = LaueGrid(diamond400, t_crystal=45e-6 / diamond400.tan_tB,
                    dx=0.25e-6, beam_width=2e-6, beam_thickness=120e-6,
                    Nz=801)
    g = grid_args(grid, diamond400)
    s = np.linspace(-6, 6, 25) / diamond400.xi_g
    out = OL.march_grid(g, s)
    got = np.abs(out[:, g["Ny"] // 2, g["Nx"] // 2]) ** 2
    want = CF.planewave_rocking(s, g["sig0"], g["sig_h"], g["sig_hbar"],
                                diamond400.cos_tB, grid.t_crystal)
    resid = got / got.max() - want / want.max()
    assert float(np.sqrt(np.mean(resid ** 2))) < 0.02


def test_uniform_shear_sets_the_coupling_sign(diamond400):
    grid = LaueGrid(diamond400, t_crystal=20e-6, dx=0.25e-6,
                    beam_width=40e-6, beam_thickness=1e-6, Nz=101)
    g = grid_args(grid, diamond400)
    Nx, Ny = g["Nx"], g["Ny"]
    q = 2 * np.pi / 10e-6
    y = (np.arange(Ny) - Ny / 2) * grid.dx
    H = np.broadcast_to((q * y)[:, None, None], (Ny, grid.Nz, Nx))
    out = OL.march_grid(g, [0.0], H=H)
    rows = slice(Ny // 2 - 60, Ny // 2 + 60)
    phi = np.unwrap(np.angle(out[0][rows, Nx // 2]))
    slope = np.polyfit(np.arange(len(phi)) * grid.dx, phi, 1)[0]
    assert slope == pytest.approx(-q, rel=0.05)


def test_winding_is_g_dot_b(diamond400, small_grid):
    Nx, Ny, dz, _, _ = small_grid.geometry()
    for sign in (+1, -1):
        seg = geometry_n(diamond400, small_grid, sign)
        H = OV.h_grid(seg, diamond400.g_vec, diamond400.nu,
                      diamond400.a_core, Nx, Ny, small_grid.Nz,
                      small_grid.dx, dz)
        plane = H[:, :, Nx // 2]
        (ra, rb, bv), = seg
        xi = (rb - ra) / np.linalg.norm(rb - ra)
        core = (ra + rb) / 2.0
        tp = (0.0 - core[0]) / xi[0]
        cy = Ny / 2 + (core[1] + tp * xi[1]) / small_grid.dx
        cz = (core[2] + tp * xi[2]) / dz
        ang = np.linspace(0, 2 * np.pi, 1440, endpoint=False)
        ys = np.clip((cy + 20 * np.cos(ang)).astype(int), 0, Ny - 1)
        zs = np.clip((cz + 20 * np.sin(ang)).astype(int), 0, small_grid.Nz - 1)
        p = plane[ys, zs]
        wind = (np.sum(np.angle(np.exp(1j * np.diff(p))))
                + np.angle(np.exp(1j * (p[0] - p[-1])))) / (2 * np.pi)
        assert wind == pytest.approx(g_dot_b(diamond400.g_vec, bv), abs=0.05)


def test_zero_burgers_and_superposition(diamond400, small_grid):
    Nx, Ny, dz, _, _ = small_grid.geometry()
    seg = geometry_n(diamond400, small_grid)
    dead = [(seg[0][0], seg[0][1], np.zeros(3))]
    args = (diamond400.g_vec, diamond400.nu, diamond400.a_core, Nx, Ny,
            small_grid.Nz, small_grid.dx, dz)
    assert np.abs(OV.h_grid(dead, *args)).max() == 0.0
    other = mixed_on_x(diamond400, small_grid)
    both = OV.h_grid(seg + other, *args)
    parts = OV.h_grid(seg, *args) + OV.h_grid(other, *args)
    assert np.abs(both - parts).max() < 1e-12 * np.abs(parts).max()


def test_closed_form_line_matches_grid_h(diamond400, small_grid):
    from dyfox_amd.dislocations import analytic_line_params
    Nx, Ny, dz, _, _ = small_grid.geometry()
    seg = mixed_on_x(diamond400, small_grid)
    lp = analytic_line_params(seg, diamond400)
    assert lp is not None
    H = OV.h_grid(seg, diamond400.g_vec, diamond400.nu, diamond400.a_core,
                  Nx, Ny, small_grid.Nz, small_grid.dx, dz)
    f = OL.analytic_h(lp, Ny, small_grid.dx, dz)
    rows = np.arange(Ny)
    for k in (0, 17, small_grid.Nz - 2):
        h = f(k, rows)[:, 0]
        ex = np.exp(1j * H[:, k, :]) - np.exp(1j * h)[:, None]
        assert np.abs(ex).max() < 1e-12


def test_paper_geometry_does_not_qualify(diamond400, small_grid):
    from dyfox_amd.dislocations import analytic_line_params
    assert analytic_line_params(geometry_n(diamond400, small_grid),
                                diamond400) is None


def test_integer_shift_interpolation_is_an_index_shift():
    F = np.arange(12.0) + 1j
    got = OL._interp(F, 1.0, 12)
    want = np.concatenate([F[1:], [0]])
    assert np.array_equal(got, want)
    got = OL._interp(F, -1.0, 12)
    assert np.array_equal(got, np.concatenate([[0], F[:-1]]))


def test_fov_rows(diamond400):
    grid = LaueGrid(diamond400, 200e-6, 0.25e-6, 120e-6)
    rows = OL.marched_rows(544, 0.25e-6, 60e-6)
    assert len(rows) == 481
    assert np.array_equal(rows, grid.marched_rows())
