"""Host-side arithmetic: crystal scalars, unit conversions and grid rules.

The expected numbers are the reference's first-principles known-answer
values (tests/test_crystal.py, test_units.py, test_grid_optics.py:19-80),
derived independently of the code.
"""

import numpy as np
import pytest

from dyfox_amd import units
from dyfox_amd.crystal import DIAMOND, CrystalParams, Material, lab_frame
from dyfox_amd.laue import LaueGrid
from dyfox_amd.units import HC_EV_M, R_E, deviation_from_urad

A = 3.567e-10


def test_wavelength_and_bragg_angle(diamond400):
    assert diamond400.lam == pytest.approx(HC_EV_M / 17.0e3, rel=1e-15)
    assert diamond400.lam * 1e10 == pytest.approx(0.72932, abs=5e-6)
    d = A / 4.0
    assert diamond400.theta_B == pytest.approx(
        np.arcsin(diamond400.lam / (2 * d)), rel=1e-15)
    assert np.degrees(diamond400.theta_B) == pytest.approx(24.137, abs=1e-3)


def test_structure_factor(diamond400):
    S = np.sum(np.exp(2j * np.pi * DIAMOND.basis @ np.array([4, 0, 0])))
    assert abs(S) == pytest.approx(8.0, abs=1e-12)
    assert abs(diamond400.F_g) == pytest.approx(12.736, abs=1e-3)
    for hkl, want in (((2, 2, 0), 8.0), ((2, 0, 0), 0.0)):
        S = np.sum(np.exp(2j * np.pi * DIAMOND.basis @ np.array(hkl)))
        assert abs(S) == pytest.approx(want, abs=1e-12)


def test_absorption_sign_and_sigma(diamond400):
    assert diamond400.sig0.real == pytest.approx(-72.45, abs=0.1)
    F0 = (6.0 + 0.003 - 0.002j) * 8
    assert diamond400.sig0 == pytest.approx(
        -1j * R_E * diamond400.lam * F0 / A ** 3, rel=1e-14)
    assert diamond400.sig_h != np.conj(diamond400.sig_hbar)
    prod = diamond400.sig_h * diamond400.sig_hbar
    assert abs(prod.imag / prod.real) > 1e-6


def test_extinction_length_and_thickness(diamond400):
    assert diamond400.xi_g * 1e6 == pytest.approx(54.47, abs=0.02)
    half = diamond400.xi_g * diamond400.cos_tB
    t = (round(270e-6 / half - 0.5) + 0.5) * half
    assert t * 1e6 == pytest.approx(273.404, abs=0.002)
    assert t / diamond400.xi_g == pytest.approx(5.019, abs=0.005)


def test_g_vector_and_frames(diamond400):
    assert np.linalg.norm(diamond400.g_vec) == pytest.approx(
        1.0 / diamond400.d_hkl, rel=1e-14)
    U = lab_frame((0, 0, 1), (1, 1, 0))
    assert np.linalg.det(U) == pytest.approx(1.0, abs=1e-12)
    turned = diamond400.with_frame(U)
    assert turned.sig_h == diamond400.sig_h
    assert turned.g_vec[1] != 0.0
    with pytest.raises(ValueError, match="Bragg"):
        lab_frame((0, 0, 1), (0, 0, 4))
    with pytest.raises(KeyError, match="no f0"):
        CrystalParams.from_reflection(DIAMOND, (1, 1, 1), 17.0)


def test_other_material():
    si = Material(name="silicon", a_lat=5.431e-10, basis=DIAMOND.basis,
                  Z=14.0, nu=0.22, f0={(4, 0, 0): 8.66})
    x = CrystalParams.from_reflection(si, (4, 0, 0), 17.0)
    assert x.b_mag == pytest.approx(5.431e-10 * np.sqrt(2) / 2)
    assert x.theta_B < 24.0 * np.pi / 180


def test_220_at_12_kev():
    x = CrystalParams.from_reflection(DIAMOND, (2, 2, 0), 12.0,
                                      f_prime=0.003, f_dprime=0.002)
    assert np.degrees(x.theta_B) == pytest.approx(24.18190, abs=1e-4)
    assert x.xi_g * 1e6 == pytest.approx(31.168, abs=0.01)


def test_unit_conversions(diamond400):
    s = np.linspace(-10, 10, 81) / 54.47e-6
    back = units.deviation_from_urad(units.urad_from_deviation(s, diamond400),
                                     diamond400)
    assert np.allclose(back, s, rtol=1e-12)
    phi = np.linspace(-150e-6, 150e-6, 121)
    exact = units.deviation_from_rad(phi, diamond400)
    folded = phi * (diamond400.sin_2tB / diamond400.lam)
    assert np.allclose(exact, folded, rtol=1e-12)
    assert np.count_nonzero(exact != folded) > 0
    s = units.deviation_from_rad(117.5e-6, diamond400)
    assert s * 54.472e-6 == pytest.approx(65.5, abs=0.2)
    assert float(deviation_from_urad(117.5, diamond400)) == pytest.approx(
        1.20243e6, rel=1e-5)


def test_exact_grid_rules(diamond400):
    g = LaueGrid(diamond400, 273.4e-6, 0.25e-6, 120e-6)
    _, _, dz, _, spx = g.geometry()
    assert g.exact and spx == 1.0
    assert dz == pytest.approx(g.dx / diamond400.tan_tB, rel=1e-9)
    assert abs(g.t_crystal - 273.4e-6) < dz
    paper = LaueGrid(diamond400, 273.4037251665087e-6, 0.25e-6, 120e-6)
    assert paper.Nz == 491 and paper.shape == (544, 1056)
    forced = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6, Nz=301)
    assert not forced.exact and forced.geometry()[4] != 1.0
    fine = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6).replace(dx=0.125e-6)
    assert fine.exact and fine.geometry()[4] == 1.0
    for t in (20e-6, 100e-6, 273e-6):
        Ny, Nx = LaueGrid(diamond400, t, 0.25e-6, 53e-6).shape
        assert Nx % 32 == 0 and Ny % 32 == 0


def test_worked_example_grids(diamond400):
    """The sizes the benchmark configurations are quoted at."""
    g2 = LaueGrid(diamond400, 200e-6, 0.25e-6, 120e-6)
    assert g2.Nz == 359 and g2.shape == (544, 800)
    assert len(g2.marched_rows()) == 481
    assert g2.pixel_steps(2) == 481 * 800 * 358 * 2
    g3 = LaueGrid(diamond400, 273e-6, 0.25e-6, 120e-6, Ny=256)
    assert g3.Nz == 490 and g3.shape == (256, 1056)
    assert g3.pixel_steps(64) == 256 * 1056 * 489 * 64


def test_grid_key(diamond400):
    a = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    assert a.key() != a.replace(beam_thickness=2e-6).key()
    assert a.key() != a.replace(grid_pad=14).key()
    assert a.key() == a.replace().key()
    assert "Ny" not in a.key()


def test_planewave_rocking_is_the_closed_form(diamond400):
    from dyfox_amd.rocking import planewave_rocking
    from oracle import closed_forms as CF
    s = deviation_from_urad(np.linspace(-30, 30, 13), diamond400)
    for absorb in (True, False):
        want = CF.planewave_rocking(
            s, complex(diamond400.sig0), complex(diamond400.sig_h),
            complex(diamond400.sig_hbar), diamond400.cos_tB, 40e-6,
            absorption=absorb)
        got = planewave_rocking(s, diamond400, 40e-6, absorption=absorb)
        assert np.allclose(got, want, rtol=1e-12, atol=0)
