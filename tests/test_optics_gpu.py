"""Exit wave -> image on the device against the NumPy restatement.

Bars: pupils equal the oracle's (the NA disc on the same pixels; phases
to 1e-14), spectra and images to 1e-12 of their maximum in FP64 (5e-6 in
FP32), for even and odd grids; the COM maps equal the oracle's sums over
the same device stacks to 1e-12.
"""

import numpy as np
import pytest

from dyfox_amd import optics as DO
from dyfox_amd.backend import FP32, FP64, MIXED, to_numpy
from dyfox_amd.laue import solve_dislocation
from dyfox_amd.optics import Optics
from dyfox_amd.rocking import com_maps
from dyfox_amd.units import urad_from_deviation
from oracle import optics as OO

from .helpers import random_segments

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture
def lens(diamond400):
    return Optics(1.0e-4, 0.274, diamond400)


@pytest.mark.parametrize("Nx,Ny", [(256, 128), (97, 60)])
def test_pupils_match_oracle(dev, lens, diamond400, Nx, Ny):
    r = lens.aperture_f
    dx = 0.25e-6
    assert np.array_equal(to_numpy(lens.pupil(Nx, Ny, dx, 0)),
                          OO.pupil(Nx, Ny, dx, 0, r))
    for ell, lv in ((2, 0), (-1, 0), (2, 8)):
        got = to_numpy(lens.spiral(Nx, Ny, dx, ell, levels=lv))
        want = OO.pupil(Nx, Ny, dx, ell, r, levels=lv)
        assert np.array_equal(got != 0, want != 0)
        assert np.abs(got - want).max() < 1e-14
    off = (2e-6, -1e-6)
    sf = (off[0] / (diamond400.lam * 0.274), off[1] / (diamond400.lam * 0.274))
    got = to_numpy(lens.spiral(Nx, Ny, dx, 2, offset_bfp=off))
    assert np.abs(got - OO.pupil(Nx, Ny, dx, 2, r, singularity_f=sf)).max() \
        < 1e-14
    dfx = diamond400.sin_2tB * 3e-4 / diamond400.lam
    got = to_numpy(lens.chromatic(Nx, Ny, dx, 2, 3e-4))
    want = OO.pupil(Nx, Ny, dx, 2, r, aperture_f=(dfx, 0.0),
                    singularity_f=(dfx, 0.0), dc_zero="never")
    assert np.array_equal(got != 0, want != 0)
    assert np.abs(got - want).max() < 1e-14
    f32 = Optics(1.0e-4, 0.274, diamond400, precision=FP32)
    p32 = to_numpy(f32.spiral(Nx, Ny, dx, 2))
    assert p32.dtype == np.complex64
    assert np.array_equal(p32, OO.pupil(Nx, Ny, dx, 2, r).astype(np.complex64))
    assert lens.spiral(Nx, Ny, dx, 2) is lens.spiral(Nx, Ny, dx, 2)


@pytest.mark.parametrize("Nx,Ny", [(64, 48), (45, 33)])
def test_spectrum_and_images_match_oracle(dev, lens, rng, Nx, Ny):
    dx = 0.25e-6
    stack = rng.normal(size=(3, Ny, Nx)) + 1j * rng.normal(size=(3, Ny, Nx))
    assert _rel(to_numpy(DO.object_spectrum(stack)),
                OO.object_spectrum(stack)) < 1e-13
    assert _rel(to_numpy(DO.object_spectrum(stack[1])),
                OO.object_spectrum(stack[1])) < 1e-13
    pups = [to_numpy(lens.spiral(Nx, Ny, dx, e)) for e in (0, 2, -2)]
    spec = OO.object_spectrum(stack)
    for p in pups:
        want = OO.image_from_spectrum(spec, p)
        assert _rel(to_numpy(DO.image_from_spectrum(spec, p)), want) < 1e-12
        assert _rel(to_numpy(DO.image_intensity(stack, p)), want) < 1e-12
    per = np.stack(pups)
    assert _rel(to_numpy(DO.image_from_spectrum(spec, per)),
                OO.image_from_spectrum(spec, per)) < 1e-12
    assert _rel(to_numpy(DO.apply_pupils(stack[0], pups)),
                OO.apply_pupils(stack[0], pups)) < 1e-12
    assert _rel(to_numpy(lens.images(stack[0], dx, ells=(0, 2, -2))),
                OO.apply_pupils(stack[0], pups)) < 1e-12
    s32 = stack.astype(np.complex64)
    lens32 = Optics(1.0e-4, 0.274, lens.xtal, precision=FP32)
    got = to_numpy(lens32.image(s32, dx, ell=2))
    assert got.dtype == np.float32
    want = OO.image_from_spectrum(OO.object_spectrum(s32.astype(complex)),
                                  OO.pupil(Nx, Ny, dx, 2, lens.aperture_f))
    assert _rel(got, want) < 5e-6


def test_the_aperture_can_only_remove_energy(dev, diamond400, small_grid,
                                             lens, rng):
    Ny, Nx = small_grid.shape
    segs = random_segments(rng, diamond400, small_grid, 2)
    stack = solve_dislocation(segs, [0.0], small_grid, diamond400)
    e_in = float((np.abs(to_numpy(stack)[0]) ** 2).sum())
    e_open = float(to_numpy(DO.apply_pupils(
        stack[0], [lens.spiral(Nx, Ny, small_grid.dx, 0)])).sum())
    e_spiral = float(to_numpy(DO.apply_pupils(
        stack[0], [lens.spiral(Nx, Ny, small_grid.dx, 2)])).sum())
    assert e_open <= e_in * 1.001
    assert e_spiral <= e_open * 1.001


def test_a_matched_plate_fills_in_a_vortex(dev, diamond400):
    lens = Optics(3e-3, 0.274, diamond400)
    n, dx = 128, 0.25e-6
    x = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(x, x)
    env = np.exp(-(X ** 2 + Y ** 2) / (2 * (3e-6) ** 2))
    c = slice(n // 2 - 2, n // 2 + 2)
    for m in (+2, -2):
        field = env * np.exp(1j * m * np.arctan2(Y, X))
        matched = to_numpy(lens.image(field, dx, ell=-m))
        opposed = to_numpy(lens.image(field, dx, ell=+m))
        assert matched[c, c].max() > 20 * opposed[c, c].max()


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_com_maps_match_the_oracle_sums(dev, diamond400, small_grid, lens,
                                        rng, prec):
    segs = random_segments(rng, diamond400, small_grid, 2)
    s = np.linspace(-4e5, 4e5, 7)
    com, rock, weight = com_maps(segs, s, small_grid, diamond400, lens,
                                 ells=(0, 2), batch=3, precision=prec)
    Ny, Nx = small_grid.shape
    stack = to_numpy(solve_dislocation(segs, s, small_grid, diamond400,
                                       precision=prec)).astype(complex)
    phi = urad_from_deviation(s, diamond400)
    for ell in (0, 2):
        imgs = OO.image_from_spectrum(OO.object_spectrum(stack),
                                      OO.pupil(Nx, Ny, small_grid.dx, ell,
                                               lens.aperture_f))
        a, b, t = OO.com_sums(imgs, phi)
        assert com[ell].shape == (Ny, Nx) and len(rock[ell]) == len(s)
        assert np.isfinite(com[ell]).all() and weight[ell].sum() > 0
        tol = 1e-12 if prec is FP64 else 5e-6
        assert _rel(weight[ell], b) < tol and _rel(rock[ell], t) < tol
        assert _rel(com[ell] * b, a) < tol



def test_incoherent_image_matches_the_oracle_sum(dev, diamond400, small_grid,
                                                 lens, rng):
    from dyfox_amd.illumination import (convergence_bandwidth_nodes,
                                        incoherent_image)
    segs = random_segments(rng, diamond400, small_grid, 2)
    s, d, w, _ = convergence_bandwidth_nodes(diamond400, 20e-6, 1e-4, 3,
                                             60e-6, 3, 3e-4)
    Ny, Nx = small_grid.shape
    stack = to_numpy(solve_dislocation(segs, s, small_grid,
                                       diamond400)).astype(complex)
    spec = OO.object_spectrum(stack)
    dfx = diamond400.sin_2tB * d / diamond400.lam
    for ell in (0, 2):
        want = sum(w[n] * OO.image_from_spectrum(
            spec[n], OO.pupil(Nx, Ny, small_grid.dx, ell, lens.aperture_f,
                              aperture_f=(dfx[n], 0.0),
                              singularity_f=(dfx[n], 0.0), dc_zero="never"))
            for n in range(len(s)))
        # four nodes per march, and (the default) every node in one march
        for batch in (4, None):
            got = to_numpy(incoherent_image(segs, small_grid, diamond400,
                                            lens, s, w, d, ell=ell,
                                            batch=batch))
            assert _rel(got, want) < 1e-12, batch
    plain = to_numpy(incoherent_image(segs, small_grid, diamond400, lens, s,
                                      w, ell=2, batch=5))
    pup = OO.pupil(Nx, Ny, small_grid.dx, 2, lens.aperture_f)
    want = sum(w[n] * OO.image_from_spectrum(spec[n], pup)
               for n in range(len(s)))
    assert _rel(plain, want) < 1e-12
