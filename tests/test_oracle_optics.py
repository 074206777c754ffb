"""Pins the NumPy imaging restatement (oracle/optics.py) with the
properties the reference's own optics tests hold (test_grid_optics.py:
83-230): the pupil's disc, winding, origin-bin rules, walk and
quantisation, and the transforms' identity, Parseval and energy
behaviour, plus a brute-force check of the COM sums."""

import numpy as np
import pytest

from oracle import optics as OO


@pytest.fixture
def radius(diamond400):
    return 1.0e-4 / diamond400.lam


def test_open_pupil_is_a_disc_of_the_right_radius(radius):
    p = OO.pupil(256, 256, 0.25e-6, 0, radius)
    f = np.fft.fftshift(np.fft.fftfreq(256, 0.25e-6))
    FX, FY = np.meshgrid(f, f)
    assert np.array_equal(p != 0, np.hypot(FX, FY) <= radius)
    assert set(np.unique(p.real)) <= {0.0, 1.0}


def test_spiral_winds_by_ell(radius):
    p = OO.pupil(256, 256, 0.25e-6, 2, radius)
    th = np.linspace(0, 2 * np.pi, 720, endpoint=False)
    ys = np.round(128 + 6 * np.sin(th)).astype(int)
    xs = np.round(128 + 6 * np.cos(th)).astype(int)
    ph = np.unwrap(np.angle(p[ys, xs]))
    assert (ph[-1] - ph[0]) / (2 * np.pi) == pytest.approx(2.0, abs=0.05)


def test_origin_bin_rules(radius, diamond400):
    assert OO.pupil(256, 128, 0.25e-6, 2, radius)[64, 128] == 0
    off = 2e-6 / (diamond400.lam * 0.274)
    assert OO.pupil(256, 128, 0.25e-6, 2, radius,
                    singularity_f=(off, 0.0))[64, 128] != 0
    assert OO.pupil(256, 128, 0.25e-6, 0, radius)[64, 128] == 1
    step = 1.0 / (256 * 0.25e-6)
    half = OO.pupil(256, 128, 0.25e-6, 2, radius, dc_zero="half_bin",
                    singularity_f=(0.3 * step, 0.0))
    assert half[64, 128] == 0
    assert OO.pupil(256, 128, 0.25e-6, 2, radius, dc_zero="never")[64, 128] != 0


def test_chromatic_walk_and_quantisation(radius, diamond400):
    dfx = diamond400.sin_2tB * 3e-4 / diamond400.lam
    walked = OO.pupil(512, 512, 0.25e-6, 0, radius, aperture_f=(dfx, 0.0),
                      singularity_f=(dfx, 0.0), dc_zero="never")
    f = np.fft.fftshift(np.fft.fftfreq(512, 0.25e-6))
    assert walked[256, np.argmin(np.abs(f - dfx))] != 0
    q = OO.pupil(256, 256, 0.25e-6, 2, radius, levels=8)
    ph = np.angle(q[q != 0]) % (2 * np.pi)
    assert len(np.unique(np.round(ph, 9))) <= 8
    assert OO.pupil(256, 128, 0.25e-6, 2, radius).shape == (128, 256)


def test_all_pass_identity_parseval_and_energy(radius, rng):
    field = rng.normal(size=(64, 48)) + 1j * rng.normal(size=(64, 48))
    out = OO.image_from_spectrum(OO.object_spectrum(field),
                                 np.ones((64, 48), complex))
    assert np.allclose(out, np.abs(field) ** 2, rtol=1e-10, atol=1e-12)
    spec = OO.object_spectrum(field)
    assert (np.abs(spec) ** 2).sum() == pytest.approx(
        (np.abs(field) ** 2).sum() * field.size, rel=1e-10)
    e_out = OO.image_from_spectrum(
        spec, OO.pupil(48, 64, 0.25e-6, 0, radius)).sum()
    assert e_out <= (np.abs(field) ** 2).sum() * (1 + 1e-9)


def test_a_matched_plate_fills_in_a_vortex(diamond400):
    r = 3e-3 / diamond400.lam
    n, dx = 128, 0.25e-6
    x = (np.arange(n) - n / 2) * dx
    X, Y = np.meshgrid(x, x)
    env = np.exp(-(X ** 2 + Y ** 2) / (2 * (3e-6) ** 2))
    for m in (+2, -2):
        field = env * np.exp(1j * m * np.arctan2(Y, X))
        spec = OO.object_spectrum(field)
        matched = OO.image_from_spectrum(spec, OO.pupil(n, n, dx, -m, r))
        opposed = OO.image_from_spectrum(spec, OO.pupil(n, n, dx, +m, r))
        c = slice(n // 2 - 2, n // 2 + 2)
        assert matched[c, c].max() > 20 * opposed[c, c].max()


def test_com_sums_by_loops(rng):
    I = rng.random((3, 4, 5))
    w = rng.normal(size=3)
    a, b, t = OO.com_sums(I, w)
    for y in range(4):
        for x in range(5):
            assert a[y, x] == pytest.approx(sum(w[i] * I[i, y, x]
                                                for i in range(3)))
            assert b[y, x] == pytest.approx(sum(I[i, y, x] for i in range(3)))
    assert np.allclose(t, [I[i].sum() for i in range(3)])
