"""Quadrature nodes for partially coherent beams (host) - the properties
the reference's tests hold (test_detector.py:75-100) and the ones the
node definitions fix."""

import numpy as np
import pytest

from dyfox_amd import illumination as IL


def test_gauss_hermite_weights_and_moments():
    _, w = IL.gauss_hermite_offsets(1.0, 8)
    assert w.sum() == pytest.approx(1.0)
    sigma = 3.5e-6
    offs, w = IL.gauss_hermite_offsets(sigma, 8)
    assert (w * offs[:, 0]).sum() == pytest.approx(0.0, abs=1e-18)
    assert (w * offs[:, 0] ** 2).sum() == pytest.approx(sigma ** 2, rel=1e-9)
    assert (w * offs[:, 1] ** 2).sum() == pytest.approx(sigma ** 2, rel=1e-9)
    # an 8-point rule integrates x^4 exactly: 3 sigma^4
    assert (w * offs[:, 0] ** 4).sum() == pytest.approx(3 * sigma ** 4,
                                                        rel=1e-9)


def test_condenser_angles():
    a, w = IL.condenser_angles(0.90, reach=75.0)
    assert 0.0 in a and a.min() == pytest.approx(-75.0)
    assert a.max() == pytest.approx(75.0)
    assert np.all(np.diff(a) > 0) and len(w) == len(a)
    core = a[np.abs(a) <= 3.6]
    assert np.max(np.diff(core)) <= 0.90 + 1e-9
    assert np.allclose(w, np.gradient(a))


def test_convergence_bandwidth_nodes(diamond400):
    x = diamond400
    s, d, w, info = IL.convergence_bandwidth_nodes(x, 20e-6, 1e-4, 7, 60e-6,
                                                   5, 3e-4)
    assert info == {"n_total": 35, "n_kept": 35, "band_fraction": 1.0}
    assert w.sum() == pytest.approx(1.0) and np.all(np.diff(np.abs(s)) >= 0)
    tb = x.theta_B
    # matched="delta": every node sits on the Bragg condition
    s2, _, _, _ = IL.convergence_bandwidth_nodes(x, 20e-6, 1e-4, 7, 60e-6,
                                                 5, 3e-4, matched="delta")
    assert np.abs(s2).max() < 1e-9 * np.abs(s).max()
    # matched="psi": the deviation depends on psi only (5 copies each)
    s3, _, _, _ = IL.convergence_bandwidth_nodes(x, 20e-6, 1e-4, 7, 60e-6,
                                                 5, 3e-4, matched="psi")
    psi = np.linspace(-60e-6, 60e-6, 7)
    want = np.sort(np.abs(np.repeat(psi * x.sin_2tB / x.lam, 5)))
    assert np.allclose(np.abs(s3), want, rtol=1e-12, atol=1e-6)
    # vignetting keeps the nodes whose carrier stays inside, and their weight
    f = 0.5 * np.abs(x.sin_2tB * d / x.lam).max()
    s4, d4, w4, i4 = IL.convergence_bandwidth_nodes(
        x, 20e-6, 1e-4, 7, 60e-6, 5, 3e-4, vignette_f=f)
    keep = np.abs(x.sin_2tB * d / x.lam) <= f
    assert i4["n_kept"] == keep.sum() < 35
    assert i4["band_fraction"] == pytest.approx(w[keep].sum())
    assert np.array_equal(s4, s[keep]) and np.array_equal(d4, d[keep])
    assert np.tan(tb) > 0
