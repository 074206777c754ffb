"""Closed forms the solvers are checked against (oracle; test use only).

    pendellosung_discrete   the march's own transfer matrix raised to
                            Nz-1: for H = 0 and a beam wider than the fan,
                            interior columns of Dg equal (M^(Nz-1))_21
    planewave_rocking       |sigma_h|^2/gam^2 sin^2(gam s) e^{2 Re s0 s},
                            gam^2 = |sigma_h sigma_hbar| + alpha^2/4
                            (the reference's analysis/rocking.py:16-34)
    bragg_discrete_perfect  -T21/T22 of the shooting step matrix^(Nz-1)
                            (bragg/solver.py:32-33, 54-57 at mode 0)
    reflectivity_semi_infinite, reflectivity_finite
                            Riccati dR/dz = -(sh + B R + sb R^2)/sin tB,
                            B = 2 s0 + i alpha, R(t) = 0: stationary root,
                            and the exact solution
                            R(0) = r1 r2 (1 - E)/(r2 - r1 E),
                            r1,2 = (-B +- q)/(2 sb), q = sqrt(B^2 - 4 sh sb),
                            E = exp(q t / sin tB)
    darwin_width_urad       2 |sqrt(chi_h chi_hbar)| / sin 2tB
"""

import numpy as np


def laue_step_matrix(s_dev, sig0, sig_h, sig_hbar, ds):
    """2x2 per-step matrix of the Laue march with H = 0."""
    alpha = -2.0 * np.pi * float(s_dev)
    gam = np.sqrt(abs(sig_h) ** 2 + alpha ** 2 / 4)
    cg, sg = np.cos(gam * ds), np.sin(gam * ds) / gam
    ep = np.exp(sig0.real * ds) * np.exp(1j * (sig0.imag + alpha / 2) * ds)
    return ep * np.array([[cg - 0.5j * alpha * sg, sig_hbar * sg],
                          [sig_h * sg, cg + 0.5j * alpha * sg]])


def pendellosung_discrete(s_dev, sig0, sig_h, sig_hbar, ds, Nz):
    """Dg after Nz-1 steps of a uniform plane wave (D0 = 1, Dg = 0)."""
    M = laue_step_matrix(s_dev, sig0, sig_h, sig_hbar, ds)
    return np.linalg.matrix_power(M, Nz - 1)[1, 0]


def planewave_rocking(s_dev, sig0, sig_h, sig_hbar, cos_tB, t_crystal,
                      absorption=True):
    s_dev = np.asarray(s_dev, float)
    alpha = -2 * np.pi * s_dev
    gam = np.sqrt(np.abs(sig_h * sig_hbar) + alpha ** 2 / 4)
    s = t_crystal / cos_tB
    rc = np.abs(sig_h) ** 2 / gam ** 2 * np.sin(gam * s) ** 2
    if absorption:
        rc = rc * np.exp(2 * sig0.real * s)
    return rc


def bragg_step_matrix(s_dev, sig0, sig_h, sig_hbar, ds):
    """Per-step matrix of the shooting march on a uniform field."""
    alpha = -2.0 * np.pi * float(s_dev)
    e0 = np.exp(sig0 * ds)
    eg = np.exp((sig0 + 1j * alpha) * ds)
    return np.array([[e0, e0 * ds * sig_hbar], [-ds * sig_h / eg, 1 / eg]])


def bragg_discrete_perfect(s_dev, sig0, sig_h, sig_hbar, ds, Nz):
    """Dg(0) = -T21/T22 for a perfect crystal under a uniform plane wave."""
    T = np.linalg.matrix_power(bragg_step_matrix(s_dev, sig0, sig_h,
                                                 sig_hbar, ds), Nz - 1)
    return -T[1, 0] / T[1, 1]


def _roots(s_dev, sig0, sig_h, sig_hbar):
    alpha = -2.0 * np.pi * np.atleast_1d(np.asarray(s_dev, float))
    B = 2.0 * sig0 + 1j * alpha
    q = np.sqrt(B * B - 4.0 * sig_h * sig_hbar + 0j)
    r1 = (-B + q) / (2.0 * sig_hbar)
    r2 = (-B - q) / (2.0 * sig_hbar)
    return q, r1, r2


def reflectivity_semi_infinite(s_dev, sig0, sig_h, sig_hbar):
    """The Riccati root with |R| <= 1."""
    _, r1, r2 = _roots(s_dev, sig0, sig_h, sig_hbar)
    return np.where(np.abs(r1) <= np.abs(r2), r1, r2)


def reflectivity_finite(s_dev, t_crystal, sig0, sig_h, sig_hbar, sin_tB):
    """Exact R(0) of the finite-thickness Riccati problem, R(t) = 0."""
    q, r1, r2 = _roots(s_dev, sig0, sig_h, sig_hbar)
    x = q * t_crystal / sin_tB
    out = np.empty_like(q)
    big = x.real > 0
    # E = e^{x}: divide through by E where it is large
    Ei = np.exp(-x[big])
    out[big] = r1[big] * r2[big] * (Ei - 1.0) / (r2[big] * Ei - r1[big])
    E = np.exp(x[~big])
    out[~big] = r1[~big] * r2[~big] * (1.0 - E) / (r2[~big] - r1[~big] * E)
    return out


def darwin_width_urad(sig_h, sig_hbar, lam, sin_2tB):
    chi = np.sqrt(sig_h * sig_hbar) * lam / (1j * np.pi)
    return float(2.0 * abs(chi) / sin_2tB * 1e6)
