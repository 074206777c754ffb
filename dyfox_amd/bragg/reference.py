"""Perfect-crystal reflectivities in closed form (host, NumPy).

Same functions and signatures as the reference's bragg/reference.py:14-67.
`reflectivity_riccati` there integrates the Riccati equation with SciPy;
here the equation is solved exactly, which gives the same answer (to the
integrator's tolerance) with no integrator:

    dR/dz = -(sigma_h + B R + sigma_hbar R^2) / sin tB,  R(t) = 0,
    B = 2 sigma_0 + i alpha,  alpha = -2 pi s
    R(0) = r1 r2 (1 - E) / (r2 - r1 E),  E = exp(q t / sin tB),
    r1,2 = (-B +- q) / (2 sigma_hbar),  q = sqrt(B^2 - 4 sigma_h sigma_hbar)
"""

import numpy as np


def _roots(s_dev, xtal):
    s0, sh, sb = complex(xtal.sig0), complex(xtal.sig_h), complex(xtal.sig_hbar)
    alpha = -2.0 * np.pi * np.atleast_1d(np.asarray(s_dev, float))
    B = 2.0 * s0 + 1j * alpha
    q = np.sqrt(B ** 2 - 4.0 * sh * sb + 0j)
    return q, (-B + q) / (2.0 * sb), (-B - q) / (2.0 * sb)


def reflectivity_semi_infinite(s_dev, xtal):
    """Amplitude reflectivity of a semi-infinite perfect crystal: the
    stationary root of the Riccati equation with |R| <= 1."""
    _, r1, r2 = _roots(s_dev, xtal)
    return np.where(np.abs(r1) <= np.abs(r2), r1, r2)


def reflectivity_riccati(s_dev, t_crystal, xtal, rtol=1e-11, atol=1e-14):
    """Finite-thickness reflectivity R(0) (exact; rtol/atol are accepted
    for signature compatibility and unused)."""
    q, r1, r2 = _roots(s_dev, xtal)
    x = q * float(t_crystal) / xtal.sin_tB
    out = np.empty_like(q)
    big = x.real > 0
    e = np.exp(-x[big])
    out[big] = r1[big] * r2[big] * (e - 1.0) / (r2[big] * e - r1[big])
    e = np.exp(x[~big])
    out[~big] = r1[~big] * r2[~big] * (1.0 - e) / (r2[~big] - r1[~big] * e)
    return out


def darwin_width_urad(xtal):
    """Full width of the total-reflection domain (urad), symmetric Bragg:
    2 |sqrt(chi_h chi_hbar)| / sin 2tB with chi = sigma lambda / (i pi)."""
    chi = (np.sqrt(complex(xtal.sig_h) * complex(xtal.sig_hbar)) * xtal.lam
           / (1j * np.pi))
    return float(2.0 * abs(chi) / xtal.sin_2tB * 1e6)
