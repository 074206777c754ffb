"""Physical constants and the rocking-angle <-> deviation-parameter maps.

The deviation parameter used everywhere in the engine is

    s = phi * sin(2 theta_B) / lambda          [1/m]

for a rocking angle phi in radians. The three conversions below each
evaluate their own fixed expression, in the order written, because
folding `sin_2tB / lam` into one constant changes the last bit of the
result (units.py:25-38 and tests/test_units.py:33-43 of the reference).
"""

import numpy as np

# Classical electron radius (m) and h*c (eV*m). Shared with crystal.py.
R_E = 2.8179403262e-15
HC_EV_M = 12398.419e-10


def param(xtal, key):
    """Look up a crystal quantity on a mapping or on an object."""
    try:
        return xtal[key]
    except (TypeError, IndexError, KeyError):
        return getattr(xtal, key)


def deviation_from_rad(phi, xtal):
    """Deviation parameter (1/m) for a rocking angle given in radians."""
    return np.asarray(phi) * param(xtal, "sin_2tB") / param(xtal, "lam")


def deviation_from_urad(phi_urad, xtal):
    """Deviation parameter (1/m) for a rocking angle in microradians."""
    return (np.asarray(phi_urad) * 1e-6 * param(xtal, "sin_2tB")
            / param(xtal, "lam"))


def urad_from_deviation(s_dev, xtal):
    """Rocking angle in microradians for a deviation parameter in 1/m."""
    return (np.asarray(s_dev) * param(xtal, "lam") / param(xtal, "sin_2tB")
            * 1e6)
