"""Materials and the two-beam diffraction scalars a reflection implies.

    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)

Mirrors the reference's crystal.py:25-200 (same attribute names, same
formulae, cubic lattice and symmetric geometry), because every march
coefficient is derived from these numbers:

    lambda   = hc / E
    theta_B  = asin(lambda / 2 d_hkl),      d_hkl = a / |hkl|
    g_vec    = U_lab @ hkl / a              (1/m, no 2 pi)
    F_0      = (Z + f' - i f'') * N_basis   (forward scattering)
    S_g      = sum_j exp(2 pi i r_j . hkl)
    F_g      = f S_g,  F_-g = f conj(S_g),  f = f0 + f' - i f''
    sigma_X  = -i r_e lambda F_X / V_cell
    xi_g     = pi / |sigma_h|
"""

from dataclasses import dataclass, replace

import numpy as np

from .units import HC_EV_M, R_E

# Conventional diamond-cubic cell: fcc lattice plus the (1/4,1/4,1/4) copy.
_FCC = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0],
                 [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
DIAMOND_CUBIC = np.concatenate([_FCC, _FCC + 0.25]) % 1.0


@dataclass(frozen=True)
class Material:
    """A substance: lattice, basis, Z, Poisson ratio and form factors.

    `f0` maps (h, k, l) to the form factor f0(q) at that reflection.
    `burgers_mag` defaults to |a/2 <110>|; `core_radius` (where the
    Volterra field is regularised) defaults to the Burgers magnitude.
    """

    name: str
    a_lat: float
    basis: np.ndarray
    Z: float
    nu: float
    f0: dict
    burgers_mag: float = None
    core_radius: float = None

    def burgers(self):
        if self.burgers_mag is not None:
            return self.burgers_mag
        return self.a_lat * np.sqrt(2) / 2

    def core(self):
        if self.core_radius is not None:
            return self.core_radius
        return self.burgers()


# Poisson ratio: Voigt-Reuss-Hill average of C11=1079, C12=124, C44=578 GPa.
DIAMOND = Material(name="diamond", a_lat=3.567e-10, basis=DIAMOND_CUBIC,
                   Z=6.0, nu=0.069,
                   f0={(4, 0, 0): 1.589, (2, 2, 0): 1.961})


def lab_frame(normal_hkl, g_hkl):
    """Rotation whose rows are lab x, y, z: z along the surface normal,
    x along g projected into the surface, y = z cross x."""
    z = np.asarray(normal_hkl, float)
    z = z / np.linalg.norm(z)
    g = np.asarray(g_hkl, float)
    x = g - np.dot(g, z) * z
    if np.linalg.norm(x) < 1e-12:
        raise ValueError("g is parallel to the surface normal: that is a "
                         "Bragg geometry, not a symmetric Laue one")
    x = x / np.linalg.norm(x)
    return np.array([x, np.cross(z, x), z])


@dataclass(frozen=True)
class CrystalParams:
    """One material at one reflection and photon energy, worked through."""

    material: Material
    hkl: tuple
    E_keV: float
    U_lab: np.ndarray
    label: str
    lam: float
    d_hkl: float
    theta_B: float
    F_g: complex
    sig0: complex
    sig_h: complex
    sig_hbar: complex
    xi_g: float
    sin_tB: float
    cos_tB: float
    tan_tB: float
    sin_2tB: float
    a_lat: float
    nu: float
    g_vec: np.ndarray
    b_mag: float
    a_core: float

    @classmethod
    def from_reflection(cls, material, hkl, E_keV, U_lab=None, f0=None,
                        f_prime=0.0, f_dprime=0.0, label=None):
        """Two-beam scalars for `hkl` at `E_keV` (crystal.py:117-165)."""
        hkl = tuple(int(v) for v in hkl)
        h = np.array(hkl)
        U_lab = np.eye(3) if U_lab is None else U_lab
        if f0 is None:
            if hkl not in material.f0:
                raise KeyError(
                    f"{material.name} has no f0 for {hkl}; pass f0=... or "
                    f"add it to the material (have {sorted(material.f0)})")
            f0 = material.f0[hkl]
        a = material.a_lat
        lam = HC_EV_M / (E_keV * 1e3)
        d_hkl = a / np.sqrt(np.sum(h ** 2))
        theta_B = np.arcsin(lam / (2 * d_hkl))
        volume = a ** 3
        f_atom = f0 + f_prime - 1j * f_dprime
        geometric = np.sum(np.exp(2j * np.pi * material.basis @ h))
        F_0 = (material.Z + f_prime - 1j * f_dprime) * len(material.basis)
        F_g = f_atom * geometric
        F_mg = f_atom * np.conj(geometric)
        sig0, sig_h, sig_hbar = (-1j * R_E * lam * F / volume
                                 for F in (F_0, F_g, F_mg))
        if label is None:
            label = (f"{material.name.capitalize()}"
                     f"({''.join(str(v) for v in hkl)})")
        return cls(material=material, hkl=hkl, E_keV=float(E_keV),
                   U_lab=U_lab, label=label, lam=lam, d_hkl=d_hkl,
                   theta_B=theta_B, F_g=F_g, sig0=sig0, sig_h=sig_h,
                   sig_hbar=sig_hbar, xi_g=np.pi / abs(sig_h),
                   sin_tB=np.sin(theta_B), cos_tB=np.cos(theta_B),
                   tan_tB=np.tan(theta_B), sin_2tB=np.sin(2 * theta_B),
                   a_lat=a, nu=material.nu, g_vec=U_lab @ (h / a),
                   b_mag=material.burgers(), a_core=material.core())

    def at_energy(self, E_keV, **kw):
        """Same material, reflection and frame at another energy."""
        return CrystalParams.from_reflection(self.material, self.hkl, E_keV,
                                             U_lab=self.U_lab,
                                             label=self.label, **kw)

    def with_frame(self, U_lab):
        """Same scalars, crystal re-mounted: only U_lab and g_vec move."""
        U_lab = np.asarray(U_lab, float)
        return replace(self, U_lab=U_lab,
                       g_vec=U_lab @ (np.array(self.hkl) / self.a_lat))

    def key(self):
        """Identity of this crystal for a cache key."""
        return {"material": self.material.name, "hkl": list(self.hkl),
                "E_keV": self.E_keV, "label": self.label}

    def __getitem__(self, name):
        return getattr(self, name)

    def report(self):
        return "\n".join([
            f"{self.label} at {self.E_keV} keV",
            f"  lambda   = {self.lam * 1e10:.4f} A",
            f"  theta_B  = {np.degrees(self.theta_B):.2f} deg",
            f"  xi_g     = {self.xi_g * 1e6:.1f} um",
            f"  |F_g|    = {abs(self.F_g):.3f}",
            f"  Re sig0  = {self.sig0.real:+.1f} /m"])
