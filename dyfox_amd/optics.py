"""The imaging stage on the device: an objective, a pupil, a spiral plate.

    lens = Optics(NA=1e-4, f_lens=0.274, xtal=xtal)
    open_img = lens.image(stack, dx, ell=0)
    spiral = lens.image(stack, dx, ell=+2)

Drop-in for the reference's optics.py (optics.py:1-207): the same class,
methods, module functions, arguments and pupil conventions.  A pupil of
charge `ell` carries e^{i ell phi} inside the NA disc, on the fftshifted
frequency grid; metres in the back focal plane map to spatial
frequencies through x_bfp = lambda f fx.  Pupils, spectra and images are
`DeviceArray`s computed by libdyfox.so (pupil kernel, batched hipFFT 2-D
transforms, fused pupil multiply and |.|^2); host (NumPy) inputs are
uploaded.  `xp` is accepted for compatibility; everything runs on the
device.  Results agree with the reference's NumPy/CuPy expressions to
rounding (a pupil's NA edge lands on the same pixels: the frequencies
are formed exactly as numpy.fft.fftfreq forms them).
"""

import ctypes as C

import numpy as np

from . import _native as N
from .backend import FP64, precision as _precision

_FP64, _FP32 = 0, 1


def _code(dtype):
    return _FP64 if np.dtype(dtype) in (np.complex128, np.float64) else _FP32


def _device(a, dtype=None):
    if isinstance(a, N.DeviceArray):
        return a
    a = np.asarray(a)
    if dtype is not None:
        a = a.astype(dtype)
    elif a.dtype not in (np.complex64, np.complex128):
        a = a.astype(np.complex128)
    return N.DeviceArray.from_host(np.ascontiguousarray(a))


class Optics:
    """An objective of given numerical aperture, viewing a given crystal.
    Pupils are cached on every argument that shapes them."""

    def __init__(self, NA, f_lens, xtal, *, precision=FP64, xp=None):
        self.NA = float(NA)
        self.f_lens = float(f_lens)
        self.xtal = xtal
        self.precision = _precision(precision)
        self.xp = xp
        self._cache = {}

    @property
    def aperture_f(self):
        """Radius of the NA disc in spatial frequency."""
        return self.NA / self.xtal.lam

    def resolution(self):
        """Rayleigh resolution, as a rough sanity number in metres."""
        return 0.61 * self.xtal.lam / self.NA

    def pupil(self, Nx, Ny, dx, ell=0, *, aperture_f=(0.0, 0.0),
              singularity_f=(0.0, 0.0), dc_zero="exact", levels=0,
              cache=True):
        """The pupil (Ny, Nx) on the fftshifted frequency grid (device).

        `aperture_f`, `singularity_f`: frequency offsets of the aperture
        stop and the spiral singularity; `levels` quantises the spiral
        phase; `dc_zero` ("exact", "half_bin", "never") decides the bin
        at the origin (optics.py:67-127)."""
        ax, ay = float(aperture_f[0]), float(aperture_f[1])
        sx, sy = float(singularity_f[0]), float(singularity_f[1])
        key = (Nx, Ny, float(dx), int(ell), ax, ay, sx, sy, dc_zero,
               int(levels), self.precision.name, self.NA, self.xtal.lam)
        if cache and key in self._cache:
            return self._cache[key]
        fx = np.fft.fftshift(np.fft.fftfreq(Nx, dx))
        zero = int(ell) != 0 and self._zero_dc(dc_zero, sx, sy, fx)
        a = N.PupilArgs(Nx=int(Nx), Ny=int(Ny), dx=float(dx),
                        aperture_f=self.aperture_f, ax=ax, ay=ay, sx=sx,
                        sy=sy, ell=int(ell), levels=int(levels),
                        zero_dc=int(zero), reserved=0)
        out = N.DeviceArray.empty((Ny, Nx), self.precision.complex)
        N.call("dyf_optics_pupil", C.byref(a), _code(self.precision.complex),
               C.c_void_p(out.ptr), None)
        if cache:
            self._cache[key] = out
        return out

    @staticmethod
    def _zero_dc(policy, sx, sy, fx):
        if policy == "never":
            return False
        if policy == "exact":
            return sx == 0.0 and sy == 0.0
        if policy == "half_bin":
            step = float(fx[1] - fx[0])
            return abs(sx) < 0.5 * step and abs(sy) < 0.5 * step
        raise ValueError(f"dc_zero is 'exact', 'half_bin' or 'never', "
                         f"not {policy!r}")

    def spiral(self, Nx, Ny, dx, ell=0, offset_bfp=(0.0, 0.0), levels=0,
               cache=True):
        """Pupil with the plate displaced by `offset_bfp` metres in the
        back focal plane (the aperture stays put)."""
        dfx = offset_bfp[0] / (self.xtal.lam * self.f_lens)
        dfy = offset_bfp[1] / (self.xtal.lam * self.f_lens)
        return self.pupil(Nx, Ny, dx, ell, singularity_f=(dfx, dfy),
                          dc_zero="exact", levels=levels, cache=cache)

    def chromatic(self, Nx, Ny, dx, ell, delta_E, levels=0,
                  dc_zero="never", cache=True):
        """Pupil seen at relative energy offset delta_E: aperture and
        singularity slide by sin(2 theta_B) delta_E / lambda."""
        dfx = self.xtal.sin_2tB * delta_E / self.xtal.lam
        return self.pupil(Nx, Ny, dx, ell, aperture_f=(dfx, 0.0),
                          singularity_f=(dfx, 0.0), dc_zero=dc_zero,
                          levels=levels, cache=cache)

    def image(self, stack, dx, ell=0, offset_bfp=(0.0, 0.0), levels=0):
        """Image intensity of a stack of exit waves through one pupil."""
        Ny, Nx = stack.shape[-2:]
        pup = self.spiral(Nx, Ny, dx, ell, offset_bfp, levels)
        return image_intensity(stack, pup)

    def images(self, field, dx, ells=(0, 2), offset_bfp=(0.0, 0.0)):
        """One exit wave through several charges, sharing one spectrum."""
        Ny, Nx = field.shape[-2:]
        pups = [self.spiral(Nx, Ny, dx, e, offset_bfp) for e in ells]
        return apply_pupils(field, pups)

    def clear_cache(self):
        """Forget the cached pupils."""
        self._cache.clear()

    def __repr__(self):
        return (f"Optics(NA={self.NA:g}, f={self.f_lens:g} m, "
                f"{self.xtal.label}, {int(self.resolution() * 1e9)} nm)")


def _images(field, pupils):
    """(n, npup, Ny, Nx) intensities of n fields through npup pupils."""
    f = _device(field)
    Ny, Nx = f.shape[-2:]
    n = 1 if f.ndim == 2 else f.shape[0]
    pdt = pupils[0].dtype if isinstance(pupils[0], N.DeviceArray) else \
        np.asarray(pupils[0]).dtype
    pdt = np.complex128 if np.dtype(pdt) == np.complex128 else np.complex64
    if len(pupils) == 1:
        P = _device(pupils[0], pdt)
    else:
        P = N.DeviceArray.empty((len(pupils), Ny, Nx), pdt)
        for i, p in enumerate(pupils):
            P[i].copy_from(_device(p, pdt))
    fc, pc = _code(f.dtype), _code(pdt)
    nb = C.c_size_t()
    N.call("dyf_optics_image_bytes", n, len(pupils), fc, pc, Nx, Ny,
           C.byref(nb))
    work = N.DeviceArray.empty((max(nb.value, 1),), np.uint8)
    odt = np.float64 if _FP64 in (fc, pc) else np.float32
    out = N.DeviceArray.empty((n, len(pupils), Ny, Nx), odt)
    N.call("dyf_optics_image", C.c_void_p(f.ptr), n, fc, C.c_void_p(P.ptr),
           len(pupils), pc, Nx, Ny, C.c_void_p(work.ptr), C.c_void_p(out.ptr),
           None)
    return out


def object_spectrum(field):
    """Spectrum of an exit wave (or a stack), on the fftshifted grid."""
    f = _device(field)
    Ny, Nx = f.shape[-2:]
    n = 1 if f.ndim == 2 else f.shape[0]
    work = N.DeviceArray.empty(f.shape, f.dtype)
    spec = N.DeviceArray.empty(f.shape, f.dtype)
    N.call("dyf_optics_spectrum", C.c_void_p(f.ptr), n, Nx, Ny,
           _code(f.dtype), C.c_void_p(work.ptr), C.c_void_p(spec.ptr), None)
    return spec


def image_from_spectrum(spec, pupil):
    """Intensity in the image plane, given a spectrum and a pupil (one
    pupil shared by a stack, or one per item)."""
    s = _device(spec)
    p = _device(pupil)
    Ny, Nx = s.shape[-2:]
    n = 1 if s.ndim == 2 else s.shape[0]
    per_item = p.ndim == 3 and p.shape[0] == n and n > 1
    sc, pc = _code(s.dtype), _code(p.dtype)
    o64 = _FP64 in (sc, pc)
    work = N.DeviceArray.empty((2 * n, Ny, Nx),
                               np.complex128 if o64 else np.complex64)
    out = N.DeviceArray.empty(s.shape, np.float64 if o64 else np.float32)
    N.call("dyf_optics_image_from_spectrum", C.c_void_p(s.ptr), n, sc,
           C.c_void_p(p.ptr), int(per_item), pc, Nx, Ny,
           C.c_void_p(work.ptr), C.c_void_p(out.ptr), None)
    return out


def apply_pupils(field, pupils):
    """One exit wave through many pupils, one forward transform.
    Returns (n_pupils, Ny, Nx)."""
    out = _images(field, list(pupils))
    return out[0]


def image_intensity(stack, pupil):
    """A stack of exit waves (or one) through one pupil, batched."""
    out = _images(stack, [pupil])
    Ny, Nx = out.shape[-2:]
    if len(stack.shape) == 2:
        return out[0][0]
    return out.reshape((out.shape[0], Ny, Nx))
