"""The imaging stage in NumPy FP64 (oracle; test use only).

Restates the reference's optics.py:

    pupil              NA disc times e^{i ell phi} on the fftshifted grid,
                       optional quantised phase and zeroed origin bin (67-127)
    spiral, chromatic  the pupil with a displaced singularity / aperture (129-150)
    object_spectrum    fftshift(fft2(fftshift(f)))                      (172-182)
    image_from_spectrum |fftshift(ifft2(ifftshift(S P)))|^2            (185-194)
    apply_pupils       one field through many pupils                    (197-207)

and the accumulation of analysis/rocking.py:66-75 (com_sums).
"""

import numpy as np


def pupil(Nx, Ny, dx, ell, aperture_radius, aperture_f=(0.0, 0.0),
          singularity_f=(0.0, 0.0), dc_zero="exact", levels=0):
    ax, ay = float(aperture_f[0]), float(aperture_f[1])
    sx, sy = float(singularity_f[0]), float(singularity_f[1])
    fx = np.fft.fftshift(np.fft.fftfreq(Nx, dx))
    fy = np.fft.fftshift(np.fft.fftfreq(Ny, dx))
    FX, FY = np.meshgrid(fx, fy)
    rad = np.sqrt((FX - ax) ** 2 + (FY - ay) ** 2)
    circ = (rad <= aperture_radius).astype(np.float64)
    if ell == 0:
        return circ.astype(np.complex128)
    ph = ell * np.arctan2(FY - sy, FX - sx)
    if levels:
        step = 2 * np.pi / levels
        ph = np.floor(np.mod(ph, 2 * np.pi) / step) * step
    p = circ * np.exp(1j * ph)
    if dc_zero == "exact":
        zero = sx == 0.0 and sy == 0.0
    elif dc_zero == "half_bin":
        step = float(fx[1] - fx[0])
        zero = abs(sx) < 0.5 * step and abs(sy) < 0.5 * step
    else:
        zero = False
    if zero:
        p[Ny // 2, Nx // 2] = 0.0
    return p


def object_spectrum(field):
    sh = np.fft.fftshift(field, axes=(-2, -1))
    return np.fft.fftshift(np.fft.fft2(sh, axes=(-2, -1)), axes=(-2, -1))


def image_from_spectrum(spec, pupil_):
    img = np.fft.ifftshift(spec * pupil_, axes=(-2, -1))
    img = np.fft.ifft2(img, axes=(-2, -1))
    img = np.fft.fftshift(img, axes=(-2, -1))
    return np.abs(img) ** 2


def apply_pupils(field, pupils):
    spec = object_spectrum(field)
    return np.stack([image_from_spectrum(spec, p) for p in pupils])


def com_sums(images, weights):
    """sum_i w_i I_i, sum_i I_i (per pixel) and per-image totals, FP64."""
    I = np.asarray(images, np.float64)
    w = np.asarray(weights, np.float64)[:, None, None]
    return (I * w).sum(axis=0), I.sum(axis=0), I.sum(axis=(1, 2))
