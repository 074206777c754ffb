"""Sampling a beam that is not a single coherent plane wave, and the
incoherent image such a beam makes.

    condenser_angles             incidence angles and weights for a
                                 focused beam
    gauss_hermite_offsets        n*n offsets and weights for a Gaussian
                                 source
    convergence_bandwidth_nodes  joint quadrature over incidence angle
                                 and energy band
    incoherent_image             sum_n w_n |image of node n|^2 on the
                                 device

The three node generators are the reference's illumination.py (16-86):
small host arrays, the same values.  `incoherent_image` is what the
nodes drive: every deviation node is one angle of a batched march, each
node's image goes through its own chromatic pupil (aperture and
singularity walked by its energy offset), and the weighted sum is
accumulated in FP64 on the device, so only the final (Ny, Nx) map
leaves it.
"""

import ctypes as C

import numpy as np

from . import _native as N
from .backend import FP64, precision as _precision

# Bytes of exit waves one incoherent_image launch may hold by default.
STACK_BUDGET = 4 << 30


def condenser_angles(w_D, reach=75.0):
    """Incidence angles for an incoherent condenser sum and their
    weights: every w_D across +-4 w_D, then every 5 units out to `reach`;
    weights are the node spacing (illumination.py:16-26)."""
    fine = np.arange(-4.0 * w_D, 4.0 * w_D + 1e-9, w_D)
    coarse = np.arange(-reach, reach + 1e-9, 5.0)
    a = np.unique(np.concatenate([fine, coarse, [0.0]]))
    return a, np.gradient(a)


def gauss_hermite_offsets(sigma, n=8):
    """n*n offsets and weights (summing to one) averaging over a
    Gaussian source of rms sigma: the probabilists' Gauss-Hermite rule in
    x and y (illumination.py:29-40)."""
    x, w = np.polynomial.hermite_e.hermegauss(n)
    X, Y = np.meshgrid(x, x)
    WX, WY = np.meshgrid(w, w)
    offs = np.stack([X.ravel() * sigma, Y.ravel() * sigma], axis=1)
    wts = (WX * WY).ravel()
    return offs, wts / wts.sum()


def convergence_bandwidth_nodes(xtal, sig_psi, sig_delta, n_psi, psi_max,
                                n_delta, del_max, *, matched=None,
                                vignette_f=None):
    """Quadrature over a beam's angular spread and energy band
    (illumination.py:43-86): an n_psi x n_delta grid with Gaussian
    weights; matched="psi" tilts each energy back onto the Bragg
    condition, matched="delta" ties the angular spread to the band;
    `vignette_f` drops nodes whose chromatic carrier leaves the
    objective.  Returns deviation parameters, chromatic arguments and
    weights sorted from the Bragg condition outward, and a dict of how
    many nodes were kept."""
    lam, theta_B, sin_2tB = xtal.lam, xtal.theta_B, xtal.sin_2tB
    tan2t = np.tan(2 * theta_B)
    psi = np.linspace(-psi_max, psi_max, n_psi)
    del_ = np.linspace(-del_max, del_max, n_delta)
    w_psi = np.exp(-0.5 * (psi / sig_psi) ** 2)
    w_del = (np.exp(-0.5 * (del_ / sig_delta) ** 2) if sig_delta > 0
             else np.ones_like(del_))
    PSI, DEL = np.meshgrid(psi, del_, indexing="ij")
    if matched == "psi":
        PSI = PSI - np.tan(theta_B) * DEL
    if matched == "delta":
        DEL = -PSI / np.tan(theta_B)
    W = np.outer(w_psi, w_del)
    W /= W.sum()
    S_DEV = (PSI + np.tan(theta_B) * DEL) * sin_2tB / lam
    D_ARG = DEL + PSI / tan2t
    order = np.argsort(np.abs(S_DEV).ravel())
    S, D, Wf = S_DEV.ravel()[order], D_ARG.ravel()[order], W.ravel()[order]
    info = {"n_total": len(S), "n_kept": len(S), "band_fraction": 1.0}
    if vignette_f is not None:
        keep = np.abs(sin_2tB * D / lam) <= vignette_f
        if not keep.all():
            S, D, Wf = S[keep], D[keep], Wf[keep]
            info["n_kept"] = int(keep.sum())
            info["band_fraction"] = float(Wf.sum())
    return S, D, Wf, info


def incoherent_image(seg_list, grid, xtal, optics, s_dev, weights,
                     d_arg=None, ell=0, *, batch=None, precision=FP64):
    """sum_n weights[n] I_n on the device, I_n the image of the exit wave
    at deviation s_dev[n] through the charge-`ell` pupil, walked by the
    chromatic argument d_arg[n] when given (Optics.chromatic).  Returns
    the (Ny, Nx) float64 map as a DeviceArray.  `batch` nodes are marched
    per launch; by default every node in one launch while their exit
    waves fit STACK_BUDGET bytes (a launch of few angles leaves the device
    partly idle and cannot march two angles per wavefront)."""
    from .laue.solver import march, prepare_phase
    from .optics import image_from_spectrum, object_spectrum
    s_dev = np.asarray(s_dev, np.float64)
    w = np.asarray(weights, np.float64)
    if w.shape != s_dev.shape:
        raise ValueError("weights must match s_dev")
    if d_arg is not None:
        d_arg = np.asarray(d_arg, np.float64)
        if d_arg.shape != s_dev.shape:
            raise ValueError("d_arg must match s_dev")
    Nx, Ny = grid.geometry()[:2]
    prec = _precision(precision)
    if batch is None:
        per_node = Ny * Nx * np.dtype(prec.complex).itemsize
        batch = int(max(1, min(len(s_dev), STACK_BUDGET // per_node)))
    total = N.DeviceArray.zeros((Ny, Nx), np.float64)
    plain = N.DeviceArray.zeros((Ny, Nx), np.float64)
    # one e^{iH} table for every node (the field does not change), then
    # each batch of nodes is one march
    src = prepare_phase(seg_list, grid, xtal, precision=prec, reuse=True)
    sd_all = N.DeviceArray.from_host(s_dev)
    for b0 in range(0, len(s_dev), batch):
        b1 = min(b0 + batch, len(s_dev))
        stack = N.DeviceArray.empty((b1 - b0, Ny, Nx), prec.complex)
        march(src, sd_all[b0:b1], grid, xtal, stack, prec)
        if d_arg is None:
            imgs = optics.image(stack, grid.dx, ell=ell)
        else:
            pups = N.DeviceArray.empty((b1 - b0, Ny, Nx),
                                       optics.precision.complex)
            for j, d in enumerate(d_arg[b0:b1]):
                pups[j].copy_from(optics.chromatic(Nx, Ny, grid.dx, ell,
                                                   float(d)))
            imgs = image_from_spectrum(object_spectrum(stack), pups)
        wd = N.DeviceArray.from_host(w[b0:b1])
        code = 0 if imgs.dtype == np.float64 else 1
        N.call("dyf_optics_com_accumulate", C.c_void_p(imgs.ptr), b1 - b0,
               code, Nx, Ny, C.c_void_p(wd.ptr), C.c_void_p(total.ptr),
               C.c_void_p(plain.ptr), None, None)
        del stack, imgs
    return total
