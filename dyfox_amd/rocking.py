"""Rocking curves (drop-in for the reference's analysis/rocking.py).

    planewave_rocking  the closed-form two-beam plane-wave solution
    com_maps           runs a scan and reduces it to one number per
                       pixel, the angle that pixel's intensity is
                       centred on

`com_maps` keeps the whole scan on the device: each batch of angles is
marched, imaged through every pupil charge (batched hipFFT), and folded
into FP64 per-pixel sums by one kernel, so only the two (Ny, Nx) sums per
charge and the per-angle totals ever come back.
"""

import ctypes as C

import numpy as np

from . import _native as N
from .backend import FP64, precision as _precision
from .units import urad_from_deviation


def planewave_rocking(s_dev, xtal, t_crystal, absorption=True):
    """|D_g|^2 against deviation parameter for a perfect crystal
    (analysis/rocking.py:16-34):

        |sigma_h|^2 / gamma^2 sin^2(gamma s) exp(2 Re(sigma_0) s)

    with gamma^2 = |sigma_h sigma_hbar| + alpha^2 / 4, alpha = -2 pi
    s_dev, s = t / cos(theta_B); absorption=False drops the exponential.
    A closed form on the host."""
    s_dev = np.asarray(s_dev)
    sig_h, sig_hbar = complex(xtal.sig_h), complex(xtal.sig_hbar)
    alpha = -2 * np.pi * s_dev
    gam = np.sqrt(np.abs(sig_h * sig_hbar) + alpha ** 2 / 4)
    spath = t_crystal / xtal.cos_tB
    rc = np.abs(sig_h) ** 2 / gam ** 2 * np.sin(gam * spath) ** 2
    if absorption:
        rc = rc * np.exp(2 * complex(xtal.sig0).real * spath)
    return rc


def com_maps(seg_list, s_dev_arr, grid, xtal, optics, ells=(0, 2), *,
             batch=16, precision=FP64):
    """Where in the rocking curve each pixel sits, per pupil charge
    (analysis/rocking.py:37-83):

        COM(x, y) = sum_s phi(s) I(x, y, s) / sum_s I(x, y, s)

    in microradians, accumulated in FP64 across batched solves.  Returns
    (com, rock, weight): the maps, the integrated rocking curve per
    charge, and the summed intensity that weighted them (host arrays)."""
    from .laue.solver import march, prepare_phase
    s_dev_arr = np.asarray(s_dev_arr, dtype=float)
    Nx, Ny = grid.geometry()[:2]
    phi_urad = np.asarray(urad_from_deviation(s_dev_arr, xtal), np.float64)
    sums_wI = {ell: N.DeviceArray.zeros((Ny, Nx), np.float64) for ell in ells}
    sums_I = {ell: N.DeviceArray.zeros((Ny, Nx), np.float64) for ell in ells}
    rock = {ell: np.zeros(len(s_dev_arr)) for ell in ells}
    # one e^{iH} table for the whole scan, then one march per batch
    prec = _precision(precision)
    src = prepare_phase(seg_list, grid, xtal, precision=prec, reuse=True)
    sd_all = N.DeviceArray.from_host(s_dev_arr)
    for bs in range(0, len(s_dev_arr), batch):
        be = min(bs + batch, len(s_dev_arr))
        stack = N.DeviceArray.empty((be - bs, Ny, Nx), prec.complex)
        march(src, sd_all[bs:be], grid, xtal, stack, prec)
        w = N.DeviceArray.from_host(phi_urad[bs:be])
        totals = N.DeviceArray.empty((be - bs,), np.float64)
        for ell in ells:
            img = optics.image(stack, grid.dx, ell=ell)
            code = 0 if img.dtype == np.float64 else 1
            N.call("dyf_optics_com_accumulate", C.c_void_p(img.ptr), be - bs,
                   code, Nx, Ny, C.c_void_p(w.ptr),
                   C.c_void_p(sums_wI[ell].ptr), C.c_void_p(sums_I[ell].ptr),
                   C.c_void_p(totals.ptr), None)
            rock[ell][bs:be] = totals.get()
            del img
        del stack
    com, weight = {}, {}
    for ell in ells:
        wI, I = sums_wI[ell].get(), sums_I[ell].get()
        com[ell] = wI / np.maximum(I, 1e-300)      # the final ratio, host
        weight[ell] = I
    return com, rock, weight
