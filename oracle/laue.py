"""The Laue depth march in NumPy FP64 (oracle; test use only).

Restates the reference's three CUDA kernels (laue/kernels.py:48-352) in
their direct form, vectorised over (s_dev, marched rows, x) with a Python
loop over depth.  Per s_dev (kernels.py:90-103):

    alpha = -2 pi s,  gam = sqrt(|sigma_h|^2 + alpha^2/4)
    cg = cos(gam ds), sg = sin(gam ds)/gam      (1, ds if gam <= 1e-30)
    ep = e^{Re s0 ds} e^{i (Im s0 + alpha/2) ds}
    M11 = ep (cg - i alpha/2 sg),  M22 = ep (cg + i alpha/2 sg)

per step k and column x (kernels.py:106-139):

    M12 = ep sg sigma_hbar e^{+iH[iy,k,x]},  M21 = ep sg sigma_h e^{-iH}
    d0 = interp(D0, x + shift), dg = interp(Dg, x - shift)   (zero inflow)
    D0' = M11 d0 + M12 dg,   Dg' = M21 d0 + M22 dg

starting from D0 = top hat |x| < sheet_hw (Gaussian for sheet_hw < 0),
Dg = 0, for k = 0 .. Nz-2; rows with |(iy - Ny/2) dx| >= fov_hy are 0.
"""

import numpy as np


def initial_beam(Nx, dx, sheet_hw):
    x = (np.arange(Nx) - Nx // 2) * dx
    if sheet_hw > 0.0:
        return (np.abs(x) < sheet_hw).astype(np.float64)
    return np.exp(-x * x / (2.0 * sheet_hw * sheet_hw))


def marched_rows(Ny, dx, fov_hy):
    y = (np.arange(Ny) - Ny // 2) * dx
    return np.nonzero(np.abs(y) < fov_hy)[0]


def _interp(F, shift, Nx):
    """F sampled at x + shift with the kernel's linear rule and zero
    inflow (kernels.py:122-135), along the last axis."""
    x = np.arange(Nx)
    src = x + shift
    i0 = np.floor(src).astype(np.int64)
    f = src - i0
    out = np.zeros_like(F)
    both = (i0 >= 0) & (i0 + 1 < Nx)
    one = (i0 >= 0) & (i0 < Nx) & ~both
    if both.any():
        out[..., both] = ((1 - f[both]) * F[..., i0[both]]
                          + f[both] * F[..., i0[both] + 1])
    if one.any():
        out[..., one] = (1 - f[one]) * F[..., i0[one]]
    return out


def coefficients(s_dev, sig0, sig_h, ds):
    """(M11, M22, ep, sg) per s_dev, as the kernels compute them."""
    s_dev = np.atleast_1d(np.asarray(s_dev, np.float64))
    alpha = -2.0 * np.pi * s_dev
    gam = np.sqrt(abs(sig_h) ** 2 + alpha * alpha * 0.25)
    big = gam > 1e-30
    cg = np.where(big, np.cos(gam * ds), 1.0)
    sg = np.where(big, np.sin(gam * ds) / np.where(big, gam, 1.0), ds)
    ep = np.exp(sig0.real * ds) * np.exp(1j * (sig0.imag + alpha * 0.5) * ds)
    M11 = ep * (cg - 1j * alpha * 0.5 * sg)
    M22 = ep * (cg + 1j * alpha * 0.5 * sg)
    return M11, M22, ep, sg


def march(s_dev, sig0, sig_h, sig_hbar, Nz, Nx, Ny, dx, ds, shift_px,
          sheet_hw, fov_hy, H=None, rows=None):
    """Exit waves (n_e, Ny, Nx) complex128.

    H: None (pristine), an array (Ny, Nz, Nx) / (nrows, Nz, Nx), or a
    callable H(k, rows) -> (nrows, Nx) or (nrows, 1).
    """
    s_dev = np.atleast_1d(np.asarray(s_dev, np.float64))
    n_e = len(s_dev)
    if rows is None:
        rows = marched_rows(Ny, dx, fov_hy)
    nr = len(rows)
    out = np.zeros((n_e, Ny, Nx), np.complex128)
    if nr == 0:
        return out
    M11, M22, ep, sg = coefficients(s_dev, sig0, sig_h, ds)
    A12 = (ep * sg * sig_hbar)[:, None, None]
    A21 = (ep * sg * sig_h)[:, None, None]
    M11 = M11[:, None, None]
    M22 = M22[:, None, None]
    D0 = np.zeros((n_e, nr, Nx), np.complex128)
    D0[:] = initial_beam(Nx, dx, sheet_hw)
    Dg = np.zeros_like(D0)
    for k in range(Nz - 1):
        if H is None:
            e = np.ones((1, 1), np.complex128)
        else:
            if callable(H):
                h = H(k, rows)
            elif H.shape[0] == Ny:
                h = H[rows, k, :]
            else:
                h = H[:, k, :]
            e = np.exp(1j * np.asarray(h, np.float64))
        M12 = A12 * e[None]
        M21 = A21 * np.conj(e)[None]
        d0 = _interp(D0, shift_px, Nx)
        dg = _interp(Dg, -shift_px, Nx)
        D0 = M11 * d0 + M12 * dg
        Dg = M21 * d0 + M22 * dg
    out[:, rows, :] = Dg
    return out


def analytic_h(line, Ny, dx, dz):
    """Callable H(k, rows) of the closed-form kernel (kernels.py:216-228)."""
    def H(k, rows):
        y = (rows - Ny // 2) * dx
        dy = y - line["y0"]
        dzv = k * dz - line["z0"]
        d1 = dy * line["e1y"] + dzv * line["e1z"]
        d2 = dy * line["e2y"] + dzv * line["e2z"]
        r2 = d1 * d1 + d2 * d2
        r2c = np.where(r2 > line["a2"], r2, line["a2"])
        h = line["C_atan"] * np.arctan2(d2, d1)
        if line["b_perp"] > 0.0:
            nu = line["nu"]
            onu = 1.0 - nu
            c1 = line["b_perp"] / (2.0 * np.pi)
            ux = c1 * d1 * d2 / (2.0 * onu * r2c)
            uy = -c1 * ((1.0 - 2.0 * nu) / (4.0 * onu) * np.log(r2c)
                        + (d1 * d1 - d2 * d2) / (4.0 * onu * r2c))
            h = h + 2.0 * np.pi * (line["ge1"] * ux + line["ge2"] * uy)
        return h[:, None]
    return H


def march_grid(grid_args, s_dev, H=None, rows=None):
    """march() with a dict of the grid/crystal scalars (see tests)."""
    g = grid_args
    return march(s_dev, g["sig0"], g["sig_h"], g["sig_hbar"], g["Nz"],
                 g["Nx"], g["Ny"], g["dx"], g["ds"], g["shift_px"],
                 g["sheet_hw"], g["fov_hy"], H=H, rows=rows)


def rel_rms(a, b):
    """sqrt(sum |a-b|^2 / sum |b|^2), the reference's error metric."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    den = np.sum(np.abs(b) ** 2)
    return float(np.sqrt(np.sum(np.abs(a - b) ** 2) / den)) if den else \
        float(np.sqrt(np.sum(np.abs(a - b) ** 2)))


def error_report(a, b):
    """Per-s_dev rel-RMS, max-abs / max|b| and intensity error."""
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    rep = []
    for j in range(a.shape[0]):
        mb = np.abs(b[j]).max() or 1.0
        rep.append(dict(rel_rms=rel_rms(a[j], b[j]),
                        max_abs=float(np.abs(a[j] - b[j]).max() / mb),
                        intensity=float(np.abs(np.abs(a[j]) ** 2
                                               - np.abs(b[j]) ** 2).max()
                                        / mb ** 2)))
    return rep
