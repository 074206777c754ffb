"""Straight dislocations and the displacement phase H = 2 pi g.u.

Segments are the reference's tuples (dislocations.py:1-15), lab-frame
metres:

    [(r_start, r_end, b), ...]

Each acts as an infinite straight line through its midpoint (the length
only fixes the direction).  H follows isotropic Volterra elasticity: the
screw and edge arctan term g.b * atan2(d2, d1) plus the smooth edge terms
regularised inside the core radius.

    compute_H_grid        (Ny, Nz, Nx) on the device      (HIP, FP64)
    H_plane               one (Nz, Nx) plane per y value   (HIP, FP64)
    analytic_line_params  the 12 closed-form scalars, or None

All arithmetic runs in libdyfox.so; this module only shapes arguments.
"""

import ctypes as C

import numpy as np

from . import _native as N
from .backend import FP64, precision as _precision


def rotate_about_z(v, deg):
    """Rotate a lab-frame vector about the surface normal (degrees)."""
    if not deg:
        return v
    c, s = np.cos(np.radians(deg)), np.sin(np.radians(deg))
    return np.array([c * v[0] - s * v[1], s * v[0] + c * v[1], v[2]])


def line_direction(U_lab, line_cryst, tilt_deg=0.0):
    """Unit lab-frame line direction of a crystal direction, optionally
    tilted about z (the Burgers vector is not tilted)."""
    xi = U_lab @ np.asarray(line_cryst, float)
    return rotate_about_z(xi / np.linalg.norm(xi), tilt_deg)


def burgers_vector(U_lab, b_cryst, b_mag, sign=+1):
    """Lab-frame Burgers vector of length `b_mag` along `b_cryst`."""
    b = U_lab @ np.asarray(b_cryst, float)
    return sign * b_mag * (b / np.linalg.norm(b))


def g_dot_b(g_vec, b):
    """g.b: the topological charge the exit-wave vortex carries."""
    return float(np.dot(np.asarray(g_vec, float), np.asarray(b, float)))


def beam_column_center(z0, y0, tan_tB):
    """Point at depth z0 on the forward characteristic: (-z0 tan tB, y0, z0)."""
    return np.array([-z0 * tan_tB, y0, z0])


def straight_segment(xi, b, center, half_length):
    """A one-segment list: the line through `center` along `xi`."""
    xi = np.asarray(xi, float)
    center = np.asarray(center, float)
    return [(center - half_length * xi, center + half_length * xi,
             np.asarray(b, float))]


def volterra_frame(xi, bv):
    """(b_screw, b_perp, e1, e2): screw part along the line, edge magnitude,
    unit edge direction, and e2 = xi x e1.  For a pure screw, e1 is the
    axis furthest from the line (dislocations.py:68-90)."""
    xi = np.asarray(xi, float)
    bv = np.asarray(bv, float)
    b_screw = float(np.dot(bv, xi))
    edge = bv - b_screw * xi
    b_perp = np.linalg.norm(edge)
    if b_perp > 1e-20:
        e1 = edge / b_perp
    elif abs(xi[2]) > 0.9:
        e1 = np.array([1.0, 0.0, 0.0])
    else:
        e1 = np.cross(xi, [1, 0, 0] if abs(xi[0]) < 0.9 else [0, 1, 0])
        e1 = e1 / np.linalg.norm(e1)
    return b_screw, b_perp, e1, np.cross(xi, e1)


def pack_segments(seg_list):
    """(n, 9) float64 array (ra, rb, b per row) for the C ABI."""
    if not seg_list:
        return np.zeros((0, 9))
    return np.ascontiguousarray(
        [np.concatenate([np.asarray(p, float).ravel() for p in s])
         for s in seg_list], dtype=np.float64)


def _volterra(seg_list, g_vec, nu, a_core, Nx, Nz, dx, dz, y_values, prec,
              out):
    segs = pack_segments(seg_list)
    g = np.ascontiguousarray(g_vec, dtype=np.float64)
    ys = np.ascontiguousarray(y_values, dtype=np.float64)
    vol = N.VolumeDesc(int(Nx), int(Nz), float(dx), float(dz))
    N.call("dyf_volterra_h", N.dptr(segs), len(segs), N.dptr(g), float(nu),
           float(a_core), C.byref(vol), N.dptr(ys), len(ys), prec,
           C.c_void_p(out.ptr), None)


def compute_H_grid(seg_list, grid, xtal, precision=FP64, chunk_y=32):
    """H over the whole (Ny, Nz, Nx) volume, on the device.

    float64 (segments summed in FP64) for FP64/MIXED, float32 (each
    segment's FP64 value summed in FP32) for FP32 - the reference's
    dislocations.py:113-168.  Launched in y-chunks of `chunk_y` rows;
    every voxel is computed independently, so the chunk size cannot
    change a bit of the result.
    """
    prec = _precision(precision)
    Nx, Ny, dz, _, _ = grid.geometry()
    Nz = grid.Nz
    out_prec = N.FP32 if prec.real == np.float32 else N.FP64
    H = N.DeviceArray.empty((Ny, Nz, Nx), prec.real)
    step = max(1, int(chunk_y))
    for ys in range(0, Ny, step):
        ye = min(ys + step, Ny)
        yv = (np.arange(ys, ye, dtype=np.float64) - Ny / 2) * grid.dx
        _volterra(seg_list, xtal.g_vec, xtal.nu, xtal.a_core, Nx, Nz,
                  grid.dx, dz, yv, out_prec, H[ys:ye])
    return H


def H_plane(seg_list, cfg, g_vec, nu, a_core, xp=np, y_plane=0.0):
    """H on the (Nz, Nx) plane of incidence at `y_plane` (FP64).

    Evaluated on the device; `xp=np` returns a NumPy copy, any other value
    the device array (dislocations.py:171-211).  `y_plane` may also be a
    sequence, giving (len(y), Nz, Nx).
    """
    ys = np.atleast_1d(np.asarray(y_plane, dtype=np.float64))
    H = N.DeviceArray.empty((len(ys), cfg.Nz, cfg.Nx), np.float64)
    _volterra(seg_list, g_vec, nu, a_core, cfg.Nx, cfg.Nz, cfg.dx, cfg.dz,
              ys, N.FP64, H)
    out = H if np.ndim(y_plane) else H[0]
    return out.get() if xp is np else out


def analytic_line_params(seg_list, xtal):
    """The closed-form kernel's 12 scalars, or None when the geometry does
    not qualify: one segment along x_lab whose frame has no x component
    (dislocations.py:214-242)."""
    if len(seg_list) != 1:
        return None
    ra, rb, bv = (np.asarray(v, float) for v in seg_list[0])
    t = rb - ra
    L = np.linalg.norm(t)
    if L < 1e-20:
        return None
    xi = t / L
    if abs(abs(xi[0]) - 1.0) > 1e-9:
        return None
    b_screw, b_perp, e1, e2 = volterra_frame(xi, bv)
    if abs(e1[0]) > 1e-9 or abs(e2[0]) > 1e-9:
        return None
    core = (ra + rb) / 2.0
    gv = xtal.g_vec
    return dict(y0=core[1], z0=core[2], e1y=e1[1], e1z=e1[2], e2y=e2[1],
                e2z=e2[2], C_atan=np.dot(gv, e1) * b_perp
                + np.dot(gv, xi) * b_screw,
                ge1=np.dot(gv, e1), ge2=np.dot(gv, e2), b_perp=b_perp,
                nu=xtal.nu, a2=xtal.a_core ** 2)
