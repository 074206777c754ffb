"""Volterra displacement phase in NumPy FP64 (oracle; test use only).

Restates the reference's dislocations.py:68-90 (frame), 113-168
(compute_H_grid) and 171-211 (H_plane), in the same expression order:

    X = x - cx, ...;  eta = X.xi;  P = X - eta xi;  d1 = P.e1;  d2 = P.e2
    r2c = max(d1^2 + d2^2, a_core^2)
    Hc  = (g.e1 b_perp + g.xi b_screw) atan2(d2, d1)
    if b_perp > 1e-20:
        Hc += 2 pi (g.e1 c1 d1 d2 / (2(1-nu) r2c)
                    - g.e2 c1 ((1-2nu)/(4(1-nu)) ln r2c
                              + (d1^2 - d2^2)/(4(1-nu) r2c))),  c1 = b_perp/2pi
    H = sum over segments in list order
"""

import numpy as np


def frame(xi, bv):
    """(b_screw, b_perp, e1, e2) per dislocations.py:68-90."""
    b_screw = float(np.dot(bv, xi))
    bp = bv - b_screw * xi
    b_perp = float(np.linalg.norm(bp))
    if b_perp > 1e-20:
        e1 = bp / b_perp
    elif abs(xi[2]) > 0.9:
        e1 = np.array([1.0, 0.0, 0.0])
    elif abs(xi[0]) < 0.9:
        e1 = np.cross(xi, [1, 0, 0])
        e1 = e1 / np.linalg.norm(e1)
    else:
        e1 = np.cross(xi, [0, 1, 0])
        e1 = e1 / np.linalg.norm(e1)
    return b_screw, b_perp, e1, np.cross(xi, e1)


def segment_h(seg, g_vec, nu, a_core, X, Y, Z):
    """One segment's H at broadcastable coordinate arrays X, Y, Z (m)."""
    ra, rb, bv = (np.asarray(v, float) for v in seg)
    t = rb - ra
    L = np.linalg.norm(t)
    if L < 1e-20:
        return None
    xi = t / L
    core = (ra + rb) / 2.0
    b_screw, b_perp, e1, e2 = frame(xi, bv)
    gv = np.asarray(g_vec, float)
    onu = 1.0 - nu
    C_atan = np.dot(gv, e1) * b_perp + np.dot(gv, xi) * b_screw
    ge1, ge2 = np.dot(gv, e1), np.dot(gv, e2)
    Xb, Yb, Zb = X - core[0], Y - core[1], Z - core[2]
    eta = Xb * xi[0] + Yb * xi[1] + Zb * xi[2]
    Px, Py, Pz = Xb - eta * xi[0], Yb - eta * xi[1], Zb - eta * xi[2]
    d1 = Px * e1[0] + Py * e1[1] + Pz * e1[2]
    d2 = Px * e2[0] + Py * e2[1] + Pz * e2[2]
    r2c = np.maximum(d1 ** 2 + d2 ** 2, a_core ** 2)
    Hc = C_atan * np.arctan2(d2, d1)
    if b_perp > 1e-20:
        c1 = b_perp / (2 * np.pi)
        ux = c1 * d1 * d2 / (2 * onu * r2c)
        uy = -c1 * ((1 - 2 * nu) / (4 * onu) * np.log(r2c)
                    + (d1 ** 2 - d2 ** 2) / (4 * onu * r2c))
        Hc = Hc + 2 * np.pi * (ge1 * ux + ge2 * uy)
    return Hc


def h_field(seg_list, g_vec, nu, a_core, X, Y, Z, dtype=np.float64):
    """Sum of segment_h over the list, accumulated in `dtype` (float32
    reproduces the reference's FP32 accumulation)."""
    shape = np.broadcast(X, Y, Z).shape
    H = np.zeros(shape, dtype=dtype)
    for seg in seg_list:
        Hc = segment_h(seg, g_vec, nu, a_core, X, Y, Z)
        if Hc is not None:
            H += np.broadcast_to(Hc, shape).astype(dtype)
    return H


def h_grid(seg_list, g_vec, nu, a_core, Nx, Ny, Nz, dx, dz, rows=None,
           dtype=np.float64):
    """(len(rows), Nz, Nx) H block of a Laue grid (x = (ix-Nx/2) dx,
    y = (iy-Ny/2) dx, z = k dz)."""
    rows = np.arange(Ny) if rows is None else np.asarray(rows)
    x = (np.arange(Nx, dtype=np.float64) - Nx / 2) * dx
    z = np.arange(Nz, dtype=np.float64) * dz
    y = (rows.astype(np.float64) - Ny / 2) * dx
    return h_field(seg_list, g_vec, nu, a_core, x[None, None, :],
                   y[:, None, None], z[None, :, None], dtype=dtype)


def h_plane(seg_list, g_vec, nu, a_core, Nx, Nz, dx, dz, y_plane):
    """(Nz, Nx) plane at y = y_plane (the reference's H_plane)."""
    x = (np.arange(Nx) - Nx / 2) * dx
    z = np.arange(Nz) * dz
    return h_field(seg_list, g_vec, nu, a_core, x[None, :],
                   np.float64(y_plane), z[:, None])
