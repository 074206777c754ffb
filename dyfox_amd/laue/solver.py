"""The transmission (Laue) solver on the device.

    stack = solve_dislocation(segs, s_dev, grid, xtal)
    stack = solve_pristine(s_dev, grid, xtal)

Same signatures, defaults and return values as the reference's
laue/solver.py:34-118: a device array (n_angles, Ny, Nx), complex128 for
FP64 and complex64 for FP32 / MIXED.  `backend.to_numpy` copies it home.

Dispatch follows the reference: `H_gpu` given -> per-voxel phase from
that grid; else one line along x_lab -> closed-form phase per (row,
depth); else the Volterra phase of the segments.  On exact grids
(shift_px == 1) the march takes the register-resident fast path and the
segment phase is evaluated straight into its e^{iH} table, so no H grid
is stored; interpolating grids take the general path on an H grid.
"""

import ctypes as C

import numpy as np

from .. import _native as N
from ..backend import FP64, precision as _precision
from ..dislocations import (analytic_line_params, compute_H_grid,
                            pack_segments)


def laue_args(grid, xtal):
    """The C ABI's dyf_laue_args for a grid and crystal."""
    Nx, Ny, dz, ds, spx = grid.geometry()
    return N.LaueArgs(
        s0r=float(np.real(xtal.sig0)), s0i=float(np.imag(xtal.sig0)),
        shr=float(np.real(xtal.sig_h)), shi=float(np.imag(xtal.sig_h)),
        sbr=float(np.real(xtal.sig_hbar)), sbi=float(np.imag(xtal.sig_hbar)),
        dz=float(dz), ds=float(ds), dx=float(grid.dx), shift_px=float(spx),
        sheet_hw=float(grid.sheet_arg()), fov_hy=float(grid.beam_width / 2.0),
        Nz=int(grid.Nz), Nx=int(Nx), Ny=int(Ny), reserved=0)


def plan(grid, xtal, precision=FP64):
    """(path, cols_per_lane, iy_lo, iy_hi) the library picks for this grid."""
    prec = _precision(precision)
    a = laue_args(grid, xtal)
    p, c, lo, hi = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    N.call("dyf_laue_plan", C.byref(a), prec.code, C.byref(p), C.byref(c),
           C.byref(lo), C.byref(hi))
    return p.value, c.value, lo.value, hi.value


def _line_struct(ap):
    if ap is None:
        return None
    return N.LineParams(y0=ap["y0"], z0=ap["z0"], e1y=ap["e1y"],
                        e1z=ap["e1z"], e2y=ap["e2y"], e2z=ap["e2z"],
                        c_atan=ap["C_atan"], ge1=ap["ge1"], ge2=ap["ge2"],
                        b_perp=ap["b_perp"], nu=ap["nu"], a2=ap["a2"])


class PhaseSource:
    """What the march reads: the e^{iH} table (fast path) or an H grid
    (general path), built once and shared by every batch of angles."""

    def __init__(self, mode, buf=None, h_prec=N.FP64, line=None):
        self.mode, self.buf, self.h_prec, self.line = mode, buf, h_prec, line

    @property
    def ptr(self):
        return None if self.buf is None else C.c_void_p(self.buf.ptr)


def _guard(nbytes, what):
    free = N.mem_info()[0]
    if nbytes > 0.7 * free:
        raise MemoryError(f"{what} needs {nbytes / 1e9:.1f} GB, "
                          f"{free / 1e9:.1f} GB free")


_table_cache = {}


def _table(nbytes, reuse):
    """A byte buffer for the e^{iH} table: the module's grow-only cached
    buffer when `reuse` (one solve at a time on the default stream, so
    the previous user has finished reading it), else a fresh one."""
    if not reuse:
        return N.DeviceArray.empty((max(nbytes, 1),), np.uint8)
    buf = _table_cache.get("phase")
    if buf is None or buf.nbytes < nbytes:
        _table_cache.pop("phase", None)
        buf = N.DeviceArray.empty((max(nbytes, 1),), np.uint8)
        _table_cache["phase"] = buf
    return buf


def release_workspaces():
    """Drop the cached phase table and coefficient workspace."""
    _table_cache.clear()
    _coef_cache.clear()


def prepare_phase(seg_list, grid, xtal, *, H_gpu=None, precision=FP64,
                  reuse=False):
    """Build the phase source a solve marches through (see module doc).
    `reuse` writes the e^{iH} table into the module's cached buffer."""
    prec = _precision(precision)
    a = laue_args(grid, xtal)
    path, _, _, _ = plan(grid, xtal, prec)
    # The reference's closed-form kernel places rows at (iy - Ny//2) dx
    # (laue/kernels.py:172) and its H grid at (iy - Ny/2) dx
    # (dislocations.py:146); they differ only for an odd Ny, which only the
    # Ny= override can give, and then the Volterra table is used.
    odd_rows = grid.geometry()[1] % 2 == 1
    ap = (None if H_gpu is not None or odd_rows
          else analytic_line_params(seg_list, xtal))
    if H_gpu is not None:
        H = N.asarray(H_gpu)
        if H.dtype not in (np.float64, np.float32):
            H = H.astype(np.float64)
        Nx, Ny = grid.geometry()[:2]
        if H.shape != (Ny, grid.Nz, Nx):
            raise ValueError(f"H_gpu has shape {H.shape}, the grid needs "
                             f"{(Ny, grid.Nz, Nx)}")
        h_prec = N.FP32 if H.dtype == np.float32 else N.FP64
        if path == N.PATH_GENERAL:
            return PhaseSource(N.MODE_GRID, H, h_prec)
        nb = C.c_size_t()
        N.call("dyf_laue_phase_bytes", C.byref(a), N.MODE_GRID, prec.code,
               C.byref(nb))
        _guard(nb.value, "the e^{iH} table")
        E = _table(nb.value, reuse)
        N.call("dyf_laue_phase_from_h", C.byref(a), prec.code,
               C.c_void_p(H.ptr), h_prec, C.c_void_p(E.ptr), None)
        return PhaseSource(N.MODE_GRID, E)
    if ap is not None:
        line = _line_struct(ap)
        if path == N.PATH_GENERAL:
            return PhaseSource(N.MODE_ANALYTIC, None, line=line)
        nb = C.c_size_t()
        N.call("dyf_laue_phase_bytes", C.byref(a), N.MODE_ANALYTIC, prec.code,
               C.byref(nb))
        E = _table(nb.value, reuse)
        N.call("dyf_laue_phase_analytic", C.byref(a), prec.code,
               C.byref(line), C.c_void_p(E.ptr), None)
        return PhaseSource(N.MODE_ANALYTIC, E, line=line)
    if path == N.PATH_GENERAL:
        Nx, Ny = grid.geometry()[:2]
        _guard(Nx * Ny * grid.Nz * prec.itemsize, "H grid")
        H = compute_H_grid(seg_list, grid, xtal, precision=prec)
        h_prec = N.FP32 if H.dtype == np.float32 else N.FP64
        return PhaseSource(N.MODE_GRID, H, h_prec)
    nb = C.c_size_t()
    N.call("dyf_laue_phase_bytes", C.byref(a), N.MODE_GRID, prec.code,
           C.byref(nb))
    _guard(nb.value, "the e^{iH} table")
    E = _table(nb.value, reuse)
    segs = pack_segments(seg_list)
    g = np.ascontiguousarray(xtal.g_vec, dtype=np.float64)
    N.call("dyf_laue_phase_from_segments", C.byref(a), prec.code,
           N.dptr(segs), len(segs), N.dptr(g), float(xtal.nu),
           float(xtal.a_core), C.c_void_p(E.ptr), None)
    return PhaseSource(N.MODE_GRID, E)


_coef_cache = {}


def _coef_workspace(a, n_e):
    nb = C.c_size_t()
    N.call("dyf_laue_coef_bytes", C.byref(a), int(n_e), C.byref(nb))
    ws = _coef_cache.get("ws")
    if ws is None or ws.nbytes < nb.value:
        ws = N.DeviceArray.empty((max(nb.value, 1),), np.uint8)
        _coef_cache["ws"] = ws
    return ws


def march(source, s_dev, grid, xtal, out, precision=FP64):
    """Launch the march of every angle in `s_dev` (a device float64 array)
    into `out` (n_e, Ny, Nx); asynchronous on the default stream."""
    prec = _precision(precision)
    a = laue_args(grid, xtal)
    n_e = len(s_dev)
    ws = _coef_workspace(a, n_e)
    line = source.line
    N.call("dyf_laue_march", C.byref(a), source.mode, prec.code,
           source.ptr, source.h_prec,
           None if line is None else C.byref(line),
           C.c_void_p(s_dev.ptr), int(n_e), C.c_void_p(ws.ptr),
           C.c_void_p(out.ptr), None)


def solve_dislocation(seg_list, s_dev_arr, grid, xtal, *, H_gpu=None,
                      batch=64, precision=FP64):
    """Exit wave of a dislocation field at each deviation parameter.

    `H_gpu` reuses a displacement grid already built (device or host
    array, float64 or float32, shape (Ny, Nz, Nx)).  `batch` sets how many
    angles one launch marches; it cannot change the result.
    """
    prec = _precision(precision)
    s_dev_arr = np.atleast_1d(np.asarray(s_dev_arr, dtype=float))
    n_e = len(s_dev_arr)
    Nx, Ny = grid.geometry()[:2]
    source = prepare_phase(seg_list, grid, xtal, H_gpu=H_gpu, precision=prec,
                           reuse=True)
    out = N.DeviceArray.empty((n_e, Ny, Nx), prec.complex)
    sd = N.DeviceArray.from_host(s_dev_arr.astype(np.float64))
    step = max(1, int(batch))
    for bs in range(0, n_e, step):
        be = min(bs + step, n_e)
        march(source, sd[bs:be], grid, xtal, out[bs:be], prec)
    return out


def solve_pristine(s_dev_arr, grid, xtal, *, precision=FP64):
    """Exit wave of a perfect crystal at each deviation parameter; one
    launch for all angles.  The closed form of the same curve is
    `analysis.rocking.planewave_rocking` in the reference."""
    prec = _precision(precision)
    s_dev_arr = np.atleast_1d(np.asarray(s_dev_arr, dtype=float))
    Nx, Ny = grid.geometry()[:2]
    out = N.DeviceArray.empty((len(s_dev_arr), Ny, Nx), prec.complex)
    sd = N.DeviceArray.from_host(s_dev_arr.astype(np.float64))
    march(PhaseSource(N.MODE_PRISTINE), sd, grid, xtal, out, prec)
    return out
