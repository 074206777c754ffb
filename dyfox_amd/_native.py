"""ctypes binding of libdyfox.so and the device array it hands out.

The library is loaded from this package directory (built in-tree by
`python -m dyfox_amd.build`).  There is no fallback: if the library or a
device is missing, the first call raises, naming what is missing.

Status codes map to the exceptions the reference raises for the same
conditions (laue/solver.py:49-54, laue/kernels.py:365-386,
backend.py:18-28): DYF_ENOMEM -> MemoryError, DYF_EINVAL -> ValueError,
DYF_EKEY -> KeyError, DYF_EHIP/DYF_ENODEV -> RuntimeError.
"""

import ctypes as C
import math
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdyfox.so")

DYF_OK, DYF_EINVAL, DYF_ENOMEM, DYF_EKEY, DYF_EHIP, DYF_ENODEV = \
    0, -1, -2, -3, -4, -5
FP64, FP32, MIXED = 0, 1, 2
MODE_GRID, MODE_ANALYTIC, MODE_PRISTINE = 0, 1, 2
PATH_FAST, PATH_GENERAL = 0, 1

vp = C.c_void_p
dp = C.POINTER(C.c_double)


class VolumeDesc(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Nz", C.c_int32),
                ("dx", C.c_double), ("dz", C.c_double)]


class LaueArgs(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("s0r", "s0i", "shr", "shi", "sbr", "sbi", "dz", "ds", "dx",
                 "shift_px", "sheet_hw", "fov_hy")] + \
               [(n, C.c_int32) for n in ("Nz", "Nx", "Ny", "reserved")]


class LineParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("y0", "z0", "e1y", "e1z", "e2y", "e2z", "c_atan", "ge1",
                 "ge2", "b_perp", "nu", "a2")]


class LaueInstance(C.Structure):
    _fields_ = [("grid", LaueArgs), ("segs", C.POINTER(C.c_double)),
                ("nseg", C.c_int32), ("n_e", C.c_int32),
                ("g_vec", C.c_double * 3), ("nu", C.c_double),
                ("a_core", C.c_double), ("s_dev", vp), ("out", vp)]


class MarchTune(C.Structure):
    _fields_ = [("waves_per_block", C.c_int32), ("blocks_per_cu", C.c_int32)]


class BraggArgs(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("s0r", "s0i", "shr", "shi", "sbr", "sbi", "ds")] + \
               [(n, C.c_int32) for n in ("Nz", "Nx", "nyc",
                                         "per_system_phase", "plane_systems",
                                         "reserved")]


class GmresOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("restart", C.c_int32),
                ("maxiter", C.c_int32)]


class GmresInfo(C.Structure):
    _fields_ = [("residual", C.c_double), ("iterations", C.c_int32),
                ("code", C.c_int32)]


class PupilArgs(C.Structure):
    _fields_ = [("Nx", C.c_int32), ("Ny", C.c_int32), ("dx", C.c_double),
                ("aperture_f", C.c_double), ("ax", C.c_double),
                ("ay", C.c_double), ("sx", C.c_double), ("sy", C.c_double),
                ("ell", C.c_int32), ("levels", C.c_int32),
                ("zero_dc", C.c_int32), ("reserved", C.c_int32)]


class FixedOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("n_iter", C.c_int32),
                ("reserved", C.c_int32)]


class FixedInfo(C.Structure):
    _fields_ = [("residual", C.c_double), ("iterations", C.c_int32),
                ("converged", C.c_int32)]


_SIGNATURES = {
    "dyf_init": [C.c_int],
    "dyf_device_count": [C.POINTER(C.c_int)],
    "dyf_malloc": [C.POINTER(vp), C.c_size_t],
    "dyf_free": [vp],
    "dyf_memcpy_h2d": [vp, vp, C.c_size_t],
    "dyf_memcpy_d2h": [vp, vp, C.c_size_t],
    "dyf_memcpy_d2d": [vp, vp, C.c_size_t],
    "dyf_memset": [vp, C.c_int, C.c_size_t],
    "dyf_memcpy2d_d2h": [vp, C.c_size_t, vp, C.c_size_t, C.c_size_t,
                         C.c_size_t],
    "dyf_host_alloc": [C.POINTER(vp), C.c_size_t],
    "dyf_host_free": [vp],
    "dyf_mem_info": [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "dyf_synchronize": [],
    "dyf_stream_sync": [vp],
    "dyf_event_create": [C.POINTER(vp)],
    "dyf_event_record": [vp, vp],
    "dyf_event_elapsed_ms": [vp, vp, C.POINTER(C.c_float)],
    "dyf_event_destroy": [vp],
    "dyf_copy_kernel": [vp, vp, C.c_size_t, vp],
    "dyf_volterra_h": [dp, C.c_int, dp, C.c_double, C.c_double,
                       C.POINTER(VolumeDesc), dp, C.c_int, C.c_int, vp, vp],
    "dyf_laue_plan": [C.POINTER(LaueArgs), C.c_int, C.POINTER(C.c_int),
                      C.POINTER(C.c_int), C.POINTER(C.c_int),
                      C.POINTER(C.c_int)],
    "dyf_laue_phase_bytes": [C.POINTER(LaueArgs), C.c_int, C.c_int,
                             C.POINTER(C.c_size_t)],
    "dyf_laue_coef_bytes": [C.POINTER(LaueArgs), C.c_int,
                            C.POINTER(C.c_size_t)],
    "dyf_laue_march_fraction": [C.POINTER(LaueArgs), C.c_int, C.c_int,
                                C.c_int, C.POINTER(C.c_double)],
    "dyf_laue_phase_from_h": [C.POINTER(LaueArgs), C.c_int, vp, C.c_int, vp,
                              vp],
    "dyf_laue_phase_from_segments": [C.POINTER(LaueArgs), C.c_int, dp,
                                     C.c_int, dp, C.c_double, C.c_double, vp,
                                     vp],
    "dyf_laue_phase_analytic": [C.POINTER(LaueArgs), C.c_int,
                                C.POINTER(LineParams), vp, vp],
    "dyf_laue_march": [C.POINTER(LaueArgs), C.c_int, C.c_int, vp, C.c_int,
                       C.POINTER(LineParams), vp, C.c_int, vp, vp, vp],
    "dyf_laue_march_noalloc": [C.POINTER(LaueArgs), C.c_int, C.c_int, vp,
                               C.c_int, C.POINTER(LineParams), vp, C.c_int,
                               vp, vp],
    "dyf_laue_batch_bytes": [C.POINTER(LaueInstance), C.c_int, C.c_int,
                             C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "dyf_laue_batch_prepare": [C.POINTER(LaueInstance), C.c_int, C.c_int, vp,
                               C.c_size_t, vp, C.c_size_t],
    "dyf_laue_batch_run": [C.POINTER(LaueInstance), C.c_int, C.c_int, vp, vp,
                           C.POINTER(MarchTune), C.c_int, vp],
    "dyf_graph_begin": [vp],
    "dyf_graph_end": [vp, C.POINTER(vp)],
    "dyf_graph_launch": [vp, vp],
    "dyf_graph_destroy": [vp],
    "dyf_stream_create": [C.POINTER(vp)],
    "dyf_stream_destroy": [vp],
    "dyf_nn_gemm": [vp, C.c_longlong, C.c_longlong, vp, C.c_longlong,
                    C.c_longlong, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                    vp, vp, C.c_longlong, C.c_int, vp],
    "dyf_nn_im2col3": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                       C.c_int, vp, vp],
    "dyf_nn_col2im3": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int,
                       C.c_int, vp],
    "dyf_nn_slice": [vp, C.c_longlong, C.c_int, C.c_int, C.c_int, vp, vp],
    "dyf_nn_relu_grad": [vp, vp, C.c_longlong, C.c_int, C.c_int, C.c_int,
                         C.c_int, vp, vp],
    "dyf_nn_bias_grad": [vp, C.c_longlong, C.c_int, vp, vp],
    "dyf_nn_pool": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                    vp, vp],
    "dyf_nn_pool_grad": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int,
                         C.c_int, vp],
    "dyf_nn_up": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, C.c_int,
                  C.c_int, vp],
    "dyf_nn_up_grad": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                       C.c_int, vp, vp],
    "dyf_nn_mse": [vp, vp, C.c_longlong, vp, vp, vp],
    "dyf_nn_adam": [vp, vp, vp, vp, vp, C.c_longlong, C.c_float, C.c_float,
                    C.c_float, C.c_float, C.c_int, vp],
    "dyf_nn_to_bf16": [vp, C.c_longlong, vp, vp],
    "dyf_nn_fill": [vp, C.c_longlong, C.c_float, vp],
    "dyf_nn_bragg_features": [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              C.c_int, C.c_int, vp, vp, vp, C.c_int, C.c_int,
                              C.c_int, vp],
    "dyf_bragg_phase": [dp, C.c_int, dp, C.c_double, C.c_double,
                        C.POINTER(VolumeDesc), dp, C.c_int, vp, vp],
    "dyf_bragg_march": [C.POINTER(BraggArgs), vp, C.c_int, vp, vp, vp, vp,
                        vp],
    "dyf_bragg_precondition": [C.POINTER(BraggArgs), vp, C.c_int, vp, vp, vp,
                               vp],
    "dyf_bragg_symbol_bytes": [C.POINTER(BraggArgs), C.c_int,
                               C.POINTER(C.c_size_t)],
    "dyf_bragg_workspace_bytes": [C.POINTER(BraggArgs), C.c_int, C.c_int,
                                  C.POINTER(C.c_size_t)],
    "dyf_bragg_solve": [C.POINTER(BraggArgs), vp, C.c_int, vp, vp,
                        C.POINTER(GmresOpts), vp, vp, vp, C.c_size_t,
                        C.POINTER(GmresInfo), vp],
    "dyf_bragg_exit_wave": [C.POINTER(BraggArgs), C.c_int, vp, vp, vp, vp],
    "dyf_optics_pupil": [C.POINTER(PupilArgs), C.c_int, vp, vp],
    "dyf_optics_spectrum": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp,
                            vp],
    "dyf_optics_image_from_spectrum": [vp, C.c_int, C.c_int, vp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, vp, vp, vp],
    "dyf_optics_image_bytes": [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_int, C.POINTER(C.c_size_t)],
    "dyf_optics_image": [vp, C.c_int, C.c_int, vp, C.c_int, C.c_int, C.c_int,
                         C.c_int, vp, vp, vp],
    "dyf_optics_com_accumulate": [vp, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                  vp, vp, vp, vp],
    "dyf_bragg_fixed_point_bytes": [C.POINTER(BraggArgs), C.c_int,
                                    C.POINTER(C.c_size_t)],
    "dyf_bragg_fixed_point": [C.POINTER(BraggArgs), vp, C.c_int, vp, vp,
                              C.POINTER(FixedOpts), vp, C.c_size_t, vp,
                              C.POINTER(FixedInfo), vp],
}

_lib = None
_lock = threading.Lock()
_initialised = False


def declared_symbols():
    """Names this binding expects the library to export."""
    return sorted(_SIGNATURES) + ["dyf_last_error",
                                  "dyf_laue_march_noalloc_max_bytes"]


def load_library(path=LIB_PATH):
    """Load libdyfox.so (no device calls).  Raises ImportError if absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"{path} is not built; run `python -m dyfox_amd.build` "
                "(hipcc, gfx950) before using the device path")
        lib = C.CDLL(path)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = C.c_int
        lib.dyf_last_error.argtypes = []
        lib.dyf_last_error.restype = C.c_char_p
        lib.dyf_laue_march_noalloc_max_bytes.argtypes = []
        lib.dyf_laue_march_noalloc_max_bytes.restype = C.c_size_t
        _lib = lib
        return lib


def _raise(code, where):
    msg = _lib.dyf_last_error().decode(errors="replace")
    text = f"{where}: {msg}"
    if code == DYF_ENOMEM:
        raise MemoryError(text)
    if code == DYF_EINVAL:
        raise ValueError(text)
    if code == DYF_EKEY:
        raise KeyError(text)
    raise RuntimeError(text)


def call(name, *args):
    """Call a library entry point; raise the mapped exception on failure."""
    lib = lib_ready()
    code = getattr(lib, name)(*args)
    if code != DYF_OK:
        _raise(code, name)


def lib_ready(device=0):
    """The loaded library with its device initialised (first call only)."""
    global _initialised
    lib = load_library()
    if not _initialised:
        with _lock:
            if not _initialised:
                code = lib.dyf_init(device)
                if code != DYF_OK:
                    _raise(code, "dyf_init")
                _initialised = True
    return lib


def device_count():
    """HIP devices visible to this process (0 without a GPU)."""
    lib = load_library()
    n = C.c_int(0)
    lib.dyf_device_count(C.byref(n))
    return n.value


def mem_info():
    """(free, total) device bytes; blocks held by the pool count as free
    (an allocation that does not fit releases them)."""
    f, t = C.c_size_t(0), C.c_size_t(0)
    call("dyf_mem_info", C.byref(f), C.byref(t))
    return f.value + _pool_bytes, t.value


def synchronize():
    call("dyf_synchronize")


class Event:
    """A hipEvent recorded on the library's (default) stream."""

    def __init__(self):
        p = vp()
        call("dyf_event_create", C.byref(p))
        self._e = p.value

    def record(self, stream=None):
        call("dyf_event_record", vp(self._e), stream)
        return self

    def elapsed_ms(self, later):
        """Milliseconds from this event to `later` (waits for `later`)."""
        ms = C.c_float()
        call("dyf_event_elapsed_ms", vp(self._e), vp(later._e), C.byref(ms))
        return float(ms.value)

    def __del__(self):
        if getattr(self, "_e", None) and _lib is not None:
            _lib.dyf_event_destroy(vp(self._e))
            self._e = None


def dptr(a):
    """ctypes double* for a C-contiguous float64 NumPy array."""
    return a.ctypes.data_as(dp)


# Blocks whose last view was dropped, kept for reuse by size class (the
# role CuPy's memory pool plays for the reference).  hipFree waits for the
# whole device, so freeing a solve's output while the next solve is being
# enqueued would drain the queue every call.  Reuse is safe in stream
# order: the library's Python calls enqueue on the null stream, and the
# streams it creates are blocking ones (ordered against the null stream);
# work on such a stream is synchronized before its buffers are dropped.
_pool = {}
_pool_bytes = 0


def _size_class(nbytes):
    """Round up to 256 B below 1 MiB, to a multiple of 1 MiB above."""
    n = max(int(nbytes), 1)
    if n < (1 << 20):
        return (n + 255) & ~255
    return (n + (1 << 20) - 1) & ~((1 << 20) - 1)


def release_pool():
    """Return every pooled block to the device (hipFree)."""
    global _pool_bytes
    if _lib is None:
        return
    for size, ptrs in _pool.items():
        for p in ptrs:
            _lib.dyf_free(vp(p))
    _pool.clear()
    _pool_bytes = 0


# Page-locked host blocks, pooled like the device blocks: a NumPy array
# handed out by pinned_empty keeps its block alive (as the array's base)
# and returns it here when the last view is gone.  At most
# _HOST_POOL_LIMIT bytes are kept; beyond that freed blocks are unpinned.
_host_pool = {}
_host_pool_bytes = 0
_HOST_POOL_LIMIT = 1 << 30


class _HostBlock:
    __slots__ = ("ptr", "nbytes", "__weakref__")

    def __init__(self, nbytes):
        global _host_pool_bytes
        size = _size_class(nbytes)
        held = _host_pool.get(size)
        if held:
            self.ptr = held.pop()
            _host_pool_bytes -= size
        else:
            p = vp()
            call("dyf_host_alloc", C.byref(p), C.c_size_t(size))
            self.ptr = p.value
        self.nbytes = size

    @property
    def __array_interface__(self):
        return {"shape": (self.nbytes,), "typestr": "|u1",
                "data": (self.ptr, False), "version": 3}

    def __del__(self):
        global _host_pool_bytes
        if self.ptr and _lib is not None:
            if _host_pool_bytes + self.nbytes <= _HOST_POOL_LIMIT:
                _host_pool.setdefault(self.nbytes, []).append(self.ptr)
                _host_pool_bytes += self.nbytes
            else:
                _lib.dyf_host_free(vp(self.ptr))
            self.ptr = None


def pinned_empty(shape, dtype):
    """An uninitialised C-contiguous NumPy array in page-locked memory."""
    dtype = np.dtype(dtype)
    shape = tuple(int(n) for n in np.atleast_1d(shape))
    count = int(np.prod(shape, dtype=np.int64))
    raw = np.asarray(_HostBlock(count * dtype.itemsize))
    return raw[:count * dtype.itemsize].view(dtype).reshape(shape)


class _Block:
    """One device allocation; returned to the pool when the last view lets
    go."""

    __slots__ = ("ptr", "nbytes")

    def __init__(self, nbytes):
        global _pool_bytes
        size = _size_class(nbytes)
        held = _pool.get(size)
        if held:
            self.ptr = held.pop()
            _pool_bytes -= size
        else:
            p = vp()
            try:
                call("dyf_malloc", C.byref(p), C.c_size_t(size))
            except MemoryError:
                if not _pool_bytes:
                    raise
                release_pool()
                call("dyf_malloc", C.byref(p), C.c_size_t(size))
            self.ptr = p.value
        self.nbytes = size

    def __del__(self):
        global _pool_bytes
        if self.ptr and _lib is not None:
            _pool.setdefault(self.nbytes, []).append(self.ptr)
            _pool_bytes += self.nbytes
            self.ptr = None


class DeviceArray:
    """A C-contiguous array in device memory.

    Stands where the reference returns a CuPy array: `.get()` copies to
    the host (so `to_numpy` works unchanged), `.shape`/`.dtype` describe
    it, and integer or slice indexing along the first axis gives a view.
    """

    def __init__(self, shape, dtype, block=None, offset=0):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) \
            if not isinstance(shape, tuple) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = math.prod(self.shape) * self.dtype.itemsize
        self._block = block if block is not None else _Block(self.nbytes)
        self._offset = int(offset)

    # ---- construction ----
    @classmethod
    def empty(cls, shape, dtype):
        return cls(tuple(shape), dtype)

    @classmethod
    def zeros(cls, shape, dtype):
        a = cls(tuple(shape), dtype)
        if a.nbytes:
            call("dyf_memset", vp(a.ptr), 0, C.c_size_t(a.nbytes))
        return a

    @classmethod
    def from_host(cls, host, dtype=None):
        host = np.ascontiguousarray(host, dtype=dtype)
        a = cls(host.shape, host.dtype)
        if a.nbytes:
            call("dyf_memcpy_h2d", vp(a.ptr), vp(host.ctypes.data),
                 C.c_size_t(a.nbytes))
        return a

    # ---- properties ----
    @property
    def ptr(self):
        return self._block.ptr + self._offset

    @property
    def size(self):
        return math.prod(self.shape)

    @property
    def ndim(self):
        return len(self.shape)

    def __len__(self):
        return self.shape[0]

    def get(self, out=None):
        """Host copy (NumPy), into `out` when given: a host array of the
        same dtype and shape that is C-contiguous, or whose first axis is
        strided over C-contiguous rows (a slice of a larger array)."""
        if out is None:
            out = np.empty(self.shape, self.dtype)
        if out.dtype != self.dtype or out.shape != self.shape:
            raise ValueError(f"out is {out.dtype} {out.shape}, the array "
                             f"{self.dtype} {self.shape}")
        if not self.nbytes:
            return out
        if out.flags.c_contiguous:
            call("dyf_memcpy_d2h", vp(out.ctypes.data), vp(self.ptr),
                 C.c_size_t(self.nbytes))
            return out
        row = self.nbytes // self.shape[0]
        if out.ndim < 2 or not out[0].flags.c_contiguous:
            raise ValueError("out must be C-contiguous or strided along its "
                             "first axis only")
        call("dyf_memcpy2d_d2h", vp(out.ctypes.data), C.c_size_t(out.strides[0]),
             vp(self.ptr), C.c_size_t(row), C.c_size_t(row),
             C.c_size_t(self.shape[0]))
        return out

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def reshape(self, *shape):
        if len(shape) == 1 and isinstance(shape[0], (tuple, list)):
            shape = tuple(shape[0])
        shape = list(shape)
        if -1 in shape:
            j = shape.index(-1)
            rest = int(np.prod([s for s in shape if s != -1]))
            shape[j] = self.size // max(rest, 1)
        if int(np.prod(shape)) != self.size:
            raise ValueError(f"cannot reshape {self.shape} to {tuple(shape)}")
        return DeviceArray(tuple(shape), self.dtype, self._block, self._offset)

    def __getitem__(self, key):
        if not self.shape:
            raise IndexError("0-d device array")
        row = math.prod(self.shape[1:]) * self.dtype.itemsize
        if isinstance(key, (int, np.integer)):
            k = int(key)
            if k < 0:
                k += self.shape[0]
            if not 0 <= k < self.shape[0]:
                raise IndexError(key)
            return DeviceArray(self.shape[1:], self.dtype, self._block,
                               self._offset + k * row)
        if isinstance(key, slice):
            start, stop, step = key.indices(self.shape[0])
            if step != 1:
                raise IndexError("device views need a unit step")
            n = max(stop - start, 0)
            return DeviceArray((n,) + self.shape[1:], self.dtype, self._block,
                               self._offset + start * row)
        raise IndexError("device arrays index the first axis only")

    def copy_from(self, other):
        """Copy a device array or host array of the same size into self."""
        if isinstance(other, DeviceArray):
            if other.nbytes != self.nbytes:
                raise ValueError("size mismatch")
            call("dyf_memcpy_d2d", vp(self.ptr), vp(other.ptr),
                 C.c_size_t(self.nbytes))
        else:
            host = np.ascontiguousarray(other, dtype=self.dtype)
            if host.nbytes != self.nbytes:
                raise ValueError("size mismatch")
            call("dyf_memcpy_h2d", vp(self.ptr), vp(host.ctypes.data),
                 C.c_size_t(self.nbytes))
        return self

    def astype(self, dtype):
        """Host round trip conversion (small arrays; not on the hot path)."""
        dtype = np.dtype(dtype)
        if dtype == self.dtype:
            return self
        return DeviceArray.from_host(self.get().astype(dtype))

    def __repr__(self):
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"


def asarray(a, dtype=None):
    """DeviceArray from a host array (or the same DeviceArray)."""
    if isinstance(a, DeviceArray):
        return a if dtype is None or np.dtype(dtype) == a.dtype \
            else a.astype(dtype)
    return DeviceArray.from_host(np.asarray(a), dtype=dtype)
