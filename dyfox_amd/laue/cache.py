"""Exit waves stored on disk, keyed by everything that produced them.

The on-disk format is the reference's (laue/cache.py:23-75), so a cache
directory written by either implementation is found by the other:

* the key is the first 16 hex digits of the SHA-1 of a JSON document
  (keys sorted, the `json` module's default separators) holding the
  segments as nested float lists, the deviations, `grid.key()`,
  `xtal.key()`, the precision's name, the stored dtype's name and the
  format version (9);
* the file is `{directory}/{precision}_{key}.npy`, or
  `{directory}/{precision}_{tag}_{key}.npy` with a tag;
* the data is the `(n_e, Ny, Nx)` stack cast to `store_dtype`
  (complex64 unless asked otherwise).

A hit comes back as a device array of the stored dtype; a miss solves,
writes the file and returns the freshly solved stack (complex128 in
FP64), so a hit is not bit-identical to a fresh FP64 solve.
"""

import hashlib
import json
import os

import numpy as np

from .. import _native as N
from ..backend import FP64, to_numpy
from .solver import solve_dislocation

VERSION = 9
KEY_HEX_DIGITS = 16


def _floats(x):
    return [float(v) for v in np.ravel(np.asarray(x, dtype=float))]


def _document(seg_list, s_dev_arr, grid, xtal, precision, store, version):
    """The canonical JSON text a key is the hash of."""
    doc = {
        "version": version,
        "store": np.dtype(store).name,
        "precision": precision.name,
        "crystal": xtal.key(),
        "grid": grid.key(),
        "s_dev": _floats(np.atleast_1d(s_dev_arr)),
        # each segment as [start, end, Burgers vector], each a float list
        "segments": [[_floats(part) for part in seg] for seg in seg_list],
    }
    return json.dumps(doc, sort_keys=True)


class WaveCache:
    """A directory of solved exit-wave stacks."""

    def __init__(self, directory, version=VERSION, store_dtype=np.complex64):
        self.directory = directory
        self.version = version
        self.store_dtype = store_dtype

    def key(self, seg_list, s_dev_arr, grid, xtal, precision=FP64):
        """16 hex digits naming the solve's inputs."""
        text = _document(seg_list, s_dev_arr, grid, xtal, precision,
                         self.store_dtype, self.version)
        return hashlib.sha1(text.encode()).hexdigest()[:KEY_HEX_DIGITS]

    def path(self, seg_list, s_dev_arr, grid, xtal, precision=FP64, tag=""):
        """The file such a solve is stored in (the tag is for the eye)."""
        parts = [precision.name]
        if tag:
            parts.append(tag)
        parts.append(self.key(seg_list, s_dev_arr, grid, xtal, precision))
        return os.path.join(self.directory, "_".join(parts) + ".npy")

    def solve(self, seg_list, s_dev_arr, grid, xtal, *, precision=FP64,
              H_gpu=None, batch=64, tag="", use_cache=True):
        """The exit-wave stack: read on a hit, solved and written on a
        miss; `use_cache=False` solves without touching the directory."""
        def fresh():
            return solve_dislocation(seg_list, s_dev_arr, grid, xtal,
                                     H_gpu=H_gpu, batch=batch,
                                     precision=precision)

        if not use_cache:
            return fresh()
        where = self.path(seg_list, s_dev_arr, grid, xtal, precision, tag)
        if os.path.exists(where):
            return N.DeviceArray.from_host(np.load(where))
        stack = fresh()
        os.makedirs(self.directory, exist_ok=True)
        np.save(where, to_numpy(stack).astype(self.store_dtype))
        return stack

    def __repr__(self):
        return "WaveCache(%r, v%s)" % (self.directory, self.version)
