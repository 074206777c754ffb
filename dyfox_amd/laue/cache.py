"""Exit waves stored on disk, keyed by everything that produced them.

On-disk format kept identical to the reference's laue/cache.py:23-75 so
a cache directory written by either implementation is found by the
other:

    key  = sha1(json.dumps({segments, s_dev, grid, crystal, precision,
                            store, version}, sort_keys=True))[:16]
    path = {directory}/{precision}_{tag}_{key}.npy   (tag optional)
    data = (n_e, Ny, Nx) stored as `store_dtype` (complex64 by default)

A hit comes back as a device array of the stored dtype; a miss returns
the freshly solved stack (complex128 in FP64), exactly as the reference.
"""

import hashlib
import json
import os

import numpy as np

from .. import _native as N
from ..backend import FP64, to_numpy
from .solver import solve_dislocation

VERSION = 9


class WaveCache:
    """A directory of solved exit waves."""

    def __init__(self, directory, version=VERSION, store_dtype=np.complex64):
        self.directory = directory
        self.version = version
        self.store_dtype = store_dtype

    def key(self, seg_list, s_dev_arr, grid, xtal, precision=FP64):
        """16 hex characters identifying the solve's inputs."""
        segs = [[[float(v) for v in np.asarray(part, float).ravel()]
                 for part in seg] for seg in seg_list]
        payload = {"segments": segs,
                   "s_dev": [float(s) for s in np.atleast_1d(s_dev_arr)],
                   "grid": grid.key(), "crystal": xtal.key(),
                   "precision": precision.name,
                   "store": np.dtype(self.store_dtype).name,
                   "version": self.version}
        blob = json.dumps(payload, sort_keys=True).encode()
        return hashlib.sha1(blob).hexdigest()[:16]

    def path(self, seg_list, s_dev_arr, grid, xtal, precision=FP64, tag=""):
        key = self.key(seg_list, s_dev_arr, grid, xtal, precision)
        stem = f"{tag}_{key}" if tag else key
        return os.path.join(self.directory, f"{precision.name}_{stem}.npy")

    def solve(self, seg_list, s_dev_arr, grid, xtal, *, precision=FP64,
              H_gpu=None, batch=64, tag="", use_cache=True):
        """The exit-wave stack: from disk on a hit, solved on a miss."""
        if not use_cache:
            return solve_dislocation(seg_list, s_dev_arr, grid, xtal,
                                     H_gpu=H_gpu, batch=batch,
                                     precision=precision)
        path = self.path(seg_list, s_dev_arr, grid, xtal, precision, tag)
        if os.path.exists(path):
            return N.DeviceArray.from_host(np.load(path))
        stack = solve_dislocation(seg_list, s_dev_arr, grid, xtal,
                                  H_gpu=H_gpu, batch=batch,
                                  precision=precision)
        os.makedirs(self.directory, exist_ok=True)
        np.save(path, to_numpy(stack).astype(self.store_dtype))
        return stack

    def __repr__(self):
        return f"WaveCache({self.directory!r}, v{self.version})"
