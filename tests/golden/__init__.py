"""Golden fixtures written by running the reference (make_reference_fixtures.py)
and the loader the tests share."""

import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def bragg():
    """(arrays, meta) of the reference's Bragg outputs."""
    arrays = np.load(os.path.join(HERE, "bragg_reference.npz"))
    with open(os.path.join(HERE, "bragg_reference.json")) as f:
        meta = json.load(f)
    return arrays, meta


def cache_keys():
    with open(os.path.join(HERE, "wavecache_keys.json")) as f:
        return json.load(f)


def segments(arr):
    """An (n, 3, 3) array back into the solvers' [(ra, rb, b), ...]."""
    return [tuple(np.asarray(p, float) for p in seg) for seg in arr]
