"""The device Bragg path against fixtures the reference itself wrote.

tests/golden/ holds what the reference's NumPy/SciPy path returned
(bragg/solver.py, dislocations.py; written by
tests/golden/make_reference_fixtures.py).  Bars: e^{iH} and the strained
march to 1e-12 relative; exit waves of converged solves to 1e-8 relative
(both sides stop at a true residual <= 1e-10, and the GMRES here
orthogonalises by CGS2 where SciPy uses MGS, so the iterates differ in
their last bits); the integer info (iterations, retries, splits,
unconverged chunks, converged, gmres_code) identical.
"""

import numpy as np
import pytest

from dyfox_amd.bragg import (BraggConfig, bragg_g_vector, solve_bragg,
                             solve_bragg_stable, solve_surface_batch)
from dyfox_amd.bragg import solver as BS
from tests import golden

pytestmark = pytest.mark.gpu

INT_KEYS = ("iterations", "retries", "splits", "unconverged_chunks",
            "converged")


@pytest.fixture(scope="module")
def fx():
    return golden.bragg()


@pytest.fixture(scope="module")
def cfg(diamond400, fx):
    g = fx[1]["grid"]
    return BraggConfig(diamond400, t_crystal=g["t_crystal"], dx=g["dx"],
                       half_x=g["half_x"])


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()
                 / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", ["edge", "screw", "fractional"])
def test_phase_matches_reference(dev, diamond400, cfg, fx, name):
    arrays, _ = fx
    segs = golden.segments(arrays[f"{name}_segments"])
    g = bragg_g_vector(diamond400)
    E = BS.chunk_phase(segs, cfg, diamond400, g, arrays["ys"][1:2]).get()
    want = np.exp(1j * arrays[f"{name}_H_plane"])
    assert _rel(E[:, 0], want) < 1e-12


def test_strained_march_matches_reference(dev, diamond400, cfg, fx):
    arrays, _ = fx
    segs = golden.segments(arrays["edge_segments"])
    g = bragg_g_vector(diamond400)
    E = BS.chunk_phase(segs, cfg, diamond400, g, arrays["ys"][:3])
    s = [float(arrays["s_dev"][2])]
    hom = BS.march(cfg, diamond400, s, E, Dg0=arrays["march_v"][None]).get()
    par = BS.march(cfg, diamond400, s, E,
                   incident=cfg.incident(n=1)[0]).get()
    assert _rel(hom[0], arrays["march_homogeneous"]) < 1e-12
    assert _rel(par[0], arrays["march_particular"]) < 1e-12


@pytest.mark.parametrize("tag", ["batch_edge", "batch_screw",
                                 "batch_fractional", "batch_edge_retry_split",
                                 "batch_edge_split",
                                 "batch_edge_retry_then_split",
                                 "batch_fractional_retry",
                                 "batch_screw_depth_limit"])
def test_surface_batch_matches_reference(dev, diamond400, cfg, fx, tag):
    """The drop-in default (joint chunks) against the reference, one
    deviation per call as the reference is called, and all three in one
    batched call."""
    arrays, meta = fx
    case = meta["cases"][tag]
    segs = golden.segments(arrays[f"{case['segments']}_segments"])
    g = bragg_g_vector(diamond400)
    want = arrays[f"{tag}_surface"]
    s_all = arrays["s_dev"]
    print(f"golden batch {tag} {case['kwargs']}", flush=True)
    outs, infos = solve_surface_batch(segs, s_all, cfg, arrays["ys"],
                                      diamond400, g, **case["kwargs"])
    for j, s in enumerate(s_all):
        one, info1 = solve_surface_batch(segs, float(s), cfg, arrays["ys"],
                                         diamond400, g, **case["kwargs"])
        ref = case["info"][j]
        for info, out in ((info1, one), (infos[j], outs[j])):
            for k in INT_KEYS:
                assert info[k] == ref[k], (tag, j, k, info[k], ref[k])
            assert info["residual"] == pytest.approx(ref["residual"],
                                                     rel=0.05, abs=1e-14)
            tol = 1e-8 if ref["converged"] else 1e-6
            assert _rel(out, want[j]) < tol, (tag, j)


def test_solve_bragg_stable_matches_reference(dev, diamond400, cfg, fx):
    arrays, meta = fx
    m = meta["stable_edge"]
    segs = golden.segments(arrays["edge_segments"])
    g = bragg_g_vector(diamond400)
    ew, info = solve_bragg_stable(segs, float(arrays["s_dev"][m["s_index"]]),
                                  cfg, diamond400, g, y_plane=m["y_plane"])
    assert info["gmres_code"] == m["info"]["gmres_code"]
    assert info["iterations"] == m["info"]["iterations"]
    assert info["residual"] < 1e-10
    assert _rel(ew, arrays["stable_edge"]) < 1e-8


def test_fixed_point_matches_reference(dev, diamond400, cfg, fx):
    arrays, meta = fx
    m = meta["fixed_point_screw"]
    segs = golden.segments(arrays["screw_segments"])
    g = bragg_g_vector(diamond400)
    ew, info = solve_bragg(segs, float(arrays["s_dev"][m["s_index"]]), cfg,
                           diamond400, g)
    assert info["converged"] == m["info"]["converged"]
    assert info["iterations"] == m["info"]["iterations"]
    assert _rel(ew, arrays["fixed_point_screw"]) < 1e-10
