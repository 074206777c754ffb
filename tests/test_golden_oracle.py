"""The oracle and the host API against fixtures the reference itself wrote.

tests/golden/make_reference_fixtures.py ran the reference's NumPy/SciPy
Bragg path (bragg/solver.py, dislocations.py) and its cache key code
(laue/cache.py) on small inputs.  These tests pin what no closed form
fixes: the Volterra edge terms on general lines, the signs of e^{-iH} in
C_h and e^{+iH} in C_hbar, the e^{-iH[0]} exit-phase restoration, the
exact integer `info` of the retry/grow/split policy, and the WaveCache
key, path and repr byte for byte.
"""

import numpy as np
import pytest

from dyfox_amd.backend import precision
from dyfox_amd.bragg import BraggConfig, bragg_g_vector
from dyfox_amd.crystal import DIAMOND, CrystalParams, lab_frame
from dyfox_amd.laue import LaueGrid, WaveCache
from oracle import bragg as OB, volterra as OV
from tests import golden

INT_KEYS = ("iterations", "retries", "splits", "unconverged_chunks",
            "converged")


@pytest.fixture(scope="module")
def fx():
    return golden.bragg()


@pytest.fixture(scope="module")
def cfg(diamond400, fx):
    g = fx[1]["grid"]
    c = BraggConfig(diamond400, t_crystal=g["t_crystal"], dx=g["dx"],
                    half_x=g["half_x"])
    assert (c.Nz, c.Nx) == (g["Nz"], g["Nx"])
    return c


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()
                 / np.abs(np.asarray(b)).max())


@pytest.mark.parametrize("name", ["edge", "screw", "fractional"])
def test_volterra_plane_matches_reference(diamond400, cfg, fx, name):
    arrays, _ = fx
    segs = golden.segments(arrays[f"{name}_segments"])
    g = bragg_g_vector(diamond400)
    H = OV.h_plane(segs, g, diamond400.nu, diamond400.a_core, cfg.Nx,
                   cfg.Nz, cfg.dx, cfg.dz, float(arrays["ys"][1]))
    want = arrays[f"{name}_H_plane"]
    assert np.abs(H - want).max() <= 1e-12 * np.abs(want).max()


def test_strained_march_matches_reference(diamond400, cfg, fx):
    """The couplings' signs: a wrong sign of e^{+-iH} in either coupling
    moves the result by O(1)."""
    arrays, _ = fx
    segs = golden.segments(arrays["edge_segments"])
    g = bragg_g_vector(diamond400)
    Ch, Cb, _ = OB.couplings(segs, cfg, diamond400, g, arrays["ys"][:3])
    e0, eg = OB.step_constants(arrays["s_dev"][2], complex(diamond400.sig0),
                               cfg.ds)
    zero = np.zeros((3, cfg.Nx), complex)
    hom = OB.march_down(zero, arrays["march_v"], Ch, Cb, e0, eg, cfg.ds,
                        cfg.Nz)
    par = OB.march_down(OB.incident(cfg, 3), zero, Ch, Cb, e0, eg, cfg.ds,
                        cfg.Nz)
    assert _rel(hom, arrays["march_homogeneous"]) < 1e-13
    assert _rel(par, arrays["march_particular"]) < 1e-13
    # planted mistake: the couplings swapped (conjugate phases) must fail
    bad = OB.march_down(zero, arrays["march_v"], Cb * complex(
        diamond400.sig_h) / complex(diamond400.sig_hbar), Ch * complex(
        diamond400.sig_hbar) / complex(diamond400.sig_h), e0, eg, cfg.ds,
        cfg.Nz)
    assert _rel(bad, arrays["march_homogeneous"]) > 1e-3


@pytest.mark.parametrize("tag", ["batch_edge", "batch_screw",
                                 "batch_fractional", "batch_edge_retry_split",
                                 "batch_edge_split",
                                 "batch_edge_retry_then_split",
                                 "batch_fractional_retry",
                                 "batch_screw_depth_limit"])
def test_oracle_surface_batch_matches_reference(diamond400, cfg, fx, tag):
    arrays, meta = fx
    case = meta["cases"][tag]
    segs = golden.segments(arrays[f"{case['segments']}_segments"])
    g = bragg_g_vector(diamond400)
    want = arrays[f"{tag}_surface"]
    for j, s in enumerate(arrays["s_dev"]):
        out, info = OB.solve_surface_batch(segs, float(s), cfg, arrays["ys"],
                                           diamond400, g, **case["kwargs"])
        ref = case["info"][j]
        for k in INT_KEYS:
            assert info[k] == ref[k], (tag, j, k)
        assert info["residual"] == pytest.approx(ref["residual"], rel=1e-3,
                                                 abs=1e-15)
        assert _rel(out, want[j]) < 1e-9, (tag, j)


def test_oracle_fixed_point_matches_reference(diamond400, cfg, fx):
    arrays, meta = fx
    m = meta["fixed_point_screw"]
    segs = golden.segments(arrays["screw_segments"])
    g = bragg_g_vector(diamond400)
    ew, (its, res, conv) = OB.fixed_point(
        segs, float(arrays["s_dev"][m["s_index"]]), cfg, diamond400, g)
    assert (its, conv) == (m["info"]["iterations"], m["info"]["converged"])
    assert _rel(ew, arrays["fixed_point_screw"]) < 1e-12


def _crystals():
    d400 = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002,
                                         label="Diamond(400)")
    d220 = CrystalParams.from_reflection(
        DIAMOND, (2, 2, 0), 12.0, U_lab=lab_frame((0, 0, 1), (2, 2, 0)),
        f_prime=0.003, f_dprime=0.002)
    return {"d400": d400, "d220": d220}


@pytest.mark.parametrize("i", [0, 1, 2])
def test_wavecache_key_and_path_match_reference(i):
    """Cache interoperability: the same inputs give the same file name as
    the reference's WaveCache, byte for byte."""
    rec = golden.cache_keys()[i]
    xt = _crystals()[rec["xtal"]]
    grid = LaueGrid(xt, **rec["grid"])
    segs = golden.segments(np.asarray(rec["segments"]).reshape(-1, 3, 3))
    wc = WaveCache(rec["directory"], version=rec["version"],
                   store_dtype=np.dtype(rec["store"]))
    prec = precision(rec["precision"])
    s = np.asarray(rec["s_dev"])
    assert wc.key(segs, s, grid, xt, prec) == rec["key"]
    assert wc.path(segs, s, grid, xt, prec, tag=rec["tag"]) == rec["path"]
    assert repr(wc) == rec["repr"]
    # the key changes with any input (a planted change of one deviation)
    assert wc.key(segs, s * (1 + 1e-12) + 1e-3, grid, xt, prec) != rec["key"]


def test_wavecache_keys_separate_the_physics(diamond400):
    """Every input that changes the exit waves changes the key (the
    reference's key-separation test, tests/test_laue_gpu.py:275-288):
    segments, deviations, grid, crystal, precision and stored dtype; the
    same inputs give the same key."""
    from dyfox_amd.dislocations import burgers_vector, straight_segment
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6)
    b = burgers_vector(np.eye(3), [1, 0, -1], diamond400.b_mag)
    segs = straight_segment(np.array([1.0, 0.0, 1.0]) / np.sqrt(2), b,
                            [0.0, grid.dx / 2, 20e-6], 400e-6)
    s = np.array([0.0, 1.2e6])
    wc = WaveCache("/tmp/waves")
    k = wc.key(segs, s, grid, diamond400)
    assert k == WaveCache("/elsewhere").key(segs, s, grid, diamond400)
    flipped = straight_segment(np.array([1.0, 0.0, 1.0]) / np.sqrt(2), -b,
                               [0.0, grid.dx / 2, 20e-6], 400e-6)
    d220 = CrystalParams.from_reflection(DIAMOND, (2, 2, 0), 12.0,
                                         f_prime=0.003, f_dprime=0.002)
    others = [
        wc.key(flipped, s, grid, diamond400),
        wc.key([], s, grid, diamond400),
        wc.key(segs, s[:1], grid, diamond400),
        wc.key(segs, s + 1.0, grid, diamond400),
        wc.key(segs, s, grid.replace(beam_thickness=2e-6), diamond400),
        wc.key(segs, s, LaueGrid(diamond400, 40e-6, 0.25e-6, 10e-6),
               diamond400),
        wc.key(segs, s, grid, d220),
        wc.key(segs, s, grid, diamond400, precision("fp32")),
        WaveCache("/tmp/waves", store_dtype=np.complex128).key(
            segs, s, grid, diamond400),
        WaveCache("/tmp/waves", version=8).key(segs, s, grid, diamond400),
    ]
    assert len(set(others + [k])) == len(others) + 1
