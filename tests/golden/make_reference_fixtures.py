"""Write the golden fixtures of tests/golden/ by running the reference.

Run in the development container only (the reference is not on the GPU
machines and never travels there in any form):

    python tests/golden/make_reference_fixtures.py /path/to/reference

It imports the reference's pure-NumPy/SciPy Bragg solver and its cache
key code (CPU only; nothing here needs CuPy) and records, for small
strained inputs, exactly what the reference returns:

* bragg_reference.npz / bragg_reference.json
    - `H_plane` on one plane for an edge, a screw and a general mixed line
      of non-integer g.b (dislocations.py:171-211): the Volterra terms;
    - `_march_down` of a random D_g(0) through a strained 3-plane chunk
      (bragg/solver.py:28-35, with the couplings of 256-261): the signs of
      e^{-iH} in C_h and e^{+iH} in C_hbar;
    - `solve_surface_batch` (bragg/solver.py:182-379, xp=numpy) for an
      edge, a screw and a fractional-g.b line over 4 planes and 3
      deviations: exit waves (with the e^{-iH[0]} restoration of 365) and
      the integer info (iterations, retries, splits, unconverged_chunks);
    - the same with a deliberately small Krylov space, so the retry and
      split branches (284-285, 314-318, 345-360) run;
    - `solve_bragg_stable` (135-179) and the fixed point `solve_bragg`
      (63-132) on one plane.
* wavecache_keys.json
    - `WaveCache.key` and `.path` (laue/cache.py:35-57) for three inputs.

Inputs are stored as numbers (segments as (n, 3, 3) arrays: start, end,
Burgers vector), so the tests rebuild them with this project's classes.
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# every Bragg fixture is solved on this grid: Nz 19, Nx 124
T_CRYSTAL, DX, HALF_X = 4e-6, 0.5e-6, 4e-6
S_URAD = [-10.0, 0.0, 15.0]


def _segs_array(seg_list):
    return np.array([[np.asarray(p, float) for p in seg] for seg in seg_list])


def _info(d):
    out = {}
    for k, v in d.items():
        if isinstance(v, (bool, np.bool_)):
            out[k] = bool(v)
        elif isinstance(v, (int, np.integer)):
            out[k] = int(v)
        else:
            out[k] = float(v)
    return out


def main(ref_root):
    sys.path.insert(0, ref_root)
    from crystal import DIAMOND, CrystalParams
    from units import deviation_from_urad
    from dislocations import H_plane, straight_segment
    from bragg.geometry import BraggConfig, bragg_g_vector, threading_segment
    from bragg import solver as RS

    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002,
                                         label="Diamond(400)")
    cfg = BraggConfig(xtal, t_crystal=T_CRYSTAL, dx=DX, half_x=HALF_X)
    g = bragg_g_vector(xtal)
    gm = float(g[2])
    s_dev = deviation_from_urad(np.array(S_URAD), xtal)
    ys = (np.arange(4) - 2) * cfg.dx + cfg.dx / 2

    # lines: threading edge and screw (g.b = 2 and -2), and a general
    # inclined line with a Burgers vector chosen so g.b = 0.7
    edge = threading_segment("Edge", +1, xtal.b_mag, gm, y0=0.3e-6,
                             x0=-0.4e-6)
    screw = threading_segment("Screw", -1, xtal.b_mag, gm, y0=-0.2e-6,
                              x0=0.6e-6, azimuth_deg=30.0)
    xi = np.array([0.3, -0.5, 0.8])
    xi /= np.linalg.norm(xi)
    bz = 0.7 / gm
    bv = np.array([0.6, 0.45, 0.0]) * xtal.b_mag + np.array([0, 0, bz])
    frac = straight_segment(xi, bv, np.array([0.25e-6, 0.1e-6, 1.7e-6]),
                            200e-6)
    cases = {"edge": edge, "screw": screw, "fractional": frac}

    arrays, meta = {}, {"grid": {"t_crystal": T_CRYSTAL, "dx": DX,
                                 "half_x": HALF_X, "Nz": cfg.Nz,
                                 "Nx": cfg.Nx},
                        "s_urad": S_URAD, "cases": {}}
    arrays["s_dev"] = s_dev
    arrays["ys"] = ys
    for name, segs in cases.items():
        arrays[f"{name}_segments"] = _segs_array(segs)
        arrays[f"{name}_H_plane"] = H_plane(segs, cfg, g, xtal.nu,
                                            xtal.a_core, xp=np,
                                            y_plane=float(ys[1]))

    # the march on a strained chunk of 3 planes, Dg(0) random, D0(0) = 0,
    # and the particular response to the unit incident wave
    rng = np.random.default_rng(20260909)
    sh, sb = complex(xtal.sig_h), complex(xtal.sig_hbar)
    Hs = np.stack([H_plane(edge, cfg, g, xtal.nu, xtal.a_core, xp=np,
                           y_plane=float(y)) for y in ys[:3]], axis=1)
    Ch = (sh * np.exp(-1j * Hs)).astype(np.complex128)
    Cb = (sb * np.exp(+1j * Hs)).astype(np.complex128)
    alpha = -2.0 * np.pi * float(s_dev[2])
    e0 = complex(np.exp(complex(xtal.sig0) * cfg.ds))
    eg = complex(np.exp((complex(xtal.sig0) + 1j * alpha) * cfg.ds))
    v = rng.normal(size=(3, cfg.Nx)) + 1j * rng.normal(size=(3, cfg.Nx))
    zero = np.zeros((3, cfg.Nx), np.complex128)
    arrays["march_v"] = v
    arrays["march_homogeneous"] = RS._march_down(zero, v, Ch, Cb, e0, eg,
                                                 cfg.ds, cfg.Nz, np)
    arrays["march_particular"] = RS._march_down(
        cfg.incident(n=3), zero, Ch, Cb, e0, eg, cfg.ds, cfg.Nz, np)

    def batch(name, segs, tag, **kw):
        outs, infos = [], []
        for s in s_dev:
            o, inf = RS.solve_surface_batch(segs, float(s), cfg, ys, xtal, g,
                                            xp=np, **kw)
            outs.append(o)
            infos.append(_info(inf))
        arrays[f"{tag}_surface"] = np.stack(outs)
        meta["cases"][tag] = {"segments": name, "kwargs": kw,
                              "info": infos}

    for name, segs in cases.items():
        batch(name, segs, f"batch_{name}")
    # a Krylov space too small for the chunk: retries then splits
    batch("edge", edge, "batch_edge_retry_split", restart=1, per_plane=0,
          maxiter=1, restart_cap=3, grow=2, tries=3)
    batch("edge", edge, "batch_edge_split", restart=2, per_plane=0,
          maxiter=2, restart_cap=2, grow=2, tries=1)
    batch("edge", edge, "batch_edge_retry_then_split", restart=1,
          per_plane=0, maxiter=1, restart_cap=3, grow=3, tries=2)
    batch("fractional", frac, "batch_fractional_retry", restart=2,
          per_plane=1, maxiter=1, restart_cap=8, grow=2, tries=2)
    batch("screw", screw, "batch_screw_depth_limit", restart=1, per_plane=0,
          maxiter=1, restart_cap=2, grow=2, tries=2, max_depth=1)

    o, inf = RS.solve_bragg_stable(edge, float(s_dev[0]), cfg, xtal, g,
                                   xp=np, y_plane=float(ys[2]))
    arrays["stable_edge"] = np.asarray(o)
    meta["stable_edge"] = {"y_plane": float(ys[2]), "s_index": 0,
                           "info": _info(inf)}
    o, inf = RS.solve_bragg(screw, float(s_dev[1]), cfg, xtal, g, xp=np)
    arrays["fixed_point_screw"] = np.asarray(o)
    meta["fixed_point_screw"] = {"s_index": 1, "info": _info(inf)}

    np.savez_compressed(os.path.join(HERE, "bragg_reference.npz"), **arrays)
    with open(os.path.join(HERE, "bragg_reference.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)

    write_cache_keys(ref_root)


def write_cache_keys(ref_root):
    from crystal import DIAMOND, CrystalParams, lab_frame
    from backend import FP32, FP64
    from dislocations import burgers_vector, straight_segment
    from laue.cache import WaveCache
    from laue.grid import LaueGrid
    from units import deviation_from_urad

    d400 = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002,
                                         label="Diamond(400)")
    d220 = CrystalParams.from_reflection(
        DIAMOND, (2, 2, 0), 12.0, U_lab=lab_frame((0, 0, 1), (2, 2, 0)),
        f_prime=0.003, f_dprime=0.002)
    xi = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    b = burgers_vector(np.eye(3), [1, 0, -1], d400.b_mag, +1)
    seg = straight_segment(xi, b, [-100e-6 * d400.tan_tB, 0.125e-6, 100e-6],
                           800e-6)
    seg2 = straight_segment([0.0, 1.0, 0.0],
                            burgers_vector(np.eye(3), [1, 1, 0], d400.b_mag,
                                           -1), [0.5e-6, 0.0, 20e-6], 400e-6)
    inputs = [
        dict(xtal="d400", segments=seg, s_urad=[0.0, 117.5],
             grid=dict(t_crystal=200e-6, dx=0.25e-6, beam_width=120e-6),
             precision="fp64", store="complex64", version=9, tag="",
             directory="output/data/waves"),
        dict(xtal="d400", segments=seg + seg2, s_urad=[-50.0],
             grid=dict(t_crystal=40e-6, dx=0.25e-6, beam_width=8e-6,
                       beam_thickness=1e-6, grid_pad=12),
             precision="fp32", store="complex64", version=9, tag="edge_+1",
             directory="/tmp/waves"),
        dict(xtal="d220", segments=[], s_urad=[-200.0, 0.0, 200.0],
             grid=dict(t_crystal=60e-6, dx=0.25e-6, beam_width=60e-6,
                       beam_sigma=0.5e-6),
             precision="fp64", store="complex128", version=8, tag="pristine",
             directory="cache"),
    ]
    xtals = {"d400": d400, "d220": d220}
    precs = {"fp64": FP64, "fp32": FP32}
    out = []
    for inp in inputs:
        xt = xtals[inp["xtal"]]
        grid = LaueGrid(xt, **inp["grid"])
        s = deviation_from_urad(np.array(inp["s_urad"]), xt)
        wc = WaveCache(inp["directory"], version=inp["version"],
                       store_dtype=np.dtype(inp["store"]))
        key = wc.key(inp["segments"], s, grid, xt, precs[inp["precision"]])
        path = wc.path(inp["segments"], s, grid, xt, precs[inp["precision"]],
                       tag=inp["tag"])
        rec = dict(inp)
        rec["segments"] = _segs_array(inp["segments"]).tolist()
        rec["s_dev"] = [float(v) for v in s]
        rec["key"] = key
        rec["path"] = path
        rec["repr"] = repr(wc)
        out.append(rec)
    with open(os.path.join(HERE, "wavecache_keys.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(os.path.abspath(sys.argv[1] if len(sys.argv) > 1
                         else "/root/reference"))
