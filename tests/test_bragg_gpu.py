"""Bragg shooting solve on the device against the oracle.

Bars: the march, the preconditioner and e^{iH} agree with the NumPy
transcriptions to 1e-12 relative; GMRES solutions (true residual <= tol
on both sides) agree with the oracle's to 1e-8 relative, with a dense
brute-force solve on a tiny grid, and with the discrete perfect-crystal
closed form; batching and chunking leave the answer unchanged.
"""

import numpy as np
import pytest

from dyfox_amd import _native as N
from dyfox_amd.bragg import (BraggConfig, bragg_g_vector, solve_bragg_stable,
                             solve_surface_batch, threading_segment)
from dyfox_amd.bragg import solver as BS
from dyfox_amd.units import deviation_from_urad
from oracle import bragg as OB, closed_forms as CF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def g(diamond400):
    return bragg_g_vector(diamond400)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max()
                 / np.abs(np.asarray(b)).max())


# Nx = 324 takes the two-wavefront march (each wavefront owns 162 = 54
# lanes x 3 columns); Nx = 334 = 2 x 167 has no column count <= 16 that
# divides it (or its half) into <= 64 lanes and takes the
# workgroup-per-plane march.
@pytest.mark.parametrize("half_x", [2e-6, 3e-6],
                         ids=["wavefront-Nx324", "workgroup-Nx334"])
def test_phase_march_and_preconditioner_match_oracle(dev, diamond400, g, rng,
                                                     half_x):
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=half_x)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = np.array([-0.3e-6, 0.1e-6, 0.5e-6])
    E = BS.chunk_phase(seg, cfg, diamond400, g, ys)
    Ch, Cb, _ = OB.couplings(seg, cfg, diamond400, g, ys)
    Eo = Cb / complex(diamond400.sig_hbar)
    assert _rel(E.get(), Eo) < 1e-12
    s = deviation_from_urad(np.array([-6.5, 12.0]), diamond400)
    v = rng.normal(size=(2, 3, cfg.Nx)) + 1j * rng.normal(size=(2, 3, cfg.Nx))
    inc = cfg.incident(n=1)[0]
    for j, sd in enumerate(s):
        e0, eg = OB.step_constants(sd, complex(diamond400.sig0), cfg.ds)
        want_h = OB.march_down(np.zeros((3, cfg.Nx), complex), v[j], Ch, Cb,
                               e0, eg, cfg.ds, cfg.Nz)
        want_p = OB.march_down(np.broadcast_to(inc, (3, cfg.Nx)).copy(),
                               np.zeros((3, cfg.Nx), complex), Ch, Cb, e0,
                               eg, cfg.ds, cfg.Nz)
        got_h = BS.march(cfg, diamond400, [sd], E, Dg0=v[j][None]).get()[0]
        got_p = BS.march(cfg, diamond400, [sd], E, incident=inc).get()[0]
        assert _rel(got_h, want_h) < 1e-12
        assert _rel(got_p, want_p) < 1e-12
        _, T22 = OB.perfect_symbol(e0, eg, cfg.ds, cfg.Nz, cfg.Nx,
                                   complex(diamond400.sig_h),
                                   complex(diamond400.sig_hbar))
        T22s = OB.floored_symbol(T22)
        want_w = np.fft.ifft(np.fft.fft(v[j], axis=-1) / T22s, axis=-1)
        got_w = BS.precondition(cfg, diamond400, [sd], v[j][None])[0]
        assert _rel(got_w, want_w) < 1e-12


# Row lengths of every kind the preconditioner's transform plans take:
# 13-smooth lengths transformed directly (324 = 2^2 3^4, 520 = 2^3 5 13,
# 2048), and lengths with a larger prime factor taken as a linear
# convolution at a 13-smooth length >= 2 Nx - 1 (124 = 4 x 31, 334 = 2 x
# 167, 1031 prime, 1948 = 4 x 487: the thickest cfg5 grid, 37 prime).
@pytest.mark.parametrize("Nx", [124, 324, 334, 520, 1031, 1948, 2048, 37])
def test_fused_preconditioner_matches_numpy_fft(dev, diamond400, rng, Nx):
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    cfg.Nx = Nx
    s = deviation_from_urad(np.array([-6.5, 0.0, 12.0]), diamond400)
    v = rng.normal(size=(3, 5, Nx)) + 1j * rng.normal(size=(3, 5, Nx))
    got = BS.precondition(cfg, diamond400, s, v)
    for j, sd in enumerate(s):
        e0, eg = OB.step_constants(sd, complex(diamond400.sig0), cfg.ds)
        _, T22 = OB.perfect_symbol(e0, eg, cfg.ds, cfg.Nz, Nx,
                                   complex(diamond400.sig_h),
                                   complex(diamond400.sig_hbar))
        want = np.fft.ifft(np.fft.fft(v[j], axis=-1) / OB.floored_symbol(T22),
                           axis=-1)
        assert _rel(got[j], want) < 1e-12, (Nx, j, _rel(got[j], want))


def test_one_and_two_wavefront_marches_agree(dev, diamond400, g, rng):
    # 41 angles x 100 planes = 4100 rows take the one-wavefront march (at
    # least 4096 rows); one angle (100 rows) takes the two-wavefront march,
    # whose owned columns do the same arithmetic
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = (np.arange(100) - 50) * 0.05e-6
    E = BS.chunk_phase(seg, cfg, diamond400, g, ys)
    s = deviation_from_urad(np.linspace(-20.0, 20.0, 41), diamond400)
    v = rng.normal(size=(41, 100, cfg.Nx)) + 1j * rng.normal(
        size=(41, 100, cfg.Nx))
    inc = cfg.incident(n=1)[0]
    all_h = BS.march(cfg, diamond400, s, E, Dg0=v).get()
    all_p = BS.march(cfg, diamond400, s, E, incident=inc).get()
    for j in (0, 17, 40):
        one_h = BS.march(cfg, diamond400, [s[j]], E, Dg0=v[j][None]).get()[0]
        one_p = BS.march(cfg, diamond400, [s[j]], E, incident=inc).get()[0]
        assert _rel(one_h, all_h[j]) < 1e-13
        assert _rel(one_p, all_p[j]) < 1e-13


def test_phase_table_integer_and_fractional_winding(dev, diamond400, g):
    """e^{iH} with g.b = +-2 (the winding taken algebraically, no atan2),
    with g.b = 0.74 (atan2), and with both in one list, against the
    oracle's exp(iH); the planes include y = 0, where a voxel lies on the
    threading line (atan2(0, 0) = 0, the core floor a^2)."""
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.2e-6, half_x=2e-6)
    b = diamond400.b_mag
    whole = (threading_segment("Edge", +1, b, float(g[2]))
             + threading_segment("Screw", -1, b, float(g[2]), x0=0.6e-6,
                                 y0=-0.2e-6, azimuth_deg=30.0))
    frac = threading_segment("Edge", +1, 0.37 * b, float(g[2]), x0=-0.4e-6,
                             y0=0.2e-6)
    ys = np.array([-0.2e-6, 0.0, 0.3e-6])
    for segs in (whole, frac, whole + frac):
        E = BS.chunk_phase(segs, cfg, diamond400, g, ys)
        _, Cb, _ = OB.couplings(segs, cfg, diamond400, g, ys)
        assert _rel(E.get(), Cb / complex(diamond400.sig_hbar)) < 1e-12


def test_surface_batch_matches_oracle(dev, diamond400, g):
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.1e-6, half_x=3e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = np.array([-0.2e-6, 0.0, 0.3e-6])
    got, info = solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    want, winfo = OB.solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    print(info, winfo)
    assert info["converged"] and info["residual"] < 1e-9
    # the default is the reference's joint chunk: the same integer info
    for key in ("iterations", "retries", "splits", "unconverged_chunks",
                "converged"):
        assert info[key] == winfo[key], key
    assert _rel(got, want) < 1e-8
    for j, y in enumerate(ys):
        one, oi = solve_bragg_stable(seg, sd, cfg, diamond400, g,
                                     y_plane=float(y))
        assert oi["residual"] < 1e-9 and oi["gmres_code"] == 0
        assert np.abs(got[j] - one).max() < 1e-7 * np.abs(one).max()


def test_fused_and_separate_arnoldi_steps_agree(dev, diamond400, g):
    # Nx 324: plane systems of 324 unknowns take the one-workgroup Arnoldi
    # step at restart 40 and the separate dot/combine/finish kernels at
    # restart 200 (h would not fit the workgroup's lanes); with fewer
    # iterations than either restart the two Krylov sequences are the same
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = np.array([-0.3e-6, 0.1e-6, 0.5e-6])
    E = BS.chunk_phase(seg, cfg, diamond400, g, ys)
    s = deviation_from_urad(np.array([-6.5, 12.0]), diamond400)
    inc = cfg.incident(n=1)[0]
    runs = [BS.gmres_batch(cfg, diamond400, s, E, inc, tol=1e-10, restart=r,
                           maxiter=20, planes=True) for r in (40, 200)]
    (x40, i40), (x200, i200) = runs
    # same sequence up to the reductions' rounding: counts within one,
    # solutions to the GMRES bar used against the oracle
    assert i40["iterations"].max() < 40
    assert np.abs(i40["iterations"] - i200["iterations"]).max() <= 1
    assert (i40["code"] == 0).all() and (i200["code"] == 0).all()
    assert _rel(x40.get(), x200.get()) < 1e-8


def test_joint_chunks_follow_scipy_and_agree_with_planes(dev, diamond400,
                                                         g):
    """planes="joint" is the reference's chunk (SciPy's iteration count);
    "independent" reaches the same surface with one Krylov space per
    plane."""
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            y0=-0.3e-6)
    ys = (np.arange(6) - 3) * cfg.dx
    joint, ij = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                    planes="joint")
    want, winfo = OB.solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    assert ij["iterations"] == winfo["iterations"]
    assert _rel(joint, want) < 1e-8
    ind, ii = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                  planes="independent")
    assert ii["converged"] and ii["residual"] <= 1e-10
    assert ii["iterations"] <= ij["iterations"]
    assert _rel(ind, want) < 1e-8
    # independent planes follow the reference's one-plane rule: at the
    # default restart 40 the Krylov size is already the cap, so no retry
    assert ii["retries"] == 0 and ii["splits"] == 0
    # below the cap (restart 2, cap min(restart_cap, 8) = 8) a plane
    # whose residual stays above `accept` is retried, never split
    few, fi = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                  planes="independent", restart=2,
                                  per_plane=2, maxiter=1, grow=2, tries=3,
                                  restart_cap=8)
    assert fi["retries"] > 0 and fi["splits"] == 0
    assert fi["converged"] == (fi["unconverged_chunks"] == 0)


def test_plane_systems_equal_single_plane_solves(dev, diamond400, g):
    """One system per (table, plane), read in place from per-instance
    tables, equals solving each plane alone on its own table."""
    cfg = BraggConfig(diamond400, t_crystal=5e-6, dx=0.2e-6, half_x=2e-6)
    ys = (np.arange(3) - 1) * cfg.dx
    segs = [threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                              y0=-0.2e-6),
            threading_segment("Screw", -1, diamond400.b_mag, float(g[2]),
                              x0=0.5e-6)]
    s = deviation_from_urad(np.array([-6.5, 9.0]), diamond400)
    inc = cfg.incident(n=1)[0]
    E = N.DeviceArray.empty((2, cfg.Nz, len(ys), cfg.Nx), np.complex128)
    for j in range(2):
        BS.chunk_phase(segs[j], cfg, diamond400, g, ys, out=E[j])
    v, info = BS.gmres_batch(cfg, diamond400, s, E, inc, tol=1e-11,
                             restart=40, maxiter=20, planes=True)
    assert len(info) == 2 * len(ys)
    vh = v.get()
    for j in range(2):
        for p, y in enumerate(ys):
            Ep = BS.chunk_phase(segs[j], cfg, diamond400, g, [y])
            one, oi = BS.gmres_batch(cfg, diamond400, [s[j]], Ep, inc,
                                     tol=1e-11, restart=40, maxiter=20)
            assert info[j * len(ys) + p][1] == oi[0][1]
            assert _rel(vh[j, p], one.get()[0, 0]) < 1e-12


def test_dense_brute_force_on_a_tiny_grid(dev, diamond400, g):
    tiny = BraggConfig(diamond400, 3e-6, 0.2e-6, 1e-6)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    sd = float(deviation_from_urad(-6.5, diamond400))
    dense, _, _ = OB.dense_shooting_solution(seg, sd, tiny, 0.1e-6,
                                             diamond400, g)
    got, info = solve_bragg_stable(seg, sd, tiny, diamond400, g, y_plane=0.1e-6)
    assert info["residual"] < 1e-10
    assert _rel(got, dense) < 1e-9


def test_perfect_crystal_is_the_discrete_closed_form(dev, diamond400, g):
    cfg = BraggConfig(diamond400, 20e-6, 0.5e-6, 32e-6)
    s = deviation_from_urad(np.linspace(-60, 60, 9), diamond400)
    out, infos = solve_surface_batch([], s, cfg, np.zeros(2), diamond400, g)
    s0, sh, sb = (complex(diamond400.sig0), complex(diamond400.sig_h),
                  complex(diamond400.sig_hbar))
    for j, sd in enumerate(s):
        want = CF.bragg_discrete_perfect(sd, s0, sh, sb, cfg.ds, cfg.Nz)
        assert abs(out[j, 0, cfg.Nx // 2] - want) < 1e-11
        assert infos[j]["converged"] and infos[j]["iterations"] <= 2


def test_batched_angles_equal_one_at_a_time(dev, diamond400, g):
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            y0=-0.5e-6)
    ys = (np.arange(4) - 2) * cfg.dx
    s = deviation_from_urad(np.array([-20.0, -6.5, 0.0, 15.0]), diamond400)
    out, infos = solve_surface_batch(seg, s, cfg, ys, diamond400, g)
    for j, sd in enumerate(s):
        one, info = solve_surface_batch(seg, float(sd), cfg, ys, diamond400, g)
        assert infos[j]["converged"] and info["converged"]
        assert _rel(out[j], one) < 1e-8


def test_big_chunk_equals_small_chunk(dev, diamond400, g):
    cfg = BraggConfig(diamond400, t_crystal=8e-6, dx=0.15e-6, half_x=3e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            y0=-1.5e-6)
    ys = (np.arange(16) - 8) * cfg.dx
    big, ib = solve_surface_batch(seg, sd, cfg, ys, diamond400, g, chunk=16)
    one, i1 = solve_surface_batch(seg, sd, cfg, ys, diamond400, g, chunk=1)
    assert ib["converged"] and i1["converged"]
    assert ib["residual"] < 1e-7 and i1["residual"] < 1e-7
    assert _rel(big, one) < 1e-6


def test_retry_and_split_policy_is_followed(dev, diamond400, g):
    """Force the policy from the host: an unattainable `accept` makes
    every attempt retry until the cap, then split down to max_depth."""
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = (np.arange(4) - 2) * cfg.dx
    out, info = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                    accept=1e-30, restart=4, restart_cap=16,
                                    per_plane=1, max_depth=1, tries=3,
                                    planes="joint")
    assert info["retries"] > 0 and info["splits"] == 1
    assert info["unconverged_chunks"] == 2 and info["converged"] is False
    want, winfo = OB.solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                         accept=1e-30, restart=4,
                                         restart_cap=16, per_plane=1,
                                         max_depth=1, tries=3)
    assert (info["retries"], info["splits"]) == (winfo["retries"],
                                                 winfo["splits"])
    assert _rel(out, want) < 1e-8


def test_sliced_arnoldi_step_follows_the_policy(dev, diamond400, g):
    """An 8-plane joint chunk (8 x 324 = 2592 unknowns, more than one
    workgroup holds: the Arnoldi step over three 1024-element slices),
    retried to the cap and then split into 4-plane halves (which take the
    one-workgroup step): the same retries, splits and surface as the
    oracle."""
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    assert 8 * cfg.Nx > 2048
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            y0=-0.3e-6)
    ys = (np.arange(8) - 4) * cfg.dx
    kw = dict(accept=1e-30, restart=4, restart_cap=16, per_plane=1,
              max_depth=1, tries=3)
    out, info = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                    planes="joint", **kw)
    want, winfo = OB.solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                         **kw)
    print(info, winfo)
    assert info["retries"] > 0 and info["splits"] == 1
    for k in ("retries", "splits", "unconverged_chunks"):
        assert info[k] == winfo[k], k
    assert _rel(out, want) < 1e-8
    # the same chunk to tol 1e-10 under the default policy: SciPy's
    # iteration count and the oracle's surface
    out, info = solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    want, winfo = OB.solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    assert info["converged"] and info["iterations"] == winfo["iterations"]
    assert _rel(out, want) < 1e-8


def test_warm_start_reaches_the_same_answer_in_fewer_iterations(dev,
                                                                diamond400, g):
    cfg = BraggConfig(diamond400, t_crystal=6e-6, dx=0.2e-6, half_x=2e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]))
    ys = (np.arange(4) - 2) * cfg.dx
    cold, ic = solve_surface_batch(seg, sd, cfg, ys, diamond400, g)
    E = BS.chunk_phase(seg, cfg, diamond400, g, ys)
    v_exact = cold * E.get()[0]                    # undo the exit phase
    noisy = v_exact * (1 + 1e-4)
    warm, iw = solve_surface_batch(seg, sd, cfg, ys, diamond400, g,
                                   warm_start=noisy)
    assert iw["converged"] and iw["iterations"] < ic["iterations"]
    assert _rel(warm, cold) < 1e-8


def test_20um_surface_with_edge_line_matches_oracle(dev, diamond400, g):
    """The full 20 um surface chunk (Nz 90, Nx 520, 32 planes) with one [101] edge
    line at mid-depth, two angles checked against the oracle."""
    from dyfox_amd.workloads import bragg_surface
    cfg, segs, ys, s, gv = bragg_surface(diamond400, "edge_line")
    pick = s[[3, 20]]
    out, infos = solve_surface_batch(segs, pick, cfg, ys, diamond400, gv)
    for j, sd in enumerate(pick):
        want, winfo = OB.solve_surface_batch(segs, float(sd), cfg, ys,
                                             diamond400, gv)
        print(infos[j], winfo)
        assert infos[j]["converged"] and infos[j]["residual"] <= 1e-10
        assert _rel(out[j], want) < 1e-8


# Nx = 432 = 54 lanes x 8 columns takes the wavefront fixed-point kernel,
# Nx = 452 = 4 x 113 the workgroup-per-plane one.
# Nx = 432 splits each row over 8 wavefronts (54 lanes x 1 column each);
# Nx = 452 = 4 x 113 has no column count <= 16 for either wavefront form
# and takes the workgroup kernel.
FP_GRIDS = pytest.mark.parametrize("half_x", [2e-6, 3e-6],
                                   ids=["wavefronts-Nx432", "workgroup-Nx452"])


def test_fixed_point_benchmark_grid_matches_oracle(dev, diamond400, g):
    # the benchmark's grid (t = 8 um, Nz 180, Nx 832): 16 wavefronts per
    # row, one column per lane
    cfg = BraggConfig(diamond400, t_crystal=8e-6, dx=0.1e-6, half_x=4e-6)
    assert (cfg.Nz, cfg.Nx) == (180, 832)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            x0=0.3e-6, y0=0.05e-6)
    s = deviation_from_urad(np.array([-20.0, 0.0, 35.0]), diamond400)
    ews, infos = BS.solve_bragg(seg, s, cfg, diamond400, g, n_iter=600,
                                tol=1e-13)
    for j in (0, 2):
        want, (its, res, conv) = OB.fixed_point(seg, float(s[j]), cfg,
                                                diamond400, g, n_iter=600,
                                                tol=1e-13)
        assert infos[j]["converged"] and conv
        assert abs(infos[j]["iterations"] - its) <= 1
        assert _rel(ews[j], want) < 1e-11


def test_fixed_point_one_and_several_wavefronts_per_row_agree(dev,
                                                              diamond400, g):
    # 1100 rows take one wavefront per row (Nx 432 = 54 lanes x 8
    # columns); one angle alone splits its row over 8 wavefronts
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.1e-6, half_x=2e-6)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            x0=0.3e-6, y0=0.05e-6)
    s = deviation_from_urad(np.linspace(-30.0, 30.0, 1100), diamond400)
    ews, infos = BS.solve_bragg(seg, s, cfg, diamond400, g, n_iter=800,
                                tol=1e-13)
    for j in (0, 377, 1099):
        one, oi = BS.solve_bragg(seg, float(s[j]), cfg, diamond400, g,
                                 n_iter=800, tol=1e-13)
        assert oi["converged"] and infos[j]["converged"]
        assert abs(oi["iterations"] - infos[j]["iterations"]) <= 1
        assert _rel(one, ews[j]) < 1e-12


@FP_GRIDS
def test_fixed_point_matches_oracle(dev, diamond400, g, half_x):
    """The round-trip fixed point with a threading line, on a small grid:
    the same sweep count as the NumPy restatement and the same exit wave;
    a batch of angles equals one angle at a time."""
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.1e-6, half_x=half_x)
    seg = threading_segment("Edge", +1, diamond400.b_mag, float(g[2]),
                            x0=0.3e-6, y0=0.05e-6)
    s = deviation_from_urad(np.array([-6.5, 2.0, 15.0]), diamond400)
    ews, infos = BS.solve_bragg(seg, s, cfg, diamond400, g, n_iter=800,
                                tol=1e-13)
    for j, sd in enumerate(s):
        want, (its, res, conv) = OB.fixed_point(seg, float(sd), cfg,
                                                diamond400, g, n_iter=800,
                                                tol=1e-13)
        assert infos[j]["converged"] and conv
        assert abs(infos[j]["iterations"] - its) <= 1
        assert _rel(ews[j], want) < 1e-11
        one, oi = BS.solve_bragg(seg, float(sd), cfg, diamond400, g,
                                 n_iter=800, tol=1e-13)
        assert np.array_equal(one, ews[j]) and oi == infos[j]


@FP_GRIDS
def test_fixed_point_perfect_and_unconverged(dev, diamond400, g, half_x):
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.1e-6, half_x=half_x)
    sd = float(deviation_from_urad(-6.5, diamond400))
    ew, info = BS.solve_bragg([], sd, cfg, diamond400, g, n_iter=800,
                              tol=1e-13)
    want = OB.fixed_point_discrete_perfect(sd, cfg, diamond400)
    assert info["converged"] and np.abs(ew - want).max() <= 1e-12 * abs(want)
    # the iteration cap is honoured and reported (tolerance forced away)
    _, capped = BS.solve_bragg([], sd, cfg, diamond400, g, n_iter=3, tol=0.0)
    assert capped == {"iterations": 3, "residual": capped["residual"],
                      "converged": False} and capped["residual"] > 0.0
    _, once = BS.solve_bragg([], sd, cfg, diamond400, g, n_iter=1, tol=0.0)
    assert np.isnan(once["residual"]) and once["iterations"] == 1


def test_shooting_agrees_with_the_fixed_point_where_both_work(dev,
                                                              diamond400, g):
    """The two independent discretisations agree to ~1e-3
    (test_bragg.py:135-152)."""
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.1e-6, half_x=3e-6)
    sd = float(deviation_from_urad(-6.5, diamond400))
    a, it = BS.solve_bragg([], sd, cfg, diamond400, g, n_iter=800, tol=1e-13)
    b, info = solve_bragg_stable([], sd, cfg, diamond400, g)
    assert it["converged"] and info["residual"] < 1e-8
    assert np.abs(a - b).max() < 2e-3 * np.abs(a).max()


@pytest.mark.parametrize("planes", ["joint", "independent"])
def test_surface_fields_equal_one_field_per_call(dev, diamond400, g, planes):
    """solve_surface_fields (one field per system, one batched GMRES) gives
    each system exactly what solve_surface_batch gives that field alone,
    integer info included, with a Krylov space small enough that some
    systems retry and split while others do not."""
    cfg = BraggConfig(diamond400, t_crystal=4e-6, dx=0.5e-6, half_x=4e-6)
    ys = (np.arange(4) - 2) * cfg.dx + cfg.dx / 2
    b = diamond400.b_mag
    fields = [threading_segment("Edge", +1, b, float(g[2]), y0=0.3e-6,
                                x0=-0.4e-6),
              [],
              threading_segment("Screw", -1, b, float(g[2]), y0=-0.2e-6,
                                x0=0.6e-6, azimuth_deg=30.0),
              threading_segment("Edge", -1, b, float(g[2]), x0=0.9e-6)]
    s = deviation_from_urad(np.array([-10.0, 4.0, 0.0, 15.0]), diamond400)
    kw = dict(restart=1, per_plane=0, maxiter=1, restart_cap=3, grow=3,
              tries=2, planes=planes)
    out, infos = BS.solve_surface_fields(fields, s, cfg, ys, diamond400, g,
                                         **kw)
    for j, segs in enumerate(fields):
        one, info = solve_surface_batch(segs, float(s[j]), cfg, ys,
                                        diamond400, g, **kw)
        for k in ("iterations", "retries", "splits", "unconverged_chunks",
                  "converged"):
            assert info[k] == infos[j][k], (j, k, info, infos[j])
        assert infos[j]["residual"] == pytest.approx(info["residual"],
                                                     rel=1e-6, abs=1e-18)
        assert _rel(out[j], one) < 1e-12
    if planes == "joint":
        assert any(i["splits"] for i in infos)
        assert any(i["retries"] for i in infos)
