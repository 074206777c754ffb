"""Laue path on the device (libdyfox.so, gfx950) against the oracle.

Bars: FP64 exit waves within 1e-10 relative (rel-RMS per s_dev and max
abs / max |ref|) of the NumPy transcription of the reference march;
mixed precision within 1e-6; bitwise invariance where the reference
asserts it (batching, H chunking).
"""

import numpy as np
import pytest

from dyfox_amd.backend import FP32, FP64, MIXED, to_numpy
from dyfox_amd.dislocations import (analytic_line_params, compute_H_grid,
                                    g_dot_b)
from dyfox_amd.laue import LaueGrid, WaveCache, solve_dislocation, solve_pristine
from dyfox_amd.laue.solver import plan
from dyfox_amd.units import deviation_from_urad
from oracle import closed_forms as CF, laue as OL, volterra as OV
from tests.helpers import (geometry_n, grid_args, mixed_on_x, oracle_exit,
                           random_segments)

pytestmark = pytest.mark.gpu


def _check(got, want, tol, label=""):
    got = np.asarray(got, np.complex128)
    for j, r in enumerate(OL.error_report(got, want)):
        print(f"{label} s[{j}] rel_rms={r['rel_rms']:.2e} "
              f"max_abs={r['max_abs']:.2e}", flush=True)
        assert r["rel_rms"] <= tol and r["max_abs"] <= tol, (label, j, r)


def _s(xtal, urad):
    return deviation_from_urad(np.asarray(urad, float), xtal)


def test_fast_fp64_segments_match_oracle(dev, diamond400, small_grid):
    assert plan(small_grid, diamond400)[0] == 0       # fast path
    seg = geometry_n(diamond400, small_grid)
    s = _s(diamond400, [0.0, 117.5, -200.0, 60.0])
    got = to_numpy(solve_dislocation(seg, s, small_grid, diamond400))
    assert got.dtype == np.complex128
    _check(got, oracle_exit(seg, s, small_grid, diamond400), 1e-10, "fast64")


def test_many_segments_match_oracle(dev, diamond400, rng):
    grid = LaueGrid(diamond400, 60e-6, 0.25e-6, 10e-6)
    segs = random_segments(rng, diamond400, grid, 6)
    s = _s(diamond400, [-150.0, 10.0, 117.5])
    got = to_numpy(solve_dislocation(segs, s, grid, diamond400))
    _check(got, oracle_exit(segs, s, grid, diamond400), 1e-10, "multi")


def test_fractional_winding_table_matches_oracle(dev, diamond400, rng):
    """The fused e^{iH} table with a segment whose g.b is not an integer
    (0.74: its winding taken with atan2) beside <110> lines of integer
    g.b (taken algebraically), FP64 and mixed, against the oracle's
    exp(iH) march."""
    grid = LaueGrid(diamond400, 50e-6, 0.25e-6, 8e-6)
    (ra, rb, b), = geometry_n(diamond400, grid, y0=grid.dx / 2)
    segs = [(ra, rb, 0.37 * b)] + random_segments(rng, diamond400, grid, 2)
    gb = [g_dot_b(diamond400.g_vec, s[2]) for s in segs]
    assert abs(abs(gb[0]) - 0.74) < 1e-9
    assert all(abs(v - round(v)) < 1e-9 for v in gb[1:])
    s = _s(diamond400, [-120.0, 0.0, 75.0])
    want = oracle_exit(segs, s, grid, diamond400)
    got = to_numpy(solve_dislocation(segs, s, grid, diamond400))
    _check(got, want, 1e-10, "frac64")
    got = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                     precision=MIXED))
    _check(got, want, 1e-6, "fracmixed")


def test_h_grid_matches_oracle(dev, diamond400, small_grid):
    seg = geometry_n(diamond400, small_grid, y0=small_grid.dx / 2)
    Nx, Ny, dz, _, _ = small_grid.geometry()
    H = to_numpy(compute_H_grid(seg, small_grid, diamond400))
    Ho = OV.h_grid(seg, diamond400.g_vec, diamond400.nu, diamond400.a_core,
                   Nx, Ny, small_grid.Nz, small_grid.dx, dz)
    assert H.shape == (Ny, small_grid.Nz, Nx) and H.dtype == np.float64
    # away from the atan2 cut the values agree; across it they may differ
    # by 2 pi g.b, which leaves e^{iH} unchanged
    gb = abs(g_dot_b(diamond400.g_vec, seg[0][2]))
    d = H - Ho
    d = d - 2 * np.pi * gb * np.round(d / (2 * np.pi * gb))
    assert np.abs(d).max() < 1e-9
    assert np.abs(np.exp(1j * H) - np.exp(1j * Ho)).max() < 1e-10


def test_h_grid_fp32_accumulates_in_float(dev, diamond400, small_grid, rng):
    segs = random_segments(rng, diamond400, small_grid, 3)
    Nx, Ny, dz, _, _ = small_grid.geometry()
    H = to_numpy(compute_H_grid(segs, small_grid, diamond400, precision=FP32))
    Ho = OV.h_grid(segs, diamond400.g_vec, diamond400.nu, diamond400.a_core,
                   Nx, Ny, small_grid.Nz, small_grid.dx, dz,
                   dtype=np.float32)
    assert H.dtype == np.float32
    ok = np.abs(H.astype(float) - Ho.astype(float)) < 1e-4 * (1 + np.abs(Ho))
    assert ok.mean() > 0.999


def test_h_gpu_path_matches_oracle(dev, diamond400, small_grid):
    seg = geometry_n(diamond400, small_grid, sign=-1)
    s = _s(diamond400, [117.5, -30.0])
    H = compute_H_grid(seg, small_grid, diamond400)
    got = to_numpy(solve_dislocation(seg, s, small_grid, diamond400, H_gpu=H))
    _check(got, oracle_exit(seg, s, small_grid, diamond400), 1e-10, "h_gpu")
    host = to_numpy(solve_dislocation([], s, small_grid, diamond400,
                                      H_gpu=to_numpy(H)))
    assert np.array_equal(host, got)


def test_chunk_size_changes_nothing(dev, diamond400, small_grid):
    seg = geometry_n(diamond400, small_grid)
    ref = to_numpy(compute_H_grid(seg, small_grid, diamond400, chunk_y=32))
    for ck in (1, 7, 16, 512):
        got = to_numpy(compute_H_grid(seg, small_grid, diamond400, chunk_y=ck))
        assert got.tobytes() == ref.tobytes()


def test_zero_burgers_vector_leaves_no_phase(dev, diamond400, small_grid):
    seg = geometry_n(diamond400, small_grid)
    dead = [(seg[0][0], seg[0][1], np.zeros(3))]
    H = to_numpy(compute_H_grid(dead, small_grid, diamond400))
    assert float(np.abs(H).max()) == 0.0
    s = _s(diamond400, [117.5])
    a = to_numpy(solve_dislocation(dead, s, small_grid, diamond400))
    b = to_numpy(solve_pristine(s, small_grid, diamond400))
    assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()


def test_pristine_is_the_discrete_pendellosung(dev, diamond400):
    grid = LaueGrid(diamond400, 30e-6, 0.25e-6, 2e-6, beam_thickness=200e-6)
    g = grid_args(grid, diamond400)
    s = _s(diamond400, [-40.0, 0.0, 25.0, 117.5])
    for prec, tol in ((FP64, 1e-11), (MIXED, 1e-6), (FP32, 1e-4)):
        got = to_numpy(solve_pristine(s, grid, diamond400, precision=prec))
        for j, sd in enumerate(s):
            want = CF.pendellosung_discrete(sd, g["sig0"], g["sig_h"],
                                            g["sig_hbar"], g["ds"], grid.Nz)
            v = got[j, g["Ny"] // 2, g["Nx"] // 2]
            assert abs(v - want) <= tol * abs(want), (prec, j, v, want)
    got = to_numpy(solve_pristine(s, grid, diamond400))
    _check(got, OL.march_grid(g, s), 1e-10, "pristine")


def test_analytic_line_matches_oracle_and_grid(dev, diamond400, small_grid):
    seg = mixed_on_x(diamond400, small_grid)
    lp = analytic_line_params(seg, diamond400)
    assert lp is not None
    s = _s(diamond400, [117.5, 0.0])
    closed = to_numpy(solve_dislocation(seg, s, small_grid, diamond400))
    g = grid_args(small_grid, diamond400)
    want = OL.march_grid(g, s, H=OL.analytic_h(lp, g["Ny"], g["dx"], g["dz"]))
    _check(closed, want, 1e-10, "analytic")
    H = compute_H_grid(seg, small_grid, diamond400)
    on_grid = to_numpy(solve_dislocation(seg, s, small_grid, diamond400,
                                         H_gpu=H))
    assert OL.rel_rms(closed, on_grid) < 1e-10


def test_core_on_a_sampled_row_changes_at_most_that_row(dev, diamond400,
                                                        small_grid):
    from dyfox_amd.dislocations import burgers_vector, straight_segment
    b = burgers_vector(np.eye(3), [1, 1, 0], diamond400.b_mag, +1)
    on_row = straight_segment([1.0, 0.0, 0.0], b,
                              [0.0, 0.0, small_grid.t_crystal / 2], 400e-6)
    s = _s(diamond400, [117.5])
    a = to_numpy(solve_dislocation(on_row, s, small_grid, diamond400))[0]
    H = compute_H_grid(on_row, small_grid, diamond400)
    c = to_numpy(solve_dislocation(on_row, s, small_grid, diamond400,
                                   H_gpu=H))[0]
    rows = np.nonzero(np.abs(a - c).max(axis=1) > 1e-12)[0]
    assert len(rows) <= 1
    assert all(r == a.shape[0] // 2 for r in rows)


def test_batching_is_bitwise_invariant(dev, diamond400, small_grid):
    seg = geometry_n(diamond400, small_grid)
    s = np.linspace(-3e5, 3e5, 9)
    H = compute_H_grid(seg, small_grid, diamond400)
    for prec in (FP64, MIXED):
        a = to_numpy(solve_dislocation(seg, s, small_grid, diamond400,
                                       H_gpu=H, batch=64, precision=prec))
        b = to_numpy(solve_dislocation(seg, s, small_grid, diamond400,
                                       H_gpu=H, batch=2, precision=prec))
        assert a.tobytes() == b.tobytes()


def test_interpolating_grid_takes_the_general_path(dev, diamond400):
    grid = LaueGrid(diamond400, t_crystal=20e-6, dx=0.25e-6,
                    beam_width=40e-6, beam_thickness=1e-6, Nz=101)
    assert plan(grid, diamond400)[0] == 1
    Nx, Ny = grid.geometry()[:2]
    q = 2 * np.pi / 10e-6
    y = (np.arange(Ny) - Ny / 2) * grid.dx
    H = np.ascontiguousarray(np.broadcast_to((q * y)[:, None, None],
                                             (Ny, grid.Nz, Nx)))
    st = to_numpy(solve_dislocation([], [0.0], grid, diamond400, H_gpu=H))
    rows = slice(Ny // 2 - 60, Ny // 2 + 60)
    phi = np.unwrap(np.angle(st[0][rows, Nx // 2]))
    slope = np.polyfit(np.arange(len(phi)) * grid.dx, phi, 1)[0]
    assert slope == pytest.approx(-q, rel=0.05)
    g = grid_args(grid, diamond400)
    _check(st, OL.march_grid(g, [0.0], H=H), 1e-10, "general-shear")
    seg = geometry_n(diamond400, grid, y0=grid.dx / 2)
    s = _s(diamond400, [117.5, -20.0])
    got = to_numpy(solve_dislocation(seg, s, grid, diamond400))
    _check(got, oracle_exit(seg, s, grid, diamond400), 1e-10, "general-seg")
    pr = to_numpy(solve_pristine(s, grid, diamond400))
    _check(pr, OL.march_grid(g, s), 1e-10, "general-pristine")


def test_pristine_matches_the_plane_wave_rocking_curve(dev, diamond400):
    grid = LaueGrid(diamond400, t_crystal=45e-6 / diamond400.tan_tB,
                    dx=0.25e-6, beam_width=20e-6, beam_thickness=120e-6,
                    Nz=801)
    Ny, Nx = grid.shape
    s = np.linspace(-6, 6, 25) / diamond400.xi_g
    got = np.abs(to_numpy(solve_pristine(s, grid, diamond400))
                 [:, Ny // 2, Nx // 2]) ** 2
    g = grid_args(grid, diamond400)
    want = CF.planewave_rocking(s, g["sig0"], g["sig_h"], g["sig_hbar"],
                                diamond400.cos_tB, grid.t_crystal)
    resid = got / got.max() - want / want.max()
    assert float(np.sqrt(np.mean(resid ** 2))) < 0.02


def test_mixed_precision_meets_1e6_on_a_deep_grid(dev, diamond400, rng):
    """t = 273 um (Nz 490, Nx 1056), 8 segments, rows subsampled for the
    oracle: the mixed mode's per-s_dev error against FP64."""
    grid = LaueGrid(diamond400, 273e-6, 0.25e-6, 4e-6, Ny=32)
    segs = random_segments(rng, diamond400, grid, 8)
    s = _s(diamond400, [-200.0, -60.0, 0.0, 117.5, 200.0])
    want = oracle_exit(segs, s, grid, diamond400)
    f64 = to_numpy(solve_dislocation(segs, s, grid, diamond400))
    _check(f64, want, 1e-10, "deep-fp64")
    mixed = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                       precision=MIXED))
    assert mixed.dtype == np.complex64
    _check(mixed, want, 1e-6, "deep-mixed")
    f32 = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                     precision=FP32))
    rel = OL.rel_rms(f32, want)
    print(f"deep-fp32 rel_rms={rel:.2e}")
    assert 0 < rel < 0.05


def test_odd_column_counts(dev, diamond400):
    """Nx whose column count per lane is padded (and a wide grid)."""
    for t, beam in ((20e-6, 3e-6), (120e-6, 3e-6), (300e-6, 2e-6)):
        grid = LaueGrid(diamond400, t, 0.25e-6, beam, grid_pad=7)
        seg = geometry_n(diamond400, grid, y0=grid.dx / 2)
        s = _s(diamond400, [117.5])
        got = to_numpy(solve_dislocation(seg, s, grid, diamond400))
        print(grid, plan(grid, diamond400))
        _check(got, oracle_exit(seg, s, grid, diamond400), 1e-10, str(grid))


def test_full_size_worked_example_rows_match_oracle(dev, diamond400):
    """The worked-example grid (Nz 359, 544 x 800, 481 marched rows,
    s_dev in {0, 117.5 urad}): every row is independent, so the oracle
    checks a spread of rows at full size."""
    grid = LaueGrid(diamond400, 200e-6, 0.25e-6, 120e-6)
    z0 = grid.t_crystal / 2
    from dyfox_amd.dislocations import burgers_vector, straight_segment
    seg = straight_segment([1 / np.sqrt(2), 0.0, 1 / np.sqrt(2)],
                           burgers_vector(np.eye(3), [1, 0, -1],
                                          diamond400.b_mag, +1),
                           [-z0 * diamond400.tan_tB, grid.dx / 2, z0], 800e-6)
    s = _s(diamond400, [0.0, 117.5])
    got = to_numpy(solve_dislocation(seg, s, grid, diamond400))
    g = grid_args(grid, diamond400)
    rows = np.array([32, 150, 271, 272, 273, 400, 511])
    H = OV.h_grid(seg, diamond400.g_vec, diamond400.nu, diamond400.a_core,
                  g["Nx"], g["Ny"], grid.Nz, g["dx"], g["dz"], rows=rows)
    want = OL.march_grid(g, s, H=H, rows=rows)
    _check(got[:, rows], want[:, rows], 1e-10, "worked-example-rows")
    outside = np.setdiff1d(np.arange(g["Ny"]), grid.marched_rows())
    assert np.abs(got[:, outside]).max() == 0.0
    mixed = to_numpy(solve_dislocation(seg, s, grid, diamond400,
                                       precision=MIXED))
    _check(mixed, got, 1e-6, "worked-example-mixed")


def test_wave_cache_round_trip(dev, diamond400, small_grid, tmp_path):
    cache = WaveCache(str(tmp_path / "waves"))
    seg = geometry_n(diamond400, small_grid)
    first = cache.solve(seg, [0.0], small_grid, diamond400, tag="edge")
    second = cache.solve(seg, [0.0], small_grid, diamond400, tag="edge")
    assert first.dtype == np.complex128 and second.dtype == np.complex64
    assert np.allclose(to_numpy(second), to_numpy(first).astype(np.complex64))


@pytest.mark.parametrize("line", ["inclined", "along_x"])
def test_odd_row_count_agrees_across_paths(dev, diamond400, line):
    """Ny = 33 (only the Ny= override gives an odd count): H is placed at
    y = (iy - Ny/2) dx with a real Ny/2 as compute_H_grid places it
    (dislocations.py:146), the field of view uses the march's integer
    Ny/2 (laue/kernels.py:63).  The fused e^{iH} table, the H_gpu= path
    and the oracle agree; a line along x (which would take the closed
    form on an even grid) takes the Volterra table here."""
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6, Ny=33)
    assert grid.shape[0] == 33
    seg = (geometry_n(diamond400, grid, y0=grid.dx / 4) if line == "inclined"
           else mixed_on_x(diamond400, grid))
    s = _s(diamond400, [0.0, 117.5, -40.0])
    fused = to_numpy(solve_dislocation(seg, s, grid, diamond400))
    H = compute_H_grid(seg, grid, diamond400)
    via_h = to_numpy(solve_dislocation(seg, s, grid, diamond400, H_gpu=H))
    want = oracle_exit(seg, s, grid, diamond400)
    _check(fused, want, 1e-10, "odd-Ny fused")
    _check(via_h, want, 1e-10, "odd-Ny H_gpu")
    # the same line on the even grid one row larger: rows move by half a
    # pixel, so the answers differ (the convention is not a no-op)
    even = LaueGrid(diamond400, 40e-6, 0.25e-6, 8e-6, Ny=34)
    other = to_numpy(solve_dislocation(seg, s, even, diamond400))
    assert np.abs(other[:, 1:33] - fused[:, 0:32]).max() > 1e-6


@pytest.mark.parametrize("beam", ["sheet", "gauss_narrow", "gauss_wide",
                                  "sheet_wide"])
def test_beam_fan_agrees_across_paths(dev, diamond400, rng, beam):
    """The fused e^{iH} table is evaluated only inside the Borrmann fan of
    the incident beam (zeros outside, where the march multiplies exact
    zeros), and the marches read no slice before a lane's first step in
    the fan.  A narrow top-hat sheet, a Gaussian whose exactly-zero tails
    leave part of the row dark, a Gaussian wider than the row and a wide
    sheet: the fused solve, the H_gpu= path (a full table) and the oracle
    agree in FP64, and mixed precision stays within 1e-6."""
    kw = {"sheet": dict(beam_thickness=1e-6),
          "gauss_narrow": dict(beam_sigma=0.5e-6),
          "gauss_wide": dict(beam_sigma=20e-6),
          "sheet_wide": dict(beam_thickness=30e-6)}[beam]
    grid = LaueGrid(diamond400, 40e-6, 0.25e-6, 4e-6, **kw)
    segs = random_segments(rng, diamond400, grid, 3)
    s = _s(diamond400, [-90.0, 0.0, 117.5])
    want = oracle_exit(segs, s, grid, diamond400)
    fused = to_numpy(solve_dislocation(segs, s, grid, diamond400))
    H = compute_H_grid(segs, grid, diamond400)
    via_h = to_numpy(solve_dislocation(segs, s, grid, diamond400, H_gpu=H))
    _check(fused, want, 1e-10, f"{beam} fused")
    _check(via_h, want, 1e-10, f"{beam} H_gpu")
    mixed = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                       precision=MIXED))
    _check(mixed, want, 1e-6, f"{beam} mixed")
    if beam == "sheet":
        # lit columns Nx/2 - 1 .. Nx/2 + 1 spread one column per step: the
        # exit wave is exactly zero outside, on the device as in the oracle
        Nx = grid.shape[1]
        lo, hi = Nx // 2 - 1 - (grid.Nz - 1), Nx // 2 + 1 + (grid.Nz - 1)
        assert lo > 0 and hi < Nx - 1
        assert not np.any(fused[..., :lo]) and not np.any(fused[..., hi + 1:])
        assert not np.any(want[..., :lo]) and not np.any(want[..., hi + 1:])


@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("n_e", [128, 127, 38, 36])
def test_pair_march_equals_single_angle_march(dev, diamond400, rng, prec,
                                              n_e):
    """A launch with at least 6 (row, angle pair) items per CU marches two
    angles per wavefront while the beam's fan fits half a wavefront (81
    rows x 64 pairs here; 81 x 19 pairs is just above the rule on a
    256-CU device, 81 x 18 just below it); a one-angle launch never pairs.
    Every angle of the launch, the unpaired last one of an odd count
    included, equals its own one-angle solve exactly (np.array_equal: up to
    the sign of zeros outside the fan), and the one-angle march is the one
    checked against the oracle above."""
    grid = LaueGrid(diamond400, 80e-6, 0.25e-6, 20e-6)
    segs = random_segments(rng, diamond400, grid, 3)
    s = _s(diamond400, np.linspace(-200.0, 200.0, n_e))
    full = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                      precision=prec))
    for j in (0, 1, n_e // 2, n_e - 2, n_e - 1):
        one = to_numpy(solve_dislocation(segs, s[j:j + 1], grid, diamond400,
                                         precision=prec))
        assert np.array_equal(full[j], one[0]), (prec.name, n_e, j)
    if prec is FP64:
        want = oracle_exit(segs, s[[0, n_e - 1]], grid, diamond400)
        _check(full[[0, n_e - 1]], want, 1e-10, f"pair{n_e}")



@pytest.mark.parametrize("prec", [FP64, MIXED])
@pytest.mark.parametrize("n_e", [6, 40])
def test_unwritten_table_entries_are_never_read(dev, diamond400, rng,
                                                monkeypatch, prec, n_e):
    """The e^{iH} table leaves the entries before each lane's first step
    in the beam's fan unwritten.  With the table's buffer filled with NaN
    bytes first, the one-angle (6 angles) and the paired (40 angles: 81 x
    20 pairs) marches give exactly the solve of a zero-filled buffer:
    no march reads an unwritten entry."""
    from dyfox_amd.laue import solver as LS
    grid = LaueGrid(diamond400, 80e-6, 0.25e-6, 20e-6)
    segs = random_segments(rng, diamond400, grid, 2)
    s = _s(diamond400, np.linspace(-120.0, 120.0, n_e))
    orig = LS._table

    def filled(byte):
        def table(nbytes, reuse):
            buf = orig(nbytes, False)
            buf.copy_from(np.full(buf.nbytes, byte, np.uint8))
            return buf
        return table

    out = {}
    for byte in (0x00, 0xFF):              # 0xFF...: NaN in FP64 and FP32
        monkeypatch.setattr(LS, "_table", filled(byte))
        out[byte] = to_numpy(solve_dislocation(segs, s, grid, diamond400,
                                               precision=prec))
    monkeypatch.setattr(LS, "_table", orig)
    assert np.isfinite(out[0xFF]).all()
    assert np.array_equal(out[0xFF], out[0x00])
