"""Seeded synthetic workloads at the benchmark configurations' sizes.

    laue_worked_example(xtal)   t = 200 um, dx = 0.25 um, beam 120 um
                                (Nz 359, 544 x 800, 481 marched rows),
                                one [101] edge line through mid-depth,
                                s_dev in {0, 117.5 urad}
    laue_rocking_scan(xtal)     t = 273 um, dx = 0.25 um, beam 120 um,
                                256 rows (Nz 490, Nx 1056), 8 segments
                                from the <110> family, 64 s_dev in
                                [-200, +200] urad
    bragg_surface(xtal, crystal)  t = 20 um, dx = 0.5 um, half_x = 32 um
                                (Nz 90, Nx 520), 32 planes, 41 s_dev in
                                [-60, 60] urad; crystal "perfect", or
                                "edge_line": one [101] edge line at
                                mid-depth

All draws use numpy.random.default_rng(20260909) unless a generator is
passed.  Cores are kept half a pixel off the sampled rows.
"""

import numpy as np

from .dislocations import burgers_vector, straight_segment
from .units import deviation_from_urad

SEED = 20260909


def family_110():
    """The 12 unit <110> directions."""
    out = []
    for i in range(3):
        for j in range(i + 1, 3):
            for s in (1.0, -1.0):
                v = np.zeros(3)
                v[i], v[j] = 1.0, s
                out += [v / np.sqrt(2), -v / np.sqrt(2)]
    return np.array(out)


def random_segments(rng, xtal, grid, n):
    """n straight lines: direction and Burgers direction drawn from the
    crystal's <110> family (taken to the lab frame by xtal.U_lab, as
    dislocations.burgers_vector does, so g.b is an integer), random sign,
    |b| = a/sqrt(2), cores uniform in the illuminated Borrmann fan (z in
    (0, t), |x| < z tan tB, y inside the field of view), snapped half a
    pixel off the sampled rows; 1 mm long."""
    dirs = family_110()
    U = np.asarray(xtal.U_lab, float)
    if not np.array_equal(U, np.eye(3)):
        dirs = dirs @ U.T
    Ny = grid.geometry()[1]
    segs = []
    for _ in range(n):
        xi = dirs[rng.integers(len(dirs))]
        bh = dirs[rng.integers(len(dirs))]
        sign = 1.0 if rng.random() < 0.5 else -1.0
        z = rng.uniform(0, grid.t_crystal)
        xw = z * xtal.tan_tB
        x = rng.uniform(-xw, xw)
        yext = min(Ny * grid.dx / 2, grid.beam_width / 2)
        y = (np.floor(rng.uniform(-yext, yext) / grid.dx) + 0.5) * grid.dx
        segs += straight_segment(xi, sign * xtal.b_mag * bh, [x, y, z], 1e-3)
    return segs


def edge_101(xtal, grid, sign=+1, z0=None, length=800e-6):
    """[101] line, b = (a/2)[10-1] (pure edge, g.b = -2 for g along x),
    through the forward characteristic at depth z0 (default t/2), half a
    pixel off the row y = 0."""
    z0 = grid.t_crystal / 2 if z0 is None else z0
    xi = np.array([1.0, 0.0, 1.0]) / np.sqrt(2)
    b = burgers_vector(np.eye(3), [1, 0, -1], xtal.b_mag, sign)
    return straight_segment(xi, b, [-z0 * xtal.tan_tB, grid.dx / 2, z0],
                            length)


def laue_worked_example(xtal):
    from .laue.grid import LaueGrid
    grid = LaueGrid(xtal, 200e-6, 0.25e-6, 120e-6)
    s = deviation_from_urad(np.array([0.0, 117.5]), xtal)
    return grid, edge_101(xtal, grid), s


def laue_rocking_scan(xtal, n_s=64, n_seg=8, rng=None):
    from .laue.grid import LaueGrid
    rng = np.random.default_rng(SEED) if rng is None else rng
    grid = LaueGrid(xtal, 273e-6, 0.25e-6, 120e-6, Ny=256)
    segs = random_segments(rng, xtal, grid, n_seg)
    s = deviation_from_urad(np.linspace(-200, 200, n_s), xtal)
    return grid, segs, s


def ensemble_batch(n=256, rng=None, dx=0.25e-6):
    """The batched-ensembles configuration: n instances, half diamond
    (400) at 17 keV (U_lab = I), half (220) at 12 keV with
    U_lab = lab_frame((0,0,1), (2,2,0)) (g along x), f' = 0.003,
    f'' = 0.002 for both; t ~ U(50, 300) um, beam ~ U(60, 120) um,
    1-40 segments of the <110> family, 16-128 angles uniform in
    [-200, +200] urad.  Returns a list of (xtal, grid, segments, s_dev)."""
    from .crystal import DIAMOND, CrystalParams, lab_frame
    from .laue.grid import LaueGrid
    rng = np.random.default_rng(SEED) if rng is None else rng
    x400 = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    x220 = CrystalParams.from_reflection(
        DIAMOND, (2, 2, 0), 12.0, U_lab=lab_frame((0, 0, 1), (2, 2, 0)),
        f_prime=0.003, f_dprime=0.002)
    out = []
    for i in range(n):
        xtal = x400 if i < n // 2 else x220
        t = rng.uniform(50e-6, 300e-6)
        beam = rng.uniform(60e-6, 120e-6)
        grid = LaueGrid(xtal, t, dx, beam)
        segs = random_segments(rng, xtal, grid, int(rng.integers(1, 41)))
        n_s = int(rng.integers(16, 129))
        s = deviation_from_urad(np.linspace(-200, 200, n_s), xtal)
        out.append((xtal, grid, segs, s))
    return out


def bragg_surface(xtal, crystal="edge_line", n_planes=32, n_s=41):
    from .bragg.geometry import BraggConfig, bragg_g_vector
    cfg = BraggConfig(xtal, t_crystal=20e-6, dx=0.5e-6, half_x=32e-6)
    ys = (np.arange(n_planes) - n_planes // 2) * cfg.dx + cfg.dx / 2
    s = deviation_from_urad(np.linspace(-60, 60, n_s), xtal)
    if crystal not in ("perfect", "edge_line"):
        raise ValueError(f"crystal is 'perfect' or 'edge_line', not "
                         f"{crystal!r}")
    segs = []
    if crystal == "edge_line":
        xi = np.array([1.0, 0.0, 1.0]) / np.sqrt(2)
        b = xtal.b_mag * np.array([1.0, 0.0, -1.0]) / np.sqrt(2)
        core = np.array([0.0, cfg.dx / 2, cfg.t_crystal / 2])
        segs = straight_segment(xi, b, core, 400e-6)
    return cfg, segs, ys, s, bragg_g_vector(xtal)
