"""Training data, training loop and evaluation of the learned warm start.

An instance is one Bragg surface solve: diamond (400) at 17 keV, dx =
0.5 um, the 128-pixel field of interest (half_x = 32 um, grid padded as
BraggConfig does), 64 y-planes (two 32-plane GMRES chunks), one s_dev,
0-12 dislocation lines drawn from the <110> family.  Instances whose
grids coincide (same Nz) are solved together: one batched GMRES per
chunk with one e^{iH} table per system.

The network predicts the dislocation part of the shooting unknown,
dv = v - v_perfect on the field of interest, where v_perfect = -T21/T22
of the uniform plane wave on a perfect crystal (closed form per s_dev
and thickness); the warm start is v_perfect everywhere plus the
prediction on the field.  Cold and warm solves run to the same true
residual (tol 1e-10).
"""

import ctypes as C
import time

import numpy as np

from .. import _native as N
from ..bragg.geometry import BraggConfig, bragg_g_vector
from ..bragg.solver import chunk_phase, exit_wave, gmres_batch
from ..dislocations import straight_segment
from ..units import deviation_from_urad
from ..workloads import family_110
from .net import H0, IN_CH, W0, WarmStartNet, to_bf16_bits

DX, HALF_X, NPLANES, CHUNK, NK = 0.5e-6, 32e-6, 64, 32, 16


def perfect_v(xtal, s_dev, cfg):
    """-T21/T22 of the uniform plane wave on a perfect crystal."""
    alpha = -2.0 * np.pi * float(s_dev)
    s0, sh, sb = complex(xtal.sig0), complex(xtal.sig_h), complex(xtal.sig_hbar)
    e0, eg = np.exp(s0 * cfg.ds), np.exp((s0 + 1j * alpha) * cfg.ds)
    M = np.array([[e0, e0 * cfg.ds * sb], [-cfg.ds * sh / eg, 1 / eg]])
    T = np.linalg.matrix_power(M, cfg.Nz - 1)
    return -T[1, 0] / T[1, 1]


def smooth_thicknesses(xtal, t_range=(20e-6, 40e-6), primes=(2, 3, 5, 7)):
    """Thicknesses in t_range whose padded grid width Nx factors into
    `primes`: the preconditioner's FFT length is Nx, and lengths with a
    large prime factor need Bluestein transforms and a plan build of
    their own.  Each value is also one batch group (same Nz)."""
    dz = DX * xtal.tan_tB
    out = []
    for nz in range(int(round(t_range[0] / dz)) + 1,
                    int(round(t_range[1] / dz)) + 2):
        nx = int(2 * (round(HALF_X / DX) + 2 * nz + 16))
        m = nx
        for p in primes:
            while m % p == 0:
                m //= p
        if m == 1:
            out.append((nz - 1) * dz)
    return np.array(out)


def thickness_grid(xtal, t_range=(20e-6, 100e-6), n=40):
    """n thicknesses evenly spread over t_range, each snapped to the Bragg
    grid's depth step (dz = dx tan(theta_B)); instances drawn from them
    share grids in groups, so each group is one batched GMRES."""
    dz = DX * xtal.tan_tB
    nz = np.unique(np.round(np.linspace(t_range[0], t_range[1], n) / dz))
    return nz * dz


def draw_instances(n, rng, xtal, t_range=(20e-6, 100e-6), max_seg=12,
                   urad=60.0, smooth=False, t_values=None):
    """n instances: thickness uniform over `t_values` (default: over the
    FFT-friendly values of t_range when `smooth`, else uniform in
    t_range), s_dev uniform in +-urad, 0..max_seg segments."""
    dirs = family_110()
    ts = smooth_thicknesses(xtal, t_range) if smooth else t_values
    out = []
    for _ in range(n):
        t = ts[rng.integers(len(ts))] if ts is not None else \
            rng.uniform(*t_range)
        cfg = BraggConfig(xtal, t_crystal=t, dx=DX, half_x=HALF_X)
        segs = []
        for _ in range(int(rng.integers(0, max_seg + 1))):
            xi = dirs[rng.integers(len(dirs))]
            b = (1.0 if rng.random() < 0.5 else -1.0) * xtal.b_mag * \
                dirs[rng.integers(len(dirs))]
            x = rng.uniform(-HALF_X, HALF_X)
            y = (np.floor(rng.uniform(-NPLANES / 2, NPLANES / 2)) + 0.5) * DX
            z = rng.uniform(0.0, cfg.t_crystal)
            segs += straight_segment(xi, b, [x, y, z], 1e-3)
        s = float(deviation_from_urad(rng.uniform(-urad, urad), xtal))
        out.append(dict(Nz=cfg.Nz, t=cfg.t_crystal, s=s, segs=segs))
    return out


def plane_ys():
    return (np.arange(NPLANES) - NPLANES // 2) * DX + DX / 2


class Dataset:
    """Device-resident inputs (N, 64, 128, 34) bf16 and targets
    (N, 64, 128, 2) fp32, plus per-instance cold-solve records."""

    def __init__(self, n):
        self.n = n
        self.X = N.DeviceArray.empty((n * H0 * W0 * IN_CH,), np.uint16)
        self.Y = np.zeros((n, H0, W0, 2), np.float32)
        self.iters = np.zeros(n, np.int64)
        self.resid = np.zeros(n)
        self.v_perf = np.zeros(n, np.complex128)
        self.scale = 1.0


def _groups(instances, max_sys):
    by = {}
    for i, inst in enumerate(instances):
        by.setdefault(inst["Nz"], []).append(i)
    for nz in sorted(by):
        idx = by[nz]
        for j in range(0, len(idx), max_sys):
            yield idx[j:j + max_sys]


def _norms(xtal, inst):
    """The network's two scalar inputs: s_dev in units of 60 urad, and the
    thickness mapped from [20, 100] um to [-1, 1]."""
    from ..units import urad_from_deviation
    s = float(urad_from_deviation(inst["s"], xtal)) / 60.0
    t = (inst["t"] - 60e-6) / 40e-6
    return s, t


def write_features(xtal, instances, ds, idx, max_sys=64):
    """Network inputs of instances idx into ds.X (their e^{iH} tables are
    built chunk by chunk; nothing is solved)."""
    g = bragg_g_vector(xtal)
    ys = plane_ys()
    for grp in _groups([instances[i] for i in idx], max_sys):
        members = [idx[k] for k in grp]
        cfg = BraggConfig(xtal, t_crystal=instances[members[0]]["t"], dx=DX,
                          half_x=HALF_X)
        x0c = (cfg.Nx - W0) // 2
        sn = np.array([_norms(xtal, instances[i])[0] for i in members],
                      np.float32)
        tn = np.array([_norms(xtal, instances[i])[1] for i in members],
                      np.float32)
        snd, tnd = N.DeviceArray.from_host(sn), N.DeviceArray.from_host(tn)
        for c0 in range(0, NPLANES, CHUNK):
            for j, i in enumerate(members):
                E = chunk_phase(instances[i]["segs"], cfg, xtal, g,
                                ys[c0:c0 + CHUNK])
                N.call("dyf_nn_bragg_features", C.c_void_p(E.ptr), 1,
                       cfg.Nz, CHUNK, cfg.Nx, x0c, W0, NK,
                       C.c_void_p(snd.ptr + 4 * j), C.c_void_p(tnd.ptr + 4 * j),
                       C.c_void_p(ds.X.ptr), H0, c0, int(i), None)
                del E


def solve_group(xtal, idx, instances, *, tol=1e-10, warm=None, ds_out=None,
                restart=None, maxiter=25, planes=True):
    """Solve the instances `idx` (same grid) chunk by chunk in batched
    GMRES; optionally write network features into `ds_out`.  `warm`:
    (len(idx), 64, Nx) initial guesses or None.  planes=True gives every
    y-plane its own Krylov space (restart 40), planes=False solves each
    32-plane chunk in one (restart 192, the reference's chunk sizing).
    Returns v (len, 64, Nx), per-instance plane-iterations (matvecs
    counted per plane: iterations x planes in the system, summed), the
    worst residual per instance and the grid."""
    cfg = BraggConfig(xtal, t_crystal=instances[idx[0]]["t"], dx=DX,
                      half_x=HALF_X)
    g = bragg_g_vector(xtal)
    ys = plane_ys()
    n = len(idx)
    s = np.array([instances[i]["s"] for i in idx])
    v_all = np.zeros((n, NPLANES, cfg.Nx), np.complex128)
    its = np.zeros(n, np.int64)
    res = np.zeros(n)
    inc = cfg.incident(n=1)[0]
    x0c = (cfg.Nx - W0) // 2
    r = restart or (40 if planes else max(40, min(256, 6 * CHUNK)))
    for c0 in range(0, NPLANES, CHUNK):
        E = N.DeviceArray.empty((n, cfg.Nz, CHUNK, cfg.Nx), np.complex128)
        for j, i in enumerate(idx):
            chunk_phase(instances[i]["segs"], cfg, xtal, g, ys[c0:c0 + CHUNK],
                        out=E[j])
        if ds_out is not None:
            sn = np.array([_norms(xtal, instances[i])[0] for i in idx],
                          np.float32)
            tn = np.array([_norms(xtal, instances[i])[1] for i in idx],
                          np.float32)
            snd, tnd = N.DeviceArray.from_host(sn), N.DeviceArray.from_host(tn)
            for j, i in enumerate(idx):
                N.call("dyf_nn_bragg_features", C.c_void_p(E[j].ptr), 1,
                       cfg.Nz, CHUNK, cfg.Nx, x0c, W0, NK,
                       C.c_void_p(snd.ptr + 4 * j), C.c_void_p(tnd.ptr + 4 * j),
                       C.c_void_p(ds_out.X.ptr), H0, c0, int(i), None)
        x0 = None
        if warm is not None:
            x0 = N.DeviceArray.from_host(
                np.ascontiguousarray(warm[:, c0:c0 + CHUNK]))
        v, info = gmres_batch(cfg, xtal, s, E, inc, tol=tol, restart=r,
                              maxiter=maxiter, x0=x0, planes=planes)
        v_all[:, c0:c0 + CHUNK] = v.get()
        per = CHUNK if planes else 1
        for j in range(n):
            for k in info[j * per:(j + 1) * per]:
                its[j] += k[1] * (1 if planes else CHUNK)
                res[j] = max(res[j], k[0])
        del E
    return v_all, its, res, cfg


def build_dataset(xtal, instances, max_sys=64, log=None, n_solve=None):
    """Cold-solve instances [0, n_solve) (default all): iterations,
    features, targets; the rest get their features only."""
    ds = Dataset(len(instances))
    n_solve = len(instances) if n_solve is None else n_solve
    if n_solve < len(instances):
        write_features(xtal, instances, ds,
                       list(range(n_solve, len(instances))), max_sys)
    x0c = None
    for idx in _groups(instances[:n_solve], max_sys):
        v, its, res, cfg = solve_group(xtal, idx, instances, ds_out=ds)
        x0c = (cfg.Nx - W0) // 2
        for j, i in enumerate(idx):
            vp = perfect_v(xtal, instances[i]["s"], cfg)
            ds.v_perf[i] = vp
            dv = v[j][:, x0c:x0c + W0] - vp
            ds.Y[i, :, :, 0] = dv.real
            ds.Y[i, :, :, 1] = dv.imag
            ds.iters[i], ds.resid[i] = its[j], res[j]
        if log:
            log(f"  solved Nz={cfg.Nz} Nx={cfg.Nx} x{len(idx)}: its "
                f"{its.min()}-{its.max()}, worst res {res.max():.1e}")
    ds.rms = np.sqrt(np.mean(ds.Y.astype(np.float64) ** 2, axis=(1, 2, 3)))
    ds.scale = (float(np.median(ds.rms[:n_solve])) or 1.0) if n_solve \
        else 1.0
    if log and n_solve:
        q = np.percentile(ds.rms[:n_solve] / ds.scale, [0, 10, 50, 90, 99, 100])
        top = np.argsort(ds.rms)[-5:][::-1]
        log("  target rms / median: min %.3g p10 %.3g p50 %.3g p90 %.3g "
            "p99 %.3g max %.3g" % tuple(q))
        for i in top:
            log(f"    largest: #{i} t {instances[i]['t'] * 1e6:.2f} um "
                f"s {instances[i]['s']:.3e} nseg {len(instances[i]['segs'])} "
                f"rms/median {ds.rms[i] / ds.scale:.3g} its {ds.iters[i]}")
    ds.Yd = N.DeviceArray.from_host(ds.Y / ds.scale)
    return ds


def _gather(ds, idx, batch):
    """Samples idx of the dataset as one contiguous device batch (inputs,
    targets)."""
    per = H0 * W0 * IN_CH
    tper = H0 * W0 * 2
    xd = N.DeviceArray.zeros((batch * per,), np.uint16)
    yd = N.DeviceArray.zeros((batch, H0, W0, 2), np.float32)
    for j, i in enumerate(idx):
        N.call("dyf_memcpy_d2d", C.c_void_p(xd.ptr + 2 * j * per),
               C.c_void_p(ds.X.ptr + 2 * int(i) * per), C.c_size_t(2 * per))
        N.call("dyf_memcpy_d2d", C.c_void_p(yd.ptr + 4 * j * tper),
               C.c_void_p(ds.Yd.ptr + 4 * int(i) * tper),
               C.c_size_t(4 * tper))
    return xd, yd


def train(ds, n_train, batch=32, epochs=20, lr=1e-3, seed=0, warmup=0.05,
          log=None, net=None, idx=None):
    """Adam on the first n_train samples, or on the samples `idx` (drawn
    into batches each epoch): linear warm-up over the first `warmup`
    fraction of the steps, then cosine decay to 2% of lr.  Returns the
    trained network and the per-epoch mean losses."""
    net = net or WarmStartNet(batch, seed=seed)
    pool = np.arange(n_train) if idx is None else np.asarray(idx)
    steps = len(pool) // batch
    rng = np.random.default_rng(seed)
    per = H0 * W0 * IN_CH
    losses = []
    total = epochs * steps
    n_warm = max(1, int(warmup * total))
    it = 0
    contiguous = idx is None
    for ep in range(epochs):
        order = rng.permutation(steps)
        perm = rng.permutation(pool)
        acc = 0.0
        for b in order:
            if contiguous:
                x = ds.X[b * batch * per:(b + 1) * batch * per]
                yd = ds.Yd[b * batch:(b + 1) * batch]
            else:
                x, yd = _gather(ds, perm[b * batch:(b + 1) * batch], batch)
            net.forward(x)
            loss = net.backward(yd)
            if it < n_warm:
                f = (it + 1) / n_warm
            else:
                frac = (it - n_warm) / max(total - n_warm - 1, 1)
                f = 0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * frac))
            net.step(lr * f)
            it += 1
            acc += float(loss.get()[0])
        losses.append(acc / max(steps, 1))
        if log:
            log(f"  epoch {ep}: loss {losses[-1]:.4e}")
    return net, losses


def predict(net, ds, idx):
    """Network output for samples idx (scaled back), (len, 64, 128) complex."""
    B = net.B
    per = H0 * W0 * IN_CH
    out = np.zeros((len(idx), H0, W0), np.complex128)
    idx = list(idx)
    for j0 in range(0, len(idx), B):
        chunk = idx[j0:j0 + B]
        # gather the samples into one contiguous batch
        xd = N.DeviceArray.zeros((B * per,), np.uint16)
        for j, i in enumerate(chunk):
            N.call("dyf_memcpy_d2d", C.c_void_p(xd.ptr + 2 * j * per),
                   C.c_void_p(ds.X.ptr + 2 * int(i) * per),
                   C.c_size_t(2 * per))
        o = net.forward(xd).get().reshape(B, H0, W0, 2)[:len(chunk)]
        out[j0:j0 + len(chunk)] = (o[..., 0] + 1j * o[..., 1]) * ds.scale
    return out


def _smooth_noise(rng, shape):
    """Unit-rms complex noise, box-smoothed over 5 pixels in y and x."""
    z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    k = np.ones(5) / 5
    for ax in (-2, -1):
        z = np.apply_along_axis(lambda a: np.convolve(a, k, "same"), ax, z)
    return z / np.sqrt(np.mean(np.abs(z) ** 2))


def evaluate(xtal, instances, ds, net, idx, max_sys=64, log=None,
             ceiling=(), seed=0):
    """Solve instances idx cold (x0 = 0), warm from the perfect-crystal
    value alone, and warm from the network (perfect value plus the
    predicted field); optionally also from the exact answer perturbed by
    smooth noise of relative size eps for eps in `ceiling` (what any
    predictor of that accuracy could gain).  Each group is first solved
    untimed at tol 1 (at most one iteration) so FFT plans and the
    workspace exist before any timed solve; then every variant runs twice,
    interleaved, and the faster run of each counts.  Returns
    per-variant solves/s, plane-iterations and residuals; the network
    variant's time includes inference."""
    sub = [instances[i] for i in idx]
    groups = list(_groups(sub, max_sys))
    N.synchronize()
    t0 = time.perf_counter()
    pred = predict(net, ds, idx)
    N.synchronize()
    infer_t = time.perf_counter() - t0
    names = ["cold", "perfect", "network"] + [f"exact_eps{e:g}"
                                             for e in ceiling]
    T = dict.fromkeys(names, 0.0)
    T["network"] = infer_t
    its = {k: np.zeros(len(idx), np.int64) for k in names}
    res = {k: np.zeros(len(idx)) for k in names}
    max_dv = 0.0
    rng = np.random.default_rng(seed)
    for g in groups:
        solve_group(xtal, g, sub, tol=1.0)    # plans + workspace, 0 its
        N.synchronize()
        t0 = time.perf_counter()
        v_c, its_c, res_c, cfg = solve_group(xtal, g, sub)
        N.synchronize()
        t_cold = time.perf_counter() - t0
        x0c = (cfg.Nx - W0) // 2
        vp = np.array([perfect_v(xtal, sub[k]["s"], cfg) for k in g])
        base = np.broadcast_to(vp[:, None, None],
                               (len(g), NPLANES, cfg.Nx)).copy()
        guesses = {"cold": None, "perfect": base}
        net_w = base.copy()
        net_w[:, :, x0c:x0c + W0] += pred[g]
        guesses["network"] = net_w
        dv_rms = np.sqrt(np.mean(np.abs(v_c - base) ** 2))
        for e in ceiling:
            guesses[f"exact_eps{e:g}"] = v_c + e * dv_rms * _smooth_noise(
                rng, v_c.shape)
        best = {"cold": t_cold}
        # every variant twice, interleaved; the faster of its runs counts
        for rep in range(2):
            for name, w in guesses.items():
                if rep == 0 and name == "cold":
                    its_w, res_w = its_c, res_c
                else:
                    t0 = time.perf_counter()
                    v_w, its_w, res_w, _ = solve_group(xtal, g, sub, warm=w)
                    N.synchronize()
                    dt = time.perf_counter() - t0
                    best[name] = min(best.get(name, np.inf), dt)
                    if name == "network":
                        max_dv = max(max_dv, float(np.abs(v_w - v_c).max()))
                for j, k in enumerate(g):
                    its[name][k], res[name][k] = its_w[j], res_w[j]
        for name in names:
            T[name] += best[name]
    per = {}
    for name in names:
        per[name] = {"solves_per_s": len(idx) / T[name],
                     "seconds": T[name],
                     "iterations_per_plane_mean":
                         float(its[name].mean() / NPLANES),
                     "worst_residual": float(res[name].max()),
                     "speedup_vs_cold": T["cold"] / T[name],
                     "iteration_ratio_vs_cold":
                         float(its["cold"].mean() / its[name].mean())}
    rep = {"n": len(idx), "variants": per,
           "cold_solves_per_s": per["cold"]["solves_per_s"],
           "warm_solves_per_s": per["network"]["solves_per_s"],
           "speedup": per["network"]["speedup_vs_cold"],
           "cold_iterations_mean": float(its["cold"].mean()),
           "warm_iterations_mean": float(its["network"].mean()),
           "cold_worst_residual": per["cold"]["worst_residual"],
           "warm_worst_residual": per["network"]["worst_residual"],
           "max_abs_diff_warm_vs_cold_v": max_dv,
           "inference_s": infer_t,
           "warm_time_includes_inference": True}
    if log:
        for name in names:
            log(f"  eval {name:16s} {per[name]}")
    return rep


BANDS_UM = ((20, 35), (35, 50), (50, 70), (70, 100))


def in_band(t_um, lo, hi):
    """Thickness band membership; thicknesses are snapped to the depth
    step, so the range's ends (20 and 100 um) land up to one step (0.23
    um) inside or outside them and are counted in the end bands."""
    lo_edge = lo - 0.5 if lo == BANDS_UM[0][0] else lo
    hi_edge = hi + 0.5 if hi == BANDS_UM[-1][1] else hi
    return (t_um >= lo_edge) & (t_um < hi_edge)


def rescale(ds, mask):
    """Normalise the targets by the median rms of the samples in `mask`
    (the converged ones: an unconverged solve is no target)."""
    ds.scale = float(np.median(ds.rms[mask])) or 1.0
    ds.Yd = N.DeviceArray.from_host(ds.Y / ds.scale)


def evaluate_joint(xtal, instances, ds, net, idx, max_sys=64, log=None,
                   tol=1e-10, planes="joint"):
    """The held-out instances idx solved with the reference's policy
    (solve_surface_fields: with planes="joint" its 32-plane chunks,
    retries and splits; "independent" one Krylov space per plane) cold,
    warm from the perfect-crystal value and warm from the network
    (perfect value plus the predicted field; inference time included).
    Instances of one grid are one batched solve per variant, timed
    after an untimed call that sets up the workspace.  Returns per
    variant and per thickness band: solves/s, mean GMRES iterations,
    retries and splits per solve, converged fraction, worst residual."""
    from ..bragg.solver import solve_surface_fields
    sub = [instances[i] for i in idx]
    N.synchronize()
    t0 = time.perf_counter()
    pred = predict(net, ds, idx)
    N.synchronize()
    infer_t = time.perf_counter() - t0
    g = bragg_g_vector(xtal)
    ys = plane_ys()
    names = ("cold", "perfect", "network")
    rec = {k: {"t": np.zeros(len(idx)), "its": np.zeros(len(idx)),
               "ret": np.zeros(len(idx)), "spl": np.zeros(len(idx)),
               "conv": np.zeros(len(idx), bool), "res": np.zeros(len(idx))}
           for k in names}
    for grp in _groups(sub, max_sys):
        cfg = BraggConfig(xtal, t_crystal=sub[grp[0]]["t"], dx=DX,
                          half_x=HALF_X)
        segs = [sub[k]["segs"] for k in grp]
        s = np.array([sub[k]["s"] for k in grp])
        x0c = (cfg.Nx - W0) // 2
        vp = np.array([perfect_v(xtal, sub[k]["s"], cfg) for k in grp])
        base = np.broadcast_to(vp[:, None, None],
                               (len(grp), NPLANES, cfg.Nx)).copy()
        net_w = base.copy()
        net_w[:, :, x0c:x0c + W0] += pred[grp]
        guesses = {"cold": None, "perfect": base, "network": net_w}
        solve_surface_fields(segs, s, cfg, ys, xtal, g, tol=1.0,
                             planes=planes)
        for name in names:
            N.synchronize()
            t1 = time.perf_counter()
            _, infos = solve_surface_fields(segs, s, cfg, ys, xtal, g,
                                            tol=tol, planes=planes,
                                            warm_start=guesses[name])
            dt = time.perf_counter() - t1
            if log:
                log(f"    {name}: Nz={cfg.Nz} x{len(grp)} {dt:.2f} s")
            if name == "network":
                dt += infer_t * len(grp) / len(idx)
            r = rec[name]
            for j, k in enumerate(grp):
                r["t"][k] = dt / len(grp)
                r["its"][k] = infos[j]["iterations"]
                r["ret"][k] = infos[j]["retries"]
                r["spl"][k] = infos[j]["splits"]
                r["conv"][k] = infos[j]["converged"]
                r["res"][k] = infos[j]["residual"]
        if log:
            log(f"  Nz={cfg.Nz} Nx={cfg.Nx} x{len(grp)}: " + ", ".join(
                f"{n} {rec[n]['t'][grp].sum():.2f} s its "
                f"{rec[n]['its'][grp].mean():.0f}" for n in names))
    t_um = np.array([x["t"] for x in sub]) * 1e6

    def summary(mask):
        out = {}
        for name in names:
            r = rec[name]
            n = int(mask.sum())
            out[name] = {"n": n,
                         "solves_per_s": n / max(r["t"][mask].sum(), 1e-12),
                         "iterations_mean": float(r["its"][mask].mean()),
                         "retries_mean": float(r["ret"][mask].mean()),
                         "splits_mean": float(r["spl"][mask].mean()),
                         "converged_fraction": float(r["conv"][mask].mean()),
                         "worst_residual": float(r["res"][mask].max())}
        for name in ("perfect", "network"):
            out[name]["speedup_vs_cold"] = (out[name]["solves_per_s"] /
                                            out["cold"]["solves_per_s"])
        return out

    policy = ("joint 32-plane chunks, retries and splits "
              "(bragg/solver.py:182-379)" if planes == "joint" else
              "one Krylov space per plane (restart 40)")
    rep = {"all": summary(np.ones(len(idx), bool)), "bands": {},
           "inference_s": infer_t, "policy": f"{policy}, tol {tol:g}"}
    for lo, hi in BANDS_UM:
        m = in_band(t_um, lo, hi)
        if m.any():
            rep["bands"][f"{lo}-{hi}um"] = summary(m)
    if log:
        for k, v in [("all", rep["all"])] + list(rep["bands"].items()):
            log(f"  eval {k}: " + "; ".join(
                f"{n} {v[n]['solves_per_s']:.2f}/s its "
                f"{v[n]['iterations_mean']:.0f} conv "
                f"{v[n]['converged_fraction']:.2f} retries "
                f"{v[n]['retries_mean']:.2f} splits {v[n]['splits_mean']:.2f}"
                for n in names))
    return rep
