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


def draw_instances(n, rng, xtal, t_range=(20e-6, 40e-6), max_seg=12,
                   urad=60.0, smooth=True):
    """n instances: thickness (uniform over the FFT-friendly values of
    t_range when `smooth`, else uniform in t_range), s_dev (uniform in
    +-urad), 0..max_seg segments."""
    dirs = family_110()
    ts = smooth_thicknesses(xtal, t_range) if smooth else None
    out = []
    for _ in range(n):
        t = ts[rng.integers(len(ts))] if smooth else rng.uniform(*t_range)
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
    from ..units import urad_from_deviation
    s = float(urad_from_deviation(inst["s"], xtal)) / 60.0
    t = (inst["t"] - 30e-6) / 10e-6
    return s, t


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


def build_dataset(xtal, instances, max_sys=64, log=None):
    """Cold-solve every instance, record iterations, features, targets."""
    ds = Dataset(len(instances))
    x0c = None
    for idx in _groups(instances, max_sys):
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
    ds.scale = float(np.sqrt(np.mean(ds.Y ** 2))) or 1.0
    ds.Yd = N.DeviceArray.from_host(ds.Y / ds.scale)
    return ds


def train(ds, n_train, batch=32, epochs=20, lr=2e-3, seed=0, log=None):
    """Adam on the first n_train samples; cosine learning-rate decay.
    Returns the trained network and the per-epoch mean losses."""
    net = WarmStartNet(batch, seed=seed)
    steps = n_train // batch
    rng = np.random.default_rng(seed)
    per = H0 * W0 * IN_CH
    per_y = H0 * W0 * 2
    losses = []
    total = epochs * steps
    it = 0
    for ep in range(epochs):
        order = rng.permutation(steps)
        acc = 0.0
        for b in order:
            x = ds.X[b * batch * per:(b + 1) * batch * per]
            yd = ds.Yd[b * batch:(b + 1) * batch]
            net.forward(x)
            loss = net.backward(yd)
            frac = it / max(total - 1, 1)
            net.step(lr * (0.02 + 0.98 * 0.5 * (1 + np.cos(np.pi * frac))))
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
                   C.c_void_p(ds.X.ptr + 2 * i * per), C.c_size_t(2 * per))
        o = net.forward(xd).get().reshape(B, H0, W0, 2)[:len(chunk)]
        out[j0:j0 + len(chunk)] = (o[..., 0] + 1j * o[..., 1]) * ds.scale
    return out


def evaluate(xtal, instances, ds, net, idx, max_sys=64, log=None):
    """Cold and warm solves of instances idx; iterations and solves/s."""
    sub = [instances[i] for i in idx]
    groups = list(_groups(sub, max_sys))
    t0 = time.perf_counter()
    pred = predict(net, ds, idx)
    infer_t = time.perf_counter() - t0
    cold_t, warm_t = 0.0, infer_t
    cold_its = np.zeros(len(idx), np.int64)
    warm_its = np.zeros(len(idx), np.int64)
    cold_res = np.zeros(len(idx))
    warm_res = np.zeros(len(idx))
    max_dv = 0.0
    for g in groups:
        t0 = time.perf_counter()
        v_c, its_c, res_c, cfg = solve_group(xtal, g, sub)
        N.synchronize()
        cold_t += time.perf_counter() - t0
        x0c = (cfg.Nx - W0) // 2
        t0 = time.perf_counter()
        warm = np.empty((len(g), NPLANES, cfg.Nx), np.complex128)
        for j, k in enumerate(g):
            vp = perfect_v(xtal, sub[k]["s"], cfg)
            warm[j] = vp
            warm[j][:, x0c:x0c + W0] += pred[k]
        v_w, its_w, res_w, _ = solve_group(xtal, g, sub, warm=warm)
        N.synchronize()
        warm_t += time.perf_counter() - t0
        for j, k in enumerate(g):
            cold_its[k], warm_its[k] = its_c[j], its_w[j]
            cold_res[k], warm_res[k] = res_c[j], res_w[j]
        max_dv = max(max_dv, float(np.abs(v_w - v_c).max()))
    rep = {"n": len(idx), "cold_solves_per_s": len(idx) / cold_t,
           "warm_solves_per_s": len(idx) / warm_t,
           "speedup": cold_t / warm_t,
           "cold_iterations_mean": float(cold_its.mean()),
           "warm_iterations_mean": float(warm_its.mean()),
           "cold_worst_residual": float(cold_res.max()),
           "warm_worst_residual": float(warm_res.max()),
           "max_abs_diff_warm_vs_cold_v": max_dv,
           "inference_s": infer_t,
           "warm_time_includes_inference": True}
    if log:
        log(f"  eval: {rep}")
    return rep
