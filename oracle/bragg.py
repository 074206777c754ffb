"""The Bragg shooting solve in NumPy/SciPy FP64 (oracle; test use only).

Restates the reference's bragg/solver.py:

    march_down         D0 <- e0 roll(D0 + ds Cb[k] Dg, +1)
                       Dg <- roll((Dg - ds Ch[k] D0)/eg, -1)     (28-35)
    perfect_symbol     mode-by-mode 2x2 chain over Nz-1 steps   (38-60)
    precon             w -> ifft(fft(w, x) / T22s), T22s floored at
                       1e-8 max|T22| by a real value            (163-168)
    solve_bragg_stable one plane, SciPy GMRES restart 40        (135-179)
    solve_surface_batch chunked planes, retry/grow/split policy (182-379)
    fixed_point        the trapezoid round-trip fixed point      (63-132)

with Ch = sigma_h e^{-iH}, Cb = sigma_hbar e^{+iH}, e0 = e^{s0 ds},
eg = e^{(s0 + i alpha) ds}, alpha = -2 pi s_dev, and the exit wave
v e^{-iH[0]}.  GMRES is scipy.sparse.linalg.gmres (left
preconditioning, true-residual stop), the reference's own dependency.
"""

import numpy as np

from . import volterra as OV


def step_constants(s_dev, sig0, ds):
    alpha = -2.0 * np.pi * float(s_dev)
    return complex(np.exp(sig0 * ds)), complex(np.exp((sig0 + 1j * alpha) * ds))


def march_down(D0, Dg, Ch, Cb, e0, eg, ds, Nz):
    for k in range(Nz - 1):
        D0n = e0 * np.roll(D0 + ds * Cb[k] * Dg, 1, axis=-1)
        Dgn = np.roll((Dg - ds * Ch[k] * D0) / eg, -1, axis=-1)
        D0, Dg = D0n, Dgn
    return Dg


def perfect_symbol(e0, eg, ds, Nz, Nx, sig_h, sig_hbar):
    m = np.fft.fftfreq(Nx) * Nx
    p = np.exp(-2j * np.pi * m / Nx)
    pi = np.conj(p)
    T11 = np.ones(Nx, complex)
    T12 = np.zeros(Nx, complex)
    T21 = np.zeros(Nx, complex)
    T22 = np.ones(Nx, complex)
    for _ in range(Nz - 1):
        a11, a12 = e0 * p, e0 * p * ds * sig_hbar
        a21, a22 = -pi * ds * sig_h / eg, pi / eg
        T11, T12, T21, T22 = (a11 * T11 + a12 * T21, a11 * T12 + a12 * T22,
                              a21 * T11 + a22 * T21, a21 * T12 + a22 * T22)
    return T21, T22


def floored_symbol(T22):
    floor = 1e-8 * np.abs(T22).max()
    return np.where(np.abs(T22) < floor, floor, T22)


def couplings(seg_list, cfg, xtal, g_vec, y_values):
    """Ch, Cb of shape (Nz, ny, Nx) and H[0] (ny, Nx)."""
    ny = len(y_values)
    if seg_list:
        H = np.stack([OV.h_plane(seg_list, g_vec, xtal.nu, xtal.a_core,
                                 cfg.Nx, cfg.Nz, cfg.dx, cfg.dz, float(y))
                      for y in y_values], axis=1)
    else:
        H = np.zeros((cfg.Nz, ny, cfg.Nx))
    Ch = complex(xtal.sig_h) * np.exp(-1j * H)
    Cb = complex(xtal.sig_hbar) * np.exp(1j * H)
    return Ch, Cb, H[0]


def incident(cfg, n):
    if cfg.beam_thickness is None:
        return np.ones((n, cfg.Nx), np.complex128)
    x = (np.arange(cfg.Nx) - cfg.Nx / 2) * cfg.dx
    top = (np.abs(x - cfg.beam_centre) < 0.5 * cfg.beam_thickness)
    return np.broadcast_to(top.astype(np.complex128), (n, cfg.Nx)).copy()


def solve_surface_batch(seg_list, s_dev, cfg, y_values, xtal, g_vec,
                        chunk=32, tol=1e-10, restart=40, maxiter=25,
                        accept=1e-7, grow=4, tries=3, restart_cap=256,
                        max_depth=6, per_plane=6, chunk_cap=32):
    """(ny, Nx) exit wave and info, the CPU path of the reference."""
    from scipy.sparse.linalg import LinearOperator, gmres
    s0, sh, sb = complex(xtal.sig0), complex(xtal.sig_h), complex(xtal.sig_hbar)
    ds = cfg.ds
    e0, eg = step_constants(s_dev, s0, ds)
    _, T22 = perfect_symbol(e0, eg, ds, cfg.Nz, cfg.Nx, sh, sb)
    T22s = floored_symbol(T22)
    y_values = np.asarray(y_values, float)
    ny = len(y_values)
    out = np.zeros((ny, cfg.Nx), np.complex128)
    st = dict(worst=0.0, its=0, retries=0, splits=0, unconverged=0)
    chunk = max(1, min(chunk, chunk_cap))

    def solve_block(j0, j1, depth=0):
        nyc = j1 - j0
        Ch, Cb, H0 = couplings(seg_list, cfg, xtal, g_vec, y_values[j0:j1])
        zero = np.zeros((nyc, cfg.Nx), np.complex128)
        d = march_down(incident(cfg, nyc), zero, Ch, Cb, e0, eg, ds,
                       cfg.Nz).ravel()

        def matvec(v):
            return march_down(zero, v.reshape(nyc, cfg.Nx), Ch, Cb, e0, eg,
                              ds, cfg.Nz).ravel()

        def precon(w):
            W = w.reshape(nyc, cfg.Nx)
            return np.fft.ifft(np.fft.fft(W, axis=-1) / T22s[None, :],
                               axis=-1).ravel()

        N = nyc * cfg.Nx
        cnt = [0]
        cap = max(restart, min(restart_cap, 8 * nyc))
        start = max(restart, min(cap, per_plane * nyc))
        A = LinearOperator((N, N), matvec=matvec, dtype=np.complex128)
        M = LinearOperator((N, N), matvec=precon, dtype=np.complex128)
        nb = np.linalg.norm(d) + 1e-300
        v, x0, r = None, None, start
        for _ in range(tries):
            v, _code = gmres(A, -d, x0=x0, M=M, rtol=tol, restart=r,
                             maxiter=maxiter,
                             callback=lambda _: cnt.__setitem__(0, cnt[0] + 1),
                             callback_type="pr_norm")
            res = float(np.linalg.norm(A.matvec(v) + d) / nb)
            if res <= accept or r >= cap:
                break
            st["retries"] += 1
            x0, r = v, min(r * grow, cap)
        if res > accept and nyc > 1 and depth < max_depth:
            st["splits"] += 1
            mid = j0 + nyc // 2
            solve_block(j0, mid, depth + 1)
            solve_block(mid, j1, depth + 1)
            return
        st["unconverged"] += res > accept
        st["worst"] = max(st["worst"], res)
        st["its"] = max(st["its"], cnt[0])
        out[j0:j1] = v.reshape(nyc, cfg.Nx) * np.exp(-1j * H0)

    for j0 in range(0, ny, chunk):
        solve_block(j0, min(j0 + chunk, ny))
    return out, {"residual": st["worst"], "iterations": st["its"],
                 "retries": st["retries"], "splits": st["splits"],
                 "unconverged_chunks": st["unconverged"],
                 "converged": st["unconverged"] == 0}


def dense_shooting_solution(seg_list, s_dev, cfg, y_plane, xtal, g_vec):
    """Brute force on a tiny grid: assemble A column by column (Nx
    matvecs), solve A v = -d densely, return the exit wave."""
    s0, ds = complex(xtal.sig0), cfg.ds
    e0, eg = step_constants(s_dev, s0, ds)
    Ch, Cb, H0 = couplings(seg_list, cfg, xtal, g_vec, [y_plane])
    zero = np.zeros((1, cfg.Nx), np.complex128)
    d = march_down(incident(cfg, 1), zero, Ch, Cb, e0, eg, ds, cfg.Nz)[0]
    A = np.empty((cfg.Nx, cfg.Nx), np.complex128)
    for j in range(cfg.Nx):
        e = np.zeros((1, cfg.Nx), np.complex128)
        e[0, j] = 1.0
        A[:, j] = march_down(zero, e, Ch, Cb, e0, eg, ds, cfg.Nz)[0]
    v = np.linalg.solve(A, -d)
    return v * np.exp(-1j * H0[0]), A, d


def fixed_point(seg_list, s_dev, cfg, xtal, g_vec, n_iter=400, tol=1e-11,
                y_plane=0.0):
    """The reference's solve_bragg (bragg/solver.py:63-132): trapezoid
    forward march of D0 with Dg frozen, backward march of Dg from
    Dg[Nz-1] = 0 with D0 frozen, uniform unit incidence, stop when
    max|Dg[0] - previous| / (max|Dg[0]| + 1e-300) < tol.  Returns the
    exit wave Dg[0] e^{-iH[0]} and (iterations, residual, converged)."""
    Ch, Cb, H = couplings(seg_list, cfg, xtal, g_vec, [float(y_plane)])
    Ch, Cb, H0 = Ch[:, 0], Cb[:, 0], H[0]
    e0, eg = step_constants(s_dev, complex(xtal.sig0), cfg.ds)
    ds, Nz, Nx = cfg.ds, cfg.Nz, cfg.Nx
    D0 = np.zeros((Nz, Nx), complex)
    Dg = np.zeros((Nz, Nx), complex)
    D0[0] = 1.0
    prev, res, its, conv = None, np.nan, n_iter, False
    for it in range(n_iter):
        for k in range(Nz - 1):
            D0[k + 1] = (e0 * np.roll(D0[k], 1)
                         + 0.5 * ds * (e0 * np.roll(Cb[k] * Dg[k], 1)
                                       + Cb[k + 1] * Dg[k + 1]))
        Dg[Nz - 1] = 0.0
        for k in range(Nz - 2, -1, -1):
            Dg[k] = (eg * np.roll(Dg[k + 1], 1)
                     + 0.5 * ds * (eg * np.roll(Ch[k + 1] * D0[k + 1], 1)
                                   + Ch[k] * D0[k]))
        cur = Dg[0].copy()
        if prev is not None:
            res = np.abs(cur - prev).max() / (np.abs(cur).max() + 1e-300)
            if res < tol:
                its, conv = it + 1, True
                break
        prev = cur
    return Dg[0] * np.exp(-1j * H0), (its, float(res), conv)


def fixed_point_discrete_perfect(s_dev, cfg, xtal):
    """The fixed point's limit for H = 0 and uniform incidence, by a dense
    solve of the discrete two-point problem it iterates on: with no x
    dependence the rolls drop out and D0[k+1] = e0 D0[k] + ds/2 (e0 sb
    Dg[k] + sb Dg[k+1]), Dg[k] = eg Dg[k+1] + ds/2 (eg sh D0[k+1] + sh
    D0[k]), D0[0] = 1, Dg[Nz-1] = 0: 2 Nz linear equations.  Returns
    Dg[0]."""
    e0, eg = step_constants(s_dev, complex(xtal.sig0), cfg.ds)
    sh, sb, h = complex(xtal.sig_h), complex(xtal.sig_hbar), 0.5 * cfg.ds
    Nz = cfg.Nz
    A = np.zeros((2 * Nz, 2 * Nz), complex)
    b = np.zeros(2 * Nz, complex)
    d0 = lambda k: k                          # unknown indices
    dg = lambda k: Nz + k
    A[0, d0(0)] = 1.0
    b[0] = 1.0
    for k in range(Nz - 1):                   # rows 1 .. Nz-1
        r = 1 + k
        A[r, d0(k + 1)] = 1.0
        A[r, d0(k)] = -e0
        A[r, dg(k)] = -h * e0 * sb
        A[r, dg(k + 1)] = -h * sb
    A[Nz, dg(Nz - 1)] = 1.0                   # back face
    for k in range(Nz - 1):                   # rows Nz+1 .. 2Nz-1
        r = Nz + 1 + k
        A[r, dg(k)] = 1.0
        A[r, dg(k + 1)] = -eg
        A[r, d0(k + 1)] = -h * eg * sh
        A[r, d0(k)] = -h * sh
    x = np.linalg.solve(A, b)
    return x[dg(0)]
