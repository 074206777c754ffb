"""Benchmark of the Laue hot path on one MI355X; prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload: the mixed-precision rocking-scan configuration - diamond (400)
at 17 keV, t = 273 um, dx = 0.25 um, beam 120 um, 256 rows (Nz 490,
Nx 1056), 64 s_dev uniform in [-200, +200] urad, 8 seeded <110>
segments: 8.460e9 pixel-steps per solve.  One step is one full
solve_dislocation call (Volterra e^{iH} table + march of all 64 angles)
with the inputs already on the host and nothing cached between steps
but the workspace buffers.

`value` is FP64 pixel-steps/s; the same line carries the mixed
(two-float FP32) rate on the same workload with its error against FP64,
the worked-example drop-in rate, the roofline of the dominant kernel
(the march, HIP events on the stream it runs on), the Bragg legs (the
reference's joint-chunk solve as the headline, independent planes
beside it) with the roofline of the Bragg matvec, and the CPU oracle's
rates on bounded samples of the same workloads (one process and all
cores).

    python bench.py --probe-laue   # the rocking-scan solves only (FP64
                                   # then mixed), for a kernel trace
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("Laue pixel-steps/s FP64 vs mixed @<=1e-6 rel err; "
          "Bragg solves/s cold vs warm")
FLOPS_PER_PIXEL_STEP = 40        # 2x2 complex matvec 28 + two couplings 12
ISSUED_FLOPS_PER_PIXEL_STEP = 32  # 16 FP64 FMA/MUL slots, 2 flops each
BRAGG_FLOPS_PER_PIXEL_STEP = 28   # the reference's _march_down step
PEAK_FP64_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0
CPU_WORKERS_MAX = 16              # the GPU box's CPU share


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU oracle baseline")
    ap.add_argument("--no-ensemble", action="store_true",
                    help="skip the batched-ensembles leg")
    ap.add_argument("--probe-bragg", action="store_true",
                    help="run only the 20 um Bragg surface batch with one edge line "
                         "one warm-up and five timed), for a kernel trace, "
                         "and exit")
    ap.add_argument("--probe-laue", action="store_true",
                    help="run only the rocking-scan solves (FP64 then "
                         "mixed; one warm-up and five timed each), for a "
                         "kernel trace, and exit")
    ap.add_argument("--probe-worked-example", action="store_true",
                    help="run only the 200 um worked example's FP64 solves "
                         "(one warm-up and five timed), for a kernel trace, "
                         "and exit")
    ap.add_argument("--probe-matvec", action="store_true",
                    help="run only the 20 um Bragg matvec timing (20 "
                         "launches of the full 1312-row march), for "
                         "counter passes, and exit")
    ap.add_argument("--probe-gemm", action="store_true",
                    help="time the warm-start network's GEMMs (forward, "
                         "data and weight gradients of every 3x3 layer at "
                         "batch 32) with HIP events, print one line each, "
                         "and exit")
    ap.add_argument("--probe-ensemble", action="store_true",
                    help="run only the 256-instance ensembles (FP64, then "
                         "mixed: autotune, capture, one replay each), for "
                         "a kernel trace, and exit")
    ap.add_argument("--probe", action="store_true",
                    help="for counter runs: one phase build and one march "
                         "per precision, then a 1 GiB copy kernel (the "
                         "byte calibration), and exit")
    return ap.parse_args()


def laue_step(N, prepare_phase, march, segs, s_dev_dev, grid, xtal, prec,
              Ny, Nx, ev):
    """One solve_dislocation, with events around the march launch."""
    ev[0].record()
    src = prepare_phase(segs, grid, xtal, precision=prec, reuse=True)
    out = N.DeviceArray.empty((len(s_dev_dev), Ny, Nx), prec.complex)
    ev[1].record()
    march(src, s_dev_dev, grid, xtal, out, prec)
    ev[2].record()
    return out


def time_laue(N, segs, s, grid, xtal, prec, steps, warmup):
    from dyfox_amd.laue.solver import march, prepare_phase
    Nx, Ny = grid.geometry()[:2]
    sd = N.DeviceArray.from_host(np.asarray(s, np.float64))
    evs = [[N.Event() for _ in range(3)] for _ in range(steps)]
    for _ in range(warmup):
        laue_step(N, prepare_phase, march, segs, sd, grid, xtal, prec, Ny,
                  Nx, evs[0])
    N.synchronize()
    t0 = time.perf_counter()
    out = None
    for k in range(steps):
        out = laue_step(N, prepare_phase, march, segs, sd, grid, xtal, prec,
                        Ny, Nx, evs[k])
    N.synchronize()
    wall = (time.perf_counter() - t0) / steps
    phase_ms = float(np.mean([e[0].elapsed_ms(e[1]) for e in evs]))
    march_ms = float(np.mean([e[1].elapsed_ms(e[2]) for e in evs]))
    return wall, phase_ms, march_ms, out


def time_bragg(xtal, crystal, repeats=3, planes="joint"):
    """20 um Bragg surface: 41 s_dev over one 32-plane chunk (Nz 90, Nx
    520), solved as one batched GMRES per repeat (cold start); returns
    solves/s and per-solve info.  planes="joint" is the reference's chunk
    (one Krylov space for the 32 planes of an angle, its iteration
    counts), "independent" one Krylov space per plane."""
    from dyfox_amd.bragg import solve_surface_batch
    from dyfox_amd.workloads import bragg_surface
    cfg, segs, ys, s, g = bragg_surface(xtal, crystal)
    solve_surface_batch(segs, s, cfg, ys, xtal, g, planes=planes)  # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        _, infos = solve_surface_batch(segs, s, cfg, ys, xtal, g,
                                       planes=planes)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    its = [i["iterations"] for i in infos]
    return {"solves_per_s": len(s) / dt, "ms_per_batch": dt * 1e3,
            "n_solves": len(s), "gmres_iterations": [min(its), max(its)],
            "worst_residual": max(i["residual"] for i in infos),
            "all_converged": all(i["converged"] for i in infos),
            "grid": f"Nz={cfg.Nz} Nx={cfg.Nx} planes={len(ys)}"}


def time_ensembles():
    """256 uneven Laue instances (two reflections, t 50-300 um,
    1-40 segments, 16-128 angles) through the persistent batched path:
    autotuned launch shape, graph captured on the first run, the second
    (replayed) run timed, FP64 then mixed."""
    from dyfox_amd.backend import FP64, MIXED
    from dyfox_amd.laue.batch import EnsembleSolver
    from dyfox_amd.workloads import ensemble_batch
    inst = ensemble_batch()
    out = {}
    for prec in (FP64, MIXED):
        t0 = time.perf_counter()
        sv = EnsembleSolver(inst, precision=prec, keep_outputs=False)
        setup = time.perf_counter() - t0
        tim = sv.autotune(repeats=1)
        sv.run()
        t0 = time.perf_counter()
        sv.run()
        dt = time.perf_counter() - t0
        ps = sv.pixel_steps
        out[prec.name] = {"value": ps / dt, "seconds": dt,
                          "pixel_steps": ps, "groups": len(sv.groups),
                          "workspace_gb": sv.work.nbytes / 1e9,
                          "tuned_waves_per_block_blocks_per_cu":
                              list(sv.tune),
                          "autotune_ms": {f"{a}x{b}": round(v, 1)
                                          for (a, b), v in tim.items()},
                          "setup_s": setup}
        sv.close()
        del sv
        print(f"[bench] ensemble-256 {prec.name} {ps / dt:.3e} pixel-steps/s "
              f"({dt:.2f} s)", file=sys.stderr, flush=True)
    out["n_instances"] = len(inst)
    return out


def time_illumination(N, segs, grid, xtal, repeats=2):
    """An incoherent image of the rocking-scan field under a convergent,
    polychromatic beam: 7 angles x 5 energies (sigma_psi 20 urad,
    sigma_delta 1e-4), each node marched as one angle, imaged through its
    own chromatic charge-2 pupil, and summed with its weight in FP64 on
    the device (illumination.incoherent_image).  Timed end to end,
    including the one e^{iH} table of the field."""
    from dyfox_amd.illumination import (convergence_bandwidth_nodes,
                                        incoherent_image)
    from dyfox_amd.optics import Optics
    lens = Optics(1e-4, 0.274, xtal)
    s, d, w, info = convergence_bandwidth_nodes(xtal, 20e-6, 1e-4, 7, 60e-6,
                                                5, 3e-4)
    incoherent_image(segs, grid, xtal, lens, s, w, d_arg=d, ell=2)
    N.synchronize()
    t0 = time.perf_counter()
    for _ in range(repeats):
        img = incoherent_image(segs, grid, xtal, lens, s, w, d_arg=d, ell=2)
    N.synchronize()
    dt = (time.perf_counter() - t0) / repeats
    del img
    return {"nodes_per_s": len(s) / dt, "ms_per_image": dt * 1e3,
            "nodes": len(s), "grid": "256 x 1056, Nz 490"}


def time_imaging(N, stack, grid, xtal, repeats=5):
    """The exit-wave stack of the rocking scan (64 x 256 x 1056 complex)
    imaged through the open and the charge-2 pupil (Optics.image: one
    batched forward 2-D FFT, pupil multiply, one batched inverse, |.|^2),
    timed with events; the oracle's NumPy imaging of 4 of the images on
    the host beside it."""
    from dyfox_amd.optics import Optics
    from oracle import optics as OO
    lens = Optics(1e-4, 0.274, xtal)
    Ny, Nx = stack.shape[-2:]
    for ell in (0, 2):
        lens.image(stack, grid.dx, ell=ell)                # plans, pupils
    N.synchronize()
    ev = [N.Event() for _ in range(2)]
    ev[0].record()
    for _ in range(repeats):
        for ell in (0, 2):
            img = lens.image(stack, grid.dx, ell=ell)
    ev[1].record()
    N.synchronize()
    ms = ev[0].elapsed_ms(ev[1]) / repeats
    n_img = 2 * stack.shape[0]
    # bytes touched per image: field read, spectrum write + read (forward),
    # product write + read (inverse), image write: 5 complex + 1 real planes
    plane = Ny * Nx
    nbytes = n_img * plane * (5 * 16 + 8)
    host = stack[:4].get().astype(complex)
    pup = OO.pupil(Nx, Ny, grid.dx, 2, lens.aperture_f)
    t0 = time.perf_counter()
    OO.image_from_spectrum(OO.object_spectrum(host), pup)
    cpu = 4 / (time.perf_counter() - t0)
    del img
    return {"images_per_s": n_img / (ms * 1e-3), "ms_per_scan_2_pupils": ms,
            "grid": f"{stack.shape[0]} x {Ny} x {Nx}",
            "effective_GB_per_s": nbytes / (ms * 1e-3) / 1e9,
            "cpu_oracle_images_per_s": cpu}


def time_fixed_point(N, xtal):
    """The reference's fixed-point Bragg solve (bragg/solver.py:63-132)
    on its own test grid (t = 8 um, dx = 0.1 um, Nz 180, Nx 832), 41
    angles in one launch, against the NumPy oracle for one angle."""
    from dyfox_amd.bragg import BraggConfig, bragg_g_vector
    from dyfox_amd.bragg.solver import solve_bragg
    from dyfox_amd.units import deviation_from_urad
    from oracle import bragg as OB
    g = bragg_g_vector(xtal)
    cfg = BraggConfig(xtal, t_crystal=8e-6, dx=0.1e-6, half_x=4e-6)
    s = deviation_from_urad(np.linspace(-60, 60, 41), xtal)
    solve_bragg([], s, cfg, xtal, g, n_iter=600, tol=1e-13)
    t0 = time.perf_counter()
    _, infos = solve_bragg([], s, cfg, xtal, g, n_iter=600, tol=1e-13)
    dt = time.perf_counter() - t0
    t0 = time.perf_counter()
    OB.fixed_point([], float(s[20]), cfg, xtal, g, n_iter=600, tol=1e-13)
    cpu = time.perf_counter() - t0
    its = [i["iterations"] for i in infos]
    return {"solves_per_s": len(s) / dt, "ms_per_batch": dt * 1e3,
            "sweeps": [min(its), max(its)],
            "all_converged": all(i["converged"] for i in infos),
            "grid": f"Nz={cfg.Nz} Nx={cfg.Nx}",
            "cpu_oracle_solves_per_s": 1.0 / cpu}


def time_bragg_matvec(N, xtal, repeats=20):
    """The GMRES matvec of the 20 um edge-line chunk on its own: one
    down-march of D_g(0) (41 angles x 32 planes = 1312 rows, Nz 90, Nx
    520; the kernel every iteration runs), HIP events around `repeats`
    launches on the stream they run on."""
    from dyfox_amd.bragg import solver as BS
    from dyfox_amd.workloads import bragg_surface
    cfg, segs, ys, s, g = bragg_surface(xtal, "edge_line")
    E = BS.chunk_phase(segs, cfg, xtal, g, ys)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(len(s), len(ys), cfg.Nx)) + 1j * rng.normal(
        size=(len(s), len(ys), cfg.Nx))
    import ctypes as C
    a = BS.bragg_args(cfg, xtal, len(ys))
    sd = N.DeviceArray.from_host(np.asarray(s, np.float64))
    vd = N.DeviceArray.from_host(v)
    out = N.DeviceArray.empty(v.shape, np.complex128)
    args = (C.byref(a), C.c_void_p(sd.ptr), len(s), C.c_void_p(E.ptr), None,
            C.c_void_p(vd.ptr), C.c_void_p(out.ptr), None)
    N.call("dyf_bragg_march", *args)                         # warm-up
    ev = [(N.Event(), N.Event()) for _ in range(repeats)]
    # a few milliseconds of copies first, so every timed launch and its
    # events are queued before the device reaches them: each event pair
    # then brackets the kernel alone, not the host's time to launch it
    nb = 1 << 30
    src = N.DeviceArray.empty((nb,), np.uint8)
    dst = N.DeviceArray.empty((nb,), np.uint8)
    N.synchronize()
    for _ in range(8):
        N.call("dyf_copy_kernel", C.c_void_p(dst.ptr), C.c_void_p(src.ptr),
               C.c_size_t(nb), None)
    for e0, e1 in ev:
        e0.record()
        N.call("dyf_bragg_march", *args)
        e1.record()
    ms = float(np.mean([e0.elapsed_ms(e1) for e0, e1 in ev]))
    del src, dst
    steps = len(s) * len(ys) * cfg.Nx * (cfg.Nz - 1)
    # e^{iH} table read once (shared by the 41 angles), D_g(0) in, D_g(t) out
    nbytes = cfg.Nz * len(ys) * cfg.Nx * 16 + 2 * len(s) * len(ys) * cfg.Nx * 16
    tfl = BRAGG_FLOPS_PER_PIXEL_STEP * steps / (ms * 1e-3) / 1e12
    gbs = nbytes / (ms * 1e-3) / 1e9
    return {"kernel": "bragg_march_halo_kernel<5, 3> (the GMRES matvec)",
            "bound": "fp64_valu", "achieved": tfl, "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s", "frac": tfl / PEAK_FP64_TFLOPS,
            "hbm_GB_per_s": gbs, "hbm_frac": gbs / PEAK_HBM_GBS,
            "algorithmic_bytes": nbytes, "pixel_steps": steps,
            "avg_launch_ms": ms,
            "note": "28 algorithmic flops per pixel-step (the reference's "
                    "_march_down step); the interaction-picture kernel "
                    "issues 16 FP64 FMA/MUL instructions per column-step "
                    "(12 FMA + 4 MUL: 28 flops, or 32 counting each "
                    "instruction as one FMA slot of the peak), plus the "
                    "halo lanes' duplicate columns (64 lanes march the "
                    "work of 52)"}


def cpu_info():
    """Host cores this process may use, the CPU model, library versions."""
    import platform
    import scipy
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"nproc": nproc, "cpu_model": model, "numpy": np.__version__,
            "scipy": scipy.__version__,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def _laue_cpu_rows(args):
    """Oracle Volterra H + direct-form march of the rocking scan on the
    given marched rows and angle indices (one process); pixel-steps."""
    pick, s_idx = args
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import laue_rocking_scan
    from oracle import laue as OL, volterra as OV
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    grid, segs, s = laue_rocking_scan(xtal)
    Nx, Ny, dz, ds, spx = grid.geometry()
    pick = np.asarray(pick)
    H = OV.h_grid(segs, xtal.g_vec, xtal.nu, xtal.a_core, Nx, Ny, grid.Nz,
                  grid.dx, dz, rows=pick)
    OL.march(np.asarray(s)[list(s_idx)], complex(xtal.sig0),
             complex(xtal.sig_h), complex(xtal.sig_hbar), grid.Nz, Nx, Ny,
             grid.dx, ds, spx, grid.sheet_arg(), grid.beam_width / 2, H=H,
             rows=pick)
    return len(pick) * len(s_idx) * Nx * (grid.Nz - 1)


_THREAD_VARS = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")


def _warm_worker(_):
    import oracle.bragg  # noqa: F401
    import oracle.laue  # noqa: F401
    import scipy.sparse.linalg  # noqa: F401
    return 0


def _bragg_cpu_angle(j):
    """The reference's CPU path (NumPy march + SciPy GMRES, joint chunk)
    for angle j of the 20 um edge-line surface; GMRES iterations."""
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import bragg_surface
    from oracle import bragg as OB
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    cfg, segs, ys, s, g = bragg_surface(xtal, "edge_line")
    _, info = OB.solve_surface_batch(segs, float(s[j]), cfg, ys, xtal, g)
    return info["iterations"]


def cpu_baseline(rows=48, n_s=8):
    """The NumPy FP64 oracle on bounded samples of the benchmark workloads,
    on the host cores, one process and all cores (a process pool, one
    row block x angle set or one Bragg angle per task).  Run before the
    GPU is initialised (the pool's workers are fresh interpreters).

    Laue: the rocking scan's Volterra H and direct-form march on `rows`
    marched rows x `n_s` angles (one process), and on W times that many
    rows over W workers.  Bragg: the 20 um edge-line surface chunk
    (reference CPU path, joint chunk) for 4 angles in one process, and
    all 41 angles over the workers."""
    import multiprocessing as mp
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import laue_rocking_scan
    info = cpu_info()
    workers = max(1, min(CPU_WORKERS_MAX, info["nproc"]))
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    grid, _, s = laue_rocking_scan(xtal)
    marched = grid.marched_rows()
    s_idx = tuple(np.linspace(0, len(s) - 1, n_s).astype(int).tolist())
    one_rows = marched[np.linspace(0, len(marched) - 1, rows).astype(int)]
    t0 = time.perf_counter()
    work1 = _laue_cpu_rows((one_rows.tolist(), s_idx))
    dt1 = time.perf_counter() - t0
    blocks = np.array_split(marched[:min(len(marched), rows * workers)],
                            workers)
    ctx = mp.get_context("spawn")
    # one thread per worker process (the workers inherit this environment)
    saved = {k: os.environ.get(k) for k in _THREAD_VARS}
    os.environ.update({k: "1" for k in _THREAD_VARS})
    try:
        pool = ctx.Pool(workers)
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    with pool:
        pool.map(_warm_worker, range(workers))          # start-up, imports
        t0 = time.perf_counter()
        workn = sum(pool.map(_laue_cpu_rows,
                             [(b.tolist(), s_idx) for b in blocks]))
        dtn = time.perf_counter() - t0
        t0 = time.perf_counter()
        its_all = pool.map(_bragg_cpu_angle, list(range(41)))
        dtb_n = time.perf_counter() - t0
    t0 = time.perf_counter()
    for j in (0, 13, 27, 40):
        _bragg_cpu_angle(j)
    dtb_1 = time.perf_counter() - t0
    return {"value": workn / dtn, "unit": "pixel-steps/s", "cores": workers,
            "kind": "port",
            "sample": f"rocking scan: {rows * workers} marched rows x {n_s} "
                      f"s_dev ({workn:.3e} pixel-steps, Volterra H of those "
                      f"rows included) over {workers} processes in "
                      f"{dtn:.1f} s",
            "single_process": {"value": work1 / dt1, "cores": 1,
                               "sample": f"{rows} rows x {n_s} s_dev "
                                         f"({work1:.3e} pixel-steps), "
                                         f"{dt1:.1f} s"},
            "bragg_surface_20um_edge_line": {
                "unit": "solves/s",
                "all_cores": 41 / dtb_n, "cores": workers,
                "single_process": 4 / dtb_1,
                "gmres_iterations": [min(its_all), max(its_all)],
                "sample": "reference CPU path (NumPy _march_down, NumPy FFT "
                          "preconditioner, SciPy GMRES, joint 32-plane "
                          "chunk): 41 angles over the workers, 4 in one "
                          "process"},
            "host": info}


def time_warm_start(n_test=32):
    """Learned warm start, measured live on a bounded held-out sample:
    n_test Bragg instances over the 20-100 um range (8 thicknesses, 0-12
    <110> lines, 64 planes, tol 1e-10) with the committed network
    (dyfox_amd/warmstart/weights_cfg5.npz, trained by
    `python -m dyfox_amd.warmstart.experiment`), solved cold, warm from
    the perfect-crystal value and warm from the network (inference
    included): one Krylov space per plane over the whole range, the
    reference's joint chunks up to 35 um."""
    from dyfox_amd.warmstart.experiment import evaluate_saved
    path = os.path.join(ROOT, "dyfox_amd", "warmstart", "weights_cfg5.npz")
    rep = evaluate_saved(path, n_test, joint_max_t=35e-6, log=lambda m: None)

    def brief(r):
        out = {}
        for band, v in [("all", r["all"])] + list(r["bands"].items()):
            out[band] = {name: {k: v[name][k] for k in
                                ("n", "solves_per_s", "iterations_mean",
                                 "converged_fraction", "retries_mean",
                                 "splits_mean") if k in v[name]}
                         for name in ("cold", "network")}
            out[band]["warm_over_cold"] = v["network"]["speedup_vs_cold"]
            out[band]["perfect_guess_over_cold"] = \
                v["perfect"]["speedup_vs_cold"]
        return out

    return {"n_test": n_test, "weights": "dyfox_amd/warmstart/"
            "weights_cfg5.npz (8,000 cold solves, 20-100 um)",
            "thicknesses_um": rep["test_thicknesses_um"],
            "independent_planes": brief(rep["independent_planes"]),
            "joint_chunks_to_35um": brief(rep["joint_chunks"]),
            "tol": 1e-10}


def march_fraction(N, grid, xtal, prec, n_e):
    """Share of the launch's pixel-steps the grid march executes
    (dyf_laue_march_fraction)."""
    import ctypes as C
    from dyfox_amd.laue.solver import laue_args
    f = C.c_double()
    N.call("dyf_laue_march_fraction", C.byref(laue_args(grid, xtal)), 0,
           prec.code, int(n_e), C.byref(f))
    return f.value


def traffic_from_profile():
    """HBM bytes per march launch from the committed PMC summary."""
    path = os.path.join(ROOT, "profiles", "laue_march_rocking_scan_273um_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("fp64_bytes_per_launch")


def probe():
    """One dispatch of each kernel for rocprofv3 --pmc passes."""
    import ctypes as C
    from dyfox_amd import _native as N
    from dyfox_amd.backend import FP64, MIXED, device
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.laue.solver import march, prepare_phase
    from dyfox_amd.workloads import laue_rocking_scan
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    grid, segs, s = laue_rocking_scan(xtal)
    Nx, Ny = grid.geometry()[:2]
    sd = N.DeviceArray.from_host(np.asarray(s, np.float64))
    for prec in (FP64, MIXED):
        src = prepare_phase(segs, grid, xtal, precision=prec)
        out = N.DeviceArray.empty((len(s), Ny, Nx), prec.complex)
        march(src, sd, grid, xtal, out, prec)
        N.synchronize()
        del src, out
    nb = 1 << 30
    a = N.DeviceArray.empty((nb,), np.uint8)
    b = N.DeviceArray.empty((nb,), np.uint8)
    N.call("dyf_copy_kernel", C.c_void_p(b.ptr), C.c_void_p(a.ptr),
           C.c_size_t(nb), None)
    N.synchronize()
    print("probe done", flush=True)


def probe_laue():
    """The rocking-scan solves alone (FP64, then mixed; one warm-up and
    five timed each), so a kernel trace's averages are this workload's."""
    from dyfox_amd import _native as N
    from dyfox_amd.backend import FP64, MIXED, device
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import laue_rocking_scan
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    grid, segs, s = laue_rocking_scan(xtal)
    for prec in (FP64, MIXED):
        w, p, m, out = time_laue(N, segs, s, grid, xtal, prec, 5, 1)
        del out
        print(f"probe-laue {prec.name}: step {w * 1e3:.2f} ms, phase "
              f"{p:.2f} ms, march {m:.2f} ms", flush=True)


def probe_worked_example():
    """The 200 um worked example's FP64 solves alone (one warm-up and
    five timed), so a kernel trace's averages are this workload's."""
    from dyfox_amd import _native as N
    from dyfox_amd.backend import FP64, device
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import laue_worked_example
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    g2, seg2, s2 = laue_worked_example(xtal)
    w, p, m, out = time_laue(N, seg2, s2, g2, xtal, FP64, 5, 1)
    del out
    print(f"probe-worked-example fp64: step {w * 1e3:.2f} ms, phase "
          f"{p:.2f} ms, march {m:.2f} ms", flush=True)


def probe_gemm(batch=32, repeats=10):
    """Every 3x3 layer's three products as the network issues them (the
    operand layouts and the weight gradient's pixel split of
    warmstart/net.py), on uninitialised operands, timed with an event
    pair around each launch."""
    import ctypes as C
    from dyfox_amd import _native as N
    from dyfox_amd.backend import device
    from dyfox_amd.warmstart.net import BUFFERS, H0, SPEC, W0
    device()
    peak_bf16 = 2500.0                       # TFLOP/s, dense
    for op in SPEC:
        if op[0] != "conv3":
            continue
        lvl = BUFFERS[op[1]][0]
        P = batch * (H0 >> lvl) * (W0 >> lvl)
        cin, cout = op[3], op[6]
        K9 = 9 * cin
        col = N.DeviceArray.empty((P * K9,), np.uint16)
        W = N.DeviceArray.empty((cout * K9,), np.uint16)
        dz = N.DeviceArray.empty((P * cout,), np.uint16)
        out = N.DeviceArray.empty((P * K9 * 2,), np.uint16)
        splits = max(1, min(256, P // 1024))
        v = C.c_void_p
        products = [
            ("fwd", P, cout, K9, (v(col.ptr), K9, 1, v(W.ptr), K9, 1, P, cout,
                                  K9, 0, 1, None, v(out.ptr), cout, 1, None)),
            ("dgrad", P, K9, cout, (v(dz.ptr), cout, 1, v(W.ptr), 1, K9, P, K9,
                                    cout, 1, 0, None, v(out.ptr), K9, 1,
                                    None)),
            ("wgrad", cout, K9, P, (v(dz.ptr), 1, cout, v(col.ptr), 1, K9,
                                    cout, K9, P, 2, 0, None, v(out.ptr), K9,
                                    splits, None))]
        for name, M, Nn, K, args in products:
            N.call("dyf_nn_gemm", *args)
            ev = [(N.Event(), N.Event()) for _ in range(repeats)]
            N.synchronize()
            for e0, e1 in ev:
                e0.record()
                N.call("dyf_nn_gemm", *args)
                e1.record()
            ms = float(np.mean([e0.elapsed_ms(e1) for e0, e1 in ev]))
            fl = 2.0 * M * Nn * K
            nb = 2.0 * (M * K + Nn * K) + (2 if name == "fwd" else 4) * M * Nn
            print(f"{name:5s} conv3 {cin:3d}->{cout:2d} level {lvl} M={M:7d} "
                  f"N={Nn:4d} K={K:7d}: {ms * 1e3:7.1f} us "
                  f"{fl / ms / 1e9:6.1f} TFLOP/s "
                  f"({fl / ms / 1e9 / peak_bf16 * 100:5.2f}% of BF16) "
                  f"{nb / ms / 1e6:6.0f} GB/s "
                  f"({nb / ms / 1e6 / PEAK_HBM_GBS * 100:5.1f}% of HBM)",
                  flush=True)


def probe_bragg():
    """The 20 um Bragg surface batch with one edge line on the
    reference's joint 32-plane chunk: one warm-up batch, five timed."""
    from dyfox_amd.backend import device
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    r = time_bragg(xtal, "edge_line", repeats=5)
    print(f"probe-bragg done: {r['ms_per_batch']:.2f} ms per batch, "
          f"iterations {r['gmres_iterations']}", flush=True)


def main():
    args = parse()
    if args.probe:
        return probe()
    if args.probe_bragg:
        return probe_bragg()
    if args.probe_laue:
        return probe_laue()
    if args.probe_worked_example:
        return probe_worked_example()
    if args.probe_gemm:
        return probe_gemm()
    if args.probe_ensemble:
        from dyfox_amd.backend import device
        device()
        r = time_ensembles()
        for name, v in r.items():
            if not isinstance(v, dict):
                continue
            print(f"probe-ensemble {name}: {v['seconds']:.3f} s, "
                  f"{v['value']:.3e} pixel-steps/s", flush=True)
        return None
    if args.probe_matvec:
        from dyfox_amd import _native as N
        from dyfox_amd.backend import device
        from dyfox_amd.crystal import DIAMOND, CrystalParams
        device()
        xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                             f_prime=0.003, f_dprime=0.002)
        r = time_bragg_matvec(N, xtal)
        print(f"probe-matvec done: {r['avg_launch_ms'] * 1e3:.1f} us per "
              f"launch", flush=True)
        return None
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path here "
              "(the work would shard over rows x s_dev with no "
              "communication between GPUs)", file=sys.stderr)
        sys.exit(2)
    from dyfox_amd import _native as N
    from dyfox_amd.backend import FP64, MIXED, device
    from dyfox_amd.crystal import DIAMOND, CrystalParams
    from dyfox_amd.workloads import laue_rocking_scan, laue_worked_example
    from oracle.laue import error_report

    # the CPU baseline first: its worker processes start before this
    # process initialises the GPU
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()
        print(f"[bench] cpu baseline {cpu['value']:.3e} pixel-steps/s on "
              f"{cpu['cores']} processes", file=sys.stderr, flush=True)
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    grid, segs, s = laue_rocking_scan(xtal)
    work = grid.pixel_steps(len(s))
    steps, warm = max(1, args.steps), max(0, args.warmup)
    print(f"[bench] rocking scan {grid} n_s={len(s)} n_seg={len(segs)} "
          f"work={work:.3e} pixel-steps", file=sys.stderr, flush=True)

    w64, p64, m64, out64 = time_laue(N, segs, s, grid, xtal, FP64, steps,
                                     warm)
    print(f"[bench] fp64  step {w64 * 1e3:.2f} ms (phase {p64:.2f}, march "
          f"{m64:.2f})", file=sys.stderr, flush=True)
    wmx, pmx, mmx, outmx = time_laue(N, segs, s, grid, xtal, MIXED, steps,
                                     warm)
    print(f"[bench] mixed step {wmx * 1e3:.2f} ms (phase {pmx:.2f}, march "
          f"{mmx:.2f})", file=sys.stderr, flush=True)
    rep = error_report(outmx.get(), out64.get())
    worst = {k: max(r[k] for r in rep) for k in rep[0]}
    imaging = time_imaging(N, out64, grid, xtal)
    print(f"[bench] imaging {imaging['images_per_s']:.0f} images/s",
          file=sys.stderr, flush=True)
    del out64, outmx
    illum = time_illumination(N, segs, grid, xtal)
    print(f"[bench] incoherent image, {illum['nodes']} nodes: "
          f"{illum['ms_per_image']:.1f} ms", file=sys.stderr, flush=True)

    g2, seg2, s2 = laue_worked_example(xtal)
    w2, p2, m2, _ = time_laue(N, seg2, s2, g2, xtal, FP64, max(3, steps // 2),
                              1)
    work2 = g2.pixel_steps(len(s2))
    print(f"[bench] worked example fp64 step {w2 * 1e3:.2f} ms", file=sys.stderr,
          flush=True)

    bragg = {c: time_bragg(xtal, c) for c in ("perfect", "edge_line")}
    bragg["edge_line_independent_planes"] = time_bragg(
        xtal, "edge_line", planes="independent")
    bragg["semantics"] = ("perfect and edge_line: the reference's joint "
                          "32-plane chunk (its iterations, retries, splits); "
                          "edge_line_independent_planes: opt-in, one Krylov "
                          "space per plane")
    print(f"[bench] bragg surface, edge line: "
          f"{bragg['edge_line']['solves_per_s']:.1f} solves/s (independent "
          f"planes {bragg['edge_line_independent_planes']['solves_per_s']:.1f})",
          file=sys.stderr, flush=True)
    bragg_roof = time_bragg_matvec(N, xtal)
    warm_start = time_warm_start()
    ws = warm_start["independent_planes"]["all"]
    print(f"[bench] learned warm start, {warm_start['n_test']} instances "
          f"20-100 um: cold {ws['cold']['solves_per_s']:.1f} solves/s, warm "
          f"{ws['network']['solves_per_s']:.1f} "
          f"({ws['warm_over_cold']:.2f}x)", file=sys.stderr, flush=True)

    # the march executes `frac` of the pixel-steps: two angles share a
    # wavefront while the beam's fan fits half of one (the rest are exact
    # zeros it does not compute), so the roofline counts what it executes
    frac = march_fraction(N, grid, xtal, FP64, len(s))
    executed = work * frac
    achieved = FLOPS_PER_PIXEL_STEP * executed / (m64 * 1e-3) / 1e12
    issued = ISSUED_FLOPS_PER_PIXEL_STEP * executed / (m64 * 1e-3) / 1e12
    roof = {"kernel": "laue_fast_pair_kernel<17, FP64>",
            "bound": "fp64_valu", "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic_from_profile(),
            "traffic_source": "profiles/laue_march_rocking_scan_273um_"
                              "traffic.md (rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE; scaled by the copy-kernel "
                              "calibration)",
            "issued_tflops": issued,
            "issued_frac": issued / PEAK_FP64_TFLOPS,
            "note": "FP64 vector-ALU bound (no MFMA: the march is not a "
                    "contraction). achieved counts the 40 algorithmic "
                    "flops per executed pixel-step; the interaction-picture "
                    "kernel issues 16 FP64 FMA/MUL instructions per "
                    "pixel-step (12 FMA + 4 MUL by the VALU counters); "
                    "issued_* counts each as one FMA slot (2 flops, the "
                    "unit of the peak), so issued_frac is the share of FP64 "
                    "issue slots used. executed_fraction: the share of the workload's "
                    "pixel-steps the march computes (two angles per "
                    "wavefront while the beam's fan fits half of one; the "
                    "others are exact zeros)",
            "executed_fraction": frac,
            "executed_pixel_steps": executed,
            "avg_launch_ms": m64}
    line = {
        "metric": METRIC, "value": work / w64, "unit": "pixel-steps/s",
        "n_gpus": 1, "steps": steps, "warmup": warm,
        "ms_per_step": w64 * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded 8-segment <110> dislocation field)",
        "config": {"workload": "laue_rocking_scan_273um_64angles",
                   "crystal": "diamond(400) 17 keV", "t_um": 273,
                   "dx_um": 0.25, "Ny": 256, "Nx": 1056, "Nz": grid.Nz,
                   "n_s_dev": len(s), "n_segments": len(segs),
                   "pixel_steps": work, "parallelism": "single GPU"},
        "fp64": {"value": work / w64, "ms_per_step": w64 * 1e3,
                 "phase_ms": p64, "march_ms": m64,
                 "march_only": work / (m64 * 1e-3),
                 "march_executed_fraction": frac},
        "mixed": {"value": work / wmx, "ms_per_step": wmx * 1e3,
                  "phase_ms": pmx, "march_ms": mmx,
                  "march_only": work / (mmx * 1e-3),
                  "speedup_vs_fp64": w64 / wmx,
                  "march_speedup_vs_fp64": m64 / mmx,
                  "err_vs_fp64_worst_s_dev": worst},
        "worked_example_fp64": {"value": work2 / w2, "ms_per_step": w2 * 1e3,
                                "phase_ms": p2, "march_ms": m2,
                                "pixel_steps": work2},
        "bragg_surface_20um_cold": bragg,
        "bragg_fixed_point": time_fixed_point(N, xtal),
        "imaging_rocking_scan_273um": imaging,
        "illumination_rocking_scan_273um": illum,
        "laue_ensemble_256": None if args.no_ensemble else time_ensembles(),
        "bragg_warm_start_cfg5": warm_start,
        "roofline": roof,
        "bragg_roofline": bragg_roof,
    }
    if cpu is not None:
        cpu["gpu_over_cpu"] = line["value"] / cpu["value"]
        cb = cpu["bragg_surface_20um_edge_line"]
        cb["gpu_over_cpu"] = bragg["edge_line"]["solves_per_s"] / \
            cb["all_cores"]
        line["cpu_baseline"] = cpu
    print(json.dumps(line))


if __name__ == "__main__":
    main()
