"""The learned warm-start experiment end to end, on one device.

    python -m dyfox_amd.warmstart.experiment --train 20000 --test 256 \\
        --weights weights.npz --out results.json

Draws Bragg instances (diamond (400), 17 keV, dx 0.5 um, 128-pixel field,
64 y-planes, 0-12 <110> lines, s_dev uniform in +-60 urad) with thickness
uniform over 40 values spread over 20-100 um for training and over 8
values for the held-out set (instances of one thickness share a grid and
are solved in one batched GMRES).  The training instances are cold-solved
on the device (one Krylov space per plane, restart 40, tol 1e-10) while
the network's input features are written; the converged ones (true
residual <= 1e-7; beyond about 50 um many strained instances do not
converge at all, cold or warm, because the shooting operator's
conditioning grows like exp(t / extinction depth)) are the training
targets, 512 of them held back to measure the predicted field's error.
The network is trained with the device kernels, and the held-out
instances are solved cold, warm from the perfect-crystal value and warm
from the network, by thickness band, to tol 1e-10: with one Krylov space
per plane over the whole range, and with the reference's policy (joint
32-plane chunks, retries, splits) up to 50 um (beyond, the joint chunks
take thousands of iterations and tens of seconds per instance, cold and
warm alike).  Progress goes to stdout with elapsed seconds; the report is
one JSON object.
"""

import argparse
import json
import sys
import time

import numpy as np

from ..backend import device
from ..crystal import DIAMOND, CrystalParams
from . import data


def run(n_train, n_test, *, epochs=16, batch=32, lr=1e-3,
        t_range=(20e-6, 100e-6), seed=20260909, n_val=512, log=print,
        weights=None, joint_max_t=50e-6):
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    rng = np.random.default_rng(seed)
    t_train = data.thickness_grid(xtal, t_range, 40)
    t_test = data.thickness_grid(xtal, t_range, 8)
    inst = (data.draw_instances(n_train, rng, xtal, t_values=t_train) +
            data.draw_instances(n_test, rng, xtal, t_values=t_test))
    t0 = time.perf_counter()
    ds = data.build_dataset(xtal, inst, log=log, n_solve=n_train)
    t_data = time.perf_counter() - t0
    conv = np.nonzero(ds.resid[:n_train] <= 1e-7)[0]
    t_um = np.array([x["t"] for x in inst[:n_train]]) * 1e6
    by_band = {}
    for lo, hi in data.BANDS_UM:
        m = data.in_band(t_um, lo, hi)
        by_band[f"{lo}-{hi}um"] = {
            "n": int(m.sum()),
            "converged_fraction": float((ds.resid[:n_train][m] <= 1e-7).mean()),
            "iterations_per_plane_mean": float(ds.iters[:n_train][m].mean()
                                               / data.NPLANES)}
    log(f"dataset: {n_train} cold solves + {n_test} feature sets in "
        f"{t_data:.1f} s; converged {len(conv)} / {n_train}; by band "
        f"{by_band}")
    data.rescale(ds, conv)
    val = conv[-n_val:] if len(conv) > 4 * n_val else conv[:0]
    fit = conv[:len(conv) - len(val)]
    t0 = time.perf_counter()
    net, losses = data.train(ds, 0, batch=batch, epochs=epochs, lr=lr,
                             seed=seed, log=log, idx=fit)
    t_fit = time.perf_counter() - t0
    log(f"trained {len(fit) // batch * batch} x {epochs} epochs in "
        f"{t_fit:.1f} s; loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    if weights:
        np.savez(weights, params=net.params(), scale=ds.scale,
                 t_range_um=np.array(t_range) * 1e6, epochs=epochs,
                 n_fit=len(fit), seed=seed)
        log(f"weights written to {weights}")
    rel_err = None
    if len(val):
        pred = data.predict(net, ds, list(val))
        tgt = ds.Y[val, ..., 0] + 1j * ds.Y[val, ..., 1]
        rel_err = float(np.sqrt(np.mean(np.abs(pred - tgt) ** 2)
                                / np.mean(np.abs(tgt) ** 2)))
        log(f"validation relative rms error of the predicted field: "
            f"{rel_err:.3f}")
    held = list(range(n_train, n_train + n_test))
    t_held = np.array([inst[i]["t"] for i in held])
    thin = [i for i, t in zip(held, t_held) if t <= joint_max_t]
    rep = {}
    t0 = time.perf_counter()
    if held:
        rep["independent_planes"] = data.evaluate_joint(
            xtal, inst, ds, net, held, log=log, planes="independent")
    log(f"evaluation, independent planes, in {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    if thin:
        rep["joint_chunks"] = data.evaluate_joint(xtal, inst, ds, net, thin,
                                                  log=log, planes="joint")
        rep["joint_chunks"]["thickness_limit_um"] = joint_max_t * 1e6
    log(f"evaluation, joint chunks, in {time.perf_counter() - t0:.1f} s")
    rep.update({"n_train_solved": n_train, "n_fit": len(fit),
                "n_test": n_test, "epochs": epochs,
                "thickness_um": [t_range[0] * 1e6, t_range[1] * 1e6],
                "train_thicknesses_um": (t_train * 1e6).round(3).tolist(),
                "test_thicknesses_um": (t_test * 1e6).round(3).tolist(),
                "cold_dataset_by_band": by_band,
                "dataset_seconds": t_data, "train_seconds": t_fit,
                "train_losses": losses,
                "validation_field_rel_rms_error": rel_err,
                "params": net.n_params})
    return rep


def evaluate_saved(weights, n_test, *, t_range=(20e-6, 100e-6),
                   seed=20260909, log=print, joint_max_t=50e-6, batch=32):
    """The held-out evaluation of `run` with a network trained earlier
    (parameters and target scale from `weights`); held-out instances are
    drawn from a generator seeded apart from the training draw."""
    from .net import WarmStartNet
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    w = np.load(weights)
    net = WarmStartNet(batch, params=w["params"])
    rng = np.random.default_rng(seed + 1)
    t_test = data.thickness_grid(xtal, t_range, 8)
    inst = data.draw_instances(n_test, rng, xtal, t_values=t_test)
    ds = data.build_dataset(xtal, inst, n_solve=0)
    ds.scale = float(w["scale"])
    held = list(range(n_test))
    thin = [i for i in held if inst[i]["t"] <= joint_max_t]
    t0 = time.perf_counter()
    rep = {"independent_planes": data.evaluate_joint(
        xtal, inst, ds, net, held, log=log, planes="independent")}
    log(f"evaluation, independent planes, in {time.perf_counter() - t0:.1f} s")
    if thin:
        t0 = time.perf_counter()
        rep["joint_chunks"] = data.evaluate_joint(xtal, inst, ds, net, thin,
                                                  log=log, planes="joint")
        rep["joint_chunks"]["thickness_limit_um"] = joint_max_t * 1e6
        log(f"evaluation, joint chunks, in {time.perf_counter() - t0:.1f} s")
    rep.update({"n_test": n_test, "weights": str(weights),
                "test_thicknesses_um": (t_test * 1e6).round(3).tolist()})
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--tmin", type=float, default=20.0, help="um")
    ap.add_argument("--tmax", type=float, default=100.0, help="um")
    ap.add_argument("--weights", default=None,
                    help="write the trained parameters here (.npz)")
    ap.add_argument("--eval-only", action="store_true",
                    help="evaluate the network in --weights on --test "
                         "held-out instances (no training)")
    ap.add_argument("--joint-max-um", type=float, default=50.0,
                    help="evaluate the reference's joint-chunk policy on "
                         "held-out instances up to this thickness")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()

    def log(m):
        print(f"{time.perf_counter() - t0:8.1f}s {m}", flush=True)

    if a.eval_only:
        rep = evaluate_saved(a.weights, a.test,
                             t_range=(a.tmin * 1e-6, a.tmax * 1e-6), log=log,
                             joint_max_t=a.joint_max_um * 1e-6)
    else:
        rep = run(a.train, a.test, epochs=a.epochs,
                  t_range=(a.tmin * 1e-6, a.tmax * 1e-6), log=log,
                  weights=a.weights, joint_max_t=a.joint_max_um * 1e-6)
    line = json.dumps(rep)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
