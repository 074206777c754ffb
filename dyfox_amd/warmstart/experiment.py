"""The learned warm-start experiment end to end, on one device.

    python -m dyfox_amd.warmstart.experiment --train 20000 --test 2000 \\
        --out results.json

Draws train + test Bragg instances (diamond (400), 17 keV, dx 0.5 um,
128-pixel field, 64 y-planes, 0-12 <110> lines, thickness 20-35 um: above
about 35 um some strained planes no longer reach 1e-10 at restart 40 and
above about 50 um the shooting solve diverges outright), cold-solves all of them
on the device to 1e-10 while writing the network's input features,
trains the network with the device kernels, and evaluates the held-out
instances cold, warm from the perfect-crystal value, warm from the
network, and warm from the exact answer perturbed at relative accuracy
1e-2 and 1e-3 (the gain available to any predictor that accurate).
Progress goes to stdout with elapsed seconds; the report is one JSON
object.
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
        t_range=(20e-6, 35e-6), seed=20260909, ceiling=(1e-2, 1e-3),
        log=print):
    device()
    xtal = CrystalParams.from_reflection(DIAMOND, (4, 0, 0), 17.0,
                                         f_prime=0.003, f_dprime=0.002)
    rng = np.random.default_rng(seed)
    n = n_train + n_test
    inst = data.draw_instances(n, rng, xtal, t_range=t_range)
    t0 = time.perf_counter()
    ds = data.build_dataset(xtal, inst, log=log)
    t_data = time.perf_counter() - t0
    log(f"dataset: {n} cold solves in {t_data:.1f} s, plane-iterations "
        f"mean {ds.iters.mean() / data.NPLANES:.1f}, worst residual "
        f"{ds.resid.max():.1e}")
    n_fit = n_train // batch * batch
    t0 = time.perf_counter()
    net, losses = data.train(ds, n_fit, batch=batch, epochs=epochs, lr=lr,
                             seed=seed, log=log)
    t_train = time.perf_counter() - t0
    log(f"trained {n_fit} x {epochs} epochs in {t_train:.1f} s; loss "
        f"{losses[0]:.4e} -> {losses[-1]:.4e}")
    held = list(range(n_train, n))
    pred = data.predict(net, ds, held)
    tgt = ds.Y[held, ..., 0] + 1j * ds.Y[held, ..., 1]
    rel_err = float(np.sqrt(np.mean(np.abs(pred - tgt) ** 2)
                            / np.mean(np.abs(tgt) ** 2)))
    log(f"held-out relative rms error of the predicted field: {rel_err:.3f}")
    t0 = time.perf_counter()
    rep = data.evaluate(xtal, inst, ds, net, held, log=log,
                        ceiling=ceiling)
    log(f"evaluation in {time.perf_counter() - t0:.1f} s")
    rep.update({"n_train": n_fit, "n_test": n_test, "epochs": epochs,
                "thickness_um": [t_range[0] * 1e6, t_range[1] * 1e6],
                "dataset_seconds": t_data, "train_seconds": t_train,
                "train_losses": losses,
                "heldout_field_rel_rms_error": rel_err,
                "params": net.n_params})
    return rep


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--train", type=int, default=20000)
    ap.add_argument("--test", type=int, default=2000)
    ap.add_argument("--epochs", type=int, default=16)
    ap.add_argument("--tmin", type=float, default=20.0, help="um")
    ap.add_argument("--tmax", type=float, default=35.0, help="um")
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    t0 = time.perf_counter()

    def log(m):
        print(f"{time.perf_counter() - t0:8.1f}s {m}", flush=True)

    rep = run(a.train, a.test, epochs=a.epochs,
              t_range=(a.tmin * 1e-6, a.tmax * 1e-6), log=log)
    line = json.dumps(rep)
    if a.out:
        with open(a.out, "w") as f:
            f.write(line + "\n")
    print(line, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
