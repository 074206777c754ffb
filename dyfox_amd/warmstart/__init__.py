"""Learned warm start for the Bragg shooting GMRES.

    from dyfox_amd.warmstart import data, net
    instances = data.draw_instances(n, rng, xtal)
    ds = data.build_dataset(xtal, instances)      # cold solves + features
    model, losses = data.train(ds, n_train)
    report = data.evaluate(xtal, instances, ds, model, held_out)

The plain solver is both the fallback and the arbiter: every warm solve
still runs GMRES to the same true residual.
"""
