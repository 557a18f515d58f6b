"""Training data and training of the rebuild-interval network.

With dynamicUpdateNL the engine checks the displacement guard on the device
and reads it once per batch of iterations; the network proposes the batch
size.  A rebuild history is the run of iterations between two rebuilds; at
every iteration s the sample is (the 16 guard features before s,
log(1 + the number of iterations from s to the one after which the next
rebuild falls due)).  The plain method (one guard read per iteration) makes
the histories.

    python -m rockable_amd.rebuild_model --out rockable_amd/data/rebuild_mlp.npy
"""
import argparse
import os
import tempfile

import numpy as np

from rockable_amd import decks as D
from rockable_amd.engine import Engine, Mlp


def collect(engine, max_steps, max_rebuilds):
    """Features before each iteration and the iterations to the next rebuild
    (samples after the last rebuild have no target and are dropped)."""
    engine.set_rebuild_model(None)
    engine.set_option("guard_batch", 1)
    feats, rebuild_at = [], []
    n0 = engine.counts()["nl_updates"]
    for s in range(max_steps):
        feats.append(engine.guard_features())
        engine.step(1)
        n1 = engine.counts()["nl_updates"]
        if n1 > n0:
            rebuild_at.append(s)
            n0 = n1
            if len(rebuild_at) >= max_rebuilds:
                break
    X, t = [], []
    r = 0
    for s in range(len(feats)):
        while r < len(rebuild_at) and rebuild_at[r] < s:
            r += 1
        if r == len(rebuild_at):
            break
        X.append(feats[s])
        t.append(np.log1p(rebuild_at[r] - s + 1))
    return np.array(X, dtype=np.float32).reshape(-1, 16), np.array(t, dtype=np.float32), len(rebuild_at)


def hull_deck(dirpath, seed, disp, angle_deg, vmax=3.0, n_bodies=600, law="Avalanche"):
    """A C2-style deck (hull polyhedra thrown at each other) with dynamicUpdateNL."""
    shapes, deck = D.hull_packing_deck(n_bodies=n_bodies, seed=seed, lattice=(10, 10, n_bodies // 100), law=law,
                                       vmax=vmax, gap_factor=1.01, interverlet_steps=10 ** 6,
                                       size_range=(5.8e-3, 6e-3))
    deck.dynamic_update_nl = 1
    deck.disp_update_nl = disp
    deck.angle_update_nl = angle_deg
    return D.write_deck(dirpath, shapes, deck)


def sphere_deck(dirpath, seed, disp, vmax=0.2):
    """A C1-style deck (512 spheres thrown onto a floor) with dynamicUpdateNL."""
    shapes, deck = D.sphere_lattice_deck(n_side=8, seed=seed, vmax=vmax, spacing_factor=2.02, dt=2e-7)
    deck.gravity = (0.0, 0.0, -98.1)
    deck.interVerlet = 1e6
    deck.dynamic_update_nl = 1
    deck.disp_update_nl = disp
    deck.angle_update_nl = 90.0
    return D.write_deck(dirpath, shapes, deck)


def training_set(workdir, histories=500, device=0, max_steps=20000):
    """Rebuild histories from C1- and C2-style runs with several guard
    thresholds and seeds, run after run until `histories` rebuilds have been
    seen."""
    runs = [("hull", 11, 2e-4, 4.0), ("hull", 12, 3e-4, 5.0), ("hull", 14, 5e-4, 8.0),
            ("sphere", 21, 2e-5, None), ("sphere", 22, 5e-5, None), ("hull", 15, 1e-4, 3.0)]
    runs += [("hull", 16 + k, 1e-4 * (1 + k % 4), 3.0 + 2 * (k % 3)) for k in range(12)]
    runs += [("sphere", 30 + k, 1e-5 * (1 + k % 5), None) for k in range(12)]
    Xs, ts, seen = [], [], 0
    for k, (kind, seed, disp, ang) in enumerate(runs):
        if seen >= histories:
            break
        d = os.path.join(workdir, "run%d" % k)
        os.makedirs(d, exist_ok=True)
        path = hull_deck(d, seed, disp, ang) if kind == "hull" else sphere_deck(d, seed, disp)
        e = Engine(path, device=device)
        e.set_option("record_geometry", 0)
        e.update_nl()
        X, t, n = collect(e, max_steps, min(100, histories - seen))
        e.close()
        Xs.append(X)
        ts.append(t)
        seen += n
    return np.concatenate(Xs), np.concatenate(ts), seen


def train(X, t, device=0, seed=1, epochs=300, batch=256, lr=2e-3):
    mlp = Mlp(device=device, seed=seed)
    loss = mlp.train(X, t, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return mlp, loss


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "data",
                                                   "rebuild_mlp.npy"))
    ap.add_argument("--histories", type=int, default=500)
    ap.add_argument("--epochs", type=int, default=300)
    a = ap.parse_args()
    work = tempfile.mkdtemp(prefix="rk_rebuild_")
    X, t, seen = training_set(work, a.histories)
    mlp, loss = train(X, t, epochs=a.epochs)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, mlp.params())
    base = float(np.mean((t - t.mean()) ** 2))
    print("rebuild histories %d, samples %d, loss %.4g (variance of the targets %.4g) -> %s" % (
        seen, len(t), loss, base, a.out), flush=True)


if __name__ == "__main__":
    main()
