"""A small batched deck (independent packings of uneven sizes, each with its
own Default-law parameters, walls first) and the map from its rows to the
rows of one packing run as a deck of its own.  Shared by the oracle test
(test_batched.py) and the GPU test (test_gpu_batched.py)."""
import os

import numpy as np

from rockable_amd import decks as D

STEPS = 3000


def small_batch(dirpath, n_packings=4, seed=5):
    """Four packings of 100-400 spheres on a lattice of spacing 2.05 R,
    random speeds up to 5 cm/s, NL rebuilt every 100 steps."""
    shapes, deck, info = D.batched_sphere_packings(n_packings=n_packings, n_range=(100, 400), seed=seed,
                                                   vmax=0.05, steps=STEPS, spacing_factor=2.05)
    return D.write_deck(dirpath, shapes, deck), shapes, deck, info


def single(dirpath, shapes, deck, info, p):
    s, d = D.single_packing_of(shapes, deck, info, p)
    sub = os.path.join(dirpath, "packing%d" % p)
    os.makedirs(sub, exist_ok=True)
    return D.write_deck(sub, s, d)


def rows_of(info, p):
    """Rows of packing p in the batched deck, in the order of its own deck
    (its 5 walls, then its spheres)."""
    f, n = info[p]["first"], info[p]["bodies"]
    return np.concatenate([np.arange(5 * p, 5 * p + 5), np.arange(f, f + n)])


def interactions_of(I, info, p):
    """Entries of packing p with body indices renumbered as in its own deck
    (the map is increasing, so the entry order is kept)."""
    rows = rows_of(info, p)
    local = np.full(rows.max() + 1, -1, dtype=np.int64)
    local[rows] = np.arange(len(rows))
    sel = np.isin(I["i"], rows)
    J = I[sel].copy()
    J["i"] = local[J["i"]]
    J["j"] = local[J["j"]]
    return J
