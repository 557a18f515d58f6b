"""The far-entry skip (skip_far, active when the step writes no geometry of
inactive entries, record_geometry 0): an inactive entry with zero history
whose gap is proven positive by the bodies' motion bounds is not evaluated.
Everything stays bit-identical to the oracle, which evaluates every listed
entry every step: bodies, active interactions, the number and steps of the
rebuilds, and the gap of every listed entry when the list is read out (the
skipped entries' gaps are recomputed then)."""
import numpy as np
import pytest

from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu


def _pair(path, steps, chunk, skip=1, **opts):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    e = Engine(path)
    e.set_option("record_geometry", 0)
    e.set_option("skip_far", skip)
    e.set_option("exact_sums", 1)  # the walls' resultants in serial order too
    for k, v in opts.items():
        e.set_option(k, v)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        e.step(n)
        o.run_steps(n)
        done += n
        print("checkpoint after %d steps, %d rebuilds" % (done, o.nl_updates), flush=True)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        assert_interactions_equal(e.interactions(), o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        assert e.counts()["nl_updates"] == o.nl_updates
    return e, o


@pytest.mark.parametrize("law", ["Default", "Avalanche"])
def test_skip_hull_packing_bit_exact(tmp_path, law):
    from rockable_amd import decks as D
    shapes, deck = D.hull_packing_deck(n_bodies=600, lattice=(10, 10, 6), law=law, vmax=10.0, gap_factor=1.01,
                                       interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = _pair(path, 1000, 250)
    assert o.nactive >= 3


def test_skip_sphere_bed_bit_exact(tmp_path):
    from tests.test_gpu_parity import _sphere_bed
    e, o = _pair(_sphere_bed(tmp_path), 800, 200)
    assert o.nactive > 1000


def test_skip_dynamic_guard_bit_exact(tmp_path):
    from tests.test_gpu_parity import _dynamic_hull_deck
    e, o = _pair(_dynamic_hull_deck(tmp_path), 1000, 250, guard_batch=16)
    assert o.nl_updates > 10


def test_skip_slope_bed_with_network_and_graphs(tmp_path):
    """The settled bed released on the slope, with the rebuild network and
    captured graphs of 8 steps: the skip changes nothing."""
    import os
    from rockable_amd.engine import Mlp
    from tests.test_gpu_slope import ROOT, settled_slope_deck
    path = settled_slope_deck(str(tmp_path))
    mlp = Mlp(device=0)
    mlp.set_params(np.load(os.path.join(ROOT, "rockable_amd", "data", "rebuild_mlp.npy")))
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    e = Engine(path)
    e.set_option("record_geometry", 0)
    e.set_option("skip_far", 1)
    e.set_option("exact_sums", 1)
    e.set_option("graph_steps", 8)
    e.set_rebuild_model(mlp)
    o = Oracle(path, threads=8)
    e.update_nl()
    o.update_nl()
    for _ in range(4):
        e.step(500)
        o.run_steps(500)
        assert_bodies_equal(e.bodies(), o.bodies())
        # (no geometry records: n, pos and vel are written only when an output is due)
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        assert_interactions_equal(e.interactions(), o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        assert e.counts()["nl_updates"] == o.nl_updates
    e.close()
    mlp.close()


def test_skip_set_bodies_resets_bounds(tmp_path):
    """Poses set from outside carry no motion bound: the next step evaluates
    every entry (same state as without the skip)."""
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    shapes, deck = D.hull_packing_deck(n_bodies=300, lattice=(10, 10, 3), vmax=5.0, gap_factor=1.01,
                                       interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    path = D.write_deck(str(tmp_path), shapes, deck)
    runs = []
    for skip in (1, 0):
        e = Engine(path)
        e.set_option("record_geometry", 0)
        e.set_option("skip_far", skip)
        e.update_nl()
        e.step(100)
        b = e.bodies()
        pos = b["pos"].copy()
        pos[5:] *= 0.97  # squeeze the packing: new contacts appear at once
        e.set_bodies(pos=pos)
        e.step(10)
        runs.append((e.bodies(), e.interactions()))
        e.close()
    assert_bodies_equal(runs[0][0], runs[1][0])
    assert_interactions_equal(runs[0][1], runs[1][1], fields=("dn", "fn", "ft", "mom", "damp", "stick"))
