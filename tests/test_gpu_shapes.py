"""Launch shapes and the autotuner: every choice of fuse_law, law_group,
geom_block, gather_lanes, gather_groups, fuse_glob, gather_waves,
body_block, persistent, graph_steps, type_sort and narrow_pairs (candidate pairs per
narrow-phase wavefront; the default for these hulls is 1) computes the
same state bit for bit (the default engine, against which they are compared, fuses the
geometry and law kernels; the first three shapes use them apart), and the
steps the autotuner takes while timing candidates are the steps rk_step
takes."""
import pytest

from tests.test_gpu_parity import _dynamic_hull_deck, assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu

FIELDS = ("dn", "fn", "ft", "mom", "damp", "stick")

SHAPES = [dict(fuse_law=0, law_group=1, geom_block=64, gather_waves=1, body_block=64),
          dict(fuse_law=0, law_group=8, geom_block=512, gather_waves=2, body_block=128),
          dict(fuse_law=0, law_group=2, geom_block=128, gather_waves=4, body_block=256, graph_steps=8),
          dict(fuse_law=1, gather_waves=2, body_block=64, graph_steps=4),
          dict(fuse_law=1, gather_lanes=8, gather_waves=1),
          dict(fuse_law=1, persistent=8, gather_lanes=8),
          dict(fuse_law=0, persistent=16, law_group=2, geom_block=128, graph_steps=4),
          dict(gather_lanes=16, gather_groups=16, fuse_glob=1, gather_waves=1),
          dict(gather_lanes=8, gather_groups=4, fuse_glob=1, persistent=8, graph_steps=4),
          dict(gather_lanes=8, gather_groups=16, fuse_glob=0, gather_waves=2),
          dict(narrow_pairs=2),
          dict(narrow_pairs=2, gather_lanes=8),
          dict(fuse_glob=2, type_sort=0),
          dict(fuse_glob=2, persistent=8, gather_groups=4, graph_steps=4),
          dict(fuse_glob=3),
          dict(fuse_glob=3, persistent=8, gather_lanes=8, graph_steps=4)]


def _hull_deck(tmp_path):
    from rockable_amd import decks as D
    shapes, deck = D.hull_packing_deck(n_bodies=600, lattice=(10, 10, 6), law="Avalanche", vmax=10.0,
                                       gap_factor=1.01, interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    return D.write_deck(str(tmp_path), shapes, deck)


@pytest.mark.parametrize("k", range(len(SHAPES)))
def test_launch_shapes_bit_exact(tmp_path, k):
    from rockable_amd.engine import Engine
    path = _hull_deck(tmp_path)
    a, b = Engine(path), Engine(path)
    for name, v in SHAPES[k].items():
        b.set_option(name, v)
        assert b.get_option(name) == v
    for e in (a, b):
        e.update_nl()
    for n in (300, 700):
        a.step(n)
        b.step(n)
        assert_bodies_equal(b.bodies(), a.bodies())
        assert_interactions_equal(b.interactions(), a.interactions(), fields=FIELDS)


def test_bad_launch_shape_rejected(tmp_path):
    from rockable_amd.engine import Engine
    e = Engine(_hull_deck(tmp_path))
    for name, v in (("geom_block", 96), ("law_group", 3), ("gather_waves", 8), ("body_block", 512),
                    ("gather_lanes", 4), ("fuse_law", 2), ("model_scale", 0.0), ("persistent", 7),
                    ("gather_groups", 3), ("fuse_glob", 4), ("narrow_pairs", 4), ("type_sort", 2)):
        with pytest.raises(RuntimeError):
            e.set_option(name, v)


@pytest.mark.parametrize("dynamic", [False, True])
def test_autotune_steps_are_rk_steps(tmp_path, dynamic):
    from rockable_amd.engine import Engine
    path = _dynamic_hull_deck(tmp_path) if dynamic else _hull_deck(tmp_path)
    a, b = Engine(path), Engine(path)
    a.update_nl()
    b.update_nl()
    n, choice = b.autotune(steps_per_trial=20)
    assert n > 20 * 19
    assert choice["law_group"] in (1, 2, 4, 8) and choice["geom_block"] in (64, 128, 256, 512)
    a.step(n)
    assert_bodies_equal(b.bodies(), a.bodies())
    assert_interactions_equal(b.interactions(), a.interactions(), fields=FIELDS)
    assert b.counts()["nl_updates"] == a.counts()["nl_updates"]
    a.step(200)
    b.step(200)
    assert_bodies_equal(b.bodies(), a.bodies())
