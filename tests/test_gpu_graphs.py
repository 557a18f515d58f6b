"""Steps replayed from captured HIP graphs ("graph_steps") give the same
state, bit for bit, as steps launched one kernel at a time: with the
periodic rebuild, with the device displacement check (fixed batches and
network-sized ones) and with conf files written mid-run."""
import glob
import os

import numpy as np
import pytest

from tests import batched_case as BC
from tests.test_gpu_parity import _dynamic_hull_deck, assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu

FIELDS = ("dn", "fn", "ft", "mom", "damp", "stick")


def _pair(path, chunks, graph_steps, setup=None):
    from rockable_amd.engine import Engine
    a, b = Engine(path), Engine(path)
    b.set_option("graph_steps", graph_steps)
    for e in (a, b):
        if setup:
            setup(e)
        e.update_nl()
    for n in chunks:
        a.step(n)
        b.step(n)
        assert_bodies_equal(b.bodies(), a.bodies())
        assert_interactions_equal(b.interactions(), a.interactions(), fields=FIELDS)
        assert b.counts()["nl_updates"] == a.counts()["nl_updates"]
    return a, b


@pytest.mark.parametrize("graph_steps", [4, 16])
def test_graph_steps_periodic_rebuild(tmp_path, graph_steps):
    """The small batched deck (NL rebuilt every 100 steps): runs of steps as
    graphs, broken by the rebuilds and by the ends of rk_step calls."""
    path = BC.small_batch(str(tmp_path))[0]
    a, b = _pair(path, (250, 37, 713), graph_steps)
    c = b.counts()
    assert c["graph_launches"] > 0 and c["graph_captures"] >= 1
    assert c["graph_captures"] < c["graph_launches"]


@pytest.mark.parametrize("batch,graph_steps", [(7, 4), (64, 8)])
def test_graph_steps_device_guard(tmp_path, batch, graph_steps):
    """dynamicUpdateNL: batches of guarded steps as graphs (and persistent
    kernels); the device keeps the step index, so a rebuild found inside a
    graph freezes the later steps and the host resumes from the right
    iteration."""
    path = _dynamic_hull_deck(tmp_path)

    def setup(e):
        e.set_option("guard_batch", batch)
        e.set_option("persistent", 16)
    a, b = _pair(path, (150, 350, 500), graph_steps, setup=setup)
    assert a.counts()["nl_updates"] > 10
    assert b.counts()["graph_launches"] > 0


def test_graph_steps_with_rebuild_network(tmp_path):
    from rockable_amd.engine import Mlp
    path = _dynamic_hull_deck(tmp_path)
    W = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rockable_amd", "data",
                     "rebuild_mlp.npy")
    mlp = Mlp(seed=1)
    mlp.set_params(np.load(W))
    a, b = _pair(path, (400, 600), 8, setup=lambda e: e.set_rebuild_model(mlp))
    assert b.counts()["graph_launches"] > 0


def test_graph_steps_integrate_conf_files(tmp_path):
    """integrate() writing conf files every 100 steps with graphs of 16
    steps: the files equal those of the plain launches line for line."""
    from oracle.rk_oracle import compare_conf
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    path = _dynamic_hull_deck(tmp_path)
    path = D.conf_with(path, str(tmp_path / "deck2.conf"), interConf=100 * 2e-7, tmax=700 * 2e-7)
    outs = []
    for g in (0, 16):
        d = tmp_path / ("g%d" % g)
        d.mkdir()
        e = Engine(path)
        e.set_option("graph_steps", g)
        e.update_nl()
        e.integrate(str(d))
        outs.append(d)
    files = sorted(glob.glob(str(outs[0] / "conf*")))
    assert len(files) >= 7
    for f in files:
        ok, k, x, y = compare_conf(str(outs[1] / os.path.basename(f)), f)
        assert ok, "%s differs at line %s" % (os.path.basename(f), k)
