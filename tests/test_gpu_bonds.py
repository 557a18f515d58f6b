"""Breakable bonds at the shape of the bonded-clusters workload: clusters of
sphero-cubes bonded face to face at their coincident vertices (one
StickedLinks interface per touching pair, lognormal strengths kept in the
interfaces) fall onto a rigid plane inclined at 35 deg at 2 m/s.  The impact
breaks part of the interfaces; the GPU follows the oracle bit for bit
through it: bodies, every listed interaction and the surviving interfaces."""
import pytest

from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nl", ["linkCells", "bruteForce"])
def test_bonded_clusters_impact_bit_exact(tmp_path, nl):
    from oracle.rk_oracle import Oracle
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    shapes, deck = D.bonded_clusters_deck(n_clusters=4, cluster_dims=(3, 3, 3), clearance=1e-4,
                                          strength_mean=20.0, nl=nl)
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = Engine(path), Oracle(path)
    assert e.counts()["ninterfaces"] == o.ninterfaces == 216
    e.update_nl()
    o.update_nl()
    for n in (800, 700, 500):
        e.step(n)
        o.run_steps(n)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
        assert e.counts()["ninterfaces"] == o.ninterfaces
    assert_interactions_equal(e.interactions(), o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    assert o.ninterfaces < 200   # the impact broke bonds
    assert e.counts()["nl_updates"] == o.nl_updates
