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


@pytest.mark.parametrize("opts", [{"guard_batch": 16}, {"graph_steps": 8, "guard_batch": 64}])
def test_bonded_clusters_deferred_steps(tmp_path, opts):
    """Bonded steps are deferred like any other: the step that breaks an
    interface freezes the steps after it on the device, the host accounts
    the broken interfaces and rebuilds when it reads the
    guard at the end of the batch.  Bit-identical to the oracle, with the
    same surviving interfaces, and the host reads the device once per batch,
    not once per step."""
    from oracle.rk_oracle import Oracle
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    shapes, deck = D.bonded_clusters_deck(n_clusters=4, cluster_dims=(3, 3, 3), clearance=1e-4, strength_mean=20.0)
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = Engine(path), Oracle(path)
    for k, v in opts.items():
        e.set_option(k, v)
    e.update_nl()
    o.update_nl()
    for n in (800, 700, 500):
        e.step(n)
        o.run_steps(n)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
        c = e.counts()
        assert c["ninterfaces"] == o.ninterfaces
        assert c["total_broken_area"] == o.total_broken_area
        assert c["nl_updates"] == o.nl_updates
    assert o.ninterfaces < 200
    c = e.counts()
    print("guard reads %d, wasted steps %d, graph launches %d, rebuilds %d" %
          (c["guard_reads"], c["guard_wasted"], c["graph_launches"], c["nl_updates"]))
    assert c["guard_reads"] < 2000 // 4
    if "graph_steps" in opts:
        assert c["graph_launches"] > 0


def test_bonded_clusters_integrate_outputs(tmp_path):
    """integrate() on the bonded clusters with conf files and the
    TrackDamage extractor: the files equal the oracle's."""
    from rockable_amd import decks as D
    from tests.test_gpu_outputs import _both, _same_files
    shapes, deck = D.bonded_clusters_deck(n_clusters=4, cluster_dims=(3, 3, 3), clearance=1e-4, strength_mean=20.0)
    path = D.write_deck(str(tmp_path), shapes, deck)
    path = D.conf_with(path, str(tmp_path / "deck2.conf"), interConf=250e-7, tmax=1500e-7)
    res = _both(path, tmp_path, ["TrackDamage damage.txt 10", "dnStat dn.txt 25"], {"guard_batch": 32})
    _same_files(res, ["damage.txt", "dn.txt", "out/kineticEnergy.txt", "out/staticBalance.txt", "out/conf6"])


@pytest.mark.parametrize("batch", [1, 16])
def test_broken_bond_active_set_step_by_step(tmp_path, batch):
    """Step by step through a bond's breakage (one step per call, then
    batches of 16 steps inside one call): the active set equals the
    oracle's after every call, the broken bond staying in it after the
    breakage's rebuild and leaving it at the next force phase."""
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    from tests import kat_decks as K
    path = K.spun_bond(str(tmp_path), w=0.4, kt=2.0e4, damp=0.035, ft0=1.0, mom0=1.0, power=2.0)
    e, o = Engine(path), Oracle(path)
    e.update_nl()
    o.update_nl()
    seen = False
    for _ in range(200 // batch):
        e.step(batch)
        o.run_steps(batch)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
        c = e.counts()
        assert (c["nactive"], c["ninterfaces"], c["nl_updates"]) == (o.nactive, o.ninterfaces, o.nl_updates)
        seen = seen or (o.ninterfaces == 0 and o.nactive == 1)
    assert o.ninterfaces == 0 and (seen or batch > 1)
