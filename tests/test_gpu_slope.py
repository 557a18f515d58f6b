"""The learned-rebuild configuration at a size the oracle covers: a bed of
3 000 hull polyhedra (the plain GPU form's shapes) settled on the GPU under
vertical gravity, then released on a 30 deg slope at 1 m/s downslope
(ForceLaw Avalanche, dt 2e-7 s), the neighbour list rebuilt by the
displacement check alone (dynamicUpdateNL, dispUpdateNL = DVerlet / 4,
angleUpdateNL 1 rad).  The engine runs it with the rebuild-interval network
sizing its batches; the oracle runs the plain check after every step.  The
two are bit-identical at every checkpoint (bodies, active interactions) and
rebuild at the same steps."""
import os

import numpy as np
import pytest

from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def settled_slope_deck(d, n=3000, tiles=1, seed=5):
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    g = 9.81
    shapes, deck = D.compact_pour_deck(n_bodies=n, seed=seed, box_in_sizes=(14, 14, 200), tilt_deg=3.0,
                                       gap_frac=0.0, vmax=0.5, law="Avalanche", dt=2e-7, interverlet_steps=20)
    deck.gravity = (0.0, 0.0, -400.0 * g)
    loose = D.write_deck(d, shapes, deck, name="loose.conf")
    e = Engine(loose)
    e.set_option("record_geometry", 0)
    e.update_nl()
    for mult, steps in ((400.0, 15000), (100.0, 3000), (25.0, 3000), (6.25, 3000), (1.0, 3000)):
        e.set_option("gravity_z", -mult * g)
        e.step(steps)
    settled = os.path.join(d, "settled.conf")
    e.save_conf(settled)
    e.close()
    pool = [s for s in shapes if s.name != "WALL"]
    return D.slope_from_settled(settled, pool, os.path.join(d, "slope"), tiles=tiles, slope_deg=30.0,
                                flow_speed=1.0, runout=0.2)


def test_settled_bed_on_slope_with_network_bit_exact(tmp_path):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine, Mlp
    path = settled_slope_deck(str(tmp_path))
    mlp = Mlp(device=0)
    mlp.set_params(np.load(os.path.join(ROOT, "rockable_amd", "data", "rebuild_mlp.npy")))
    e = Engine(path)
    e.set_option("exact_sums", 1)  # the base's resultant in serial order too
    e.set_rebuild_model(mlp)
    o = Oracle(path, threads=8)
    e.update_nl()
    o.update_nl()
    st = None
    for _ in range(4):
        e.step(500)
        o.run_steps(500)
        if st is None:
            st = e.stats()
            print("released bed after 500 steps: %d bodies, coordination %.2f, %d contacts" %
                  (st["nfree"], st["coordination"], st["ncontacts"]), flush=True)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
        assert e.counts()["nl_updates"] == o.nl_updates
    c = e.counts()
    print("2000 steps: %d rebuilds, %d guard reads, %d polls" % (c["nl_updates"], c["guard_reads"],
                                                                  c["guard_polls"]))
    assert o.nl_updates >= 5
    assert c["guard_reads"] < 2000 // 4
    assert st["ncontacts"] > 1000   # a flowing bed, not a dilute spray
    e.close()
    mlp.close()
