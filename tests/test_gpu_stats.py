"""Device statistics (rk_get_stats) against the same quantities computed
with numpy from the oracle's state, and a long run checked statistically:
kinetic energy, coordination number and broken-bond count of the GPU run
within 1 % of the oracle's (the runs are bit-identical, so the figures agree
to rounding of the reductions)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def numpy_stats(o, ndriven):
    b = o.bodies()
    m, I = o.mass()
    free = slice(ndriven, None)
    v, w, Q = b["vel"][free], b["vrot"][free], b["Q"][free]
    qw, qv = Q[:, 0], Q[:, 1:]
    # body-frame angular velocity: conj(Q) w Q
    t = 2.0 * np.cross(-qv, w)
    wb = w + qw[:, None] * t + np.cross(-qv, t)
    I_, m_ = I[free], m[free]
    A = o.interactions(active=True)
    contacts = A[A["dn"] < 0.0]
    pairs = len(set(zip(A["i"].tolist(), A["j"].tolist())))
    nfree = len(m_)
    return {"kinetic_translational": 0.5 * np.sum(m_ * np.sum(v * v, axis=1)),
            "kinetic_rotational": 0.5 * np.sum(np.sum(I_ * wb * wb, axis=1)),
            "vmax": np.sqrt(np.max(np.sum(v * v, axis=1))), "nfree": nfree, "nactive": len(A),
            "ncontacts": len(contacts), "active_pairs": pairs, "coordination": 2.0 * pairs / nfree,
            "dn_mean": contacts["dn"].mean() if len(contacts) else 0.0,
            "dn_min": contacts["dn"].min() if len(contacts) else 0.0, "ninterfaces": o.ninterfaces}


def _cluster_deck(tmp_path):
    from rockable_amd import decks as D
    shapes, deck = D.bonded_clusters_deck(n_clusters=4, cluster_dims=(3, 3, 3), clearance=1e-4, strength_mean=20.0)
    return D.write_deck(str(tmp_path), shapes, deck)


def test_stats_match_numpy_on_oracle_state(tmp_path):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    path = _cluster_deck(tmp_path)
    e, o = Engine(path), Oracle(path)
    e.update_nl()
    o.update_nl()
    e.step(1600)
    o.run_steps(1600)
    g, r = e.stats(), numpy_stats(o, 1)
    for k in ("nfree", "nactive", "ncontacts", "active_pairs", "ninterfaces"):
        assert g[k] == r[k], k
    assert g["dn_min"] == r["dn_min"] and g["vmax"] == r["vmax"]
    for k in ("kinetic_translational", "kinetic_rotational", "coordination", "dn_mean"):
        assert abs(g[k] - r[k]) <= 1e-12 * abs(r[k]), (k, g[k], r[k])
    assert g["broken_interfaces"] == 216 - o.ninterfaces > 0


def test_long_run_statistics_within_one_percent(tmp_path):
    """20 000 steps of 600 hulls thrown at each other (Avalanche, linkCells
    every 20 steps): kinetic energy, coordination number and broken
    interfaces of the GPU run within 1 % of the oracle's at five points."""
    from oracle.rk_oracle import Oracle
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    shapes, deck = D.hull_packing_deck(n_bodies=600, lattice=(10, 10, 6), law="Avalanche", vmax=10.0,
                                       gap_factor=1.01, interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = Engine(path), Oracle(path)
    e.set_option("record_geometry", 0)
    e.update_nl()
    o.update_nl()
    for _ in range(5):
        e.step(4000)
        o.run_steps(4000)
        g, r = e.stats(), numpy_stats(o, deck.nDriven)
        for k in ("kinetic_translational", "kinetic_rotational", "coordination"):
            assert abs(g[k] - r[k]) <= 0.01 * abs(r[k]) + 1e-300, (k, g[k], r[k])
        assert g["ninterfaces"] == r["ninterfaces"]
    assert r["coordination"] > 0.0
