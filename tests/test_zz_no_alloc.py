"""rk_get_stats_noalloc (the statistics with their scratch in a fixed device
array, no allocation) through the C ABI against rk_get_stats, bit for bit,
on two calls in a row: a large packing, then a smaller one on the same
device (the second call finds the first's partial sums in the shared array
and must not read them)."""
import ctypes
import os

import numpy as np
import pytest

from tests.conftest import DECKS

pytestmark = pytest.mark.gpu

FIELDS = ("kinetic_translational", "kinetic_rotational", "vmax", "nfree", "nactive", "ncontacts", "active_pairs",
          "coordination", "dn_mean", "dn_min", "ninterfaces", "broken_interfaces", "total_broken_area")


def _noalloc(e):
    from rockable_amd.engine import RkStats
    st = RkStats()
    rc = e.L.rk_get_stats_noalloc(e.h, ctypes.byref(st))
    assert rc == 0, e.L.rk_last_error(e.h).decode()
    return {k: getattr(st, k) for k in FIELDS}


def _same(a, b):
    for k in FIELDS:
        x, y = a[k], b[k]
        assert x == y or (x != x and y != y), "%s: noalloc %r, rk_get_stats %r" % (k, x, y)


def test_noalloc_stats_equal_allocating_stats(tmp_path):
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    big = Engine(os.path.join(DECKS, "redirection_test_518_poly_bruteforce_default_noperiod.conf"),
                 driving=os.path.join(DECKS, "input", "518_poly", "drivingSystem.txt"))
    big.update_nl()
    big.step(50)
    first = _noalloc(big)
    _same(first, big.stats())
    assert first["nactive"] > 1000 and big.counts()["nlisted"] > 10000

    shapes, deck = D.hull_packing_deck(n_bodies=64, lattice=(4, 4, 4), vmax=2.0, gap_factor=1.01,
                                       interverlet_steps=25)
    small = Engine(D.write_deck(str(tmp_path), shapes, deck))
    small.update_nl()
    small.step(20)
    second = _noalloc(small)
    _same(second, small.stats())
    assert second["nfree"] == 64 and small.counts()["nlisted"] < big.counts()["nlisted"]
    assert second["kinetic_translational"] != first["kinetic_translational"]
    # and the large packing again, after the small one used the array
    _same(_noalloc(big), big.stats())
    big.close()
    small.close()
