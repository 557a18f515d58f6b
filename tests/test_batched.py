"""The batched deck is the union of its packings: on the oracle, each
packing of a batch follows bit for bit the trajectory it follows as a deck
of its own (bodies of two packings never meet, the body renumbering is
increasing, so every per-body sum keeps its order)."""
import numpy as np

from tests import batched_case as BC
from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal


def test_batched_oracle_equals_each_packing(oracle_lib, tmp_path):
    path, shapes, deck, info = BC.small_batch(str(tmp_path))
    o = oracle_lib.Oracle(path)
    o.update_nl()
    o.run_steps(BC.STEPS)
    B, I = o.bodies(), o.interactions()
    contacts = 0
    for p in range(len(info)):
        q = oracle_lib.Oracle(BC.single(str(tmp_path), shapes, deck, info, p))
        q.update_nl()
        q.run_steps(BC.STEPS)
        rows = BC.rows_of(info, p)
        S = q.bodies()
        assert_bodies_equal({k: B[k][rows] for k in S}, S)
        assert_interactions_equal(BC.interactions_of(I, info, p), q.interactions(),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        contacts += q.nactive
    assert contacts >= 5 and len(I) > 1000
