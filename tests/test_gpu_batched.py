"""Batched packings on the GPU: one engine runs every packing of the batch
in each launch; the batch is bit-identical to the oracle, and each packing
to the same packing run alone on the GPU."""
import numpy as np
import pytest

from tests import batched_case as BC
from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu


def test_batched_bit_exact_against_oracle_and_single_packings(tmp_path):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    path, shapes, deck, info = BC.small_batch(str(tmp_path))
    e = Engine(path)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    for n in (1000, 1000, 1000):
        e.step(n)
        o.run_steps(n)
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
    B, I = e.bodies(), e.interactions()
    assert_interactions_equal(I, o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    for p in range(len(info)):
        s = Engine(BC.single(str(tmp_path), shapes, deck, info, p))
        s.update_nl()
        s.step(BC.STEPS)
        rows = BC.rows_of(info, p)
        S = s.bodies()
        assert_bodies_equal({k: B[k][rows] for k in S}, S)
        assert_interactions_equal(BC.interactions_of(I, info, p), s.interactions(),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        s.close()
    e.close()


def test_load_batch_of_separate_decks(tmp_path):
    """rk_load_batch: the packings as separate decks, loaded into one engine,
    give the same bodies as the batched deck (same layout: every deck's
    walls first) and as each deck run alone, bit for bit."""
    from rockable_amd.engine import Engine
    path, shapes, deck, info = BC.small_batch(str(tmp_path))
    singles = [BC.single(str(tmp_path), shapes, deck, info, p) for p in range(len(info))]
    e = Engine()
    e.load_batch(singles)
    c = Engine(path)
    for p in range(len(info)):
        assert np.array_equal(e.packing_rows(p), BC.rows_of(info, p))
    for x in (e, c):
        x.update_nl()
        x.step(BC.STEPS)
    B, I = e.bodies(), e.interactions()
    assert_bodies_equal(B, c.bodies())
    assert_interactions_equal(I, c.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    for p in range(len(info)):
        s = Engine(singles[p])
        s.update_nl()
        s.step(BC.STEPS)
        rows = e.packing_rows(p)
        S = s.bodies()
        assert_bodies_equal({k: B[k][rows] for k in S}, S)
        assert_interactions_equal(BC.interactions_of(I, info, p), s.interactions(),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        s.close()
    e.close()
    c.close()


def test_load_batch_rejects_mismatched_decks(tmp_path):
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    path, shapes, deck, info = BC.small_batch(str(tmp_path))
    a = BC.single(str(tmp_path), shapes, deck, info, 0)
    b = D.conf_with(BC.single(str(tmp_path), shapes, deck, info, 1), str(tmp_path / "dt.conf"), dt=2e-6)
    e = Engine()
    with pytest.raises(RuntimeError, match="dt"):
        e.load_batch([a, b])
