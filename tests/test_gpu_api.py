"""Error behaviour of the C ABI on a device: null paths and invalid decks
return an error code with a message, never abort."""
import ctypes
import os

import pytest

from tests.conftest import DECKS

pytestmark = pytest.mark.gpu

DECK_518 = os.path.join(DECKS, "redirection_test_518_poly_bruteforce_default_noperiod.conf")


def test_null_driving_path_is_an_error():
    from rockable_amd.engine import Engine
    e = Engine(DECK_518)
    assert e.L.rk_load_driving_system(e.h, None) < 0
    assert b"no path" in e.L.rk_last_error(e.h)
    assert e.L.rk_save_conf(e.h, None) < 0
    e.close()


def test_bad_group_is_an_error(tmp_path):
    from rockable_amd import decks as D
    from rockable_amd.engine import Engine
    D.write_shp(str(tmp_path / "shapes.shp"), [D.sphere_shape("S", 0.5)])
    deck = D.Deck(density={0: 1000.0}, params=D.contact_params(1e5, 8e4, 0.25, 0.3))
    deck.bodies = [D.Body("S", group=0), D.Body("S", group=300, pos=(3.0, 0.0, 0.0))]
    D.write_conf(str(tmp_path / "deck.conf"), deck)
    e = Engine(None)
    with pytest.raises(RuntimeError, match="group 300"):
        e.load_conf(str(tmp_path / "deck.conf"))
    e.close()


def test_unknown_law_falls_back_to_default():
    """Rockable.cpp:774-780: an unknown law warns and Default is used."""
    from rockable_amd.engine import Engine
    e = Engine(DECK_518)
    assert e.set_force_law("NoSuchLaw") == 1
    assert b"unknown" in e.L.rk_last_error(e.h)
    e.close()
