"""The deck reader on the host (rk_inspect_conf: no device): invalid group ids
are rejected with a message, and a deck of half a million breakable
interfaces loads in linear time (duplicates of an (i, j) pair dropped, the
first kept, as the reference's std::set of interfaces keeps it,
Rockable.cpp:956-1002)."""
import os
import time

import numpy as np
import pytest

from rockable_amd import decks as D
from rockable_amd.engine import inspect_conf


def _spheres_deck(dirpath, n, groups=None, interfaces=()):
    shapes = [D.sphere_shape("S", 0.5)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=1e-6, density={0: 1000.0}, params=D.contact_params(1e5, 8e4, 0.25, 0.3))
    deck.bodies = [D.Body("S", group=(groups[k] if groups else 0), pos=(2.0 * k, 0.0, 0.0)) for k in range(n)]
    deck.interfaces = list(interfaces)
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    return path


def test_inspect_counts(tmp_path):
    info = inspect_conf(_spheres_deck(str(tmp_path), 7))
    assert info["nbodies"] == 7 and info["nshapes"] == 1 and info["ninterfaces"] == 0
    assert info["law"] == 0 and info["linkcells"] == 0


@pytest.mark.parametrize("bad", [256, 1000, -1])
def test_group_out_of_range_is_an_error(tmp_path, bad):
    """A body group outside [0, 256) would index the density and parameter
    tables out of range (mass 0, then NaN accelerations): the reader
    refuses the deck and names the body."""
    path = _spheres_deck(str(tmp_path), 4, groups=[0, 0, bad, 0])
    with pytest.raises(RuntimeError, match="body 2: group"):
        inspect_conf(path)


def test_group_255_is_accepted(tmp_path):
    info = inspect_conf(_spheres_deck(str(tmp_path), 3, groups=[0, 255, 0]))
    assert info["nbodies"] == 3


def test_interface_index_out_of_range_is_an_error(tmp_path):
    path = _spheres_deck(str(tmp_path), 3, interfaces=[(0, 7, 0.0, None, [])])
    with pytest.raises(RuntimeError, match="interface body index"):
        inspect_conf(path)


def test_half_a_million_interfaces_load_in_linear_time(tmp_path):
    """500 000 interfaces over the pairs of 1 001 bodies, then every 100th
    pair repeated: the repeats are dropped and the load stays far below a
    quadratic scan's minutes."""
    n = 1001
    ii, jj = np.triu_indices(n, k=1)
    pairs = list(zip(ii[:500000].tolist(), jj[:500000].tolist()))
    dups = pairs[::100]
    ifc = [(i, j, 0.0, None, []) for (i, j) in pairs + dups]
    path = _spheres_deck(str(tmp_path), n, interfaces=ifc)
    t0 = time.perf_counter()
    info = inspect_conf(path)
    elapsed = time.perf_counter() - t0
    assert info["ninterfaces"] == 500000
    assert elapsed < 20.0, elapsed


def test_reader_takes_tabs_crlf_and_a_last_token_without_newline(tmp_path):
    """The reader parses numbers in place in the file's buffer: a deck with
    CRLF line ends, tabs between fields and no newline after its last token
    reads as the plain one does."""
    path = _spheres_deck(str(tmp_path), 5)
    plain = inspect_conf(path)
    text = open(path).read()
    odd = text.replace(" ", "\t").replace("\n", "\r\n").rstrip()
    odd_path = os.path.join(str(tmp_path), "odd.conf")
    with open(odd_path, "w", newline="") as f:
        f.write(odd)
    assert inspect_conf(odd_path) == plain
