"""Pins the CPU oracle against Rockable's own regression goldens and stored
conf records (reference test/regression/release, test/input).

- Golden runs: the reference's CTest recipe (test/CMakeLists.txt:41-45) --
  load the redirection deck, initial UpdateNL, integrate, then compareConf
  (src/Apps/run.cpp:38-92) of the saved conf against the golden file.
- Stored-record pins: the input decks' "Interactions" records (n, dn, pos) and
  body accelerations were written by the reference from the very body poses
  and forces stored beside them, so recomputing them is an exact, step-free
  check of the kinematics (Interaction.cpp:147-326, Particle.cpp:49-65) and of
  the resultant/acceleration stage (Rockable.cpp:2752-2789, 3419-3552).
- Planted mistakes: each mutant below is a plausible slip (dropped term,
  transposed operand, wrong sign or index) planted in the oracle; at least one
  pin must catch every one of them.
"""
import decimal
import os
import tempfile

import numpy as np
import pytest

from tests.conftest import DECKS


def _run_golden(rk, redirection, out_name, golden, mutant=0):
    rk.set_variant(7, mutant)
    try:
        d = tempfile.mkdtemp()
        o = rk.Oracle(os.path.join(DECKS, redirection))
        o.update_nl()
        o.integrate(d)
        ok, k, a, b = rk.compare_conf(os.path.join(d, out_name), os.path.join(DECKS, "regression", golden))
        o.close()
        return ok, k, a, b
    finally:
        rk.set_variant(7, 0)


CASES = {
    "512": ("redirection_test_512_particles_bruteforce_LawDefault_noPeriodicity.conf", "conf101",
            "regression_512_particles_LawDefault_noPeriodicity.conf"),
    "518": ("redirection_test_518_poly_bruteforce_default_noperiod.conf", "conf3001",
            "regression_518_poly_default_noperiod.conf"),
}


@pytest.mark.parametrize("case", ["512", "518"])
def test_golden_line_for_line(oracle_lib, case):
    ok, k, a, b = _run_golden(oracle_lib, *CASES[case])
    assert ok, "first difference at line %s:\nREF %s\nNEW %s" % (k, a, b)


def test_unknown_law_name_falls_back_to_default(oracle_lib, tmp_path):
    """'forceLaw Avalanches' is not a registered name (Rockable.cpp:154) and
    falls back to Default (Rockable.cpp:774-780); the reference's two
    'LawAvalanche' goldens are therefore the Default goldens."""
    src = os.path.join(DECKS, "input", "512_particles")
    deck = open(os.path.join(src, "512_particles_bruteforce_LawDefault_noPeriodicity.conf")).read()
    deck = deck.replace("forceLaw Default", "forceLaw Avalanches")
    (tmp_path / "deck.conf").write_text(deck)
    (tmp_path / "shape.shp").write_text(open(os.path.join(src, "shape.shp")).read())
    o = oracle_lib.Oracle(str(tmp_path / "deck.conf"))
    o.update_nl()
    out = tmp_path / "out"
    out.mkdir()
    o.integrate(str(out))
    ok, k, a, b = oracle_lib.compare_conf(str(out / "conf101"), os.path.join(
        DECKS, "regression", "regression_512_particles_LawDefault_noPeriodicity.conf"))
    assert ok, (k, a, b)


def _trunc16(x):
    """%.16e under FE_TOWARDZERO, as saveConf prints n (Rockable.cpp:572-574)."""
    x = float(x)
    if x == 0.0:
        return "%.16e" % x
    d = decimal.Decimal(x)
    with decimal.localcontext() as c:
        c.rounding = decimal.ROUND_DOWN
        c.prec = 17
        d = +d
    sign, digits, exp = d.as_tuple()
    digits = "".join(map(str, digits)).ljust(17, "0")[:17]
    e10 = exp + len(d.as_tuple().digits) - 1
    return "%s%s.%se%s%02d" % ("-" if sign else "", digits[0], digits[1:], "-" if e10 < 0 else "+", abs(e10))


def _parse_records(path):
    lines = open(path).read().split("\n")
    k = next(i for i, l in enumerate(lines) if l.startswith("Particles"))
    nb = int(lines[k].split()[1])
    bodies = np.array([[float(x) for x in l.split()[4:]] for l in lines[k + 2:k + 2 + nb]])
    k2 = next(i for i, l in enumerate(lines) if l.startswith("Interactions"))
    recs = {}
    raw = {}
    for l in lines[k2 + 1:k2 + 1 + int(lines[k2].split()[1])]:
        t = l.split()
        recs[tuple(map(int, t[:5]))] = np.array([float(x) for x in t[5:]])
        raw[tuple(map(int, t[:5]))] = t[5:8]
    return bodies, recs, raw


KIN_DECKS = ["512_particles/512_particles_bruteforce_LawDefault_noPeriodicity.conf",
             "518_poly/518_poly_bruteforce_default_noperiod.conf"]


def _kinematics_mismatches(rk, deck, mutant=0):
    rk.set_variant(7, mutant)
    try:
        path = os.path.join(DECKS, "input", deck)
        _, recs, raw = _parse_records(path)
        o = rk.Oracle(path, fake_step=False)
        o.update_kinematics()
        got = {(int(r["i"]), int(r["j"]), int(r["type"]), int(r["isub"]), int(r["jsub"])): r
               for r in o.interactions()}
        bad = 0
        for key, ref in recs.items():
            r = got[key]
            n_ok = all(_trunc16(r["n"][c]) == raw[key][c] for c in range(3))
            if not (n_ok and r["dn"] == ref[3] and np.array_equal(r["pos"], ref[4:7])):
                bad += 1
        o.close()
        return bad, len(recs)
    finally:
        rk.set_variant(7, 0)


@pytest.mark.parametrize("deck", KIN_DECKS)
def test_kinematics_reproduce_stored_records(oracle_lib, deck):
    bad, tot = _kinematics_mismatches(oracle_lib, deck)
    assert tot > 100 and bad == 0, "%d of %d records differ" % (bad, tot)


def _resultant_mismatches(rk, deck, mutant=0):
    rk.set_variant(7, mutant)
    try:
        path = os.path.join(DECKS, "input", deck)
        bodies, _, _ = _parse_records(path)
        o = rk.Oracle(path, fake_step=False)
        o.update_kinematics()
        o.resultants_only()
        b = o.bodies()
        nd = o.ndriven
        acc_ok = np.all(b["acc"][nd:] == bodies[nd:, 6:9], axis=1)
        arot_ok = np.all(b["arot"][nd:] == bodies[nd:, 16:19], axis=1)
        o.close()
        return int((~(acc_ok & arot_ok)).sum()), len(acc_ok)
    finally:
        rk.set_variant(7, 0)


@pytest.mark.parametrize("deck", KIN_DECKS)
def test_resultants_reproduce_stored_accelerations(oracle_lib, deck):
    bad, tot = _resultant_mismatches(oracle_lib, deck)
    assert tot == 512 and bad == 0, "%d of %d bodies differ" % (bad, tot)


MUTANTS = {
    1: "viscous normal term dropped (ForceLaw_Default.cpp:85-86)",
    2: "cross-product operands transposed in the relative velocity (Interaction.cpp:162)",
    3: "Ri+Ri instead of Ri+Rj in the vertex-vertex gap (Interaction.cpp:157)",
    4: "wrong sign of the j-side contact moment (Rockable.cpp:2764)",
    5: "inertia indices swapped in Euler's equation (Rockable.cpp:3546)",
    6: "full dt instead of dt/2 in the closing kick (Rockable.cpp:2153-2154)",
    7: "face radius instead of vertex radius in the vertex-face contact point (Interaction.cpp:259)",
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_is_caught(oracle_lib, mutant, tmp_path):
    from tests.test_oracle_kat import asymmetric_top_ok, rolling_moment_ok, unequal_spheres_ok
    caught = [not unequal_spheres_ok(oracle_lib, tmp_path, mutant),
              not rolling_moment_ok(oracle_lib, tmp_path, mutant),
              not asymmetric_top_ok(oracle_lib, tmp_path, mutant)]
    ok, _, _, _ = _run_golden(oracle_lib, *CASES["518"], mutant=mutant)
    caught.append(not ok)
    if not caught[-1]:
        ok, _, _, _ = _run_golden(oracle_lib, *CASES["512"], mutant=mutant)
        caught.append(not ok)
    if not any(caught):
        bad, _ = _kinematics_mismatches(oracle_lib, KIN_DECKS[1], mutant)
        caught.append(bad > 0)
    assert any(caught), "mutant %d (%s) passed every pin" % (mutant, MUTANTS[mutant])
