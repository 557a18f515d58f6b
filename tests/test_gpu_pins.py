"""The decks of the oracle's known-answer pins (tests/test_oracle_pins.py) on
the GPU: the Default friction cap with damping on, critical damping for
en2 <= 0, the Avalanche rolling-moment cap, bonds breaking on their
tangential, viscous and moment terms (power 2 and 1.5), the StickedLinks
contact cap, a compressed bond breaking into a capped contact, the biaxial
servo and moment controls.  The engine is bit-identical to the oracle on every one, at
every checkpoint, with the same live interfaces.

PreferredDirection is the exception: its restoring angle is acos of a
cosine, and the device's acos and glibc's may differ in the last bit.  The
bar there is the tolerance of the path's contract: body moments within 1e-10
of the oscillation's moment amplitude at every checkpoint, orientations and
spins within 1e-8 of theirs, over 128 000 steps."""
import numpy as np
import pytest

from tests import kat_decks as K
from tests import test_oracle_pins as P
from tests.test_gpu_parity import assert_bodies_equal, assert_interactions_equal

pytestmark = pytest.mark.gpu


def _case(name, d):
    if name == "default_cap":
        return P.default_cap_deck(d)
    if name.startswith("critical_damping"):
        return P.critical_damping_deck(d, 0.0 if name.endswith("0") else -1.0)
    if name == "rolling_normal":
        return P.avalanche_rolling_deck(d, (0, 0, 1))
    if name == "rolling_oblique":
        return P.avalanche_rolling_deck(d, (0.5, 0.0, 0.8660254037844386))
    if name == "rolling_parallel":
        return P.avalanche_rolling_deck(d, (1, 0, 0))
    if name == "bond_viscous":
        return K.spun_bond(d, w=0.4, kt=2.0e4, damp=0.035, ft0=1.0, mom0=1.0, power=2.0), None, 200
    if name == "bond_moment":
        return K.spun_bond(d, w=0.4, kt=1.0, kr=3.1e2, ft0=1e6, mom0=0.25, power=2.0), None, 2200
    if name == "bond_power":
        return K.spun_bond(d, w=0.4, kt=1.5e4, damp=0.02, kr=1.3e2, ft0=1.0, mom0=0.5, power=1.5), None, 300
    if name == "sticked_contact_cap":
        import os
        from rockable_amd import decks as Dk
        path, drv, steps = P.default_cap_deck(d)
        return Dk.conf_with(path, os.path.join(d, "sticked.conf"), forceLaw="StickedLinks"), drv, 2 * steps
    if name == "bond_compressed":
        return K.spun_bond(d, w=0.4, kt=2e4, ft0=1.0, mom0=1.0, power=2.0, overlap=1e-5), None, 600
    if name == "biaxial_servo":
        return P.servo_box_deck(d, "tritriBiaxialCompression 50.0 0.2")
    if name in ("moment_x", "moment_xyz"):
        return P.moment_control_deck(d, name.split("_")[1])
    raise ValueError(name)


BIT_EXACT = ["default_cap", "critical_damping_0", "critical_damping_m1", "rolling_normal", "rolling_oblique",
             "rolling_parallel", "bond_viscous", "bond_moment", "bond_power", "sticked_contact_cap", "bond_compressed",
             "biaxial_servo", "moment_x", "moment_xyz"]


@pytest.mark.parametrize("name", BIT_EXACT)
def test_pinned_decks_bit_exact(tmp_path, name):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    path, drv, steps = _case(name, str(tmp_path))
    e = Engine(path, driving=drv)
    o = Oracle(path, driving=drv)
    e.update_nl()
    o.update_nl()
    done, chunk = 0, max(1, steps // 4)
    while done < steps:
        n = min(chunk, steps - done)
        e.step(n)
        o.run_steps(n)
        done += n
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(), o.interactions(),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        assert e.counts()["ninterfaces"] == o.ninterfaces
    if name.startswith("bond"):
        assert o.ninterfaces == 0, "the bond must have broken within the run"


@pytest.mark.parametrize("antiparallel", [False, True])
def test_preferred_direction_within_tolerance(tmp_path, antiparallel):
    from oracle.rk_oracle import Oracle
    from rockable_amd.engine import Engine
    path, _, steps = P.preferred_direction_deck(str(tmp_path), antiparallel)
    e = Engine(path)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    # amplitudes of the oscillation: moment Kr theta0, spin theta0 w
    import math
    R, momentMax, theta0 = 0.5, 2.0e3, 0.4
    Kr = 2.0 * momentMax / math.pi
    w = math.sqrt(Kr / (0.4 * R * R * K.sphere_mass(R)))
    differing = 0
    for _ in range(8):
        e.step(steps // 8)
        o.run_steps(steps // 8)
        g, r = e.bodies(), o.bodies()
        assert np.abs(g["moment"] - r["moment"]).max() <= 1e-10 * Kr * theta0
        assert np.abs(g["Q"] - r["Q"]).max() <= 1e-8
        assert np.abs(g["vrot"] - r["vrot"]).max() <= 1e-8 * theta0 * w
        differing += int(not np.array_equal(g["Q"], r["Q"]))
    print("checkpoints with orientations differing in some bit: %d of 8" % differing)
