"""Known-answer pins of the oracle parts that no Rockable golden deck covers:
the Avalanche and StickedLinks laws, bond breakage, linkCells, the body
forces, the driving controls and dynamicUpdateNL.  Each checks a fact the
mathematics or physics fixes independently of the oracle's code."""
import math
import os

import numpy as np
import pytest

from tests import kat_decks as K


def _rebound_ratio(o, v0, steps):
    o.update_nl()
    o.run_steps(steps)
    b = o.bodies()
    return (b["vel"][1, 0] - b["vel"][0, 0]) / v0


@pytest.mark.parametrize("en2", [0.25, 0.64])
def test_avalanche_restitution_is_sqrt_en2(oracle_lib, tmp_path, en2):
    """ForceLaw_Avalanche.cpp:83-96: loading with stiffness kn from the
    first contact (fn += -kn d(dn)), unloading with kn en2 (fn = -kn en2 dn),
    no viscosity.  The energy given back is en2 times the energy stored, so
    the rebound/approach speed ratio is sqrt(en2) (to O(dt / contact time))."""
    R, v0, kn, dt = 0.5, 0.2, 1.0e5, 1e-6
    path = K.two_spheres(str(tmp_path), R, (-R - 1e-3, 0, 0), (R + 1e-3, 0, 0), v1=(v0 / 2, 0, 0),
                         v2=(-v0 / 2, 0, 0), kn=kn, en2=en2, law="Avalanche", dt=dt)
    meff = 0.5 * K.sphere_mass(R)
    tc = math.pi * math.sqrt(meff / kn)
    e = _rebound_ratio(oracle_lib.Oracle(path), v0, int(3 * tc / dt) + int(2e-3 / v0 / dt))
    assert abs(e / math.sqrt(en2) - 1.0) < 2e-3, (e, math.sqrt(en2))


def test_avalanche_friction_saturates_at_mu_fn(oracle_lib, tmp_path):
    """Two overlapping spheres of huge mass slide tangentially at constant
    speed: ft grows by kt vt dt per step (ForceLaw_Avalanche.cpp:100-103)
    until |ft| exceeds mu fn, then stays at |ft| = mu fn (108-111), with
    fn = kn |dn| from loading."""
    R, kn, mu, dt, vt, delta = 0.5, 1.0e5, 0.4, 1e-6, 1e-2, 1e-4
    path = K.two_spheres(str(tmp_path), R, (0, 0, 0), (2 * R - delta, 0, 0), v2=(0, vt, 0), kn=kn, mu=mu,
                         law="Avalanche", dt=dt, density=1e15)
    o = oracle_lib.Oracle(path)
    o.update_nl()
    o.run_steps(1)
    I = o.interactions(active=True)
    assert len(I) == 1
    fn = I["fn"][0]
    assert abs(fn / (kn * delta) - 1.0) < 1e-9
    assert abs(np.linalg.norm(I["ft"][0]) / (0.8 * kn * vt * dt) - 1.0) < 1e-9
    o.run_steps(int(2 * mu * kn * delta / (0.8 * kn * vt * dt)))
    I = o.interactions(active=True)
    assert abs(np.linalg.norm(I["ft"][0]) / (mu * I["fn"][0]) - 1.0) < 1e-12
    # along the sliding direction, up to the turn of n while sliding (ft is
    # not re-projected on the tangent plane without FT_CORR)
    assert abs(I["ft"][0][1]) / np.linalg.norm(I["ft"][0]) > 1.0 - 1e-3


def _bond_motion(R, kn):
    meff = 0.5 * K.sphere_mass(R)
    return math.sqrt(kn / meff)


@pytest.mark.parametrize("breaks", [True, False])
def test_sticked_bond_breaks_at_fn0(oracle_lib, tmp_path, breaks):
    """An undamped bond between two spheres pulled apart at speed v
    (ForceLaw_StickedLinks.cpp:118-150): fn = -kn u with u = dn - dn0 the
    elongation, u(t) = (v / w) sin(w t), w^2 = kn / meff.  The yield
    function -fn / fn0 - 1 (ft = mom = 0) turns positive when kn u > fn0:
    at t* = asin(fn0 w / (kn v)) / w if v / w > fn0 / kn, never otherwise.
    Breaking erases the interface (Rockable.cpp:3350-3413)."""
    R, kn, fn0, dt = 0.5, 1.0e5, 10.0, 1e-6
    w = _bond_motion(R, kn)
    umax_break = fn0 / kn
    v = 1.5 * umax_break * w if breaks else 0.5 * umax_break * w
    o = oracle_lib.Oracle(K.bonded_spheres(str(tmp_path), R, v=v, kn=kn, fn0=fn0, dt=dt))
    assert o.ninterfaces == 1
    if breaks:
        tstar = math.asin(fn0 * w / (kn * v)) / w
        k = int(tstar / dt)
        o.run_steps(k - 3)
        assert o.ninterfaces == 1
        o.run_steps(6)
        assert o.ninterfaces == 0
        o.run_steps(int(2 * math.pi / w / dt))
        b = o.bodies()
        assert b["vel"][1, 0] - b["vel"][0, 0] > 0.0  # the freed bodies separate
    else:
        period = 2 * math.pi / w
        umax = 0.0
        for _ in range(40):
            o.run_steps(int(period / 20 / dt))
            b = o.bodies()
            umax = max(umax, b["pos"][1, 0] - b["pos"][0, 0] - 2 * R)
        assert o.ninterfaces == 1
        assert abs(umax / (v / w) - 1.0) < 2e-2, (umax, v / w)


def test_linkcells_list_equals_bruteforce(oracle_lib, tmp_path):
    """UpdateNL_linkCells (Rockable.cpp:1667-1809) finds every pair
    UpdateNL_bruteForce (1614-1662) finds when the cells are larger than any
    body's box plus DVerlet: the two lists (keys and histories) are equal
    after the first rebuild and after 300 and 1000 steps of hulls thrown at
    each other with rebuilds every 20."""
    cells, brute = K.hull_pair_decks(str(tmp_path))
    a, b = oracle_lib.Oracle(cells), oracle_lib.Oracle(brute)
    a.update_nl()
    b.update_nl()
    for step in (0, 300, 700):
        a.run_steps(step)
        b.run_steps(step)
        Ia, Ib = a.interactions(), b.interactions()
        assert len(Ia) == len(Ib)
        for f in Ia.dtype.names:
            assert np.array_equal(Ia[f], Ib[f], equal_nan=Ia[f].dtype.kind == "f"), f
    assert len(Ia) > 100 and a.nactive > 3


def test_attracting_point_constant_acceleration(oracle_lib, tmp_path):
    """BodyForce AttractingPoint (BodyForce_AttractingPoint.cpp:51-59): a
    force m a towards the point, so a body at rest covers a t^2 / 2 along the
    line in time t (velocity Verlet integrates a constant force exactly)."""
    a, L, dt = 2.0, 3.0, 1e-4
    path = K.single_sphere(str(tmp_path), pos=(L, 0, 0), body_force="AttractingPoint 0 0 0 %r" % a, dt=dt)
    o = oracle_lib.Oracle(path)
    n = 5000
    o.run_steps(n)
    t = n * dt
    b = o.bodies()
    assert abs((L - b["pos"][0, 0]) / (0.5 * a * t * t) - 1.0) < 1e-9
    assert np.abs(b["pos"][0, 1:]).max() == 0.0


def test_viscous_fluid_terminal_velocity(oracle_lib, tmp_path):
    """BodyForce ViscousFluid (BodyForce_ViscousFluid.cpp:51-62): the drag
    0.5 * 0.47 * rho_f * S * v^2 with S = V^(2/3) balances the weight of a
    falling sphere at v_t = sqrt(2 m g / (0.47 rho_f S))."""
    R, rho_f, g, dt = 0.5, 1000.0, 9.81, 1e-4
    path = K.single_sphere(str(tmp_path), R=R, gravity=(0, 0, -g), body_force="ViscousFluid %r" % rho_f, dt=dt)
    o = oracle_lib.Oracle(path)
    m = K.sphere_mass(R)
    S = (4.0 / 3.0 * math.pi * R ** 3) ** (2.0 / 3.0)
    vt = math.sqrt(2 * m * g / (0.47 * rho_f * S))
    o.run_steps(int(8 * vt / g / dt))
    vz = o.bodies()["vel"][0, 2]
    assert abs(-vz / vt - 1.0) < 1e-4, (vz, vt)


def test_velocity_and_force_controls(oracle_lib, tmp_path):
    """drivingSystem.txt controls (DrivingSystem.cpp:44-268; Rockable.cpp
    1856-1900, 3450-3514): an imposed x velocity moves the plate by v t; an
    imposed z force gives a constant acceleration, so from rest the plate
    covers z = v t / 2 (velocity Verlet is exact for a constant force) with
    v growing linearly; the free body far away does not move.  The
    accelerations are recomputed once the controls are read (the load-time
    "fake step" runs before them)."""
    vx, Fz, dt, n = 0.5, 3.0e3, 1e-5, 4000
    path, drv = K.driven_plate(str(tmp_path), ["Control _x_Vel_ 0 %r" % vx, "Control _z_For_ 0 %r" % Fz], dt=dt)
    o = oracle_lib.Oracle(path, driving=drv)
    o.accelerations()
    b0 = o.bodies()
    o.run_steps(n // 2)
    vz_half = o.bodies()["vel"][0, 2]
    o.run_steps(n - n // 2)
    b = o.bodies()
    t = n * dt
    assert abs((b["pos"][0, 0] - b0["pos"][0, 0]) / (vx * t) - 1.0) < 1e-12
    vz = b["vel"][0, 2]
    assert vz > 0.0
    assert abs((b["pos"][0, 2] - b0["pos"][0, 2]) / (0.5 * vz * t) - 1.0) < 1e-9
    assert abs(vz_half / (0.5 * vz) - 1.0) < 1e-9
    assert np.array_equal(b["pos"][1], b0["pos"][1])


def test_dynamic_update_nl_counts(oracle_lib, tmp_path):
    """dynamicCheckUpdateNL (Rockable.cpp:1571-1596): with a body moving at
    speed v and no periodic rebuild, the list is rebuilt each time the
    accumulated displacement reaches dispUpdateNL, i.e. floor(v t / disp)
    times."""
    v, disp, dt, n = 0.3, 0.01, 1e-4, 2000
    path = K.single_sphere(str(tmp_path), vel=(v, 0, 0), dt=dt, dynamic=(disp, 45.0))
    o = oracle_lib.Oracle(path)
    before = o.nl_updates
    o.run_steps(n)
    expected = int(v * n * dt / disp + 1e-9)
    assert abs((o.nl_updates - before) - expected) <= 1, (o.nl_updates - before, expected)
