"""Known-answer pins of the oracle functions that no reference golden vector
reaches (the 512- and 518-body goldens run frictionless Default contacts).
Each pin compares the oracle with a result the mathematics fixes on its own
-- a saturation level, a release speed of a solved ODE, a break step counted
in closed form, the period of a harmonic oscillator -- and each is shown to
fail on a planted mistake in the oracle (a plausible slip: the wrong force in
a threshold, a dropped branch, a dropped term).

  Default Coulomb cap on the elastic normal force   ForceLaw_Default.cpp:109-117
  critical damping for en2 <= 0                     Rockable.cpp:1372-1373
  Avalanche rolling-moment cap                      ForceLaw_Avalanche.cpp:118-132
  StickedLinks yield: tangential, moment and the
    cumulative viscous term                         ForceLaw_StickedLinks.cpp:118-150
  PreferredDirection restoring moment               BodyForce_PreferredDirection.cpp:54-75
"""
import math
import os

import numpy as np
import pytest

from rockable_amd import decks as D
from tests import kat_decks as K

MUTANT = 7  # rko_set_variant slot of the planted mistakes


def _with_mutant(rk, mutant, fn):
    rk.set_variant(MUTANT, mutant)
    try:
        return fn()
    finally:
        rk.set_variant(MUTANT, 0)


# --- Default: friction capped at mu |fne|, fne = -kn dn (elastic part) ---------
def default_cap_deck(d):
    """(conf, driving file, steps) of the Default friction-cap pin."""
    R, kn, mu, dt, vt, delta, en2 = 0.5, 1.0e5, 0.4, 1e-6, 0.1, 1e-4, 0.25
    rho = 1e15
    m = rho * 4.0 / 3.0 * math.pi * R ** 3
    logen = 0.5 * math.log(en2)
    damp = -logen / math.sqrt(logen * logen + math.pi ** 2) * 2.0 * math.sqrt(kn * 0.5 * m)
    vn = kn * delta / damp  # fnv = fne
    w = vt / (R - 0.5 * delta)
    path = K.two_spheres(d, R, (0, 0, 0), (2 * R - delta, 0, 0), v2=(-vn, 0, 0), kn=kn, mu=mu,
                         en2=en2, law="Default", dt=dt, density=rho, w2=(0.0, 0.0, w))
    return path, None, int(3 * mu * kn * delta / (0.8 * kn * vt * dt))


def default_cap_ok(rk, tmp_path, mutant=0):
    """Two spheres of huge mass, overlapping by delta, body j approaching
    along the line of centres at a tiny vn and spinning about z, so that the
    contact point slides at vt = (R - delta / 2) w while the normal stays
    fixed: the damping term damp vn doubles the normal force (fn = fne + fnv
    with fnv = fne), but the friction threshold is mu |fne|
    (ForceLaw_Default.cpp:109-117, without ALLOW_NEGATIVE_THRESHOLDS): once
    saturated, |ft| = mu kn |dn| exactly (to rounding), half of mu fn."""
    path, _, steps = default_cap_deck(str(tmp_path))
    kn, mu = 1.0e5, 0.4

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        o.run_steps(steps)
        return o.interactions(active=True)
    I = _with_mutant(rk, mutant, run)
    if len(I) != 1:
        return False
    fne = -kn * I["dn"][0]
    ft = np.linalg.norm(I["ft"][0])
    return (abs(I["fn"][0] / fne - 2.0) < 1e-6        # the damping term is as large as the elastic one
            and abs(ft / (mu * fne) - 1.0) < 1e-12)   # capped on the elastic part


def test_default_friction_cap_on_elastic_force(oracle_lib, tmp_path):
    assert default_cap_ok(oracle_lib, tmp_path)


def test_default_friction_cap_catches_total_force(oracle_lib, tmp_path):
    assert not default_cap_ok(oracle_lib, tmp_path, mutant=8)


# --- critical damping: en2 <= 0 gives Damp = 2 sqrt(kn meff) --------------------
def _clamped_release_speed(kn, c, meff):
    """Release/approach speed ratio of meff d'' = -max(0, kn d + c d') from
    d = 0, d' = 1 (see test_oracle_kat)."""
    from tests.test_oracle_kat import _clamped_dashpot_restitution
    return _clamped_dashpot_restitution(kn, c, meff)


def critical_damping_deck(d, en2):
    R, v0, kn, dt = 0.5, 0.2, 1.0e5, 1e-6
    path = K.two_spheres(d, R, (-R - 1e-3, 0, 0), (R + 1e-3, 0, 0), v1=(v0 / 2, 0, 0),
                         v2=(-v0 / 2, 0, 0), kn=kn, en2=en2, dt=dt)
    tc = math.pi * math.sqrt(0.5 * K.sphere_mass(R) / kn)
    return path, None, int(3 * tc / dt) + int(2e-3 / v0 / dt)


def critical_damping_ok(rk, tmp_path, en2, mutant=0):
    """en2 <= 0 selects critical damping, c = 2 sqrt(kn meff)
    (Rockable.cpp:1372-1373).  For the clamped spring-dashpot the bodies
    separate when kn d + c d' = 0: with d = t exp(-w t), at w t = 2, leaving
    with e^-2 = 0.135 of the approach speed (solved numerically to 1e-12
    here too), to within the time-step error (0.5 %)."""
    R, v0, kn = 0.5, 0.2, 1.0e5
    path, _, steps = critical_damping_deck(str(tmp_path), en2)
    meff = 0.5 * K.sphere_mass(R)

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        o.run_steps(steps)
        b = o.bodies()
        return (b["vel"][1, 0] - b["vel"][0, 0]) / v0, o.interactions()
    e, I = _with_mutant(rk, mutant, run)
    c = 2.0 * math.sqrt(kn * meff)
    e_ode = _clamped_release_speed(kn, c, meff)
    damp_ok = len(I) == 0 or abs(I["damp"][0] / c - 1.0) < 1e-15
    return abs(e_ode - math.exp(-2.0)) < 1e-6 and abs(e / e_ode - 1.0) < 5e-3 and damp_ok


@pytest.mark.parametrize("en2", [0.0, -1.0])
def test_critical_damping_for_nonpositive_en2(oracle_lib, tmp_path, en2):
    assert critical_damping_ok(oracle_lib, tmp_path, en2)


def test_critical_damping_catches_elastic_fallback(oracle_lib, tmp_path):
    assert not critical_damping_ok(oracle_lib, tmp_path, 0.0, mutant=9)


# --- Avalanche rolling moment -------------------------------------------------
def avalanche_rolling_deck(d, omega_dir):
    Ri, Rj, kn, kr, mur, dt, delta, w = 0.3, 0.5, 1.0e5, 200.0, 0.05, 1e-6, 1e-4, 3.0
    shapes = [D.sphere_shape("SI", Ri), D.sphere_shape("SJ", Rj)]
    D.write_shp(os.path.join(d, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e-4, interConf=1e3, DVerlet=0.05, dVerlet=0.025,
                  density={0: 1e15}, gravity=(0.0, 0.0, 0.0), force_law="Avalanche",
                  params=D.contact_params(kn, 0.8 * kn, 0.25, 0.0, kr=kr, mur=mur))
    wi = w * np.asarray(omega_dir, float) / np.linalg.norm(omega_dir)
    deck.bodies = [D.Body("SI", pos=(0.0, 0.0, 0.0), vrot=tuple(float(x) for x in wi)),
                   D.Body("SJ", pos=(Ri + Rj - delta, 0.0, 0.0))]
    path = os.path.join(d, "deck.conf")
    D.write_conf(path, deck)
    return path, None, 400


def avalanche_rolling_ok(rk, tmp_path, omega_dir, mutant=0):
    """Two overlapping spheres of huge mass and unequal radii; body i (free,
    so the branch is taken from it) spins at w along omega_dir, body j is at
    rest.  The rolling moment grows by kr (wj - wi) dt per evaluation
    (ForceLaw_Avalanche.cpp:116) and is capped at mur |b_perp| fn, b_perp the
    branch pos - xi without its component along wi (118-132): for wi normal
    to the branch that is mur (Ri + dn/2) fn; for wi along it 0 up to the
    rounding of the projection (the moment is cut to that size); at 60 deg,
    sin 60 of the first."""
    Ri, kn, kr, mur, dt, delta, w = 0.3, 1.0e5, 200.0, 0.05, 1e-6, 1e-4, 3.0
    path, _, _ = avalanche_rolling_deck(str(tmp_path), omega_dir)
    wi = w * np.asarray(omega_dir, float) / np.linalg.norm(omega_dir)

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        o.run_steps(400)
        return o.interactions(active=True)
    I = _with_mutant(rk, mutant, run)
    if len(I) != 1:
        return False
    dn, fn, mom = I["dn"][0], I["fn"][0], I["mom"][0]
    branch = np.array([Ri + 0.5 * dn, 0.0, 0.0])
    u = wi / np.linalg.norm(wi)
    perp = np.linalg.norm(branch - np.dot(branch, u) * u)
    cap = mur * perp * fn
    grown = kr * w * dt * 401  # unbounded growth after the load's evaluation and 400 steps
    if perp < 1e-12 * np.linalg.norm(branch):
        # the branch along w leaves a perpendicular part of rounding size only
        return (np.linalg.norm(mom) < 1e-12 * mur * np.linalg.norm(branch) * fn
                and abs(fn / (kn * delta) - 1.0) < 1e-9)
    return (cap < grown and abs(np.linalg.norm(mom) / cap - 1.0) < 1e-12
            and np.allclose(mom / np.linalg.norm(mom), -u, atol=1e-12)
            and abs(fn / (kn * delta) - 1.0) < 1e-9)


@pytest.mark.parametrize("omega_dir", [(0, 0, 1), (1, 0, 0), (0.5, 0.0, math.sqrt(3) / 2)])
def test_avalanche_rolling_moment_cap(oracle_lib, tmp_path, omega_dir):
    assert avalanche_rolling_ok(oracle_lib, tmp_path, omega_dir)


@pytest.mark.parametrize("omega_dir", [(1, 0, 0), (0.5, 0.0, math.sqrt(3) / 2)])
def test_avalanche_rolling_catches_unprojected_branch(oracle_lib, tmp_path, omega_dir):
    assert not avalanche_rolling_ok(oracle_lib, tmp_path, omega_dir, mutant=10)


# --- StickedLinks yield: ft, mom and the cumulative viscous term ---------------
def spun_bond_break_step(rk, tmp_path, kt, damp, kr, ft0, mom0, power, mutant=0, max_steps=4000):
    """Two touching spheres (dn = dn0 = 0) of huge mass joined by one bond;
    body j spins at w about z.  The contact point moves at vt = R w along -y
    (Interaction.cpp:162) with vn = 0, so per evaluation ft grows by
    (kt dt + damp) vt -- the elastic increment and the viscous term damp vt,
    which the law adds to ft and keeps (ForceLaw_StickedLinks.cpp:124-137) --
    mom by kr w dt, and fn stays 0.  The yield function
    (|ft| / ft0)^p + (|mom| / mom0)^p - fn / fn0 - 1 first turns positive at
    evaluation k* = the least k with k^p (a^p + b^p) > 1; the load's
    accelerations() is evaluation 1, so the interface breaks during step
    k* - 1.  Returns (break step, k*)."""
    R, w, dt = 0.5, 0.4, 1e-6
    path = K.spun_bond(str(tmp_path), R=R, w=w, kn=1e5, kt=kt, kr=kr, fn0=10.0, ft0=ft0, mom0=mom0, power=power,
                       damp=damp, dt=dt)
    a = (kt * dt + damp) * R * w / ft0
    b = kr * dt * w / mom0
    s = a ** power + b ** power
    x = (1.0 / s) ** (1.0 / power)
    kstar = int(math.floor(x)) + 1
    assert min(x - math.floor(x), math.ceil(x) - x) > 1e-3, "crossing too close to an evaluation"

    def run():
        o = rk.Oracle(path)
        if o.ninterfaces != 1:
            return -1
        o.update_nl()
        for step in range(1, max_steps + 1):
            o.run_steps(1)
            if o.ninterfaces == 0:
                return step
        return None
    return _with_mutant(rk, mutant, run), kstar


@pytest.mark.parametrize("case", [
    # (kt, damp, kr, ft0, mom0, power): viscous and elastic tangential terms
    (2.0e4, 0.035, 0.0, 1.0, 1.0, 2.0),
    # the moment term alone (tangential terms far from ft0)
    (1.0, 0.0, 3.1e2, 1e6, 0.25, 2.0),
    # both, another power
    (1.5e4, 0.02, 1.3e2, 1.0, 0.5, 1.5),
])
def test_bond_breaks_at_closed_form_step(oracle_lib, tmp_path, case):
    step, kstar = spun_bond_break_step(oracle_lib, tmp_path, *case)
    assert kstar > 20
    assert step == kstar - 1, (step, kstar)


def test_bond_pin_catches_dropped_viscous_term(oracle_lib, tmp_path):
    step, kstar = spun_bond_break_step(oracle_lib, tmp_path, 2.0e4, 0.035, 0.0, 1.0, 1.0, 2.0, mutant=11)
    assert step != kstar - 1


def test_bond_pin_catches_dropped_moment_term(oracle_lib, tmp_path):
    step, kstar = spun_bond_break_step(oracle_lib, tmp_path, 1.0, 0.0, 3.1e2, 1e6, 0.25, 2.0, mutant=12)
    assert step != kstar - 1


# --- PreferredDirection: a harmonic rotor -----------------------------------------
def preferred_direction_deck(d, antiparallel):
    R, momentMax, theta0, dt = 0.5, 2.0e3, 0.4, 1e-5
    axis = np.array([math.cos(theta0), math.sin(theta0), 0.0])
    if antiparallel:
        axis = -axis
    bf = "PreferredDirection 1 0 0 %r %r %r %r" % (float(axis[0]), float(axis[1]), float(axis[2]), momentMax)
    return K.single_sphere(d, R=R, body_force=bf, dt=dt), None, 8 * 16000


def preferred_direction_ok(rk, tmp_path, antiparallel, mutant=0):
    """A free sphere (isotropic inertia I) whose body x axis starts at theta0
    from the preferred axis (or from its opposite), no gravity, no contact.
    The moment Kr sgn u angle, Kr = 2 momentMax / pi, is linear in the angle
    on both branches (BodyForce_PreferredDirection.cpp:54-75), so the body
    turns about z as phi'' = (Kr / I)(theta0 - phi): phi(t) = theta0 (1 -
    cos(w t)), w = sqrt(Kr / I).  Checked at eight instants over a period, to
    the velocity-Verlet error (1e-5 of theta0)."""
    R, momentMax, theta0, dt = 0.5, 2.0e3, 0.4, 1e-5
    path, _, _ = preferred_direction_deck(str(tmp_path), antiparallel)
    I = 0.4 * R * R * K.sphere_mass(R)
    w = math.sqrt(2.0 * momentMax / math.pi / I)
    period = 2.0 * math.pi / w
    n_per = int(round(period / 8.0 / dt))

    def rot(q, x):
        qw, u = q[0], q[1:]
        t = 2.0 * np.cross(u, x)
        return x + qw * t + np.cross(u, t)

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        errs = []
        for k in range(1, 9):
            o.run_steps(n_per)
            t = k * n_per * dt
            xb = rot(o.bodies()["Q"][0], np.array([1.0, 0.0, 0.0]))
            phi = math.atan2(xb[1], xb[0])
            errs.append(abs(phi - theta0 * (1.0 - math.cos(w * t))) / theta0)
            if abs(xb[2]) > 1e-12:
                return 1.0
        return max(errs)
    return _with_mutant(rk, mutant, run) < 1e-5


@pytest.mark.parametrize("antiparallel", [False, True])
def test_preferred_direction_harmonic_rotor(oracle_lib, tmp_path, antiparallel):
    assert preferred_direction_ok(oracle_lib, tmp_path, antiparallel)


def test_preferred_direction_catches_missing_antiparallel_branch(oracle_lib, tmp_path):
    assert not preferred_direction_ok(oracle_lib, tmp_path, True, mutant=13)


# --- tritriBiaxialCompression servo ---------------------------------------------
def servo_box_deck(d, servo, r=0.01, dt=1e-5, nsteps=4000):
    """Six driven walls of an empty 1 x 0.6 x 0.8 box (floor, x0, x1, y0, y1,
    lid) under the servo line `servo` (its name and values), walls in the
    order xmin xmax ymin ymax zmin zmax = 1 2 3 4 0 5."""
    lo, hi = np.zeros(3), np.array([1.0, 0.6, 0.8])
    shapes, walls = D.box_walls(lo, hi, r, top=True)
    D.write_shp(os.path.join(d, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e3, interConf=1e3, DVerlet=0.01, dVerlet=0.005,
                  density={1: 1000.0}, gravity=(0.0, 0.0, 0.0), nDriven=6,
                  params=D.contact_params(1e5, 8e4, 0.25, 0.3, groups=((1, 1),)))
    deck.bodies = list(walls)
    path = os.path.join(d, "deck.conf")
    D.write_conf(path, deck)
    drv = os.path.join(d, "drivingSystem.txt")
    name, _, values = servo.partition(" ")
    with open(drv, "w") as fh:
        fh.write("Servo %s 1 2 3 4 0 5 %s\n" % (name, values))
    return path, drv, nsteps


def biaxial_servo_error(rk, tmp_path, mutant=0):
    """Six driven walls of an empty box (no contact), servo
    tritriBiaxialCompression p v (DrivingSystem.cpp:273-301): ymax moves at
    -v, xmax and zmax are force-driven by -p surfX = -p dY dZ and
    -p surfZ = -p dY dX, the gaps measured between the walls' inner faces
    (position -/+ Minkowski radius), the min walls fixed.  With no
    resistance the two force-driven walls follow
        mx x'' = -p dY(t) dZ(t),  mz z'' = -p dY(t) dX(t),  dY(t) = dY0 - v t,
    which scipy integrates to 1e-12; the walls must follow it to within the
    one-step lag of the servo (relative 1e-3 of the displacements).
    Returns the largest relative error."""
    from scipy.integrate import solve_ivp
    r, p, v, dt = 0.01, 50.0, 0.2, 1e-5
    path, drv, nsteps = servo_box_deck(str(tmp_path), "tritriBiaxialCompression %r %r" % (p, v))

    def run():
        o = rk.Oracle(path, driving=drv)
        o.update_nl()
        b0 = o.bodies()
        o.run_steps(nsteps)
        return b0, o.bodies(), o.mass()[0]
    b0, b1, m = _with_mutant(rk, mutant, run)
    gap = lambda b, a, k: (b["pos"][k, a] - r) - (b["pos"][{0: 1, 1: 3, 2: 0}[a], a] + r)
    dX0, dY0, dZ0 = gap(b0, 0, 2), gap(b0, 1, 4), gap(b0, 2, 5)
    mx, mz = m[2], m[5]

    def rhs(t, y):
        dX, dZ = y[0], y[2]
        dY = dY0 - v * t
        return [y[1], -p * dY * dZ / mx, y[3], -p * dY * dX / mz]
    T = nsteps * dt
    sol = solve_ivp(rhs, (0.0, T), [dX0, 0.0, dZ0, 0.0], rtol=1e-12, atol=1e-15)
    dX, dZ = sol.y[0][-1], sol.y[2][-1]
    err_x = abs(gap(b1, 0, 2) - dX) / abs(dX - dX0)
    err_z = abs(gap(b1, 2, 5) - dZ) / abs(dZ - dZ0)
    err_y = abs(gap(b1, 1, 4) - (dY0 - v * T)) / (v * T)
    fixed = max(np.abs(b1["pos"][k] - b0["pos"][k]).max() for k in (0, 1, 3))
    return max(err_x, err_z, err_y, fixed)


def test_biaxial_servo_follows_its_ode(oracle_lib, tmp_path):
    assert biaxial_servo_error(oracle_lib, tmp_path) < 1e-3


def test_biaxial_servo_catches_swapped_surface(oracle_lib, tmp_path):
    assert biaxial_servo_error(oracle_lib, tmp_path, mutant=14) > 1e-2


# --- moment controls of a driven body -------------------------------------------------
def _moment_axis(kind):
    """Initial orientation and spin axis of the moment-control pin: body x
    along global x, or body y turned onto d = (1, 1, 1) / sqrt(3)."""
    if kind == "x":
        return np.array([1.0, 0.0, 0.0, 0.0]), np.array([1.0, 0.0, 0.0])
    d = np.ones(3) / math.sqrt(3.0)
    ey = np.array([0.0, 1.0, 0.0])
    ax = np.cross(ey, d)
    ang = math.acos(float(np.dot(ey, d)))
    ax /= np.linalg.norm(ax)
    return np.concatenate([[math.cos(ang / 2)], math.sin(ang / 2) * ax]), d


def moment_control_deck(dirpath, kind, M0=0.7, dt=1e-4, n=2000):
    v = np.array([[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)], dtype=float)
    top = D.Shape("TOP", 0.01, v * np.array([0.3, 0.2, 0.1]), volume=1.0, inertia_mass=(1.0, 2.0, 3.0))
    ball = D.sphere_shape("S", 0.1)
    D.write_shp(os.path.join(dirpath, "shapes.shp"), [top, ball])
    q0, d = _moment_axis(kind)
    if kind == "x":
        ctrl = "Control _xrot_Mom_ 0 %r" % M0
    else:
        Md = M0 * d
        ctrl = "Control _xyzrot_Mom_ 0 %r %r %r" % (float(Md[0]), float(Md[1]), float(Md[2]))
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e3, interConf=1e3, DVerlet=0.01, dVerlet=0.005,
                  density={0: 1.0}, gravity=(0.0, 0.0, 0.0), nDriven=1,
                  params=D.contact_params(1e5, 8e4, 0.25, 0.3))
    deck.bodies = [D.Body("TOP", Q=tuple(float(x) for x in q0)), D.Body("S", pos=(100.0, 0.0, 0.0))]
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    drv = os.path.join(dirpath, "drivingSystem.txt")
    with open(drv, "w") as fh:
        fh.write(ctrl + "\n")
    return path, drv, n


def moment_control_error(rk, tmp_path, kind, mutant=0):
    """A driven body of principal inertia (1, 2, 3) m under a constant
    controlled moment M0 about one of its principal axes, no contact
    (Rockable.cpp:3450-3514, 1988-2033): it spins up about that axis at
    alpha = M0 / I_axis.  Velocity Verlet kicks a constant acceleration
    exactly, but the deck's "fake step" runs before the driving system is read
    (console_run, Rockable.cpp:1232-1238), so the body's arot is still 0 for
    the first half kick: w(t) = alpha (t - dt / 2) exactly; the angle is
    alpha t^2 / 2 to O(dt / t).  kind "x": _xrot_Mom_ about
    global x = body x; kind "xyz": _xyzrot_Mom_ along d, the body turned so
    that its y axis lies along d.  Returns the larger relative error of w
    and of the angle."""
    Ib = np.array([1.0, 2.0, 3.0])
    dt, M0 = 1e-4, 0.7
    path, drv, n = moment_control_deck(str(tmp_path), kind)
    q0, d = _moment_axis(kind)
    axis_inertia = Ib[0] if kind == "x" else Ib[1]

    def run():
        o = rk.Oracle(path, driving=drv)
        o.update_nl()
        o.run_steps(n)
        return o.bodies()
    b = _with_mutant(rk, mutant, run)
    t = n * dt
    alpha = M0 / axis_inertia  # body mass 1 (density 1, volume 1)
    w = b["vrot"][0]
    q = b["Q"][0]
    # relative rotation q q0^-1: its vector part lies along d
    q0c = np.array([q0[0], -q0[1], -q0[2], -q0[3]])
    rw = q[0] * q0c[0] - np.dot(q[1:], q0c[1:])
    rv = q[0] * q0c[1:] + q0c[0] * q[1:] + np.cross(q[1:], q0c[1:])
    angle = 2.0 * math.atan2(float(np.dot(rv, d)), float(rw))
    err_w = np.linalg.norm(w - alpha * (t - 0.5 * dt) * d) / (alpha * t)
    err_a = abs(angle / (0.5 * alpha * t * t) - 1.0)
    off_axis = np.linalg.norm(rv - np.dot(rv, d) * d)
    return max(err_w, err_a if err_a > 3.0 / n else 0.0, off_axis)


@pytest.mark.parametrize("kind", ["x", "xyz"])
def test_moment_control_spins_up_at_M_over_I(oracle_lib, tmp_path, kind):
    assert moment_control_error(oracle_lib, tmp_path, kind) < 1e-9


def test_moment_control_catches_unrotated_moment(oracle_lib, tmp_path):
    assert moment_control_error(oracle_lib, tmp_path, "xyz", mutant=15) > 1e-3


# --- StickedLinks: the contact branch and the breakage of a compressed bond --------
def sticked_contact_cap_ok(rk, tmp_path, mutant=0):
    """The Default-cap deck under StickedLinks (no interface: every entry is a
    contact): that branch caps the friction at mu |fn| with fn the total
    normal force, damping included (ForceLaw_StickedLinks.cpp:185-197), so
    |ft| saturates at mu (fne + fnv) = 2 mu fne here -- twice Default's."""
    path, _, steps = default_cap_deck(str(tmp_path))
    from rockable_amd import decks as Dk
    path = Dk.conf_with(path, os.path.join(str(tmp_path), "sticked.conf"), forceLaw="StickedLinks")
    kn, mu = 1.0e5, 0.4

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        o.run_steps(2 * steps)
        return o.interactions(active=True)
    I = _with_mutant(rk, mutant, run)
    if len(I) != 1:
        return False
    fne = -kn * I["dn"][0]
    fn = I["fn"][0]
    return abs(fn / fne - 2.0) < 1e-6 and abs(np.linalg.norm(I["ft"][0]) / (mu * fn) - 1.0) < 1e-12


def test_sticked_contact_cap_on_total_force(oracle_lib, tmp_path):
    assert sticked_contact_cap_ok(oracle_lib, tmp_path)


def test_sticked_contact_cap_catches_elastic_threshold(oracle_lib, tmp_path):
    assert not sticked_contact_cap_ok(oracle_lib, tmp_path, mutant=16)


def compressed_bond_break_ok(rk, tmp_path, mutant=0):
    """A bond pre-compressed by 1e-5 (dn = -1e-5, dn0 = 0, so fn = kn 1e-5)
    and sheared by the spin of body j breaks on its tangential term at the
    step the closed form gives; the broken interface's contact keeps fn and
    has its friction cut to mu |fn| (Rockable.cpp:3366-3380: dn < 0), not
    zeroed as a bond in tension is."""
    R, w, dt, kn, kt, mu = 0.5, 0.4, 1e-6, 1e5, 2e4, 0.3
    path = K.spun_bond(str(tmp_path), R=R, w=w, kn=kn, kt=kt, fn0=10.0, ft0=1.0, mom0=1.0, power=2.0, damp=0.0,
                       dt=dt, overlap=1e-5)

    def run():
        o = rk.Oracle(path)
        o.update_nl()
        for step in range(1, 2000):
            o.run_steps(1)
            if o.ninterfaces == 0:
                return step, o.interactions()
        return None, None
    step, I = _with_mutant(rk, mutant, run)
    if step is None or len(I) != 1:
        return False
    fn = I["fn"][0]
    return (I["stick"][0] == -1 and abs(fn / (kn * 1e-5) - 1.0) < 1e-6
            and abs(np.linalg.norm(I["ft"][0]) / (mu * fn) - 1.0) < 1e-12)


def test_compressed_bond_breaks_to_a_capped_contact(oracle_lib, tmp_path):
    assert compressed_bond_break_ok(oracle_lib, tmp_path)


def test_compressed_bond_pin_catches_zeroed_contact(oracle_lib, tmp_path):
    assert not compressed_bond_break_ok(oracle_lib, tmp_path, mutant=17)
