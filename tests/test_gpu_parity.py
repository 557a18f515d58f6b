"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle and
the reference goldens.  Integer and index work (contact sets, keys, stick
links) must be identical; FP64 state must be bit-identical, since both sides
evaluate the same expressions in the same order without fused multiply-add."""
import os
import tempfile

import numpy as np
import pytest

from tests.conftest import DECKS

pytestmark = pytest.mark.gpu

CASES = {
    "512": ("redirection_test_512_particles_bruteforce_LawDefault_noPeriodicity.conf", "conf101",
            "regression_512_particles_LawDefault_noPeriodicity.conf"),
    "518": ("redirection_test_518_poly_bruteforce_default_noperiod.conf", "conf3001",
            "regression_518_poly_default_noperiod.conf"),
}


def _engine(path, **kw):
    from rockable_amd.engine import Engine
    return Engine(path, **kw)


def assert_bodies_equal(a, b, keys=("pos", "vel", "Q", "vrot", "acc", "arot", "force", "moment"), rows=None):
    for k in keys:
        x, y = a[k], b[k]
        if rows is not None:
            x, y = x[rows], y[rows]
        bad = np.argwhere(~((x == y) | (np.isnan(x) & np.isnan(y))))
        assert bad.size == 0, "%s differs at %s: gpu %r oracle %r" % (k, bad[:3].tolist(), x[tuple(bad[0])],
                                                                      y[tuple(bad[0])])


def assert_interactions_equal(g, o, fields=("dn", "fn", "ft", "mom", "damp", "stick", "n", "pos", "vel")):
    keys = ("i", "j", "type", "isub", "jsub")
    assert len(g) == len(o), "listed/active counts differ: gpu %d oracle %d" % (len(g), len(o))
    for k in keys:
        assert np.array_equal(g[k], o[k]), "key column %s differs" % k
    for f in fields:
        x, y = g[f], o[f]
        same = (x == y) | (np.isnan(x) & np.isnan(y)) if x.dtype.kind == "f" else (x == y)
        bad = np.argwhere(~same)
        assert bad.size == 0, "interaction field %s differs at %s" % (f, bad[:3].tolist())


@pytest.mark.parametrize("case", ["512", "518"])
def test_golden_line_for_line_on_gpu(case):
    from oracle.rk_oracle import compare_conf
    red, out, golden = CASES[case]
    d = tempfile.mkdtemp()
    e = _engine(os.path.join(DECKS, red))
    e.update_nl()
    e.integrate(d)
    ok, k, a, b = compare_conf(os.path.join(d, out), os.path.join(DECKS, "regression", golden))
    assert ok, "first difference at line %s:\nREF %s\nGPU %s" % (k, a, b)


def test_fake_step_and_nl_match_oracle_518():
    from oracle.rk_oracle import Oracle
    path = os.path.join(DECKS, CASES["518"][0])
    e = _engine(path)
    o = Oracle(path)
    assert_bodies_equal(e.bodies(), o.bodies())
    assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
    e.update_nl()
    o.update_nl()
    g, r = e.interactions(), o.interactions()
    assert_interactions_equal(g, r, fields=("dn", "fn", "ft", "mom", "damp", "stick"))


def _run_pair(path, steps, chunk=None):
    from oracle.rk_oracle import Oracle
    e = _engine(path)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    chunk = chunk or steps
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        e.step(n)
        o.run_steps(n)
        done += n
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(active=True), o.interactions(active=True))
    ge, go = e.interactions(), o.interactions()
    assert_interactions_equal(ge, go, fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    return e, o


def test_sphere_lattice_trajectory_bit_exact(tmp_path):
    """512 spheres on a tight lattice thrown onto a floor (Default law,
    friction, brute-force NL rebuilt every 100 steps): 3000 steps
    bit-identical to the oracle, with ~100 active contacts at the end."""
    from rockable_amd import decks as D
    shapes, deck = D.sphere_lattice_deck(n_side=8, vmax=0.2, spacing_factor=2.02, dt=2e-7)
    deck.gravity = (0.0, 0.0, -98.1)
    for b in deck.bodies[deck.nDriven:]:
        b.vel = (b.vel[0], b.vel[1], b.vel[2] - 0.1)
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = _run_pair(path, 3000, chunk=1000)
    assert o.nactive > 50


def test_cpu_runnable_configuration_verbatim(tmp_path):
    """The CPU-runnable configuration exactly as specified: 512 spheres
    (R = 0.5 mm, density 2700) on an 8 x 8 x 8 lattice of spacing 2.2 R,
    velocities uniform in [-0.01, 0.01] m/s from seed 1, dropping under
    1 g onto a fixed floor inside four fixed walls; bruteForce neighbour list
    (rebuilt every 100 steps), ForceLaw Default with kn = 1e6 N/m,
    kt = 0.8 kn, en = 0.5 (en2 0.25), mu = 0.4; dt = 1e-6 s, 10 000 steps.
    Bit-identical to the oracle at every 1000-step checkpoint (bodies and
    active interactions), the bottom layers in contact with the floor and
    each other by the end."""
    from rockable_amd import decks as D
    shapes, deck = D.sphere_lattice_deck(n_side=8, R=0.5e-3, spacing_factor=2.2, seed=1, vmax=0.01, kn=1e6,
                                         en=0.5, mu=0.4, dt=1e-6, steps=10000, nl="bruteForce")
    assert deck.gravity == (0.0, 0.0, -9.81) and len(deck.bodies) == 517 and deck.nDriven == 5
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = _run_pair(path, 10000, chunk=1000)
    I = o.interactions(active=True)
    assert len(I) > 40 and np.any(I["j"] < 5) and np.any(I["j"] >= 5)
    assert o.nl_updates == e.counts()["nl_updates"] >= 100


@pytest.mark.parametrize("law", ["Default", "Avalanche"])
def test_hull_packing_linkcells_bit_exact(tmp_path, law):
    """600 convex sphero-polyhedra (hulls of 8-14 points) thrown at each
    other, linkCells NL rebuilt every 20 steps, all four contact types: 1000
    steps bit-identical to the oracle."""
    from rockable_amd import decks as D
    shapes, deck = D.hull_packing_deck(n_bodies=600, lattice=(10, 10, 6), law=law, vmax=10.0, gap_factor=1.01,
                                       interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    path = D.write_deck(str(tmp_path), shapes, deck)
    e, o = _run_pair(path, 1000, chunk=250)
    I = o.interactions()
    assert o.nactive >= 3 and len(I) > 100
    assert set(I["type"].tolist()) == {0, 1, 2, 3}


def _sphere_bed(tmp_path):
    """3200 spheres in two layers on the floor, thrown down at 0.05 m/s under
    100 g: the floor holds ~1600 listed entries (four 512-slot chunks of the
    block-wide sum)."""
    from rockable_amd import decks as D
    shapes, deck = D.sphere_lattice_deck(dims=(40, 40, 2), vmax=0.01, spacing_factor=2.02, dt=1e-6, nl="linkCells")
    deck.gravity = (0.0, 0.0, -981.0)
    for b in deck.bodies[deck.nDriven:]:
        b.vel = (b.vel[0], b.vel[1], b.vel[2] - 0.05)
    return D.write_deck(str(tmp_path), shapes, deck)


def _wall_scale(o, nwalls):
    """Per wall, the sum of |terms| of its active interactions: the scale
    against which a reordered sum is compared."""
    I = o.interactions(active=True)
    f = np.abs(I["fn"][:, None] * I["n"] + I["ft"])
    sc = np.zeros((nwalls, 3))
    for w in range(nwalls):
        m = (I["i"] == w) | (I["j"] == w)
        sc[w] = f[m].sum(axis=0) if m.any() else 0.0
    return sc


def test_heavy_bodies_serial_sum_bit_exact(tmp_path):
    """Bodies with more slots than heavy_min summed block-wide in serial
    order (exact_sums): every body bit-identical to the oracle, walls
    included, over 800 steps, with the floor's sum spanning four chunks and
    the free bodies (5-6 slots) also on the block-wide path."""
    from oracle.rk_oracle import Oracle
    path = _sphere_bed(tmp_path)
    e = _engine(path)
    e.set_option("heavy_min", 4)
    e.set_option("exact_sums", 1)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    I = o.interactions()
    slots = np.bincount(I["i"], minlength=o.nbodies) + np.bincount(I["j"], minlength=o.nbodies)
    assert slots[:5].max() > 1024
    for _ in range(4):
        e.step(200)
        o.run_steps(200)
        assert_bodies_equal(e.bodies(), o.bodies())
    assert o.nactive > 1000


def test_heavy_passive_walls_tree_sum(tmp_path):
    """Default mode: fixed walls with many slots are summed by a fixed tree.
    All motion stays bit-identical (their resultants feed nothing back); the
    walls' force and moment agree with the serial sums to 1e-12 of the sum of
    |terms|."""
    from oracle.rk_oracle import Oracle
    path = _sphere_bed(tmp_path)
    e = _engine(path)
    e.set_option("heavy_min", 16)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    for _ in range(4):
        e.step(200)
        o.run_steps(200)
        g, r = e.bodies(), o.bodies()
        nw = 5
        assert_bodies_equal(g, r, keys=("pos", "vel", "Q", "vrot", "acc", "arot"))
        assert_bodies_equal(g, r, keys=("force", "moment"), rows=slice(nw, None))
        sc = _wall_scale(o, nw)
        assert np.all(np.abs(g["force"][:nw] - r["force"][:nw]) <= 1e-12 * sc.sum(axis=1)[:, None] + 1e-300)
        lever = 0.1  # box size bound on the moment arms (m)
        assert np.all(np.abs(g["moment"][:nw] - r["moment"][:nw]) <= 1e-12 * lever * sc.sum(axis=1)[:, None] + 1e-300)
    assert np.abs(r["force"][0]).max() > 0.0


def _kat_case(name, d):
    """(conf path, driving file or None, steps, recompute accelerations after
    the controls) of the known-answer decks of test_oracle_laws.py."""
    import math
    from tests import kat_decks as K
    if name == "avalanche_restitution":
        return K.two_spheres(d, 0.5, (-0.501, 0, 0), (0.501, 0, 0), v1=(0.1, 0, 0), v2=(-0.1, 0, 0),
                             law="Avalanche", en2=0.25), None, 60000, False
    if name == "avalanche_friction":
        return K.two_spheres(d, 0.5, (0, 0, 0), (0.9999, 0, 0), v2=(0, 1e-2, 0), mu=0.4, law="Avalanche",
                             density=1e15), None, 3000, False
    if name in ("bond_breaks", "bond_holds"):
        w = math.sqrt(1e5 / (0.5 * K.sphere_mass(0.5)))
        v = (1.5 if name == "bond_breaks" else 0.5) * 1e-4 * w
        return K.bonded_spheres(d, 0.5, v=v, kn=1e5, fn0=10.0), None, 150000, False
    if name == "attracting_point":
        return K.single_sphere(d, pos=(3.0, 0, 0), body_force="AttractingPoint 0 0 0 2.0"), None, 5000, False
    if name == "viscous_fluid":
        return K.single_sphere(d, gravity=(0, 0, -9.81), body_force="ViscousFluid 1000.0"), None, 5000, False
    if name == "controls":
        path, drv = K.driven_plate(d, ["Control _x_Vel_ 0 0.5", "Control _z_For_ 0 3000.0"])
        return path, drv, 4000, True
    if name == "dynamic_nl":
        return K.single_sphere(d, vel=(0.3, 0, 0), dynamic=(0.01, 45.0)), None, 2000, False
    if name == "moment_controls":
        # a plate spun by imposed moments about one axis, then about all three
        # (the rotational branch of the driven bodies' accelerations and kick)
        path, drv = K.driven_plate(d, ["Control _zrot_Mom_ 0 2.0", "Control _x_For_ 0 50.0",
                                       "Control _xrot_Vel_ 0 0.3"])
        return path, drv, 3000, True
    if name == "moment_xyz":
        path, drv = K.driven_plate(d, ["Control _xyzrot_Mom_ 0 1.0 0.5 -0.25", "Control _y_Vel_ 0 0.2"])
        return path, drv, 3000, True
    raise ValueError(name)


@pytest.mark.parametrize("name", ["avalanche_restitution", "avalanche_friction", "bond_breaks", "bond_holds",
                                  "attracting_point", "viscous_fluid", "controls", "dynamic_nl", "moment_controls",
                                  "moment_xyz"])
def test_known_answer_decks_bit_exact(tmp_path, name):
    """The decks that pin the Avalanche and StickedLinks laws, breakage, the
    body forces, the driving controls and dynamicUpdateNL against physics:
    the GPU engine is bit-identical to the oracle on them, including the
    count of live interfaces and of list rebuilds."""
    from oracle.rk_oracle import Oracle
    path, drv, steps, reacc = _kat_case(name, str(tmp_path))
    e = _engine(path, driving=drv)
    o = Oracle(path, driving=drv)
    if reacc:
        e.accelerations()
        o.accelerations()
    e.update_nl()
    o.update_nl()
    chunk = max(1, steps // 4)
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        e.step(n)
        o.run_steps(n)
        done += n
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(), o.interactions(),
                                  fields=("dn", "fn", "ft", "mom", "damp", "stick"))
        c = e.counts()
        assert c["ninterfaces"] == o.ninterfaces
        assert c["nl_updates"] == o.nl_updates
    if name == "bond_breaks":
        assert o.ninterfaces == 0
    if name == "bond_holds":
        assert o.ninterfaces == 1


def test_linkcells_equals_bruteforce_on_gpu(tmp_path):
    """The same 300-hull deck with linkCells and with bruteForce gives
    bit-identical lists and bodies on the GPU after 1000 steps (rebuilds
    every 20), and both equal the oracle."""
    from oracle.rk_oracle import Oracle
    from tests import kat_decks as K
    cells, brute = K.hull_pair_decks(str(tmp_path))
    a, b, o = _engine(cells), _engine(brute), Oracle(cells)
    for x in (a, b, o):
        x.update_nl()
    a.step(1000)
    b.step(1000)
    o.run_steps(1000)
    assert_bodies_equal(a.bodies(), b.bodies())
    assert_bodies_equal(a.bodies(), o.bodies())
    assert_interactions_equal(a.interactions(), b.interactions())
    assert_interactions_equal(a.interactions(), o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    assert len(o.interactions()) > 100


def _dynamic_hull_deck(tmp_path, disp=1e-3, angle_deg=15.0):
    from rockable_amd import decks as D
    shapes, deck = D.hull_packing_deck(n_bodies=600, lattice=(10, 10, 6), law="Avalanche", vmax=10.0,
                                       gap_factor=1.01, interverlet_steps=100000, size_range=(5.8e-3, 6e-3))
    deck.dynamic_update_nl = 1
    deck.disp_update_nl = disp
    deck.angle_update_nl = angle_deg
    return D.write_deck(str(tmp_path), shapes, deck)


@pytest.mark.parametrize("batch", [1, 7, 64])
def test_device_guard_bit_exact(tmp_path, batch):
    """dynamicUpdateNL with the displacement check on the device and the
    host reading it once per batch of iterations: every rebuild happens at
    the reference's iteration (same count as the oracle, whose check runs on
    the host each iteration) and the state stays bit-identical, for batches
    of 1, 7 and 64 iterations, across several rk_step calls."""
    from oracle.rk_oracle import Oracle
    path = _dynamic_hull_deck(tmp_path)
    e = _engine(path)
    e.set_option("guard_batch", batch)
    o = Oracle(path)
    e.update_nl()
    o.update_nl()
    for n in (150, 350, 500):
        e.step(n)
        o.run_steps(n)
        assert e.counts()["nl_updates"] == o.nl_updates
        assert_bodies_equal(e.bodies(), o.bodies())
        assert_interactions_equal(e.interactions(), o.interactions(), fields=("dn", "fn", "ft", "mom", "damp", "stick"))
    assert o.nl_updates > 10


def test_device_guard_integrate_conf_files(tmp_path):
    """integrate() with dynamicUpdateNL and conf outputs: the conf files
    written by the engine (batches of 16) equal the oracle's line for line."""
    import glob
    from oracle.rk_oracle import Oracle, compare_conf
    from rockable_amd import decks as D
    path = _dynamic_hull_deck(tmp_path)
    path = D.conf_with(path, str(tmp_path / "deck2.conf"), interConf=100 * 2e-7, tmax=700 * 2e-7)
    ge, go = tmp_path / "gpu", tmp_path / "cpu"
    ge.mkdir()
    go.mkdir()
    e = _engine(path)
    e.update_nl()
    n = e.integrate(str(ge))
    o = Oracle(path)
    o.update_nl()
    o.integrate(str(go))
    files = sorted(glob.glob(str(go / "conf*")))
    assert n >= 699 and len(files) >= 7
    for f in files:
        ok, k, a, b = compare_conf(str(ge / os.path.basename(f)), f)
        assert ok, "%s differs at line %s:\nCPU %s\nGPU %s" % (os.path.basename(f), k, b, a)
    assert o.nl_updates == e.counts()["nl_updates"] and o.nl_updates > 5
