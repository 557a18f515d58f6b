"""Small decks with known answers, shared by the oracle pins
(test_oracle_laws.py) and the GPU parity tests (test_gpu_parity.py)."""
import math
import os

import numpy as np

from rockable_amd import decks as D

RHO = 1000.0


def sphere_mass(R):
    return RHO * 4.0 / 3.0 * math.pi * R ** 3


def two_spheres(dirpath, R, x1, x2, v1=(0, 0, 0), v2=(0, 0, 0), kn=1e5, en2=0.25, mu=0.0, law="Default",
                dt=1e-6, density=RHO, kr=0.0, mur=0.0, w1=(0, 0, 0), w2=(0, 0, 0)):
    """Two spheres of radius R (angular velocities w1, w2), no gravity, one group."""
    shapes = [D.sphere_shape("S", R)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e-4, interConf=1e3, DVerlet=0.1 * R, dVerlet=0.05 * R,
                  density={0: density}, gravity=(0.0, 0.0, 0.0), force_law=law,
                  params=D.contact_params(kn, 0.8 * kn, en2, mu, kr=kr, mur=mur))
    deck.bodies = [D.Body("S", pos=tuple(x1), vel=tuple(v1), vrot=tuple(w1)),
                   D.Body("S", pos=tuple(x2), vel=tuple(v2), vrot=tuple(w2))]
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    return path


def bonded_spheres(dirpath, R=0.5, v=0.1, kn=1e5, fn0=10.0, dt=1e-6):
    """Two touching spheres (dn0 = 0) joined by one StickedLinks bond with
    its parameters in the interface (ParamsInInterfaces 1): kn, kt = 0.8 kn,
    kr = 0, fn0, ft0 = mom0 = 1, power 2; the bonded vertex-vertex entry has
    zero damping.  Body 1 leaves at speed v along +x."""
    shapes = [D.sphere_shape("S", R)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e-4, interConf=1e3, DVerlet=0.1 * R, dVerlet=0.05 * R,
                  density={0: RHO}, gravity=(0.0, 0.0, 0.0), force_law="StickedLinks",
                  params=D.contact_params(kn, 0.8 * kn, 0.25, 0.3))
    x0, x1 = np.zeros(3), np.array([2 * R, 0.0, 0.0])
    deck.bodies = [D.Body("S", pos=tuple(x0)), D.Body("S", pos=tuple(x1), vel=(v, 0.0, 0.0))]
    n = (x0 - x1) / np.linalg.norm(x0 - x1)
    pos = x0 - n * R
    rec = (0, 1, 0, 0, 0) + tuple(n) + (0.0,) + tuple(pos) + (0.0, 0.0, 0.0) + (0.0,) + (0.0,) * 6 + (0.0,)
    deck.interactions = [rec]
    deck.params_in_interfaces = 1
    par = (kn, 0.8 * kn, 0.0, fn0, 1.0, 1.0, 2.0, 0.0)
    deck.interfaces = [(0, 1, 0.0, par, [(0, 0, 0)])]
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    return path


def single_sphere(dirpath, R=0.5, pos=(0, 0, 0), vel=(0, 0, 0), gravity=(0.0, 0.0, 0.0), body_force=None,
                  dt=1e-4, dynamic=None, interverlet=1e3):
    """One free sphere; optional BodyForce line and dynamicUpdateNL
    (dynamic = (dispUpdateNL, angleUpdateNL in degrees))."""
    shapes = [D.sphere_shape("S", R)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=interverlet, interConf=1e3, DVerlet=0.1 * R, dVerlet=0.05 * R,
                  density={0: RHO}, gravity=tuple(gravity), params=D.contact_params(1e5, 8e4, 0.25, 0.3))
    deck.bodies = [D.Body("S", pos=tuple(pos), vel=tuple(vel))]
    deck.body_force = body_force
    if dynamic is not None:
        deck.dynamic_update_nl = 1
        deck.disp_update_nl, deck.angle_update_nl = dynamic
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    return path


def driven_plate(dirpath, controls, R=0.5, dt=1e-5):
    """A driven square plate (body 0, nDriven 1) far from a free sphere,
    no gravity; `controls` are drivingSystem.txt lines."""
    plate_shapes, plates = D.box_walls((-1.0, -1.0, 0.0), (1.0, 1.0, 1.0), 0.05)
    shapes = plate_shapes + [D.sphere_shape("S", R)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e3, interConf=1e3, DVerlet=0.1 * R, dVerlet=0.05 * R,
                  density={0: RHO, 1: RHO}, gravity=(0.0, 0.0, 0.0), nDriven=1,
                  params=D.contact_params(1e5, 8e4, 0.25, 0.3, groups=((0, 0), (0, 1))))
    deck.bodies = [plates[0], D.Body("S", pos=(0.0, 0.0, 50.0))]
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    drv = os.path.join(dirpath, "drivingSystem.txt")
    with open(drv, "w") as fh:
        fh.write("\n".join(controls) + "\n")
    return path, drv


def hull_pair_decks(dirpath, n_bodies=300, law="Default"):
    """The same 300-hull deck twice, linkCells (cells larger than any body's
    box plus DVerlet) and bruteForce."""
    shapes, deck = D.hull_packing_deck(n_bodies=n_bodies, lattice=(10, 10, 3), law=law, vmax=10.0,
                                       gap_factor=1.01, interverlet_steps=20, size_range=(5.8e-3, 6e-3))
    cells = D.write_deck(dirpath, shapes, deck, name="cells.conf")
    brute = D.conf_with(cells, os.path.join(dirpath, "brute.conf"), UpdateNL="bruteForce")
    return cells, brute


def spun_bond(dirpath, R=0.5, w=2.0, kn=1e5, kt=8e4, kr=0.0, fn0=10.0, ft0=1.0, mom0=1.0, power=2.0, damp=0.0,
              dt=1e-6, density=1e15, overlap=0.0):
    """Two touching spheres of huge mass (dn = dn0 = 0; or overlapping by
    `overlap`, dn0 still 0: a bond in compression) joined by one StickedLinks
    bond with its parameters in the interface (ParamsInInterfaces 1); the
    bonded vertex-vertex entry carries the damping coefficient damp.  Body 1
    spins at w about z, body 0 is at rest."""
    shapes = [D.sphere_shape("S", R)]
    D.write_shp(os.path.join(dirpath, "shapes.shp"), shapes)
    deck = D.Deck(dt=dt, tmax=1e3, interVerlet=1e-4, interConf=1e3, DVerlet=0.1 * R, dVerlet=0.05 * R,
                  density={0: density}, gravity=(0.0, 0.0, 0.0), force_law="StickedLinks",
                  params=D.contact_params(kn, 0.8 * kn, 0.25, 0.3))
    x0, x1 = np.zeros(3), np.array([2 * R - overlap, 0.0, 0.0])
    deck.bodies = [D.Body("S", pos=tuple(x0)), D.Body("S", pos=tuple(x1), vrot=(0.0, 0.0, w))]
    n = (x0 - x1) / np.linalg.norm(x0 - x1)
    pos = x0 - n * R
    rec = (0, 1, 0, 0, 0) + tuple(n) + (-overlap,) + tuple(pos) + (0.0, 0.0, 0.0) + (0.0,) + (0.0,) * 6 + (damp,)
    deck.interactions = [rec]
    deck.params_in_interfaces = 1
    deck.interfaces = [(0, 1, 0.0, (kn, kt, kr, fn0, ft0, mom0, power, 0.0), [(0, 0, 0)])]
    path = os.path.join(dirpath, "deck.conf")
    D.write_conf(path, deck)
    return path
