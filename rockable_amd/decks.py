"""Writers for Rockable's text inputs: shape libraries (.shp) and conf decks.

The formats follow Shape::read / Shape::write (Shape.cpp:69-171) and the conf
keywords of Rockable::initParser / saveConf (Rockable.cpp:440-610, 643-1069).
Also holds the seeded generators of the synthetic benchmark packings
(sphere lattices, convex-hull sphero-polyhedra poured in a box, bonded
clusters, batched packings).
"""
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np


@dataclass
class Shape:
    name: str
    radius: float
    vertex: np.ndarray                      # (nv, 3) body frame
    edge: List[Tuple[int, int]] = field(default_factory=list)
    face: List[List[int]] = field(default_factory=list)
    volume: float = 1.0
    inertia_mass: Tuple[float, float, float] = (1.0, 1.0, 1.0)
    obb_extent: Optional[Tuple[float, float, float]] = None
    obb_center: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    obb_axes: Tuple = ((1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, 0.0, 1.0))

    def extent(self):
        if self.obb_extent is not None:
            return self.obb_extent
        v = np.asarray(self.vertex, dtype=float)
        c = np.asarray(self.obb_center)
        return tuple(float(x) for x in (np.abs(v - c).max(axis=0) + self.radius))


def _f(x):
    return "%.16e" % float(x)


def _v(v):
    return " ".join(_f(x) for x in v)


def write_shp(path, shapes: Sequence[Shape]):
    """Shape library in the Shape::write layout (Shape.cpp:130-171)."""
    out = []
    for s in shapes:
        out.append("<")
        out.append("name " + s.name)
        out.append("radius " + _f(s.radius))
        out.append("preCompDone y")
        out.append("volume " + _f(s.volume))
        out.append("I/m " + _v(s.inertia_mass))
        out.append("obb.extent " + _v(s.extent()))
        out.append("obb.center " + _v(s.obb_center))
        out.append("obb.e1 " + _v(s.obb_axes[0]))
        out.append("obb.e2 " + _v(s.obb_axes[1]))
        out.append("obb.e3 " + _v(s.obb_axes[2]))
        out.append("OBBtreeLevel 0")
        out.append("position 0 0 0")
        out.append("orientation 1 0 0 0")
        out.append("nv %d" % len(s.vertex))
        out.extend(_v(v) for v in s.vertex)
        out.append("ne %d" % len(s.edge))
        out.extend("%d %d" % (a, b) for a, b in s.edge)
        out.append("nf %d" % len(s.face))
        out.extend(" ".join(str(x) for x in [len(f)] + list(f)) for f in s.face)
        out.append(">")
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")


def sphere_shape(name, radius):
    """A single-vertex sphero-polyhedron; OBB extent = radius (not 1)."""
    return Shape(name, radius, np.zeros((1, 3)), volume=4.0 / 3.0 * np.pi * radius ** 3,
                 inertia_mass=(0.4 * radius ** 2,) * 3)


def square_plane_shape(name, half, radius):
    """A square plate (4 vertices, 4 edges, 1 face) in the body xy plane."""
    v = np.array([[half, half, 0.0], [half, -half, 0.0], [-half, -half, 0.0], [-half, half, 0.0]])
    edge = [(0, 1), (0, 3), (1, 2), (2, 3)]
    area = (2 * half) ** 2
    vol = area * 2 * radius
    return Shape(name, radius, v, edge, [[0, 1, 2, 3]], volume=vol,
                 inertia_mass=((2 * half) ** 2 / 12.0, (2 * half) ** 2 / 12.0, (2 * half) ** 2 / 6.0),
                 obb_extent=(half + radius, half + radius, radius))


def polyhedron_mass_properties(vertex, faces):
    """Volume, centroid and inertia tensor per unit mass of a closed convex
    polyhedron (faces oriented or not), by decomposition into tetrahedra from
    the centroid of its vertices."""
    v = np.asarray(vertex, dtype=float)
    o = v.mean(axis=0)
    vol = 0.0
    c = np.zeros(3)
    cov = np.zeros((3, 3))
    canon = np.array([[2, 1, 1], [1, 2, 1], [1, 1, 2]]) / 120.0
    for f in faces:
        for k in range(1, len(f) - 1):
            a, b, cc = v[f[0]] - o, v[f[k]] - o, v[f[k + 1]] - o
            A = np.array([a, b, cc]).T
            d = abs(np.linalg.det(A))
            vol += d / 6.0
            c += d / 6.0 * (a + b + cc) / 4.0
            cov += d * A @ canon @ A.T
    c = c / vol
    cov = cov - vol * np.outer(c, c)
    inertia = np.trace(cov) * np.eye(3) - cov
    return vol, c + o, inertia / vol


def hull_shape(name, points, radius, rng=None):
    """Convex hull of points, recentred at its centroid and rotated into its
    principal inertia frame (the frame Rockable shapes are expressed in)."""
    from scipy.spatial import ConvexHull
    pts = np.asarray(points, dtype=float)
    hull = ConvexHull(pts)
    used = sorted(set(hull.simplices.flatten()))
    remap = {old: k for k, old in enumerate(used)}
    v = pts[used]
    faces = [[remap[int(a)] for a in tri] for tri in hull.simplices]
    vol, cen, Im = polyhedron_mass_properties(v, faces)
    w, R = np.linalg.eigh(Im)
    if np.linalg.det(R) < 0:
        R[:, 0] = -R[:, 0]
    v = (v - cen) @ R
    # orient faces outward (for tidiness; the engine is orientation-agnostic)
    for f in faces:
        n = np.cross(v[f[1]] - v[f[0]], v[f[2]] - v[f[0]])
        if np.dot(n, v[f[0]]) < 0:
            f[1], f[2] = f[2], f[1]
    edges = set()
    for f in faces:
        for k in range(len(f)):
            a, b = f[k], f[(k + 1) % len(f)]
            edges.add((min(a, b), max(a, b)))
    lo, hi = v.min(axis=0), v.max(axis=0)
    center = 0.5 * (lo + hi)
    ext = 0.5 * (hi - lo) + radius
    return Shape(name, radius, v, sorted(edges), faces, volume=vol, inertia_mass=tuple(float(x) for x in w),
                 obb_extent=tuple(float(x) for x in ext), obb_center=tuple(float(x) for x in center))


@dataclass
class Body:
    shape: str
    group: int = 0
    cluster: int = 0
    homothety: float = 1.0
    pos: Sequence[float] = (0.0, 0.0, 0.0)
    vel: Sequence[float] = (0.0, 0.0, 0.0)
    acc: Sequence[float] = (0.0, 0.0, 0.0)
    Q: Sequence[float] = (1.0, 0.0, 0.0, 0.0)
    vrot: Sequence[float] = (0.0, 0.0, 0.0)
    arot: Sequence[float] = (0.0, 0.0, 0.0)


@dataclass
class Deck:
    shape_file: str = "shapes.shp"
    t: float = 0.0
    tmax: float = 1.0
    dt: float = 1e-6
    interVerlet: float = 1e-4
    interConf: float = 1.0
    DVerlet: float = 0.0
    dVerlet: float = 0.0
    density: dict = field(default_factory=lambda: {0: 2700.0})
    gravity: Sequence[float] = (0.0, 0.0, -9.81)
    force_law: str = "Default"
    update_nl: str = "bruteForce"
    cell_min_sizes: Sequence[float] = (1.0, 1.0, 1.0)
    nDriven: int = 0
    params: list = field(default_factory=list)   # (name, g1, g2, value)
    bodies: List[Body] = field(default_factory=list)
    interactions: list = field(default_factory=list)  # raw record tuples
    interfaces: list = field(default_factory=list)    # (i, j, dn0, params or None, [(type,isub,jsub)])
    params_in_interfaces: int = 0
    precision: int = 16
    iconf: int = 0
    extra: list = field(default_factory=list)         # extra raw lines
    body_force: Optional[str] = None
    dynamic_update_nl: int = 0
    disp_update_nl: float = 1.0
    angle_update_nl: float = 1.0


def _r(x):
    return repr(float(x))


def write_conf(path, deck: Deck):
    """A conf deck readable by Rockable::loadConf (Rockable.cpp:1085-1145)."""
    o = ["Rockable 21-08-2022"]
    o.append("AddOrRemoveInteractions bruteForce")
    o.append("UpdateNL " + deck.update_nl)
    o.append("forceLaw " + deck.force_law)
    o.append("t %s" % _r(deck.t))
    o.append("tmax %s" % _r(deck.tmax))
    o.append("dt %s" % _r(deck.dt))
    o.append("interVerlet %s" % _r(deck.interVerlet))
    o.append("interConf %s" % _r(deck.interConf))
    o.append("DVerlet %s" % _r(deck.DVerlet))
    o.append("dVerlet %s" % _r(deck.dVerlet))
    for g, d in sorted(deck.density.items()):
        o.append("density %d %s" % (g, _r(d)))
    o.append("gravity %s %s %s" % tuple(_r(x) for x in deck.gravity))
    if deck.body_force:
        o.append("BodyForce " + deck.body_force)
    o.append("ParamsInInterfaces %d" % deck.params_in_interfaces)
    o.append("dynamicUpdateNL %d" % deck.dynamic_update_nl)
    if deck.dynamic_update_nl:
        o.append("dispUpdateNL %s" % _r(deck.disp_update_nl))
        o.append("angleUpdateNL %s" % _r(deck.angle_update_nl))
    for name, g1, g2, v in deck.params:
        o.append("%s %d %d %s" % (name, g1, g2, _r(v)))
    o.append("cellMinSizes %s %s %s" % tuple(_r(x) for x in deck.cell_min_sizes))
    o.append("iconf %d" % deck.iconf)
    o.append("nDriven %d" % deck.nDriven)
    o.extend(deck.extra)
    o.append("shapeFile " + deck.shape_file)
    o.append("precision %d" % deck.precision)
    o.append("Particles %d" % len(deck.bodies))
    for b in deck.bodies:
        vals = [b.homothety] + list(b.pos) + list(b.vel) + list(b.acc) + list(b.Q) + list(b.vrot) + list(b.arot)
        o.append("%s %d %d %s" % (b.shape, b.group, b.cluster, " ".join(_f(x) for x in vals)))
    o.append("Interactions %d" % len(deck.interactions))
    for rec in deck.interactions:
        o.append(" ".join(str(int(x)) for x in rec[:5]) + " " + " ".join(_f(x) for x in rec[5:]))
    o.append("Interfaces %d" % len(deck.interfaces))
    for (i, j, dn0, par, bonds) in deck.interfaces:
        line = "%d %d %d %s" % (i, j, len(bonds), _f(dn0))
        if deck.params_in_interfaces:
            line += " " + " ".join(_f(x) for x in par)
        for (ty, a, b) in bonds:
            line += " %d %d %d" % (ty, a, b)
        o.append(line)
    with open(path, "w") as fh:
        fh.write("\n".join(o) + "\n")


def contact_params(kn, kt, en2, mu, kr=0.0, mur=0.0, groups=((0, 0),)):
    p = []
    for g1, g2 in groups:
        p += [("knContact", g1, g2, kn), ("ktContact", g1, g2, kt), ("en2Contact", g1, g2, en2),
              ("muContact", g1, g2, mu), ("krContact", g1, g2, kr), ("murContact", g1, g2, mur)]
    return p
