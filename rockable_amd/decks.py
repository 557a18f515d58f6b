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
    box_for_linkcells_opt: int = 0
    dynamic_update_nl: int = 0
    disp_update_nl: float = 1.0
    angle_update_nl: float = 1.0


def _r(x):
    return repr(float(x))


def _body_block_text(block):
    """Particles lines of many bodies at once: block holds arrays "names"
    (one per body, or a single str), "group", "h", "pos", "vel", "Q" (and
    optionally "cluster").  Columns holding only exact integers are written
    with %d (read back exactly), the others with %.16e."""
    import io
    n = len(block["pos"])
    z3 = np.zeros((n, 3))
    cols = np.column_stack([np.asarray(block["group"], float),
                            np.asarray(block.get("cluster", np.zeros(n)), float),
                            np.asarray(block["h"], float), np.asarray(block["pos"], float),
                            np.asarray(block["vel"], float), z3, np.asarray(block["Q"], float), z3, z3])
    fmt = []
    for k in range(cols.shape[1]):
        c = cols[:, k]
        fmt.append("%d" if (k < 2 or (np.all(c == np.round(c)) and np.abs(c).max(initial=0) < 2 ** 52)) else "%.16e")
    buf = io.StringIO()
    np.savetxt(buf, cols, fmt=fmt, delimiter=" ")
    lines = buf.getvalue().rstrip("\n").split("\n")
    names = block["names"]
    if isinstance(names, str):
        return "\n".join(names + " " + ln for ln in lines)
    return "\n".join(nm + " " + ln for nm, ln in zip(names, lines))


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
    o.append("boxForLinkCellsOpt %d" % deck.box_for_linkcells_opt)
    o.append("iconf %d" % deck.iconf)
    o.append("nDriven %d" % deck.nDriven)
    o.extend(deck.extra)
    o.append("shapeFile " + deck.shape_file)
    o.append("precision %d" % deck.precision)
    block = getattr(deck, "body_block", None)
    nblock = 0 if block is None else len(block["pos"])
    o.append("Particles %d" % (len(deck.bodies) + nblock))
    for b in deck.bodies:
        vals = [b.homothety] + list(b.pos) + list(b.vel) + list(b.acc) + list(b.Q) + list(b.vrot) + list(b.arot)
        o.append("%s %d %d %s" % (b.shape, b.group, b.cluster, " ".join(_f(x) for x in vals)))
    if nblock:
        o.append(_body_block_text(block))
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


# ---------------------------------------------------------------------------
# seeded synthetic workloads
# ---------------------------------------------------------------------------
def box_walls(lo, hi, radius, names=("WALLZ0", "WALLX0", "WALLX1", "WALLY0", "WALLY1"), top=False):
    """Driven square plates bounding the box [lo, hi]: floor (and lid) plus four
    side walls, each a 4-vertex face (its quaternion turns the body xy plane
    onto the wall plane).  Returns (shapes, bodies)."""
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    c = 0.5 * (lo + hi)
    L = hi - lo
    half = 0.5 * float(L.max()) + 2 * radius
    shp = square_plane_shape("WALL", half, radius)
    s2 = np.sqrt(0.5)
    q_x = (s2, 0.0, s2, 0.0)     # 90 deg about y: body z -> global x
    q_y = (s2, -s2, 0.0, 0.0)    # -90 deg about x: body z -> global y
    bodies = [Body("WALL", group=1, pos=(c[0], c[1], lo[2] - radius))]
    bodies += [Body("WALL", group=1, pos=(lo[0] - radius, c[1], c[2]), Q=q_x),
               Body("WALL", group=1, pos=(hi[0] + radius, c[1], c[2]), Q=q_x),
               Body("WALL", group=1, pos=(c[0], lo[1] - radius, c[2]), Q=q_y),
               Body("WALL", group=1, pos=(c[0], hi[1] + radius, c[2]), Q=q_y)]
    if top:
        bodies.append(Body("WALL", group=1, pos=(c[0], c[1], hi[2] + radius)))
    return [shp], bodies


def sphere_lattice_deck(n_side=8, R=0.5e-3, spacing_factor=2.2, seed=1, vmax=0.01, kn=1e6, en=0.5, mu=0.4,
                        dt=1e-6, steps=10000, nl="bruteForce", dims=None):
    """Plain method, CPU-runnable case: n_side^3 single-vertex spheres of
    radius R on a lattice of spacing spacing_factor * R, dropping under gravity
    into a fixed floor and four walls (5 driven plates first), velocities
    uniform in [-vmax, vmax] from the seed; Default law.  dims = (nx, ny, nz)
    replaces the cubic lattice by an nx x ny x nz one."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = dims if dims is not None else (n_side, n_side, n_side)
    a = spacing_factor * R
    gap = a - 2 * R
    lo = np.array([0.0, 0.0, 0.0])
    hi = np.array([nx * a, ny * a, nz * a + 4 * a])
    wall_shapes, walls = box_walls(lo, hi, 0.1 * R)
    shapes = wall_shapes + [sphere_shape("SPHERE", R)]
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=100 * dt,
                interConf=10.0, DVerlet=0.2 * R, dVerlet=0.1 * R, density={0: 2700.0, 1: 2700.0},
                gravity=(0.0, 0.0, -9.81), update_nl=nl, nDriven=len(walls),
                params=contact_params(kn, 0.8 * kn, en * en, mu, groups=((0, 0), (0, 1))))
    deck.bodies = list(walls)
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                p = lo + np.array([R + 0.5 * gap + ix * a, R + 0.5 * gap + iy * a, R + 0.5 * gap + iz * a])
                v = rng.uniform(-vmax, vmax, size=3)
                deck.bodies.append(Body("SPHERE", group=0, pos=tuple(p), vel=tuple(v)))
    deck.cell_min_sizes = (2.0 * a, 2.0 * a, 2.0 * a)
    return shapes, deck


def hull_shape_pool(n_shapes, seed, nmin=8, nmax=14, axes=(1.0, 0.8, 0.6), radius_frac=0.05, min_volume_frac=0.25):
    """Unit-size convex sphero-polyhedra: hull of nmin..nmax random points on an
    ellipsoid of semi-axes 0.5 * axes, Minkowski radius radius_frac.  Hulls
    whose volume is below min_volume_frac of the ellipsoid's are drawn again:
    such slivers have a tenth of the median rotational inertia, and with
    several stiff contacts (kn = 1e7 N/m) their rotation outruns an explicit
    step of dt = 2e-7 s."""
    rng = np.random.default_rng(seed)
    vmin = min_volume_frac * (4.0 / 3.0) * np.pi * np.prod(0.5 * np.asarray(axes))
    out = []
    for k in range(n_shapes):
        n = int(rng.integers(nmin, nmax + 1))
        while True:
            u = rng.normal(size=(n, 3))
            u /= np.linalg.norm(u, axis=1)[:, None]
            pts = u * (0.5 * np.asarray(axes))
            try:
                s = hull_shape("HULL%03d" % k, pts, radius_frac)
            except Exception:
                continue
            if s.volume >= vmin:
                out.append(s)
                break
    return out


def random_quaternions(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1)[:, None]
    q[q[:, 0] < 0] *= -1
    return q


def small_rotations(rng, n, max_angle):
    """Quaternions of rotations by angles uniform in [0, max_angle] about
    random axes."""
    ax = rng.normal(size=(n, 3))
    ax /= np.linalg.norm(ax, axis=1)[:, None]
    ang = rng.uniform(0.0, max_angle, size=n)
    return np.concatenate([np.cos(0.5 * ang)[:, None], np.sin(0.5 * ang)[:, None] * ax], axis=1)


def hull_packing_deck(n_bodies=50000, seed=2, n_shapes=64, size_range=(4e-3, 6e-3), box_in_sizes=(40, 40, 200),
                      kn=1e7, en=0.3, mu=0.6, dt=2e-7, steps=20000, vmax=0.5, law="Avalanche", gap_factor=1.05,
                      interverlet_steps=200, lattice=None, aligned=False, tilt_deg=10.0):
    """Plain GPU form: convex sphero-polyhedra (hulls of 8-14 points on an
    ellipsoid 1 : 0.8 : 0.6, size uniform in size_range, Minkowski radius 5 %
    of the size) inside a box of box_in_sizes mean sizes under gravity, with
    velocities uniform in [-vmax, vmax]; linkCells neighbour list; the given
    force law.  Bodies sit on a lattice: with aligned=False each gets a random
    orientation and the spacing is the largest circumscribed diameter times
    gap_factor; with aligned=True (a compact pour) each body keeps its principal
    axes along the lattice axes up to a random tilt of at most tilt_deg, and the
    spacing along each axis is the largest extent along it times gap_factor."""
    rng = np.random.default_rng(seed)
    pool = hull_shape_pool(n_shapes, seed)
    smean = 0.5 * (size_range[0] + size_range[1])
    box = np.array(box_in_sizes, float) * smean
    if aligned:
        ext = np.max([np.asarray(s.extent()) for s in pool], axis=0)
        t = np.radians(tilt_deg)
        tilted = np.array([ext[0] * np.cos(t) + max(ext[1], ext[2]) * np.sin(t),
                           ext[1] * np.cos(t) + max(ext[0], ext[2]) * np.sin(t),
                           ext[2] * np.cos(t) + max(ext[0], ext[1]) * np.sin(t)])
        a3 = gap_factor * 2.0 * tilted * size_range[1]
    else:
        bound = max(float(np.linalg.norm(np.asarray(s.vertex), axis=1).max()) + s.radius for s in pool)
        a3 = np.full(3, gap_factor * 2.0 * bound * size_range[1])
    nx, ny = max(1, int(box[0] // a3[0])), max(1, int(box[1] // a3[1]))
    nz = int(np.ceil(n_bodies / (nx * ny)))
    if lattice is not None:
        nx, ny, nz = lattice
    lo = np.zeros(3)
    hi = np.array([nx * a3[0], ny * a3[1], max(box[2], nz * a3[2])])
    wall_shapes, walls = box_walls(lo, hi, 0.05 * smean)
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=interverlet_steps * dt,
                interConf=1e3, DVerlet=0.1 * smean, dVerlet=0.05 * smean, density={0: 2700.0, 1: 2700.0},
                gravity=(0.0, 0.0, -9.81), update_nl="linkCells", nDriven=len(walls), force_law=law,
                params=contact_params(kn, 0.8 * kn, en * en, mu, groups=((0, 0), (0, 1))))
    amax = float(a3.max())
    deck.cell_min_sizes = (amax, amax, amax)
    deck.box_for_linkcells_opt = 1   # cell grid over the free bodies; the walls are oversized
    deck.bodies = list(walls)
    q = small_rotations(rng, n_bodies, np.radians(tilt_deg)) if aligned else random_quaternions(rng, n_bodies)
    k = 0
    for iz in range(nz):
        for iy in range(ny):
            for ix in range(nx):
                if k >= n_bodies:
                    break
                p = lo + a3 * (np.array([ix, iy, iz]) + 0.5)
                size = rng.uniform(*size_range)
                v = rng.uniform(-vmax, vmax, size=3)
                deck.bodies.append(Body(pool[int(rng.integers(len(pool)))].name, group=0, homothety=size,
                                        pos=tuple(p), vel=tuple(v), Q=tuple(q[k])))
                k += 1
    return wall_shapes + pool, deck


def quat_matrix(q):
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def compact_pour_deck(n_bodies=50000, seed=2, n_shapes=64, size_range=(4e-3, 6e-3), box_in_sizes=(40, 40, 200),
                      kn=1e7, en=0.3, mu=0.6, dt=2e-7, steps=20000, vmax=0.5, law="Avalanche", gap_frac=0.02,
                      tilt_deg=15.0, interverlet_steps=100, width_in_sizes=None):
    """Plain GPU form as a compact pour: the convex sphero-polyhedra of
    hull_packing_deck (hulls of 8-14 points on an ellipsoid 1 : 0.8 : 0.6,
    size uniform in size_range, Minkowski radius 5 %), principal axes along
    the box axes up to a random tilt of at most tilt_deg, sorted by size
    (largest at the bottom) and packed row by row, layer by layer from the
    floor of a box of box_in_sizes mean sizes, each body's own rotated
    bounding box apart from its neighbours by gap_frac of the mean size;
    velocities uniform in [-vmax, vmax]; gravity; linkCells; the given law."""
    rng = np.random.default_rng(seed)
    pool = hull_shape_pool(n_shapes, seed)
    smean = 0.5 * (size_range[0] + size_range[1])
    box = np.array(box_in_sizes, float) * smean
    if width_in_sizes is not None:
        box[0] = box[1] = width_in_sizes * smean
    gap = gap_frac * smean
    sizes = np.sort(rng.uniform(*size_range, size=n_bodies))[::-1]
    shape_idx = rng.integers(len(pool), size=n_bodies)
    q = small_rotations(rng, n_bodies, np.radians(tilt_deg))
    vel = rng.uniform(-vmax, vmax, size=(n_bodies, 3))
    pos = np.zeros((n_bodies, 3))
    x = y = z = 0.0
    row_h = 0.0   # max y extent in the current row
    lay_h = 0.0   # max z extent in the current layer
    row_start = 0
    layer_start = 0
    placed_rows = []
    for k in range(n_bodies):
        s = pool[shape_idx[k]]
        R = quat_matrix(q[k])
        e = np.asarray(s.extent()) * sizes[k]
        c = R @ (np.asarray(s.obb_center) * sizes[k])
        half = np.abs(R) @ e
        if x + 2 * half[0] > box[0]:
            x = 0.0
            y += row_h + gap
            row_h = 0.0
        if y + 2 * half[1] > box[1]:
            y = 0.0
            z += lay_h + gap
            lay_h = 0.0
            row_h = 0.0
        pos[k] = np.array([x + half[0], y + half[1], z + half[2]]) - c
        x += 2 * half[0] + gap
        row_h = max(row_h, 2 * half[1])
        lay_h = max(lay_h, 2 * half[2])
    top = z + lay_h
    lo = np.zeros(3)
    hi = np.array([box[0], box[1], max(box[2], top)])
    wall_shapes, walls = box_walls(lo, hi, 0.05 * smean)
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=interverlet_steps * dt,
                interConf=1e3, DVerlet=0.1 * smean, dVerlet=0.05 * smean, density={0: 2700.0, 1: 2700.0},
                gravity=(0.0, 0.0, -9.81), update_nl="linkCells", nDriven=len(walls), force_law=law,
                params=contact_params(kn, 0.8 * kn, en * en, mu, groups=((0, 0), (0, 1))))
    cell = 1.1 * (2 * max(float(np.max(s.extent())) for s in pool) * size_range[1]) + 0.1 * smean
    deck.cell_min_sizes = (cell, cell, cell)
    deck.box_for_linkcells_opt = 1
    deck.bodies = list(walls)
    for k in range(n_bodies):
        deck.bodies.append(Body(pool[shape_idx[k]].name, group=0, homothety=float(sizes[k]), pos=tuple(pos[k]),
                                vel=tuple(vel[k]), Q=tuple(q[k])))
    return wall_shapes + pool, deck


def quat_matrices(q):
    """Rotation matrices of an (n, 4) array of unit quaternions (w, x, y, z)."""
    w, x, y, z = (q[:, k] for k in range(4))
    R = np.empty((len(q), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - w * z)
    R[:, 0, 2] = 2 * (x * z + w * y)
    R[:, 1, 0] = 2 * (x * y + w * z)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - w * x)
    R[:, 2, 0] = 2 * (x * z - w * y)
    R[:, 2, 1] = 2 * (y * z + w * x)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def compact_layout(pool, shape_idx, sizes, q, floor, gap):
    """Positions packing bodies row by row (along x), layer by layer from
    z = 0, each body's rotated OBB box `gap` apart from its neighbours, on a
    floor of floor = (Lx, Ly).  Returns (pos, top of the last layer)."""
    ext = np.array([s.extent() for s in pool])[shape_idx] * sizes[:, None]
    cen = np.array([s.obb_center for s in pool])[shape_idx] * sizes[:, None]
    R = quat_matrices(q)
    c = np.einsum("nij,nj->ni", R, cen)
    half = np.einsum("nij,nj->ni", np.abs(R), ext)
    pos = np.zeros((len(sizes), 3))
    x = y = z = row_h = lay_h = 0.0
    Lx, Ly = floor
    hx, hy, hz = (half[:, k].tolist() for k in range(3))
    for k in range(len(sizes)):
        if x + 2 * hx[k] > Lx:
            x = 0.0
            y += row_h + gap
            row_h = 0.0
        if y + 2 * hy[k] > Ly:
            y = 0.0
            z += lay_h + gap
            lay_h = row_h = 0.0
        pos[k] = (x + hx[k], y + hy[k], z + hz[k])
        x += 2 * hx[k] + gap
        row_h = max(row_h, 2 * hy[k])
        lay_h = max(lay_h, 2 * hz[k])
    return pos - c, z + lay_h


def avalanche_deck(n_bodies=1000000, seed=3, n_shapes=64, size_range=(4e-3, 6e-3), floor_in_sizes=(400, 100),
                   slope_deg=30.0, kn=1e7, en=0.3, mu=0.6, dt=2e-7, steps=20000, vmax=0.5, flow_speed=1.0,
                   law="Avalanche", gap_frac=0.02, tilt_deg=15.0, disp_frac=0.25, angle_deg=57.29577951308232):
    """Learned-rebuild workload: a polyhedral avalanche of the plain GPU form's
    bodies (hulls of 8-14 points on an ellipsoid 1 : 0.8 : 0.6, size uniform
    in size_range, Minkowski radius 5 %) packed compactly on the floor of a
    box floor_in_sizes mean sizes long (x) and wide (y), principal axes along
    the box axes up to tilt_deg, sent down a slope of slope_deg (gravity
    tilted about y, downslope along +x) at flow_speed plus velocities uniform
    in [-vmax, vmax]; linkCells; the neighbour list rebuilt by the
    displacement check only (dynamicUpdateNL with dispUpdateNL = disp_frac
    DVerlet and angleUpdateNL = angle_deg, by default 1 rad as in Rockable;
    interVerlet beyond the run)."""
    rng = np.random.default_rng(seed)
    pool = hull_shape_pool(n_shapes, seed)
    smean = 0.5 * (size_range[0] + size_range[1])
    floor = np.array(floor_in_sizes, float) * smean
    sizes = np.sort(rng.uniform(*size_range, size=n_bodies))[::-1]
    shape_idx = rng.integers(len(pool), size=n_bodies)
    q = small_rotations(rng, n_bodies, np.radians(tilt_deg))
    vel = rng.uniform(-vmax, vmax, size=(n_bodies, 3))
    vel[:, 0] += flow_speed
    pos, top = compact_layout(pool, shape_idx, sizes, q, floor, gap_frac * smean)
    wall_shapes, walls = box_walls(np.zeros(3), np.array([floor[0], floor[1], top + 4 * size_range[1]]),
                                   0.05 * smean)
    th = np.radians(slope_deg)
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=10.0 * steps * dt,
                interConf=1e3, DVerlet=0.1 * smean, dVerlet=0.05 * smean, density={0: 2700.0, 1: 2700.0},
                gravity=(9.81 * np.sin(th), 0.0, -9.81 * np.cos(th)), update_nl="linkCells", nDriven=len(walls),
                force_law=law, params=contact_params(kn, 0.8 * kn, en * en, mu, groups=((0, 0), (0, 1))))
    cell = 1.1 * (2 * max(float(np.max(s.extent())) for s in pool) * size_range[1]) + 0.1 * smean
    deck.cell_min_sizes = (cell, cell, cell)
    deck.box_for_linkcells_opt = 1
    deck.dynamic_update_nl = 1
    deck.disp_update_nl = disp_frac * deck.DVerlet
    deck.angle_update_nl = angle_deg
    deck.bodies = list(walls)
    names = np.array([s.name for s in pool])[shape_idx]
    deck.body_block = {"names": names, "group": np.zeros(n_bodies), "h": sizes, "pos": pos, "vel": vel, "Q": q}
    return wall_shapes + pool, deck


def cube_shape(name, side, radius):
    """A sphero-cube: cube of the given side (vertex to vertex) in its body
    frame, 8 vertices, 12 edges, 6 square faces, Minkowski radius; volume of
    the sphero-cube, inertia per unit mass of the cube of side side + 2 radius."""
    h = 0.5 * side
    v = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)])
    edges = sorted({(a, b) for a in range(8) for b in range(a + 1, 8)
                    if np.count_nonzero(v[a] != v[b]) == 1})
    faces = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]
    r = radius
    vol = side ** 3 + 6 * side ** 2 * r + 3 * np.pi * side * r ** 2 + 4.0 / 3.0 * np.pi * r ** 3
    im = (side + 2 * r) ** 2 / 6.0
    return Shape(name, radius, v, edges, faces, volume=vol, inertia_mass=(im, im, im),
                 obb_extent=(h + r, h + r, h + r))


def bonded_clusters_deck(n_clusters=2000, cluster_dims=(5, 5, 4), side=5e-3, radius_frac=0.05, seed=3,
                         strength_mean=50.0, strength_cv=0.3, speed=2.0, plane_deg=35.0, clearance=2e-3,
                         kn=1e7, en2_bond=0.25, dt=1e-7, steps=10000, nl="linkCells", cluster_gap=2e-3):
    """Breakable-bonds workload: n_clusters clusters of prod(cluster_dims)
    sphero-cubes (side, Minkowski radius radius_frac side) stacked face to
    face, each touching pair of a cluster bonded at its coincident vertex
    pairs (the stickVerticesInClusters rule: vertices closer than
    Ri + Rj + epsilon, here 1 % of the radius; one StickedLinks interface
    per pair, its vertex-vertex bonds at rest), the interface strengths fn0
    and ft0 drawn lognormal (mean strength_mean N, coefficient of variation
    strength_cv) and kept in the interfaces (ParamsInInterfaces 1; kn, kt =
    0.8 kn, no rotational stiffness, mom0 1e10, power 2).  The clusters sit
    `clearance` above a rigid plane inclined at plane_deg (body 0, driven)
    and fall onto it at `speed` under gravity; contacts follow the
    StickedLinks contact branch (kn, kt = 0.8 kn, en2 0.25, mu 0.5).
    Returns (shapes, deck)."""
    rng = np.random.default_rng(seed)
    r = radius_frac * side
    pitch = side + 2 * r
    cube = cube_shape("CUBE", side, r)
    nx, ny, nz = cluster_dims
    cl = np.array([nx, ny, nz], float) * pitch
    ncx = int(np.ceil(np.sqrt(n_clusters)))
    ncy = int(np.ceil(n_clusters / ncx))
    span = np.array([ncx * (cl[0] + cluster_gap), ncy * (cl[1] + cluster_gap)])
    th = np.radians(plane_deg)
    half = 0.5 * float(max(span[0] / np.cos(th), span[1])) + 2 * cl.max()
    plane = square_plane_shape("PLANE", half, r)
    qp = (float(np.cos(th / 2)), 0.0, float(np.sin(th / 2)), 0.0)   # plane normal (sin th, 0, cos th)
    nrm = np.array([np.sin(th), 0.0, np.cos(th)])
    bodies = [Body("PLANE", group=1, pos=(0.0, 0.0, 0.0), Q=qp)]
    # cube centres of one cluster, relative to its lowest corner
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    local = (np.column_stack([ix.ravel(), iy.ravel(), iz.ravel()]) + 0.5) * pitch
    vtx = np.asarray(cube.vertex)
    interactions, interfaces = [], []
    meff = 0.5 * 2700.0 * cube.volume
    logen = 0.5 * np.log(en2_bond)
    damp = -logen / np.sqrt(logen * logen + np.pi ** 2) * 2.0 * np.sqrt(kn * meff)
    s2 = np.log(1.0 + strength_cv ** 2)
    mu_ln = np.log(strength_mean) - 0.5 * s2
    eps = 0.01 * r
    # the bonds of one cluster (every cluster has the same): touching pairs
    # (a, b) in order, and per pair its vertex pairs (iv, jv) closer than
    # 2 r + eps
    tpl = []
    for a in range(len(local)):
        for b in range(a + 1, len(local)):
            if np.abs(local[a] - local[b]).sum() > pitch * 1.5:
                continue
            gi, gj = local[a] + vtx, local[b] + vtx
            vv = [(iv, int(jv)) for iv in range(8)
                  for jv in np.nonzero(np.linalg.norm(gj - gi[iv], axis=1) < eps + 2 * r)[0]]
            if vv:
                tpl.append((a, b, vv))
    ta = np.array([a for a, b, vv in tpl for _ in vv])
    tb = np.array([b for a, b, vv in tpl for _ in vv])
    tiv = np.array([iv for a, b, vv in tpl for iv, jv in vv])
    tjv = np.array([jv for a, b, vv in tpl for iv, jv in vv])
    tbonds = [[(0, iv, jv) for iv, jv in vv] for a, b, vv in tpl]
    zeros = (0.0, 0.0, 0.0) + (0.0,) + (0.0,) * 6 + (damp,)
    for c in range(n_clusters):
        x0 = (c % ncx) * (cl[0] + cluster_gap) - 0.5 * span[0]
        y0 = (c // ncx) * (cl[1] + cluster_gap) - 0.5 * span[1]
        # lowest corner of the cluster's box just `clearance` above the plane
        corners = np.array([[x0 + a * cl[0], y0 + b * cl[1], 0.0] for a in (0, 1) for b in (0, 1)])
        z0 = max(-np.dot(p, nrm) / nrm[2] for p in corners) + (clearance + r) / nrm[2]
        first = len(bodies)
        pos = np.array([x0, y0, z0]) + local
        for p in pos:
            bodies.append(Body("CUBE", group=0, cluster=c, pos=tuple(p), vel=(0.0, 0.0, -speed)))
        # the bonded vertex pairs at rest: normal, gap and contact point
        gi = pos[ta] + vtx[tiv]
        n = gi - (pos[tb] + vtx[tjv])
        dist = np.sqrt((n * n).sum(axis=1))
        n = n / dist[:, None]
        dn = dist - 2 * r
        cp = gi - n * (r + 0.5 * dn)[:, None]
        for k in range(len(ta)):
            interactions.append((first + int(ta[k]), first + int(tb[k]), 0, int(tiv[k]), int(tjv[k]))
                                + tuple(n[k].tolist()) + (float(dn[k]),) + tuple(cp[k].tolist()) + zeros)
        for (a, b, vv), bonds in zip(tpl, tbonds):
            fn0, ft0 = np.exp(mu_ln + np.sqrt(s2) * rng.standard_normal(2))
            par = (kn, 0.8 * kn, 0.0, float(fn0), float(ft0), 1e10, 2.0, 0.0)
            interfaces.append((first + a, first + b, 0.0, par, list(bonds)))
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=100 * dt, interConf=1e3,
                DVerlet=0.1 * side, dVerlet=0.05 * side, density={0: 2700.0, 1: 2700.0},
                gravity=(0.0, 0.0, -9.81), update_nl=nl, nDriven=1, force_law="StickedLinks",
                params=contact_params(kn, 0.8 * kn, 0.25, 0.5, groups=((0, 0), (0, 1))))
    cell = 2.0 * pitch
    deck.cell_min_sizes = (cell, cell, cell)
    deck.box_for_linkcells_opt = 1
    deck.bodies = bodies
    interactions.sort(key=lambda t: t[:5])
    deck.interactions = interactions
    deck.interfaces = interfaces
    deck.params_in_interfaces = 1
    return [plane, cube], deck


def write_deck(dirpath, shapes, deck, name="deck.conf"):
    import os
    os.makedirs(dirpath, exist_ok=True)
    write_shp(os.path.join(dirpath, deck.shape_file), shapes)
    write_conf(os.path.join(dirpath, name), deck)
    return os.path.join(dirpath, name)


def conf_with(src, dst, **header):
    """Copy a conf file, replacing header keywords (the lines before
    "Particles"): scalar keywords take a value or a tuple (gravity), pair
    parameters are given as name=[(g1, g2, value), ...]."""
    with open(src) as fh:
        lines = fh.read().split("\n")
    out = []
    in_header = True
    for ln in lines:
        tok = ln.split()
        if in_header and tok and tok[0] == "Particles":
            in_header = False
        if in_header and tok and tok[0] in header:
            v = header[tok[0]]
            if isinstance(v, list):
                for g1, g2, x in v:
                    if int(tok[1]) == g1 and int(tok[2]) == g2:
                        ln = "%s %d %d %s" % (tok[0], g1, g2, _r(x))
            elif isinstance(v, (tuple, np.ndarray)):
                ln = tok[0] + " " + " ".join(_r(x) for x in v)
            else:
                ln = "%s %s" % (tok[0], v if isinstance(v, str) else _r(v))
        out.append(ln)
    with open(dst, "w") as fh:
        fh.write("\n".join(out))
    return dst


def conf_drop_bodies(src, dst, keep):
    """Copy a conf file without the bodies where keep is False: their
    Particles lines go, interactions and interfaces that involve them go, and
    the body indices of the rest are renumbered."""
    keep = np.asarray(keep, bool)
    new_index = np.cumsum(keep) - 1
    with open(src) as fh:
        lines = fh.read().split("\n")
    out = []
    k = 0
    while k < len(lines):
        ln = lines[k]
        tok = ln.split()
        if tok and tok[0] in ("Particles", "Interactions", "Interfaces") and len(tok) == 2:
            n = int(tok[1])
            k += 1
            head, body = [], []
            while k < len(lines) and lines[k].startswith("#"):  # column comments
                head.append(lines[k])
                k += 1
            for q in range(n):
                ln = lines[k + q]
                if tok[0] == "Particles":
                    if keep[q]:
                        body.append(ln)
                    continue
                t = ln.split()
                i, j = int(t[0]), int(t[1])
                if keep[i] and keep[j]:
                    t[0], t[1] = str(new_index[i]), str(new_index[j])
                    body.append(" ".join(t))
            out.append("%s %d" % (tok[0], len(body)))
            out.extend(head)
            out.extend(body)
            k += n
            continue
        out.append(ln)
        k += 1
    with open(dst, "w") as fh:
        fh.write("\n".join(out))
    return dst


def batched_sphere_packings(n_packings=256, n_range=(500, 20000), seed=4, R=0.5e-3, spacing_factor=2.2, kn=1e6,
                            mu_range=(0.2, 0.8), en_range=(0.1, 0.9), dt=1e-6, steps=5000, vmax=0.01,
                            interverlet_steps=100):
    """Batched form: n_packings independent packings of the plain CPU-runnable
    configuration (spheres of radius R on a lattice of spacing
    spacing_factor R dropping into a floor and four walls), of uneven sizes
    (body counts log-uniform in n_range), each with its own Default-law
    parameters (mu uniform in mu_range, en in en_range, en2Contact = en^2)
    as group pair (p, p).  The packings sit on a square grid far enough apart
    that no pair of bodies of two packings is ever near; all walls come first
    (the driven bodies, nDriven = 5 n_packings).  Returns (shapes, deck, info)
    where info lists each packing's body range, walls and parameters."""
    rng = np.random.default_rng(seed)
    a = spacing_factor * R
    counts = np.round(np.exp(rng.uniform(np.log(n_range[0]), np.log(n_range[1]), size=n_packings))).astype(int)
    mus = rng.uniform(*mu_range, size=n_packings)
    ens = rng.uniform(*en_range, size=n_packings)
    dims = []
    for n in counts:
        nx = int(np.ceil(n ** (1.0 / 3.0)))
        nz = int(np.ceil(n / (nx * nx)))
        dims.append((nx, nx, nz))
    ext = max(max(d[0], d[2] + 4) for d in dims) * a
    pitch = ext + 20 * a
    gridn = int(np.ceil(np.sqrt(n_packings)))
    wall_shapes = None
    walls, blocks, info = [], [], []
    for p in range(n_packings):
        nx, ny, nz = dims[p]
        origin = np.array([(p % gridn) * pitch, (p // gridn) * pitch, 0.0])
        lo = origin
        hi = origin + np.array([nx * a, ny * a, nz * a + 4 * a])
        ws, wb = box_walls(lo, hi, 0.1 * R)
        wall_shapes = ws
        for b in wb:
            b.group = p
        walls.append(wb)
        ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        idx = np.column_stack([ix.ravel(), iy.ravel(), iz.ravel()])[:counts[p]]
        pos = origin + R + 0.5 * (a - 2 * R) + idx * a
        vel = rng.uniform(-vmax, vmax, size=(counts[p], 3))
        blocks.append((pos, vel))
        info.append({"bodies": int(counts[p]), "mu": float(mus[p]), "en": float(ens[p])})
    shapes = wall_shapes + [sphere_shape("SPHERE", R)]
    params = []
    for p in range(n_packings):
        params += contact_params(kn, 0.8 * kn, ens[p] ** 2, mus[p], groups=((p, p),))
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=steps * dt * (1 - 1e-9), interVerlet=interverlet_steps * dt,
                interConf=1e3, DVerlet=0.2 * R, dVerlet=0.1 * R, density={p: 2700.0 for p in range(n_packings)},
                gravity=(0.0, 0.0, -9.81), update_nl="linkCells", nDriven=5 * n_packings, params=params)
    deck.bodies = [b for wb in walls for b in wb]
    first = deck.nDriven
    for p in range(n_packings):
        info[p]["first"] = int(first)
        info[p]["walls"] = list(range(5 * p, 5 * p + 5))
        first += counts[p]
    pos = np.concatenate([b[0] for b in blocks])
    vel = np.concatenate([b[1] for b in blocks])
    groups = np.concatenate([np.full(counts[p], p) for p in range(n_packings)])
    deck.body_block = {"names": "SPHERE", "group": groups, "h": np.ones(len(pos)), "pos": pos, "vel": vel,
                       "Q": np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))}
    deck.cell_min_sizes = (2.0 * a, 2.0 * a, 2.0 * a)
    deck.box_for_linkcells_opt = 1
    return shapes, deck, info


def single_packing_of(shapes, deck, info, p):
    """Packing p of a batched deck as a deck of its own (its 5 walls, then its
    spheres), with the same parameters: the batched run must reproduce it."""
    import copy
    d = copy.copy(deck)
    d.bodies = [copy.copy(b) for b in deck.bodies[5 * p:5 * p + 5]]
    f, n = info[p]["first"] - deck.nDriven, info[p]["bodies"]
    blk = deck.body_block
    d.body_block = {"names": "SPHERE", "group": blk["group"][f:f + n], "h": blk["h"][f:f + n],
                    "pos": blk["pos"][f:f + n], "vel": blk["vel"][f:f + n], "Q": blk["Q"][f:f + n]}
    d.nDriven = 5
    return shapes, d


def read_conf_bodies(path):
    """The header scalars and the Particles block of a conf file: a dict of
    header keyword -> list of tokens (the first occurrence), and arrays of
    the bodies' names, group, cluster, homothety, pos, vel, acc, Q, vrot,
    arot (Rockable.cpp:895-909 column order)."""
    with open(path) as fh:
        lines = fh.read().split("\n")
    header = {}
    k = 0
    while k < len(lines):
        tok = lines[k].split()
        k += 1
        if not tok or tok[0].startswith(("#", "!", "/")):
            continue
        if tok[0] == "Particles":
            n = int(tok[1])
            break
        header.setdefault(tok[0], tok[1:])
    rows = []
    while len(rows) < n and k < len(lines):
        ln = lines[k]
        k += 1
        if not ln.strip() or ln.lstrip().startswith("#"):
            continue
        rows.append(ln.split())
    names = np.array([r[0] for r in rows])
    num = np.array([[float(x) for x in r[1:23]] for r in rows])
    return header, {"names": names, "group": num[:, 0].astype(int), "cluster": num[:, 1].astype(int),
                    "h": num[:, 2], "pos": num[:, 3:6], "vel": num[:, 6:9], "acc": num[:, 9:12],
                    "Q": num[:, 12:16], "vrot": num[:, 16:19], "arot": num[:, 19:22]}


def slope_from_settled(settled_conf, pool, dirpath, tiles=1, slope_deg=30.0, flow_speed=1.0, runout=1.0,
                       disp_frac=0.25, angle_deg=57.29577951308232, tile_gap=None, name="slope.conf"):
    """An avalanche released from a settled bed: the free bodies of
    settled_conf (a packing settled under vertical gravity between the walls
    of a box, its walls the first nDriven bodies) tiled `tiles` times along
    x (each copy one box length plus tile_gap further), set on a floor
    tilted by slope_deg (gravity (g sin, 0, -g cos); downslope is +x) with
    flow_speed added downslope to every body's velocity.  New walls: the
    floor, the upslope wall and the two side walls of the tiled bed, and a
    downslope wall `runout` metres past its end.  The settled contacts'
    histories are dropped (the first neighbour-list update finds the
    contacts again).  The neighbour list follows the displacement check
    alone (dynamicUpdateNL, dispUpdateNL = disp_frac DVerlet, angleUpdateNL
    angle_deg), with the settled deck's law, parameters, time step and cell
    size.  pool: the free bodies' shapes (written with the walls' shape)."""
    import os
    hdr, B = read_conf_bodies(settled_conf)
    nd = int(hdr["nDriven"][0])
    wpos = B["pos"][:nd]
    wr = float(B["h"][0])  # walls: homothety 1 -> the Minkowski radius is the shape's
    free = slice(nd, len(B["names"]))
    # the settled box's inner faces: floor (0), x0 (1), x1 (2), y0 (3), y1 (4)
    walls_r = {}
    with open(os.path.join(os.path.dirname(settled_conf), hdr["shapeFile"][0])) as fh:
        shp_text = fh.read()
    for blk in shp_text.split("<")[1:]:
        t = blk.split()
        if "name" in t and "radius" in t:
            walls_r[t[t.index("name") + 1]] = float(t[t.index("radius") + 1])
    r = walls_r.get(str(B["names"][0]), 0.0) * wr
    x0 = wpos[1, 0] + r
    x1 = wpos[2, 0] - r
    y0 = wpos[3, 1] + r
    y1 = wpos[4, 1] - r
    z0 = wpos[0, 2] + r
    L = x1 - x0
    gap = float(tile_gap) if tile_gap is not None else 0.0
    n1 = len(B["names"]) - nd
    reps = lambda a: np.concatenate([a[free]] * tiles)
    pos = reps(B["pos"]).copy()
    for k in range(tiles):
        pos[k * n1:(k + 1) * n1, 0] += k * (L + gap)
    vel = reps(B["vel"]).copy()
    vel[:, 0] += flow_speed
    top = float(pos[:, 2].max()) + 0.1 * L
    hi_x = x0 + tiles * L + (tiles - 1) * gap + runout
    lo = np.array([x0, y0, z0])
    hi = np.array([hi_x, y1, top])
    c = 0.5 * (lo + hi)
    half = 0.5 * float((hi - lo).max()) + 2 * r
    wall = square_plane_shape("SLOPEWALL", half, r)
    s2 = np.sqrt(0.5)
    q_x = (s2, 0.0, s2, 0.0)
    q_y = (s2, -s2, 0.0, 0.0)
    walls = [Body("SLOPEWALL", group=1, pos=(c[0], c[1], lo[2] - r)),
             Body("SLOPEWALL", group=1, pos=(lo[0] - r, c[1], c[2]), Q=q_x),
             Body("SLOPEWALL", group=1, pos=(hi[0] + r, c[1], c[2]), Q=q_x),
             Body("SLOPEWALL", group=1, pos=(c[0], lo[1] - r, c[2]), Q=q_y),
             Body("SLOPEWALL", group=1, pos=(c[0], hi[1] + r, c[2]), Q=q_y)]
    th = np.radians(slope_deg)
    g = 9.81
    dt = float(hdr["dt"][0])
    DV = float(hdr["DVerlet"][0])
    params = []
    with open(settled_conf) as fh:
        for ln in fh:
            t = ln.split()
            if t and t[0].endswith("Contact") and len(t) == 4:
                params.append((t[0], int(t[1]), int(t[2]), float(t[3])))
            if t and t[0] == "Particles":
                break
    deck = Deck(shape_file="shapes.shp", dt=dt, tmax=1e3, interVerlet=1e3, interConf=1e3, DVerlet=DV,
                dVerlet=float(hdr["dVerlet"][0]), density={0: 2700.0, 1: 2700.0},
                gravity=(g * np.sin(th), 0.0, -g * np.cos(th)), update_nl="linkCells", nDriven=len(walls),
                force_law=hdr["forceLaw"][0], params=params)
    with open(settled_conf) as fh:
        dens = {}
        for ln in fh:
            t = ln.split()
            if t and t[0] == "density":
                dens[int(t[1])] = float(t[2])
            if t and t[0] == "Particles":
                break
    if dens:
        deck.density = dens
    cs = [float(x) for x in hdr["cellMinSizes"]]
    deck.cell_min_sizes = tuple(cs)
    deck.box_for_linkcells_opt = 1
    deck.dynamic_update_nl = 1
    deck.disp_update_nl = disp_frac * DV
    deck.angle_update_nl = angle_deg
    deck.bodies = walls
    deck.body_block = {"names": reps(B["names"]), "group": reps(B["group"]), "cluster": reps(B["cluster"]),
                       "h": reps(B["h"]), "pos": pos, "vel": vel, "Q": reps(B["Q"])}
    os.makedirs(dirpath, exist_ok=True)
    write_shp(os.path.join(dirpath, "shapes.shp"), [wall] + list(pool))
    write_conf(os.path.join(dirpath, name), deck)
    return os.path.join(dirpath, name)
