"""The narrow phase skips every sub-element test of a candidate pair whose
vertices a plane normal to the line of centres separates by more than the
tests reach: sum = Ri + Rj + dVerlet, plus dVerlet times the longest edge
(VertexIsNearEdge accepts a foot r in [-dVerlet, 1 + dVerlet] along an edge),
plus the largest offset of a face vertex from the plane of the face's first
two and last vertices (VertexIsNearFace projects on that plane), plus a
rounding margin (Particle.cpp:88-209).  The GPU kernel applies that bound
(and the same slack to single vertices, edges and faces); the GPU parity tests
check its lists bit for bit.  Here the bound itself is held against the
oracle's own tests: on random pairs of hulls, cubes and non-planar hexahedra
placed near contact, a pair the bound calls apart gets no listed interaction
from the oracle's rebuild, and a good share of the pairs are close to the
bound on either side.  A third of the pairs are cubes face to face along the
line of centres, where the bound is tight: a bound 0.04 short of this one
(the sum's dVerlet and the edge term left out) lets pairs with listed
interactions through and fails here."""
import numpy as np

from rockable_amd import decks as D
from rockable_amd.decks import Body, Deck, Shape


def hexahedron(name, lift, radius):
    """A box of unit side with one top vertex lifted: its quad faces through
    that vertex are not planar."""
    v = np.array([[x, y, z] for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)], dtype=float)
    v[7, 2] += lift
    edges = sorted({(a, b) for a in range(8) for b in range(a + 1, 8)
                    if np.count_nonzero(np.array([a, b]) >= 0) and bin(a ^ b).count("1") == 1})
    faces = [[0, 1, 3, 2], [4, 5, 7, 6], [0, 1, 5, 4], [2, 3, 7, 6], [0, 2, 6, 4], [1, 3, 7, 5]]
    ext = tuple(float(x) for x in np.abs(v).max(axis=0) + radius)
    return Shape(name, radius, v, edges, faces, volume=1.0, inertia_mass=(0.2, 0.2, 0.2), obb_extent=ext)


def reach_of(s):
    """Longest edge and largest face-vertex offset from the face's plane
    (first two and last vertices), as the kernel's per-shape table."""
    v = np.asarray(s.vertex, dtype=float)
    lmax = max(np.linalg.norm(v[b] - v[a]) for a, b in s.edge) if s.edge else 0.0
    off = 0.0
    for f in s.face:
        a, b, c = v[f[0]], v[f[1]], v[f[-1]]
        v1 = (b - a) / np.linalg.norm(b - a)
        n = np.cross(v1, c - a)
        n /= np.linalg.norm(n)
        off = max(off, max(abs(np.dot(v[k] - a, n)) for k in f))
    return lmax, off


def test_separation_bound_never_hides_a_near_test(tmp_path):
    from oracle.rk_oracle import Oracle
    rng = np.random.default_rng(11)
    shapes = D.hull_shape_pool(6, 7) + [D.cube_shape("CUBE", 1.0, 0.05), hexahedron("HEXA", 0.15, 0.05)]
    reach = [reach_of(s) for s in shapes]
    dverlet = 0.03
    apart = listed = near_bound = 0
    for trial in range(240):
        a, b = (int(x) for x in rng.integers(0, len(shapes), size=2))
        qa, qb = D.random_quaternions(rng, 2)
        u = rng.normal(size=3)
        gap = rng.uniform(-0.05, 0.35)
        if trial % 3 == 0:
            # cubes face to face along the line of centres, where the bound is
            # tight: their gap is within 0.02 of the tests' reach
            a = b = len(shapes) - 2
            qa = qb = np.array([1.0, 0.0, 0.0, 0.0])
            u = np.array([1.0, 0.0, 0.0]) + 1e-3 * rng.normal(size=3)
            gap = 0.13 + rng.uniform(-0.02, 0.02)
        va = np.asarray(shapes[a].vertex) @ D.quat_matrix(qa).T
        vb0 = np.asarray(shapes[b].vertex) @ D.quat_matrix(qb).T
        u /= np.linalg.norm(u)
        # centre distance for that gap of the vertex sets along u
        c = gap - (vb0 @ u).min() + (va @ u).max()
        pb = c * u
        vb = vb0 + pb
        deck = Deck(dt=1e-6, DVerlet=1.0, dVerlet=dverlet, gravity=(0.0, 0.0, 0.0), update_nl="bruteForce",
                    params=D.contact_params(1e4, 1e4, 0.25, 0.3))
        deck.bodies = [Body(shape=shapes[a].name, Q=tuple(float(x) for x in qa)),
                       Body(shape=shapes[b].name, pos=tuple(float(x) for x in pb), Q=tuple(float(x) for x in qb))]
        path = D.write_deck(str(tmp_path / ("t%d" % trial)), shapes, deck)
        o = Oracle(path)
        o.update_nl()
        n_listed = len(o.interactions())
        o.close()
        # the kernel's bound, restated
        axis = pb / np.linalg.norm(pb)
        hi_i, lo_j = (va @ axis).max(), (vb @ axis).min()
        ssum = shapes[a].radius + shapes[b].radius + dverlet
        xmax = max(np.abs(va).max(), np.abs(vb).max()) + ssum
        margin = 1e-9 * (1.0 + xmax + ssum)
        slack = ssum + dverlet * max(reach[a][0], reach[b][0]) + max(reach[a][1], reach[b][1]) + margin
        sep = lo_j - hi_i
        if sep > slack:
            apart += 1
            assert n_listed == 0, (trial, shapes[a].name, shapes[b].name, sep, slack, n_listed)
        listed += n_listed > 0
        near_bound += abs(sep - slack) < 0.1
    print("%d pairs apart, %d with listed interactions, %d within 0.1 of the bound" % (apart, listed, near_bound))
    assert apart >= 40 and listed >= 40 and near_bound >= 40
