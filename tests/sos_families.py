"""Point families that stop Simulation of Simplicity at a chosen depth (test data generator).

The lifted orientation of five points p0..p4 (rows sorted by point index) is resolved by
SoS through the minors of the reference's depth table (GD/GDelaunay/GDelPredDevice.h:
194-660).  Its first thirteen entries are
    1 rows 1234 xyz    2 1234 xyl    3 1234 xzl    4 1234 yzl    5 0234 xyz
    6 234 xy           7 234 xz      8 234 yz      9 0234 xyl   10 234 xl
   11 234 yl          12 0234 xzl   13 234 zl
(l = x^2 + y^2 + z^2).  Each family below makes every earlier minor vanish and its own minor
nonzero, by construction:
  1   cospherical, p1..p4 not coplanar
  2-4 all five in a plane z=c / y=c / x=c, p1..p4 not concyclic in it
  5   p1..p4 concyclic in a plane z=c, p0 on a sphere through them off that plane
  6-8 all five in a plane z=c / y=c / x=c, p1..p4 concyclic in it
  9   p1..p4 on a line along x, p0 off it with y0 != 0
 10   p1..p4 on a line along x, p0 off it in the plane y = line's y
 11   p1..p4 on a line along y, p0 off it in the plane x = line's x
 12   p1..p4 on a line along z, p0 with x0 != line's x
 13   p1..p4 on a line along z, p0 with x0 = line's x, y0 != line's y
No depth beyond 13 is reachable with distinct points: depths 6-8 vanish only when p2, p3, p4
are collinear, and then the depth 10, 11 and 13 minors are d_x, d_y, d_z times a nonzero
Vandermonde factor for the line direction d, which cannot all vanish.
Coordinates are small integers (exact in FP64); indices are distinct and sorted with the
rows, then points and indices are shuffled together (SoS sorts by index).
"""
import numpy as np

# integer points on spheres x^2 + y^2 + z^2 = r^2 (r = 5, 3 and 7 have such points)
_SPHERE = {}


def _sphere_points(r):
    if r not in _SPHERE:
        pts = [(x, y, z) for x in range(-r, r + 1) for y in range(-r, r + 1) for z in range(-r, r + 1)
               if x * x + y * y + z * z == r * r]
        _SPHERE[r] = np.array(pts, np.int64)
    return _SPHERE[r]


def _orient3d_int(a, b, c, d):
    m = np.array([b - a, c - a, d - a], dtype=object)
    return (m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0])
            + m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]))


def _incircle_int(p, q):
    """p: 4 points (k, 2) in a plane; nonzero iff not concyclic (and not collinear)."""
    rows = [[int(a), int(b), int(a) * int(a) + int(b) * int(b), 1] for a, b in p[:, q]]

    def det(m):
        if len(m) == 1:
            return m[0][0]
        return sum((-1) ** j * m[0][j] * det([r[:j] + r[j + 1:] for r in m[1:]]) for j in range(len(m)) if m[0][j])
    return det(rows)


def _shuffle(rng, p):
    idx = np.sort(rng.choice(1000, 5, replace=False)).astype(np.int32)
    order = rng.permutation(5)
    return p[order].astype(np.float64), idx[order]


def _random_plane_points(rng, axis, c, count, span=6):
    p = rng.integers(-span, span + 1, size=(count, 3))
    p[:, axis] = c
    return p


def _concyclic_in_plane(rng, axis, c, count):
    """count distinct integer points on a circle of radius 5 in the plane coordinate[axis] = c."""
    free = [a for a in range(3) if a != axis]
    circ = np.array([(x, y) for x in range(-5, 6) for y in range(-5, 6) if x * x + y * y == 25])
    sel = circ[rng.choice(circ.shape[0], count, replace=False)]
    off = rng.integers(-3, 4, size=2)
    p = np.zeros((count, 3), np.int64)
    p[:, free[0]] = sel[:, 0] + off[0]
    p[:, free[1]] = sel[:, 1] + off[1]
    p[:, axis] = c
    return p


def family(rng, depth):
    """Five points (5, 3) float64 and indices (5,) int32 whose SoS resolution stops at `depth`."""
    while True:
        if depth == 1:
            s = _sphere_points(5)
            p = s[rng.choice(s.shape[0], 5, replace=False)] + rng.integers(-3, 4, size=3)
            if _orient3d_int(p[1], p[2], p[3], p[4]) == 0:
                continue
        elif depth in (2, 3, 4):
            axis = {2: 2, 3: 1, 4: 0}[depth]
            p = _random_plane_points(rng, axis, int(rng.integers(-3, 4)), 5)
            free = [a for a in range(3) if a != axis]
            if len({tuple(r) for r in p}) < 5 or _incircle_int(p[1:], free) == 0:
                continue
        elif depth == 5:
            s = _sphere_points(5)
            ring = s[s[:, 2] == 0]
            cap = s[s[:, 2] != 0]
            p = np.concatenate([cap[rng.choice(cap.shape[0], 1)], ring[rng.choice(ring.shape[0], 4, replace=False)]])
            p = p + rng.integers(-3, 4, size=3)
        elif depth in (6, 7, 8):
            axis = {6: 2, 7: 1, 8: 0}[depth]
            c = int(rng.integers(-3, 4))
            p = np.concatenate([_random_plane_points(rng, axis, c, 1), _concyclic_in_plane(rng, axis, c, 4)])
            if len({tuple(r) for r in p}) < 5:
                continue
        else:
            # p1..p4 on an axis-parallel line, p0 placed to make depths 9..12 vanish as needed
            axis = {9: 0, 10: 0, 11: 1, 12: 2, 13: 2}[depth]
            base = rng.integers(-3, 4, size=3)
            t = np.sort(rng.choice(np.arange(-6, 7), 4, replace=False))
            line = np.repeat(base[None], 4, axis=0)
            line[:, axis] = t
            p0 = base + rng.integers(-3, 4, size=3)
            if depth == 9:
                if p0[1] == base[1]:
                    continue
            elif depth == 10:
                p0[1] = base[1]
            elif depth == 11:
                p0[0] = base[0]
            elif depth == 12:
                if p0[0] == base[0]:
                    continue
            elif depth == 13:
                p0[0] = base[0]
                if p0[1] == base[1]:
                    continue
            if (p0[[a for a in range(3) if a != axis]] == base[[a for a in range(3) if a != axis]]).all():
                continue  # p0 on the line
            p = np.concatenate([p0[None], line])
            if len({tuple(r) for r in p}) < 5:
                continue
        return _shuffle(rng, p)
