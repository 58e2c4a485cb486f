"""Exact rational restatements used as independent checkers in tests (pure Python).

* det / lifted determinants over Fractions (FP64 inputs are exact rationals);
* a brute-force Simulation of Simplicity: the sign of the perturbed 5x5 lifted determinant
  det[[x+e, y+e, z+e, l+e, 1]] with e_{i,c} = eps^(2^(4i+3-c)) over the index-sorted rows
  (columns x=0 y=1 z=2 lift=3), found as the first nonzero coefficient over all rook
  placements of perturbed entries, ordered by their eps exponent.  This is the
  Edelsbrunner-Muecke scheme the reference's 49-depth table enumerates
  (GDelaunay/GDelaunay/GDelPredDevice.h:142-669); it shares no code with either
  implementation under test.
"""
import itertools
from fractions import Fraction
from functools import lru_cache


def det(m):
    n = len(m)
    if n == 1:
        return m[0][0]
    if n == 2:
        return m[0][0] * m[1][1] - m[0][1] * m[1][0]
    tot = 0
    for j in range(n):
        if m[0][j] == 0:
            continue
        sub = [row[:j] + row[j + 1:] for row in m[1:]]
        tot += (-1) ** j * m[0][j] * det(sub)
    return tot


def sgn(v):
    return (v > 0) - (v < 0)


def fr(p):
    return [Fraction(float(c)) for c in p]


def orient3d(a, b, c, d):
    return sgn(det([fr(p) + [1] for p in (a, b, c, d)]))


def lift_row(p):
    q = fr(p)
    return q + [q[0] * q[0] + q[1] * q[1] + q[2] * q[2], 1]


def insphere(a, b, c, d, e):
    return sgn(det([lift_row(p) for p in (a, b, c, d, e)]))


def minor_sign(rows, cols):
    """sign det[[c_0 .. c_{k-2} 1]] with c in {0,1,2,3(lift)}."""
    m = []
    for p in rows:
        q = lift_row(p)
        m.append([q[c] for c in cols] + [1])
    return sgn(det(m))


def _inversions(seq):
    return sum(1 for i in range(len(seq)) for j in range(i + 1, len(seq)) if seq[i] > seq[j])


@lru_cache(maxsize=1)
def _placements():
    out = []
    for k in range(0, 5):
        for I in itertools.combinations(range(5), k):
            for J in itertools.permutations(range(4), k):
                e = sum(1 << (4 * i + 3 - j) for i, j in zip(I, J))
                out.append((e, I, J))
    out.sort()
    return tuple(out)


def sos_orient4(pts, idx):
    """Brute-force SoS sign of the lifted orientation of 5 distinct points."""
    order = sorted(range(5), key=lambda i: idx[i])
    parity = _inversions(order) & 1
    M = [lift_row(pts[i]) for i in order]
    for e, I, J in _placements():
        rows = [r for r in range(5) if r not in I]
        cols = [c for c in range(5) if c not in J]
        sub = [[M[r][c] for c in cols] for r in rows]
        d = det(sub) if sub else 1
        if d == 0:
            continue
        s = (-1) ** ((sum(I) + sum(J)) & 1) * (-1) ** (_inversions(J) & 1)
        v = sgn(d) * s
        return -v if parity else v
    raise AssertionError("no nonzero SoS term")
