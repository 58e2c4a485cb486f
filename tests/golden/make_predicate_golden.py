"""Writes tests/golden/predicates_golden.npz: orient3d / insphere sign vectors computed by the
reference's own host predicates (GDelaunay/Common/predicates.c, built from the reference tree
into oracle/_ref by oracle/build.py with -DREAL_TYPE_DOUBLE) on random, near-degenerate and
exactly degenerate FP64 tuples.  Run here (the reference tree is not on the GPU box); the
committed .npz is the fixture.

Usage: python tests/golden/make_predicate_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import build as obuild  # noqa: E402
from oracle import orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "predicates_golden.npz")


def tuples(rng, k, count):
    out = []
    # random, non-degenerate
    out.append(rng.random((count, k, 3)))
    # near-cospherical / near-coplanar: points on a sphere (rounded to FP64)
    v = rng.normal(size=(count, k, 3))
    v /= np.linalg.norm(v, axis=2, keepdims=True)
    out.append(0.5 + 0.37 * v)
    # exactly degenerate: small dyadic lattice
    out.append(rng.integers(0, 4, size=(count, k, 3)).astype(np.float64) / 8.0)
    # lattice scaled by a non-dyadic step (rounded), as in the 80^3 lattice workload
    out.append(rng.integers(0, 80, size=(count, k, 3)).astype(np.float64) / 80.0)
    # one-ulp perturbations of exact degeneracies
    base = rng.integers(0, 4, size=(count, k, 3)).astype(np.float64) / 8.0 + 0.25
    i = rng.integers(0, k, size=count)
    c = rng.integers(0, 3, size=count)
    sgn = np.where(rng.random(count) < 0.5, -np.inf, np.inf)
    base[np.arange(count), i, c] = np.nextafter(base[np.arange(count), i, c], sgn)
    out.append(base)
    # tiny coordinates next to unit-size ones (wide exponent range)
    w = rng.random((count, k, 3))
    w[:, 0, :] *= 2.0 ** -30
    out.append(w)
    # exact zeros: k corners of an axis-aligned box (cospherical, and 4 of them on a face
    # are coplanar), with box edges from a rounded non-dyadic lattice
    lo = rng.integers(0, 70, size=(count, 3)).astype(np.float64)
    ext = rng.integers(1, 9, size=(count, 3)).astype(np.float64)
    corners = np.array([[i & 1, (i >> 1) & 1, (i >> 2) & 1] for i in range(8)], np.float64)
    box = []
    for j in range(count):
        pick = rng.permutation(8)[:k] if k == 5 else np.array([0, 1, 2, 3]) if j % 2 else rng.permutation(8)[:k]
        box.append((lo[j] + corners[pick] * ext[j]) / 80.0)
    out.append(np.array(box))
    t = np.concatenate(out)
    # keep tuples of distinct points
    keep = [len({tuple(r) for r in t[j]}) == k for j in range(t.shape[0])]
    return t[np.array(keep)]


def main():
    obuild.build_ref()
    ref = orc.ref_lib()
    if ref is None:
        raise SystemExit("reference predicates not built (reference tree absent)")
    rng = np.random.default_rng(20261202)
    o3 = tuples(rng, 4, 400)
    is_ = tuples(rng, 5, 400)

    def p(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        return a, a.ctypes.data_as(orc._dp)

    o3_sign = np.empty(o3.shape[0], np.int8)
    for j in range(o3.shape[0]):
        args = [p(o3[j, q]) for q in range(4)]
        v = ref.orient3d(*[a[1] for a in args])
        o3_sign[j] = (v > 0) - (v < 0)
    is_sign = np.empty(is_.shape[0], np.int8)
    for j in range(is_.shape[0]):
        args = [p(is_[j, q]) for q in range(5)]
        v = ref.insphere(*[a[1] for a in args])
        is_sign[j] = (v > 0) - (v < 0)
    np.savez_compressed(OUT, o3_pts=o3, o3_sign=o3_sign, is_pts=is_, is_sign=is_sign)
    print(OUT, o3.shape, is_.shape, "zeros:", int((o3_sign == 0).sum()), int((is_sign == 0).sum()))


if __name__ == "__main__":
    main()
