"""CPU oracle predicates pinned against the reference's own predicates.c (golden vectors), exact
rational arithmetic and a brute-force SoS expansion."""
import os

import numpy as np
import pytest

import exact_ref as X

GOLD = os.path.join(os.path.dirname(__file__), "golden", "predicates_golden.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def test_orient3d_matches_reference_predicates(orc, gold):
    pts, want = gold["o3_pts"], gold["o3_sign"]
    got = np.array([orc.orient3d(*pts[j]) for j in range(pts.shape[0])])
    got_exact = np.array([orc.orient3d(*pts[j], exact=True) for j in range(pts.shape[0])])
    assert (got == want).all()
    assert (got_exact == want).all()
    assert (want == 0).sum() > 100 and (want > 0).sum() > 100 and (want < 0).sum() > 100


def test_insphere_matches_reference_predicates(orc, gold):
    pts, want = gold["is_pts"], gold["is_sign"]
    got = np.array([orc.insphere(*pts[j]) for j in range(pts.shape[0])])
    got_exact = np.array([orc.insphere(*pts[j], exact=True) for j in range(pts.shape[0])])
    assert (got == want).all()
    assert (got_exact == want).all()
    assert (want == 0).sum() > 100


def test_golden_signs_match_rational_arithmetic(gold):
    # pins the fixture itself: the reference's adaptive predicates agree with exact rationals
    rng = np.random.default_rng(1)
    for key, k, f in (("o3", 4, X.orient3d), ("is", 5, X.insphere)):
        pts, want = gold[key + "_pts"], gold[key + "_sign"]
        sel = rng.choice(pts.shape[0], 150, replace=False)
        for j in sel:
            assert f(*pts[j]) == want[j]


def test_lifted_minors_match_rational_arithmetic(orc):
    rng = np.random.default_rng(2)
    for _ in range(300):
        k = int(rng.integers(2, 5))
        cols = list(rng.choice(4, k - 1, replace=False))
        if rng.random() < 0.5:
            rows = rng.integers(0, 3, size=(k, 3)).astype(np.float64) / 4.0
        else:
            rows = rng.random((k, 3))
        assert orc.sign_minor(list(rows), cols) == X.minor_sign(rows, cols)


def _lattice_tuples(rng, count, span=3, denom=4.0):
    out = []
    while len(out) < count:
        p = rng.integers(0, span, size=(5, 3)).astype(np.float64) / denom
        if len({tuple(r) for r in p}) == 5:
            out.append(p)
    return out


def _deep_degenerate_tuples(rng, count):
    """Collinear / coplanar / cocircular families that push SoS past the first depths."""
    out = []
    while len(out) < count:
        kind = int(rng.integers(0, 4))
        if kind == 0:  # all on one axis-parallel line
            base = rng.integers(0, 3, size=3).astype(np.float64)
            ax = int(rng.integers(0, 3))
            p = np.repeat(base[None], 5, axis=0)
            p[:, ax] = rng.permutation(8)[:5]
        elif kind == 1:  # all in one axis-parallel plane
            p = rng.integers(0, 3, size=(5, 3)).astype(np.float64)
            p[:, int(rng.integers(0, 3))] = float(rng.integers(0, 3))
        elif kind == 2:  # four on a line, one free
            base = rng.integers(0, 3, size=3).astype(np.float64)
            ax = int(rng.integers(0, 3))
            p = np.repeat(base[None], 5, axis=0)
            p[:4, ax] = rng.permutation(6)[:4]
            p[4] = rng.integers(0, 3, size=3)
        else:  # on a general lattice line
            d = rng.integers(-1, 2, size=3)
            if not d.any():
                continue
            t = rng.permutation(7)[:5] - 3
            p = (rng.integers(0, 2, size=3)[None] + t[:, None] * d[None]).astype(np.float64)
        p = p / 4.0
        if len({tuple(r) for r in p}) == 5:
            out.append(p)
    return out


def test_sos_matches_bruteforce_perturbation(orc):
    rng = np.random.default_rng(3)
    depths = set()
    for p in _lattice_tuples(rng, 120) + _deep_degenerate_tuples(rng, 200):
        idx = [int(i) for i in rng.permutation(50)[:5]]
        s, d = orc.orient4_sos(p, idx)
        depths.add(d)
        assert s == X.sos_orient4(p, idx), (p.tolist(), idx, d)
    # the check has power over the table: many different depths (and so many sign flips)
    # were exercised
    assert len(depths - {0}) >= 6, depths


def test_sos_antisymmetric_under_transposition(orc):
    rng = np.random.default_rng(4)
    for p in _lattice_tuples(rng, 200, span=2, denom=2.0):
        idx = [int(i) for i in rng.permutation(20)[:5]]
        s, _ = orc.orient4_sos(p, idx)
        i, j = rng.choice(5, 2, replace=False)
        q = p.copy()
        q[[i, j]] = q[[j, i]]
        idx2 = list(idx)
        idx2[i], idx2[j] = idx2[j], idx2[i]
        s2, _ = orc.orient4_sos(q, idx2)
        assert s != 0 and s2 == -s


def test_sos_equals_exact_when_nonzero(orc):
    rng = np.random.default_rng(5)
    for _ in range(200):
        p = rng.random((5, 3))
        s, d = orc.orient4_sos(p, [0, 1, 2, 3, 4])
        assert d == 0 and s == X.insphere(*p)


def test_sos_reaches_every_depth_1_to_13(orc):
    """Families built to stop SoS at each depth 1..13 (every depth reachable with distinct
    points): the oracle stops exactly there and its sign equals the brute-force expansion."""
    import sos_families as F
    rng = np.random.default_rng(31)
    for depth in range(1, 14):
        for _ in range(6):
            p, idx = F.family(rng, depth)
            s, d = orc.orient4_sos(p, [int(i) for i in idx])
            assert d == depth, (depth, p.tolist(), idx.tolist())
            assert s == X.sos_orient4(p, [int(i) for i in idx]), (depth, p.tolist(), idx.tolist())
