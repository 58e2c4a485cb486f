"""Pins the oracle's Voronoi-neighbour pair rule (GD/GDelaunay/GDelKernels.cu:181-290).

Two independent checks:

* Hand-worked cubes.  Each fixture is one 2x2x2 cube of voxel labels (corner c at offset
  (c & 1, c >> 1 & 1, c >> 2), Marker = -1) whose expected pair list was written out by hand
  from the reference's clauses: root corner 0, opposite corner 7, link cycle 6, 2, 3, 1, 5, 4;
  for consecutive link labels a != b with neither equal to root or opposite emit root-a,
  opp-a, root-b, opp-b, a-b (grabPair drops a pair whose first value is Marker); root-opp
  once when such a quad exists and neither end is Marker.  Each fixture is chosen so that
  one plausible mistake in one clause changes its list.

* What the geometry fixes.  The six (root, a, b, opp) quadruples with a, b consecutive on the
  link cycle are exactly the Freudenthal (Kuhn) simplices of the cube: the chains
  0 -> e_i -> e_i + e_j -> 7 over the six orders (i, j, k) of the axes.  So the rule is "every
  label pair spanned by a Freudenthal tetrahedron with four distinct labels, Marker pairs
  dropped", which `freudenthal_pairs` below restates from that description alone (no link
  table) over whole grids with the outer Marker layer, compared against orc.grid_pairs.
"""
import itertools

import numpy as np

M = -1


def P(*pairs):
    return sorted({(min(a, b), max(a, b)) for a, b in pairs})


def cube(**kw):
    """Cube labels by corner name c0..c7 (defaults: corner c -> 10 + c)."""
    v = [10 + c for c in range(8)]
    for k, val in kw.items():
        v[int(k[1:])] = val
    return v


def uniq(pairs):
    return sorted(set(pairs))


def test_all_distinct_labels(orc):
    # root 10, opp 17, link cycle labels 16, 12, 13, 11, 15, 14 (corners 6, 2, 3, 1, 5, 4)
    v = cube()
    got = orc.cube_pairs(v)
    assert len(got) == 6 * 5 + 1  # every consecutive link pair is a quad, plus root-opp
    want = P((10, 16), (10, 12), (10, 13), (10, 11), (10, 15), (10, 14),
             (17, 16), (17, 12), (17, 13), (17, 11), (17, 15), (17, 14),
             (16, 12), (12, 13), (13, 11), (11, 15), (15, 14), (14, 16),
             (10, 17))
    assert uniq(got) == want
    # the link edges are those of the hexagon 6-2-3-1-5-4; e.g. corners 1 and 2 (labels 11,
    # 12) and corners 3 and 6 (13, 16) are not consecutive on it
    assert (11, 12) not in got and (13, 16) not in got


def test_only_singleton_pairs_distinct_emits_nothing(orc):
    # corners 1 and 2 carry new labels, every other link corner repeats root or opp: on the
    # cycle 6,2,3,1,5,4 each of 1 and 2 sits between corners labelled root/opp, so no quad
    v = [0, 21, 22, 1, 0, 1, 0, 1]
    assert orc.cube_pairs(v) == []


def test_one_quad(orc):
    # only (6, 2) qualifies: 2 -> 3 hits opp, 4 -> 6 hits root
    v = [0, 1, 3, 1, 0, 1, 2, 1]
    assert uniq(orc.cube_pairs(v)) == P((0, 2), (1, 2), (0, 3), (1, 3), (2, 3), (0, 1))


def test_root_equals_opp_emits_nothing(orc):
    v = cube(c7=10)
    assert orc.cube_pairs(v) == []


def test_no_quad_no_root_opp(orc):
    # root != opp but the link holds one label: no quad, so root-opp is not emitted either
    v = [0, 2, 2, 2, 2, 2, 2, 1]
    assert orc.cube_pairs(v) == []
    # two link labels alternating with root: 6=2, 2=0, 3=3, 1=0, 5=2, 4=0 -> no a,b both free
    v = [0, 0, 0, 3, 0, 2, 2, 1]
    assert orc.cube_pairs(v) == []


def test_marker_root(orc):
    # root Marker: pairs led by root are skipped, opp-x and a-b kept, no root-opp
    v = cube(c0=M)
    want = P((17, 16), (17, 12), (17, 13), (17, 11), (17, 15), (17, 14),
             (16, 12), (12, 13), (13, 11), (11, 15), (15, 14), (14, 16))
    got = orc.cube_pairs(v)
    assert uniq(got) == want
    assert len(got) == 6 * 3


def test_marker_opp(orc):
    v = cube(c7=M)
    want = P((10, 16), (10, 12), (10, 13), (10, 11), (10, 15), (10, 14),
             (16, 12), (12, 13), (13, 11), (11, 15), (15, 14), (14, 16))
    assert uniq(orc.cube_pairs(v)) == want


def test_marker_on_link_is_never_emitted(orc):
    # with root and opp inner every corner is inner; but a Marker on the link (as the outer
    # layer gives when both ends are Marker) only blocks its own quads
    v = [M, M, 12, 13, M, M, 16, M]
    # root = opp = Marker -> nothing at all
    assert orc.cube_pairs(v) == []
    v = [M, M, 12, 13, M, M, 16, 17]
    # root Marker, opp 17: link 6=16, 2=12, 3=13, 1=M, 5=M, 4=M; quads (16,12), (12,13)
    assert uniq(orc.cube_pairs(v)) == P((17, 16), (17, 12), (17, 13), (16, 12), (12, 13))


def freudenthal_pairs(grid):
    """Label pairs of the Freudenthal tetrahedra with four distinct labels, over every 2x2x2
    cube of the grid padded with one layer of Marker (index [z, y, x])."""
    g = grid.shape[0]
    pad = np.full((g + 2, g + 2, g + 2), M, np.int64)
    pad[1:-1, 1:-1, 1:-1] = grid
    out = set()
    axes = ((0, 0, 1), (0, 1, 0), (1, 0, 0))  # unit steps in (z, y, x) index order: x, y, z
    for order in itertools.permutations(range(3)):
        e1 = np.array(axes[order[0]])
        e2 = e1 + np.array(axes[order[1]])
        corners = [np.zeros(3, int), e1, e2, np.ones(3, int)]
        lab = [pad[c[0]:c[0] + g + 1, c[1]:c[1] + g + 1, c[2]:c[2] + g + 1].reshape(-1) for c in corners]
        L = np.stack(lab, 1)
        distinct = np.ones(L.shape[0], bool)
        for i, j in itertools.combinations(range(4), 2):
            distinct &= L[:, i] != L[:, j]
        for row in np.unique(L[distinct], axis=0):
            for i, j in itertools.combinations(range(4), 2):
                a, b = int(row[i]), int(row[j])
                if a != M and b != M:
                    out.add((min(a, b), max(a, b)))
    return sorted(out)


def test_grid_rule_equals_freudenthal_restatement(orc):
    rng = np.random.default_rng(11)
    for g in (2, 3, 4, 5):
        for nlab in (3, 4, 6, 12):
            for _ in range(4):
                grid = rng.integers(0, nlab, size=(g, g, g)).astype(np.int32)
                got = orc.grid_pairs(grid)
                got = sorted((int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF))) for k in got)
                assert got == freudenthal_pairs(grid), (g, nlab)


def test_grid_rule_on_a_pba_diagram(orc):
    rng = np.random.default_rng(12)
    g = 10
    vox = np.unique(rng.integers(0, g, size=(30, 3)), axis=0)
    perm = np.lexsort((vox[:, 0], vox[:, 1], vox[:, 2]))  # any order: labels are point indices
    grid = orc.pba(vox[perm], g)
    got = orc.grid_pairs(grid)
    got = sorted((int(k >> np.uint64(32)), int(k & np.uint64(0xFFFFFFFF))) for k in got)
    assert len(got) > 30
    assert got == freudenthal_pairs(grid)
