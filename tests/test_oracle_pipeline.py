"""CPU oracle pipeline: Morton order, PBA, pairs and the full star-splaying triangulation,
pinned by brute force, scipy/Qhull and an empty-sphere validator."""
import itertools
import json
import os

import numpy as np
import pytest

from gstar4d_amd import datasets

GOLD = os.path.join(os.path.dirname(__file__), "golden", "pipeline_golden.json")


def _morton_np(vox, bits=10):
    code = np.zeros(vox.shape[0], np.int64)
    for b in range(bits):
        for c in range(3):
            code |= ((vox[:, c] >> b) & 1).astype(np.int64) << (3 * b + c)
    return code


def test_morton_order_is_stable_sort_of_codes(orc):
    rng = np.random.default_rng(0)
    x = rng.random((4000, 3))
    x[1000:1100] = x[0]  # shared voxels but distinct points: perturb slightly
    x[1000:1100, 0] += np.arange(100) * 1e-9
    perm, vox = orc.morton_order(x, 64)
    # voxel scaling: (G-3)(x-min)/(max-min)+1, truncated (Application::_scalePoint)
    mn, mx = x.min(), x.max()
    want_vox = np.clip(((61.0 * (x - mn)) / (mx - mn) + 1.0).astype(np.int64), 1, 62)
    assert (vox == want_vox).all()
    code = _morton_np(want_vox)
    assert (perm == np.argsort(code, kind="stable")).all()


def test_morton_order_on_wide_grids(orc):
    """Grids above 1024 need an eleventh bit per axis; the order is still the stable sort of the
    interleaved codes, and below 1024 the wider code orders exactly as the 10-bit one."""
    rng = np.random.default_rng(4)
    x = rng.random((3000, 3))
    for g in (1100, 2048):
        perm, vox = orc.morton_order(x, g)
        mn, mx = x.min(), x.max()
        want_vox = np.clip(((float(g - 3) * (x - mn)) / (mx - mn) + 1.0).astype(np.int64), 1, g - 2)
        assert (vox == want_vox).all() and vox.max() >= 1024
        assert (perm == np.argsort(_morton_np(want_vox, 11), kind="stable")).all()
        assert not (perm == np.argsort(_morton_np(want_vox & 1023, 10), kind="stable")).all()
    perm, vox = orc.morton_order(x, 1000)
    assert (perm == np.argsort(_morton_np(vox, 10), kind="stable")).all()


def test_pba_labels_do_not_change_the_diagram(orc):
    """orc.pba sweeps point-index labels, orc.pba(codes=True) sweeps site codes; the id grid is
    the code grid mapped through the voxel leaders (lowest Morton index in a voxel)."""
    rng = np.random.default_rng(6)
    for g, ns in ((24, 300), (40, 60)):
        vox = rng.integers(0, g, size=(ns, 3))  # repeated voxels: several points per site
        vs = vox[np.argsort(_morton_np(vox), kind="stable")]
        ids = orc.pba(vs, g)
        codes = orc.pba(vs, g, codes=True)
        leader = {}
        for i in range(vs.shape[0] - 1, -1, -1):
            leader[(int(vs[i, 0]) << 20) | (int(vs[i, 1]) << 10) | int(vs[i, 2])] = i
        want = np.vectorize(lambda c: leader[int(c)])(codes)
        assert (ids == want).all()


def test_pba_is_exact_nearest_site(orc):
    rng = np.random.default_rng(1)
    for g, ns in ((16, 30), (24, 200), (20, 3)):
        vox = np.unique(rng.integers(0, g, size=(ns, 3)), axis=0)
        code = _morton_np(vox)
        vs = vox[np.argsort(code, kind="stable")]
        grid = orc.pba(vs, g)  # [z, y, x] -> site index
        zz, yy, xx = np.meshgrid(np.arange(g), np.arange(g), np.arange(g), indexing="ij")
        P = np.stack([xx, yy, zz], -1).reshape(-1, 3)
        d = ((P[:, None, :] - vs[None, :, :]) ** 2).sum(-1)
        best = d.min(1)
        got = d[np.arange(P.shape[0]), grid.reshape(-1)]
        assert (got == best).all()


def test_grid_pairs_come_from_touching_voxels(orc):
    rng = np.random.default_rng(2)
    g = 12
    vox = np.unique(rng.integers(0, g, size=(40, 3)), axis=0)
    vs = vox[np.argsort(_morton_np(vox), kind="stable")]
    grid = orc.pba(vs, g)
    pairs = orc.grid_pairs(grid)
    assert len(pairs) > 0
    a = (pairs >> np.uint64(32)).astype(np.int64)
    b = (pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)
    assert (a < b).all() and len(np.unique(pairs)) == len(pairs)
    # every pair's Voronoi cells touch within one 2x2x2 voxel cube
    cells = {i: np.argwhere(grid == i) for i in np.unique(grid)}
    for i, j in zip(a[:200], b[:200]):
        ci, cj = cells[i], cells[j]
        dmax = np.abs(ci[:, None, :] - cj[None, :, :]).max(-1)
        assert dmax.min() <= 1


def _brute_dt(orc, x):
    """All 4-subsets whose lifted facet has every other point above it (SoS orient4 with
    Morton ranks as indices), minus flat ones: the canonical output by enumeration."""
    n = x.shape[0]
    perm, _ = orc.morton_order(x, 32)
    rank = np.empty(n, np.int64)
    rank[perm] = np.arange(n)
    out = []
    for q in itertools.combinations(range(n), 4):
        o3 = orc.orient3d(*x[list(q)])
        if o3 == 0:
            continue
        ok = True
        for e in range(n):
            if e in q:
                continue
            s, _ = orc.orient4_sos([x[q[0]], x[q[1]], x[q[2]], x[q[3]], x[e]],
                                   [rank[q[0]], rank[q[1]], rank[q[2]], rank[q[3]], rank[e]])
            if s != -o3:
                ok = False
                break
        if ok:
            out.append(q)
    return np.array(sorted(out), np.int32).reshape(-1, 4)


@pytest.mark.parametrize("kind", ["random", "lattice", "cospherical"])
def test_pipeline_matches_bruteforce(orc, kind):
    rng = np.random.default_rng(3)
    if kind == "random":
        x = rng.random((22, 3))
    elif kind == "lattice":
        x = np.stack(np.meshgrid(*[np.arange(3)] * 3, indexing="ij"), -1).reshape(-1, 3) / 3.0
    else:
        v = rng.integers(-2, 3, size=(60, 3)).astype(np.float64)
        v = v[np.abs(v).sum(1) > 0]
        x = np.unique(v[(v ** 2).sum(1) == 5], axis=0) / 4.0  # integer points on a sphere
        x = np.concatenate([x, [[0.05, 0.02, 0.01]]])
    tets, perm, st = orc.triangulate(x, 32)
    want = _brute_dt(orc, x)
    assert tets.shape == want.shape and (tets == want).all()


def test_pipeline_matches_qhull_random(orc):
    scipy_spatial = pytest.importorskip("scipy.spatial")
    x = datasets.uniform_cube(3000, seed=5)
    tets, _, st = orc.triangulate(x, 64)
    ref = orc.canonicalize(scipy_spatial.Delaunay(x).simplices.astype(np.int32))
    assert (tets == ref).all()
    assert st["splay_rounds"] > 0


def test_c1_full_size_validates_and_matches_golden(orc):
    g = json.load(open(GOLD))["c1_uniform_100k_g128"]
    x = datasets.make("c1_uniform_100k_g128")
    tets, _, st = orc.triangulate(x, 128)
    import hashlib
    assert tets.shape[0] == g["num_tets"]
    assert hashlib.sha256(np.ascontiguousarray(tets, "<i4").tobytes()).hexdigest() == g["sha256"]
    flags, counts = orc.validate(x, tets, full_empty=True)
    assert flags == 0, counts


def test_hard_mix_sample_validates(orc):
    x = datasets.hard_mix(20000, seed=1)
    tets, _, st = orc.triangulate(x, 128)
    assert st["sos_calls"] > 1000  # the degenerate lattice really goes through SoS
    flags, counts = orc.validate(x, tets, full_empty=True)
    assert flags == 0, counts


def test_adjacency_is_symmetric(orc):
    x = datasets.uniform_cube(2000, seed=6)
    tets, _, _ = orc.triangulate(x, 32)
    opp, bad = orc.adjacency(tets, x.shape[0])
    assert bad == 0
    for t in range(0, tets.shape[0], 7):
        for i in range(4):
            u = opp[t, i]
            if u < 0:
                continue
            face = set(tets[t]) - {tets[t, i]}
            assert face <= set(tets[u]) and t in opp[u]


def test_min_workset_does_not_change_output(orc):
    x = datasets.uniform_cube(1500, seed=7)
    a, _, sa = orc.triangulate(x, 32, min_workset=8)
    b, _, sb = orc.triangulate(x, 32, min_workset=24)
    assert (a == b).all()
    assert sb["num_pumped"] > sa["num_pumped"]


@pytest.mark.parametrize("kind", ["uniform", "hard_mix", "lattice", "cospherical"])
def test_insertion_order_does_not_change_the_stars(orc, kind):
    """The GPU's grouped star kernel inserts each star's list nearest-first instead of in
    ascending id.  A star's link is the unique hull of the lifted points seen from the star
    point (SoS decides degenerate cases consistently), so the order must change neither the
    tetrahedra nor the splaying rounds and insertions (drowned flags and pairs are set
    properties of the final stars).  Here the oracle itself runs both orders, on inputs whose
    degenerate cases go through SoS."""
    rng = np.random.default_rng(9)
    if kind == "uniform":
        x, g = datasets.uniform_cube(20000, seed=11), 32
    elif kind == "hard_mix":
        x, g = datasets.hard_mix(20000, seed=2), 64
    elif kind == "lattice":
        x = np.stack(np.meshgrid(*[np.arange(9)] * 3, indexing="ij"), -1).reshape(-1, 3) / 9.0
        x, g = x + 0.0, 16
    else:
        v = rng.integers(-6, 7, size=(4000, 3)).astype(np.float64)
        v = np.unique(v[(v ** 2).sum(1) == 29], axis=0) / 8.0  # integer points on one sphere
        x, g = np.concatenate([v, rng.random((300, 3)) - 0.5]), 16
    a, _, sa = orc.triangulate(x, g, insert_order=0)
    b, _, sb = orc.triangulate(x, g, insert_order=1)
    assert a.shape == b.shape and (a == b).all()
    assert sa["splay_rounds"] == sb["splay_rounds"]
    assert sa["splay_insertions"] == sb["splay_insertions"]
    if kind in ("hard_mix", "lattice", "cospherical"):
        assert sb["sos_calls"] > 0
