"""Batched mode on the GPU: every set of a batch must equal its own single-set triangulation
(same tetrahedra in the same order, same adjacency) and the oracle's canonical set."""
import numpy as np
import pytest

import gstar4d_amd as g
from gstar4d_amd import datasets

pytestmark = pytest.mark.gpu


def _mixed_sets(seed, sizes):
    rng = np.random.default_rng(seed)
    sets = []
    for k, n in enumerate(sizes):
        if k % 3 == 0:
            x = rng.random((n, 3))
        elif k % 3 == 1:
            x = 0.5 + 0.1 * rng.normal(size=(n, 3))
        else:
            x = rng.random((n, 3)) * np.array([1.0, 0.2, 3.0]) - 7.0  # offset, anisotropic
        sets.append(x)
    return sets


def _run_batch(sets, grids=None):
    xyz = np.concatenate(sets)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    with g.Triangulator(grid_size=64) as t:
        return t.triangulate_batch(xyz, off, grids)


def test_batch_matches_single_and_oracle(orc):
    sizes = [5, 6, 40, 300, 2000, 137, 1000, 64]
    sets = _mixed_sets(1, sizes)
    grids = [16, 16, 24, 48, 64, 32, 96, 20]
    tets, opp, toff, gs = _run_batch(sets, grids)
    assert gs.tolist() == grids
    assert toff[0] == 0 and toff[-1] == tets.shape[0]
    for k, x in enumerate(sets):
        mine, mo = tets[toff[k]:toff[k + 1]], opp[toff[k]:toff[k + 1]]
        with g.Triangulator(grid_size=grids[k]) as t1:
            st, so, _ = t1.triangulate(x)
        assert np.array_equal(mine, st), k
        assert np.array_equal(mo, so), k
        ref, _, _ = orc.triangulate(x, grids[k])
        assert np.array_equal(g.canonical(mine), ref), k


def test_batch_default_grids_and_many_sets(orc):
    rng = np.random.default_rng(2)
    sizes = rng.integers(5, 400, size=300).tolist()
    sets = _mixed_sets(3, sizes)
    tets, opp, toff, gs = _run_batch(sets)
    assert gs.tolist() == [g.batch_grid_size(n) for n in sizes]
    for k in (0, 1, 2, 57, 150, 299):
        x = sets[k]
        with g.Triangulator(grid_size=int(gs[k])) as t1:
            st, so, _ = t1.triangulate(x)
        assert np.array_equal(tets[toff[k]:toff[k + 1]], st), k
        assert np.array_equal(opp[toff[k]:toff[k + 1]], so), k
    # every set's tetrahedra index only its own points; adjacency stays inside the set
    for k in range(len(sizes)):
        sl = tets[toff[k]:toff[k + 1]]
        assert sl.min() >= 0 and sl.max() < sizes[k]
        o = opp[toff[k]:toff[k + 1]]
        assert o.min() >= -1 and o.max() < sl.shape[0]


def test_batch_single_set_equals_plain(orc):
    x = _mixed_sets(4, [3000])[0]
    tets, opp, toff, gs = _run_batch([x], [64])
    with g.Triangulator(grid_size=64) as t1:
        st, so, _ = t1.triangulate(x)
    assert np.array_equal(tets, st) and np.array_equal(opp, so)


def test_batch_rejects_bad_sets():
    x = np.random.default_rng(5).random((50, 3))
    with g.Triangulator(grid_size=64) as t:
        with pytest.raises(g.G4DError):
            t.triangulate_batch(x, np.array([0, 3, 50]), None)  # a set below 5 points
        with pytest.raises(g.G4DError):
            t.triangulate_batch(x, np.array([0, 20, 50]), np.array([64, 512]))  # batched grid above 256
        with pytest.raises(g.G4DError):
            t.triangulate_batch(x, np.array([0, 30, 30, 50]), None)  # empty set


def test_c5_batch_4096_sets(orc):
    """The 4096-set batched workload (sizes log-uniform in [1000, 200000], uniform or one
    Gaussian each, about 153.5 M points): batched calls over consecutive sets of at most 24 M
    points; every set's tetrahedra and adjacency equal its own single-set call, in order;
    every set passes the GPU validator with the full empty-sphere test; the 64 sets with
    oracle hashes (sizes spanning the range) match them."""
    import hashlib
    import json
    import os
    from gstar4d_amd import datasets
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "batch_golden.json")))
    want = {e["set"]: e for e in gold["sets"]}
    sets = datasets.batch_sets(gold["nsets"], seed=gold["seed"])
    sizes = np.array([x.shape[0] for x in sets], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    print("c5: %d sets, %d points" % (len(sets), int(off[-1])), flush=True)
    chunks, a = [], 0
    while a < len(sets):
        b = a + 1
        while b < len(sets) and off[b + 1] - off[a] <= 24_000_000:
            b += 1
        chunks.append((a, b))
        a = b
    hashed = checked = 0
    with g.Triangulator(grid_size=64) as t:
        for a, b in chunks:
            xyz = np.concatenate(sets[a:b])
            tets, opp, toff, gs = t.triangulate_batch(xyz, off[a:b + 1] - off[a])
            for k in range(a, b):
                mine = tets[toff[k - a]:toff[k - a + 1]]
                mo = opp[toff[k - a]:toff[k - a + 1]]
                st, so, stoff, _ = t.triangulate_batch(sets[k], np.array([0, sizes[k]]), gs[k - a:k - a + 1])
                assert np.array_equal(mine, st) and np.array_equal(mo, so), k
                rep = t.check_mesh(sets[k], mine, mo, full_empty_sphere=True)
                assert rep["euler"] == 1 and rep["orientation_failures"] == 0 and rep["flat_tets"] == 0, k
                assert rep["insphere_failures"] == 0 and rep["adjacency_failures"] == 0, k
                assert rep["empty_sphere_failures"] == 0 and rep["empty_sphere_tests"] > 0, k
                checked += 1
                if k in want:
                    c = g.canonical(mine)
                    assert c.shape[0] == want[k]["num_tets"], k
                    assert hashlib.sha256(np.ascontiguousarray(c, "<i4").tobytes()).hexdigest() == want[k]["sha256"], k
                    hashed += 1
            print("c5: sets %d..%d (%d points, grids %d..%d) equal to single-set calls, validated" %
                  (a, b - 1, int(off[b] - off[a]), int(gs.min()), int(gs.max())), flush=True)
    assert checked == len(sets) and hashed == len(want) == 64


def test_grid_table_and_star_kernel_variants(orc):
    """The autotuner's hooks: a per-size grid table decides the grids of sets passed without
    explicit grids (the same output as passing those grids), and every variant of the 32-vertex
    star kernel gives the oracle's tetrahedra with consistent adjacency."""
    from gstar4d_amd import autotune
    sets = datasets.batch_sets(24, seed=5, lo=1000, hi=30000)
    sizes = np.array([s.shape[0] for s in sets], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    xyz = np.concatenate(sets)
    table = {"buckets": [{"max_points": 5000, "voxels_per_point": 8.0}, {"max_points": 200000, "voxels_per_point": 20.0}],
             "star_kernel": 1}
    with g.Triangulator(grid_size=64) as t:
        autotune.apply(t, table)
        gr = autotune.grids(t, sizes)
        for n, gg in zip(sizes, gr):
            vpp = 8.0 if n <= 5000 else 20.0
            assert gg % 8 == 0 and 32 <= gg <= 256
            assert gg == 256 or gg ** 3 >= vpp * n
            assert gg == 32 or (gg - 8) ** 3 < vpp * n
        a, ao, aoff, ags = t.triangulate_batch(xyz, off)
        assert np.array_equal(ags, gr)
        b, bo, boff, _ = t.triangulate_batch(xyz, off, gr)
        assert np.array_equal(a, b) and np.array_equal(ao, bo) and np.array_equal(aoff, boff)
        # the variants may list a star's tetrahedra in another order (their link triangles
        # take other slots); the canonical set is the contract
        x = datasets.uniform_cube(30000, seed=8)
        ref, _, _ = orc.triangulate(x, 64)
        for v in range(5):
            t.set_star_kernel(v)
            tets, _, _ = t.triangulate(x)
            assert (g.canonical(tets) == ref).all(), v
            rep = t.check()
            assert rep["adjacency_failures"] == 0 and rep["euler"] == 1, v
        t.set_star_kernel(-1)


def test_stage_graph_replays_match_eager(orc):
    """The grid stage (fill, site scatter, PBA passes) as a captured hipGraph: captured on the
    first call, replayed on the later ones, with eager work of another context and device
    synchronisations between the replays; every replay's result equals the eager one's."""
    sets = datasets.batch_sets(16, seed=12, lo=1000, hi=20000)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    xyz = np.concatenate(sets)
    single = datasets.uniform_cube(30000, seed=13)
    with g.Triangulator(grid_size=64) as eager:
        eager.set_stage_graphs(0)
        ref = eager.triangulate_batch(xyz, off)
        ref1 = eager.triangulate(single)
        assert eager.stats()["graph_launches"] == 0
        with g.Triangulator(grid_size=64) as t:
            t.set_stage_graphs(1)
            for rep in range(4):
                got = t.triangulate_batch(xyz, off)
                st = t.stats()
                assert st["graph_launches"] == 1 and st["graph_captures"] == (1 if rep == 0 else 0), (rep, st)
                for a, b in zip(got, ref):
                    assert np.array_equal(a, b), rep
                got1 = t.triangulate(single)
                assert t.stats()["graph_launches"] == 1
                assert np.array_equal(got1[0], ref1[0]) and np.array_equal(got1[1], ref1[1]), rep
                # eager work on the other context between replays
                e1 = eager.triangulate(single)
                assert np.array_equal(e1[0], ref1[0]), rep
    oracle, _, _ = orc.triangulate(single, 64)
    assert (g.canonical(ref1[0]) == oracle).all()
