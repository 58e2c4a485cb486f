"""Batched mode on the GPU: every set of a batch must equal its own single-set triangulation
(same tetrahedra in the same order, same adjacency) and the oracle's canonical set."""
import numpy as np
import pytest

import gstar4d_amd as g

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
