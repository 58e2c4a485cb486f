"""The pair-key sort of the pipeline (directed keys a << 32 | b, a and b below 2^nb) against
numpy's sort: uniform keys, duplicates, buckets beyond the LDS capacity (one id with a very
long list, as for a star on a lattice or a dense cluster), few ids, runs of ids in input
order, workset-like lists, a list of a few thousand keys, and the short lists and wide ids that
take the one-work-group and LSD sorts.  The pipeline's choice (method 0), the LSD radix sort (1)
and the bucket sort forced at any length (3) must give numpy's order."""
import numpy as np
import pytest

import gstar4d_amd as g

pytestmark = pytest.mark.gpu


def _keys(a, b):
    return (np.asarray(a, np.uint64) << np.uint64(32)) | np.asarray(b, np.uint64)


def _cases():
    rng = np.random.default_rng(7)
    nb = 22
    n = 1 << nb
    m = 3_000_000
    yield "uniform", _keys(rng.integers(0, n, m), rng.integers(0, n, m)), nb
    # every key twice, and runs of equal a (worksets)
    a = np.repeat(rng.integers(0, n, 200_000), 20)
    b = rng.integers(0, n, a.shape[0])
    k = _keys(a, b)
    yield "duplicates", np.concatenate([k, k[::-1]]), nb
    # ids near the top of the range and one id with 300k keys (a bucket beyond LDS)
    a = np.concatenate([rng.integers(n - 5000, n, 400_000), np.full(300_000, 12345)])
    b = rng.integers(0, n, a.shape[0])
    yield "long_list", rng.permutation(_keys(a, b)), nb
    # few ids (one bucket of everything), short id width
    yield "few_ids", _keys(rng.integers(0, 40, 100_000), rng.integers(0, 1000, 100_000)), 10
    # clustered ids in input order (the pipeline's lists come in runs)
    a = np.sort(rng.integers(0, 1 << 20, 1_000_000))
    yield "runs", _keys(a, rng.integers(0, 1 << 20, a.shape[0])), 20
    # a workset-like list: about 20 neighbours per id, near it in id order, both directions
    a = np.repeat(np.arange(1 << 21), 10)
    b = np.clip(a + rng.integers(-3000, 3000, a.shape[0]), 0, (1 << 21) - 1)
    w = np.concatenate([_keys(a, b), _keys(b, a)])
    yield "workset", w, 21
    # a list of a few thousand keys (bucket sort), then the fallbacks: a list short enough for
    # the one-work-group sort, and ids wider than 23 bits (LSD sort)
    yield "mid", _keys(rng.integers(0, n, 9000), rng.integers(0, n, 9000)), nb
    yield "short", _keys(rng.integers(0, n, 3000), rng.integers(0, n, 3000)), nb
    yield "wide_ids", _keys(rng.integers(0, 1 << 25, 200_000), rng.integers(0, 1 << 25, 200_000)), 25


@pytest.mark.parametrize("name,keys,nb", list(_cases()), ids=[c[0] for c in _cases()])
def test_sort_pairs_matches_numpy(name, keys, nb):
    ref = np.sort(keys)
    with g.Triangulator(grid_size=16) as t:
        t.diag_sort_pairs(keys[:20000], nb, method=3)  # loads the kernels
        t.diag_sort_pairs(keys[:20000], nb, method=1)
        got0, ms0, sh0 = t.diag_sort_pairs(keys, nb, method=0)
        got1, ms1, _ = t.diag_sort_pairs(keys, nb, method=1)
        got3, ms3, sh3 = t.diag_sort_pairs(keys, nb, method=3)
    print("%s: %d keys, nb %d: pipeline (shift %d) %.3f ms, LSD %.3f ms, bucket (shift %d) %.3f ms" %
          (name, keys.shape[0], nb, sh0, ms0, ms1, sh3, ms3), flush=True)
    assert np.array_equal(got1, ref)
    assert np.array_equal(got0, ref)
    assert np.array_equal(got3, ref)
    if name in ("short", "wide_ids"):
        assert sh3 == -1
    else:
        assert sh3 >= 0


def test_sort_drowned_keys_matches_numpy():
    """Drowned-insertion keys min << 33 | max << 1 | direction, as the splaying rounds sort them."""
    rng = np.random.default_rng(11)
    nb = 22
    a = rng.integers(0, 1 << nb, 6_000_000)
    b = rng.integers(0, 1 << nb, a.shape[0])
    mn, mx = np.minimum(a, b), np.maximum(a, b)
    keys = (mn.astype(np.uint64) << np.uint64(33)) | (mx.astype(np.uint64) << np.uint64(1)) | \
        (a > b).astype(np.uint64)
    ref = np.sort(keys)
    with g.Triangulator(grid_size=16) as t:
        got2, ms2, sh2 = t.diag_sort_pairs(keys, nb, method=2)
        got4, ms4, sh4 = t.diag_sort_pairs(keys, nb, method=4)
    print("drowned keys: %d, pipeline (shift %d) %.3f ms, bucket (shift %d) %.3f ms" %
          (keys.shape[0], sh2, ms2, sh4, ms4), flush=True)
    assert sh4 >= 0
    assert np.array_equal(got2, ref)
    assert np.array_equal(got4, ref)
