import numpy as np

from gstar4d_amd import datasets as D


def test_uniform_stream_is_counter_based_and_in_range():
    a = D.uniform01(7, 1000)
    b = D.uniform01(7, 400, start=600)
    assert (a[600:] == b).all()
    assert a.min() >= 0.0 and a.max() < 1.0
    # multiples of 2^-53 exactly
    assert (np.floor(a * 2.0 ** 53) == a * 2.0 ** 53).all()


def test_generators_are_deterministic_and_unique():
    for f in (lambda: D.uniform_cube(5000, 3), lambda: D.hard_mix(8000, 3), lambda: D.clustered_mix(6000, 3)):
        x, y = f(), f()
        assert (x == y).all()
        assert np.unique(x, axis=0).shape[0] == x.shape[0]
        assert x.min() >= 0.0 and x.max() < 1.0


def test_hard_mix_composition():
    x = D.hard_mix(40000, 0)
    lat = x[np.all(np.abs(x * 80 - np.round(x * 80)) < 1e-12, axis=1)]
    assert abs(lat.shape[0] - 10000) < 200
    r = np.linalg.norm(x - 0.5, axis=1)
    assert (np.abs(r - 0.4) < 1e-12).sum() >= 9000


def test_batch_sizes_log_uniform():
    sets = D.batch_sets(64, seed=1)
    sizes = np.array([s.shape[0] for s in sets])
    assert sizes.min() >= 900 and sizes.max() <= 200000
