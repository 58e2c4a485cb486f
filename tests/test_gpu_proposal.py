"""Learned proposal on the GPU: the MFMA network against its CPU reference (forward, gradient),
training, the exported training labels against the oracle's Delaunay edges, and the rule that
the proposal never changes the tetrahedra."""
import numpy as np
import pytest

import gstar4d_amd as g
import mlp_ref as R
from gstar4d_amd import datasets
from gstar4d_amd import proposal as P

pytestmark = pytest.mark.gpu


def test_mlp_forward_matches_reference(tri):
    rng = np.random.default_rng(0)
    p = R.init_params(0)
    for n in (1, 31, 32, 1000, 4133):
        x, _ = R.random_samples(rng, n)
        got = tri.mlp_forward(p, x)
        want, _ = R.forward(p, x, bf16_ops=True)
        # same BF16 operands, FP32 accumulation in another order: a hidden unit can round to a
        # neighbouring BF16 value, which moves a logit by far less than 1e-2
        assert np.allclose(got, want, rtol=1e-2, atol=1e-2), (n, np.abs(got - want).max())
        assert np.abs(got - want).mean() < 1e-3


def test_mlp_gradient_matches_reference(tri):
    rng = np.random.default_rng(1)
    p = R.init_params(1)
    x, y = R.random_samples(rng, 3000)
    g_gpu, l_gpu = tri.mlp_grad(p, x, y)
    l_cpu, g_cpu, _ = R.forward_backward(p, x, y, bf16_ops=True)
    assert abs(l_gpu - l_cpu) <= 1e-4 * abs(l_cpu)
    # gradient: relative error of the whole vector, and per block (W1, b1, W2, b2, W3, b3)
    assert np.linalg.norm(g_gpu - g_cpu) <= 1e-2 * np.linalg.norm(g_cpu)
    edges = [0, R.OFF_B1, R.OFF_W2, R.OFF_B2, R.OFF_W3, R.OFF_B3, R.NPARAM]
    for a, b in zip(edges[:-1], edges[1:]):
        assert np.linalg.norm(g_gpu[a:b] - g_cpu[a:b]) <= 2e-2 * np.linalg.norm(g_cpu[a:b]) + 1e-6, (a, b)
    assert np.all(g_gpu[R.F * R.H:R.IN * R.H] == 0)
    # and against central finite differences of the float64 reference on a few parameters
    pd = p.astype(np.float64)
    for i in (R.OFF_B3 + 3, R.OFF_B3 + 10, R.OFF_W3 + 77, R.OFF_B1 + 5, 17):
        q = pd.copy()
        q[i] += 1e-4
        lp, _, _ = R.forward_backward(q, x, y, bf16_ops=False)
        q[i] -= 2e-4
        lm, _, _ = R.forward_backward(q, x, y, bf16_ops=False)
        fd = (lp - lm) / 2e-4
        assert abs(g_gpu[i] - fd) <= 0.03 * abs(fd) + 0.05, (i, g_gpu[i], fd)


def test_mlp_training_reduces_loss(tri):
    """Labels from a fixed teacher network: training the student from scratch must fit them."""
    rng = np.random.default_rng(2)
    teacher = R.init_params(7)
    x, _ = R.random_samples(rng, 40000)
    z, _ = R.forward(teacher, x, bf16_ops=False)
    y = np.zeros(x.shape[0], np.uint32)
    for j in range(8):
        y |= (z[:, j] > 0).astype(np.uint32) << j
        y |= np.uint32(1) << (8 + j)
    y |= (z[:, 8:].argmax(1).astype(np.uint32)) << 16
    p0 = R.init_params(3)
    l0, _, _ = R.forward_backward(p0, x, y, bf16_ops=True)
    p, losses = tri.mlp_train(p0, x, y, epochs=10, batch=2048, lr=3e-3, seed=1)
    assert losses[-1] < losses[0] and losses[-1] < 0.5 * l0 / x.shape[0], (l0 / x.shape[0], losses)
    l1, _, z1 = R.forward_backward(p, x, y, bf16_ops=True)
    assert abs(l1 / x.shape[0] - losses[-1]) < 0.2 * losses[-1] + 0.05
    assert ((z1[:, :8] > 0) == (z[:, :8] > 0)).mean() > 0.9


def _dt_edges(tets):
    t = np.asarray(tets, np.int64)
    e = [np.minimum(t[:, i], t[:, j]) << 32 | np.maximum(t[:, i], t[:, j]) for i in range(4) for j in range(i + 1, 4)]
    return np.unique(np.concatenate(e))


def test_exported_labels_are_final_links(orc):
    """Label bit j of point i says candidate j is in star i's final link: for points inside
    the convex hull that is exactly a Delaunay edge (i, candidate) of the oracle's
    tetrahedralisation, and the size bin is that of the Delaunay degree."""
    x = datasets.clustered_mix(30000, seed=4)
    with g.Triangulator(grid_size=64, index_space="morton") as t:
        t.set_sample_export(True)
        tets, _, perm = t.triangulate(x)
        f, y, c = t.get_samples(candidates=True)
    n = x.shape[0]
    assert f.shape == (n, 12) and np.isfinite(f).all()
    ref, _, _ = orc.triangulate(x, 64, canonical=False)  # Morton index space
    assert (g.canonical(tets) == g.canonical(ref)).all()
    edges = _dt_edges(ref)
    E = set(edges.tolist())
    deg = np.bincount(np.concatenate([edges >> 32, edges & 0xFFFFFFFF]).astype(np.int64), minlength=n)
    # hull vertices: those of faces with one tetrahedron
    faces = np.sort(np.concatenate([ref[:, [1, 2, 3]], ref[:, [0, 2, 3]], ref[:, [0, 1, 3]], ref[:, [0, 1, 2]]]), 1)
    u, cnt = np.unique(faces, axis=0, return_counts=True)
    hull = np.zeros(n, bool)
    hull[np.unique(u[cnt == 1])] = True
    checked = 0
    for i in range(0, n, 7):
        for j in range(8):
            cj = int(c[i, j])
            has = bool((y[i] >> (8 + j)) & 1)
            assert has == (cj >= 0)
            if cj < 0:
                continue
            lab = bool((y[i] >> j) & 1)
            edge = (min(i, cj) << 32 | max(i, cj)) in E
            if not hull[i]:
                assert lab == edge, (i, j, cj)
                checked += 1
            elif edge:
                assert lab
        if not hull[i]:
            b = int((y[i] >> 16) & 7)
            edges_b = [10, 13, 16, 19, 22, 26, 32]
            want = next((k for k, e in enumerate(edges_b) if deg[i] <= e), 7)
            assert b == want, (i, deg[i], b)
    assert checked > 10000
    # candidates are the nearest among the window: distances ascending
    assert (np.diff(f[:, 5:], axis=1) >= -1e-6)[(f[:, 5:] < 8).all(1)].all()


def _weights(tri_quick=None):
    if P.available():
        params, thr, _ = P.load_weights()
        return params, thr
    # no trained file: a quick network trained here on exported samples
    with g.Triangulator(grid_size=64, adjacency=False) as t:
        t.set_sample_export(True)
        xs, ys = [], []
        for k in range(4):
            x = datasets.clustered_mix(60000, seed=100 + k)
            t.triangulate(x)
            f, y = t.get_samples(stride=2)
            xs.append(f)
            ys.append(y)
        x, y = np.concatenate(xs), np.concatenate(ys)
        mean, std = x.mean(0), x.std(0) + 1e-6
        p, _ = t.mlp_train(P.init_params(0), ((x - mean) / std).astype(np.float32), y, epochs=3, batch=4096, lr=3e-3)
        return P.fold_normalisation(p, mean, std), 0.0


@pytest.mark.parametrize("kind", ["clustered", "hard", "uniform"])
def test_proposal_never_changes_the_tetrahedra(orc, kind):
    params, thr = _weights()
    if kind == "clustered":
        x, grid = datasets.clustered_mix(150000, seed=9), 128
    elif kind == "hard":
        x, grid = datasets.hard_mix(60000, seed=9), 128
    else:
        x, grid = datasets.uniform_cube(80000, seed=9), 96
    with g.Triangulator(grid_size=grid) as t:
        plain, popp, _ = t.triangulate(x)
        st0 = t.stats()
        t.set_proposal(params, thr)
        prop, propp, _ = t.triangulate(x)
        st1 = t.stats()
        rep = t.check(full_empty_sphere=True)
        # an untrained network (arbitrary proposals) cannot change it either
        t.set_proposal(R.init_params(5), -0.5)
        rnd, _, _ = t.triangulate(x)
        t.set_proposal(enable=False)
    ref, _, _ = orc.triangulate(x, grid)
    assert (g.canonical(plain) == ref).all()
    assert (g.canonical(prop) == ref).all()
    assert (g.canonical(rnd) == ref).all()
    assert st1["prop_enabled"] == 1 and st1["prop_pairs"] > 0 and st0["prop_enabled"] == 0
    assert rep["euler"] == 1 and rep["adjacency_failures"] == 0 and rep["empty_sphere_failures"] == 0
    print("%s: rounds %d -> %d, insertions %d -> %d, repacks %d -> %d, %.1f -> %.1f ms" % (
        kind, st0["splay_rounds"], st1["splay_rounds"], st0["splay_insertions"], st1["splay_insertions"],
        st0["store_repacks"], st1["store_repacks"], st0["t_total_ms"], st1["t_total_ms"]))
    if kind == "clustered":
        assert st1["splay_rounds"] <= st0["splay_rounds"]
        assert st1["splay_insertions"] < st0["splay_insertions"]


def test_proposal_in_batched_mode(orc):
    params, thr = _weights()
    rng = np.random.default_rng(3)
    sets = [datasets.clustered_mix(int(n), seed=int(s)) for n, s in zip(rng.integers(2000, 20000, 12), range(12))]
    xyz = np.concatenate(sets)
    off = np.concatenate([[0], np.cumsum([s.shape[0] for s in sets])]).astype(np.int64)
    with g.Triangulator(grid_size=64) as t:
        a, ao, aoff, gs = t.triangulate_batch(xyz, off)
        t.set_proposal(params, thr)
        b, bo, boff, _ = t.triangulate_batch(xyz, off)
    for k in range(len(sets)):
        ta, tb = a[aoff[k]:aoff[k + 1]], b[boff[k]:boff[k + 1]]
        assert (g.canonical(ta) == g.canonical(tb)).all(), k
    ref, _, _ = orc.triangulate(sets[5], int(gs[5]))
    assert (g.canonical(b[boff[5]:boff[6]]) == ref).all()


def test_c4_learned_proposal_matches_golden():
    """C4 (8M clustered mix, G = 512) with the committed trained weights: the tetrahedra equal the
    oracle's hash (the proposal changes only the splaying work), every star is consistent, and the
    trained network needs fewer splaying insertions and rounds than the plain pipeline."""
    import hashlib
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "pipeline_golden.json")))
    want = gold["c4_clustered_8m_g512"]
    params, thr, rec = P.load_weights()
    x = datasets.make("c4_clustered_8m_g512")
    with g.Triangulator(grid_size=want["grid"], adjacency=False) as t:
        d, n = t.upload(x)
        t.triangulate_device(d, n)
        plain = t.stats()
        t.set_proposal(params, thr)
        tets, _, _ = t.triangulate(x)
        st = t.stats()
        bad, tot = t.check_stars()
    print("c4 learned: %d tetrahedra, rounds %d -> %d, insertions %d -> %d, %.1f -> %.1f ms" % (
        tets.shape[0], plain["splay_rounds"], st["splay_rounds"], plain["splay_insertions"], st["splay_insertions"],
        plain["t_total_ms"], st["t_total_ms"]), flush=True)
    c = g.canonical(tets)
    assert c.shape[0] == want["num_tets"]
    assert hashlib.sha256(np.ascontiguousarray(c, "<i4").tobytes()).hexdigest() == want["sha256"]
    assert bad == 0 and tot >= 4 * want["num_tets"]
    assert st["prop_enabled"] == 1 and st["prop_pairs"] > 0
    assert st["splay_insertions"] < plain["splay_insertions"] and st["splay_rounds"] <= plain["splay_rounds"]
