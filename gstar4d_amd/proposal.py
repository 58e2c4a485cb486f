"""Learned proposal: weights file, feature normalisation and the GPU training driver.

The network and its kernels live in libgstar4d.so (g4d_set_proposal, g4d_mlp_train, ...;
include/gstar4d.h).  Training data comes from the plain pipeline itself: each training set
is triangulated with sample export on, and every sampled point contributes its 12 features
and a label word saying which of its 8 candidates ended in its final link (the difference
between the initial stars and the final links) and the bin of its final link size.

Training (on the GPU, through the library's own forward/backward/Adam kernels):

    python -m gstar4d_amd.proposal train --sets 200 --points 1000000 --grid 256

writes gstar4d_amd/proposal_weights.npz (parameters with the feature normalisation folded
into the first layer, the decision threshold and the training record).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "proposal_weights.npz")

IN, H, OUT, F = 32, 64, 16, 12
OFF_B1 = IN * H
NPARAM = IN * H + H + H * H + H + H * OUT + OUT


def init_params(seed=0, slot_prior=0.75):
    """He-initialised parameters; slot biases start at the logit of `slot_prior`."""
    rng = np.random.default_rng(seed)
    W1 = np.zeros((IN, H), np.float32)
    W1[:F] = rng.normal(0, np.sqrt(2.0 / F), size=(F, H))
    b1 = np.zeros(H, np.float32)
    W2 = rng.normal(0, np.sqrt(2.0 / H), size=(H, H)).astype(np.float32)
    b2 = np.zeros(H, np.float32)
    W3 = rng.normal(0, np.sqrt(1.0 / H), size=(H, OUT)).astype(np.float32)
    b3 = np.zeros(OUT, np.float32)
    b3[:8] = np.log(slot_prior / (1 - slot_prior))
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def fold_normalisation(params, mean, std):
    """Parameters acting on raw features x equal to `params` acting on (x - mean) / std."""
    p = np.array(params, np.float32, copy=True)
    W1 = p[:OFF_B1].reshape(IN, H)
    b1 = p[OFF_B1:OFF_B1 + H]
    s = np.asarray(std, np.float64)
    m = np.asarray(mean, np.float64)
    W1f = W1[:F].astype(np.float64) / s[:, None]
    b1[:] = (b1.astype(np.float64) - (m[:, None] * W1f).sum(0)).astype(np.float32)
    W1[:F] = W1f.astype(np.float32)
    return p


def save_weights(params, threshold, record, path=WEIGHTS_PATH):
    np.savez(path, params=np.asarray(params, np.float32), threshold=np.float32(threshold),
             record=np.frombuffer(json.dumps(record).encode(), np.uint8))


def load_weights(path=WEIGHTS_PATH):
    """(params, threshold, record) of a weights file written by save_weights."""
    with np.load(path, allow_pickle=False) as d:
        params = d["params"].astype(np.float32)
        thr = float(d["threshold"])
        rec = json.loads(bytes(d["record"]).decode())
    if params.shape != (NPARAM,):
        raise ValueError("weights file holds %s parameters, expected %d" % (params.shape, NPARAM))
    return params, thr, rec


def available(path=WEIGHTS_PATH):
    return os.path.exists(path)


def slot_metrics(logits, labels, threshold):
    """Precision / recall of the slot decisions at `threshold` over the existing candidates,
    and the accuracy of the link-size bin."""
    j = np.arange(8)
    valid = ((labels[:, None] >> (8 + j)) & 1).astype(bool)
    pos = ((labels[:, None] >> j) & 1).astype(bool) & valid
    pred = (logits[:, :8] > threshold) & valid
    tp = (pred & pos).sum()
    prec = tp / max(pred.sum(), 1)
    rec = tp / max(pos.sum(), 1)
    bins = (labels >> 16) & 7
    acc = (logits[:, 8:].argmax(1) == bins).mean()
    return {"precision": float(prec), "recall": float(rec), "bin_accuracy": float(acc),
            "proposed_per_point": float(pred.sum() / max(labels.shape[0], 1)),
            "positive_rate": float(pos.sum() / max(valid.sum(), 1))}


def collect_samples(tri, sets, points, grid, stride, seed0, log=print):
    """Triangulates `sets` clustered-mix sets with sample export on; every stride-th point of
    each is a sample."""
    from . import datasets
    tri.set_grid(grid)
    tri.set_sample_export(True)
    xs, ys = [], []
    t0 = time.time()
    for k in range(sets):
        x = datasets.clustered_mix(points, seed=seed0 + k)
        d_xyz, n = tri.upload(x)
        tri.triangulate_device(d_xyz, n)
        f, y = tri.get_samples(stride=stride, offset=k % stride)
        xs.append(f)
        ys.append(y)
        if k % 20 == 19 or k == sets - 1:
            log("samples: %d / %d sets, %d samples, %.1f s" % (k + 1, sets, sum(a.shape[0] for a in xs),
                                                               time.time() - t0))
    tri.set_sample_export(False)
    return np.concatenate(xs), np.concatenate(ys)


def train(sets=200, points=1_000_000, grid=256, stride=8, epochs=6, batch=8192, lr=2e-3, seed=0, holdout=0.05,
          log=print, path=WEIGHTS_PATH):
    import gstar4d_amd as g
    t_all = time.time()
    with g.Triangulator(grid_size=grid, adjacency=False) as tri:
        x, y = collect_samples(tri, sets, points, grid, stride, seed0=1000 + seed, log=log)
        rng = np.random.default_rng(seed)
        perm = rng.permutation(x.shape[0])
        nh = max(1, int(holdout * x.shape[0]))
        hx, hy = x[perm[:nh]], y[perm[:nh]]
        tx, ty = x[perm[nh:]], y[perm[nh:]]
        mean = tx.mean(0).astype(np.float64)
        std = tx.std(0).astype(np.float64)
        std[std < 1e-6] = 1.0
        txn = ((tx - mean) / std).astype(np.float32)
        p0 = init_params(seed)
        t0 = time.time()
        p, losses = tri.mlp_train(p0, txn, ty, epochs=epochs, batch=batch, lr=lr, seed=seed)
        t_train = time.time() - t0
        log("trained %d epochs on %d samples in %.1f s: loss per epoch %s" % (epochs, txn.shape[0], t_train,
                                                                           np.round(losses, 4).tolist()))
        params = fold_normalisation(p, mean, std)
        z = tri.mlp_forward(params, hx)
        # threshold: the logit whose decisions on held-out samples maximise F1 of the slots
        best = (0.0, -1.0)
        for thr in np.linspace(-2.0, 2.0, 41):
            m = slot_metrics(z, hy, thr)
            f1 = 2 * m["precision"] * m["recall"] / max(m["precision"] + m["recall"], 1e-9)
            if f1 > best[1]:
                best = (float(thr), f1)
        thr = best[0]
        metrics = slot_metrics(z, hy, thr)
        record = {"sets": sets, "points": points, "grid": grid, "stride": stride, "epochs": epochs, "batch": batch,
                  "lr": lr, "seed": seed, "samples": int(tx.shape[0]), "holdout": int(nh),
                  "epoch_loss": [float(v) for v in losses], "train_seconds": round(t_train, 2),
                  "total_seconds": round(time.time() - t_all, 1), "threshold": thr, "holdout_metrics": metrics,
                  "feature_mean": mean.tolist(), "feature_std": std.tolist(),
                  "data": "clustered mix: 70% from 256 Gaussian clusters (sigma log-uniform 0.002-0.05) + 30% uniform"}
        log("held-out: %s (threshold %.2f)" % (json.dumps(metrics), thr))
        save_weights(params, thr, record, path)
        log("wrote %s" % path)
    return params, thr, record


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--sets", type=int, default=200)
    t.add_argument("--points", type=int, default=1_000_000)
    t.add_argument("--grid", type=int, default=256)
    t.add_argument("--stride", type=int, default=8)
    t.add_argument("--epochs", type=int, default=6)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--out", default=WEIGHTS_PATH)
    a = ap.parse_args(argv)
    if a.cmd == "train":
        train(a.sets, a.points, a.grid, a.stride, a.epochs, a.batch, a.lr, a.seed,
              log=lambda m: print(m, flush=True), path=a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
