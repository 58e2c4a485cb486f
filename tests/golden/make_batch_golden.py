"""Writes tests/golden/batch_golden.json: oracle hashes for 64 sets of the 4096-set batched
workload (datasets.batch_sets(4096, seed=0)), chosen at evenly spaced size ranks so they span
1,000 to 200,000 points, both distributions.

Each entry holds the set's index, size, distribution, the grid the oracle ran at and the
SHA-256 of its canonical tetrahedron set (input index space, quadruples sorted, rows sorted,
int32 little endian).  The sets are in general position (FP64 uniform or Gaussian draws), so
their Delaunay tetrahedralisation is unique and does not depend on the PBA grid width.

Usage: python tests/golden/make_batch_golden.py
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from gstar4d_amd import datasets  # noqa: E402
from oracle import orc  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_golden.json")
NSETS, SEED, PICK = 4096, 0, 64


def grid_for(n):
    """About 24 voxels per point, multiples of 16 in [32, 256]."""
    g = 32
    while g < 256 and g ** 3 < 24 * n:
        g += 16
    return g


def main():
    sets = datasets.batch_sets(NSETS, seed=SEED)
    sizes = np.array([x.shape[0] for x in sets])
    order = np.argsort(sizes, kind="stable")
    picks = sorted(int(order[int(round(r))]) for r in np.linspace(0, NSETS - 1, PICK))
    coin = datasets.uniform01(datasets._stream(SEED, 42), NSETS) < 0.5
    out = {"nsets": NSETS, "seed": SEED, "sets": []}
    t0 = time.time()
    for k in picks:
        x = sets[k]
        grid = grid_for(x.shape[0])
        tets, _, st = orc.triangulate(x, grid, threads=0)
        out["sets"].append({"set": k, "n": int(x.shape[0]), "grid": grid,
                            "distribution": "uniform" if coin[k] else "gaussian",
                            "num_tets": int(tets.shape[0]),
                            "sha256": hashlib.sha256(np.ascontiguousarray(tets, "<i4").tobytes()).hexdigest(),
                            "sos_calls": int(st["sos_calls"])})
        print(out["sets"][-1], flush=True)
    out["oracle_seconds"] = round(time.time() - t0, 1)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
