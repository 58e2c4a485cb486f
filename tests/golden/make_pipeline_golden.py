"""Writes tests/golden/pipeline_golden.json: SHA-256 of the oracle's canonical tetrahedron set
(input index space, quadruples sorted, rows sorted, int32 little endian) for full-size
benchmark workloads, so GPU parity tests at those sizes need no oracle run on the GPU box.

Usage: python tests/golden/make_pipeline_golden.py [config ...]
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "pipeline_golden.json")


def canonical_hash(tets):
    t = np.ascontiguousarray(tets, dtype="<i4")
    return hashlib.sha256(t.tobytes()).hexdigest()


def main(names):
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        cfg = datasets.CONFIGS[name]
        x = datasets.make(name, seed=0)
        t0 = time.time()
        tets, perm, st = orc.triangulate(x, cfg["grid"], threads=0)
        dt = time.time() - t0
        data[name] = {"n": int(x.shape[0]), "grid": cfg["grid"], "seed": 0,
                      "num_tets": int(tets.shape[0]), "sha256": canonical_hash(tets),
                      "splay_rounds": int(st["splay_rounds"]), "oracle_seconds": round(dt, 1),
                      "first_tets": tets[:4].tolist()}
        print(name, data[name], flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["c1_uniform_100k_g128"])
