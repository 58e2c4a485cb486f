"""Autotuner for the batched mode and the star kernel.

The reference fixes its PBA grid per run (Config::_gridSize) and its kernel shapes at compile
time.  Batched mode has to choose a grid for every set, and the best density of points per
voxel depends on the set's size (small sets pay per-stage overhead, large ones per-voxel
work), so it is measured rather than guessed:

  * grid: for each set-size bucket, batches of sets of that size (half uniform, half one
    Gaussian, like the 4096-set workload) are triangulated with every candidate density
    (voxels per point); the fastest density wins.  The library turns a density into a grid
    (the smallest multiple of 8 in [32, 256] with that many voxels per point).
  * star kernel: every variant of the 32-vertex star kernel on the same batches; the fastest
    is kept.  (Variants change only the speed: the tests hold every one to the same output.)
  * PBA bands: the y and x passes with whole columns or 2 or 4 bands (same diagram).

    python -m gstar4d_amd.autotune            # writes gstar4d_amd/autotune_table.json

The table is applied with apply(tri) (g4d_set_grid_table, g4d_set_star_kernel,
g4d_set_pba_bands); grids(sizes)
returns the grid of each set under it.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE_PATH = os.path.join(HERE, "autotune_table.json")

# upper bounds of the set-size buckets (the batched workload spans 1,000 .. 200,000 points)
BUCKETS = (2_000, 5_000, 12_000, 30_000, 75_000, 200_000)
DENSITIES = (8.0, 12.0, 16.0, 24.0, 32.0, 48.0, 64.0, 96.0)
STAR_VARIANTS = (0, 1, 2, 3, 4)
PBA_BANDS = (1, 2, 4)


def _bucket_sets(lo, hi, total, seed):
    """Sets like the 4096-set workload's (uniform or one Gaussian, sigma 0.1, by seeded coin
    flip) with sizes log-uniform in [lo, hi], about `total` points in all."""
    from . import datasets
    mean = (hi - lo) / np.log(hi / lo)
    return datasets.batch_sets(max(2, int(round(total / mean))), seed=seed, lo=lo, hi=hi)


def _time_batch(tri, d_xyz, off, grids, reps=2):
    tri.triangulate_batch_device(d_xyz, off, grids)  # warm-up
    t0 = time.perf_counter()
    for _ in range(reps):
        tri.triangulate_batch_device(d_xyz, off, grids)
    return (time.perf_counter() - t0) / reps


def _grid(tri, n, vpp):
    tri.set_grid_table([1 << 40], [vpp])
    return tri.grid_for_set(n)


def tune(total_points=3_000_000, seed=0, log=print, path=TABLE_PATH):
    import gstar4d_amd as g
    rec = {"buckets": [], "star_kernel": None, "total_points_per_bucket": total_points, "seed": seed}
    lo = 1_000
    with g.Triangulator(grid_size=64) as tri:
        best_vpp = []
        for b, hi in enumerate(BUCKETS):
            sets = _bucket_sets(lo, hi, total_points, seed + b)
            sizes = np.array([s.shape[0] for s in sets], np.int64)
            off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
            d_xyz, n = tri.upload(np.concatenate(sets))
            times = {}
            for vpp in DENSITIES:
                grids = np.array([_grid(tri, int(k), vpp) for k in sizes], np.int32)
                times[vpp] = _time_batch(tri, d_xyz, off, grids)
                log("bucket <= %d (%d sets, %d points): %.0f voxels/point -> %.2f M points/s" %
                    (hi, len(sets), n, vpp, n / times[vpp] / 1e6))
            vpp = min(times, key=times.get)
            best_vpp.append(vpp)
            rec["buckets"].append({"max_points": hi, "voxels_per_point": vpp, "sets": len(sets), "points": int(n),
                                   "points_per_s": {str(k): round(n / v, 1) for k, v in times.items()}})
            lo = hi
        # star kernel variants on the whole mix with the tuned grids
        tri.set_grid_table(list(BUCKETS), best_vpp)
        sets = []
        lo = 1_000
        for b, hi in enumerate(BUCKETS):
            sets += _bucket_sets(lo, hi, total_points // 2, seed + 100 + b)
            lo = hi
        sizes = np.array([s.shape[0] for s in sets], np.int64)
        off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        d_xyz, n = tri.upload(np.concatenate(sets))
        grids = np.array([tri.grid_for_set(int(k)) for k in sizes], np.int32)
        vt = {}
        for v in STAR_VARIANTS:
            tri.set_star_kernel(v)
            vt[v] = _time_batch(tri, d_xyz, off, grids)
            log("star kernel variant %d: %.2f M points/s" % (v, n / vt[v] / 1e6))
        best = min(vt, key=vt.get)
        tri.set_star_kernel(best)
        rec["star_kernel"] = best
        rec["star_kernel_points_per_s"] = {str(k): round(n / v, 1) for k, v in vt.items()}
        # PBA bands of the y and x passes on the same mix
        bt = {}
        for nb in PBA_BANDS:
            tri.set_pba_bands(nb)
            bt[nb] = _time_batch(tri, d_xyz, off, grids)
            log("PBA bands %d: %.2f M points/s" % (nb, n / bt[nb] / 1e6))
        tri.set_pba_bands(0)
        tri.set_star_kernel(-1)
        rec["pba_bands"] = min(bt, key=bt.get)
        rec["pba_bands_points_per_s"] = {str(k): round(n / v, 1) for k, v in bt.items()}
    with open(path, "w") as f:
        json.dump(rec, f, indent=1)
    log("wrote %s: densities %s, star kernel %d, PBA bands %d" % (path, best_vpp, best, rec["pba_bands"]))
    return rec


def load(path=TABLE_PATH):
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def apply(tri, rec=None):
    """Applies a tuning record (default: the committed table) to a Triangulator; returns it."""
    rec = rec if rec is not None else load()
    if rec is None:
        return None
    tri.set_grid_table([b["max_points"] for b in rec["buckets"]], [b["voxels_per_point"] for b in rec["buckets"]])
    if rec.get("star_kernel") is not None:
        tri.set_star_kernel(int(rec["star_kernel"]))
    if rec.get("pba_bands") is not None:
        tri.set_pba_bands(int(rec["pba_bands"]))
    return rec


def grids(tri, sizes):
    """Grid of each set under the table applied to tri."""
    return np.array([tri.grid_for_set(int(n)) for n in sizes], np.int32)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=3_000_000, help="points per size bucket")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=TABLE_PATH)
    a = ap.parse_args(argv)
    tune(a.points, a.seed, log=lambda m: print(m, flush=True), path=a.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
