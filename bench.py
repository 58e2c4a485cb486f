"""Benchmark: points/s to the exact 3D Delaunay tetrahedralisation on one MI355X.

Workload (default): 4,194,304 FP64 points i.i.d. uniform in [0,1)^3 (SplitMix64 stream, seed 0),
PBA grid 512^3, tetrahedra with face adjacency.  One step = one full triangulation of the
HBM-resident points (Morton sort, PBA, pairs and worksets, star creation, splaying rounds,
lower-hull output, adjacency), result left in HBM.

Prints one JSON line.  Roofline: the three PBA axis-pass kernels (HBM bound), algorithmic
bytes 24 G^3 per triangulation (each pass reads and writes the int32 grid once), timed with
HIP events recorded on the library's stream around those launches.  CPU baseline: the
oracle's C restatement of the same pipeline, single thread, on a 524,288-point sample at the
same density (G = 256), timed on this host.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "points/s to exact 3D DT on 1 MI355X: plain vs learned-proposal+batched"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(seed=0):
    from gstar4d_amd import datasets
    from oracle import orc
    n, grid = 524288, 256
    x = datasets.uniform_cube(n, seed=seed)
    t0 = time.perf_counter()
    tets, _, st = orc.triangulate(x, grid, canonical=False)
    dt = time.perf_counter() - t0
    return {"value": round(x.shape[0] / dt, 1), "unit": "points/s", "cores": 1, "kind": "port",
            "sample": "524288 uniform points, G=256 (same points-per-voxel as 4M at 512^3), "
                      "oracle C restatement single-threaded, %.1f s" % dt,
            "seconds": round(dt, 2), "num_tets": int(tets.shape[0])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_uniform_4m_g512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X; multi-GPU runs are not part of this build "
              "(sharding is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    import gstar4d_amd as g
    from gstar4d_amd import datasets

    cfg = datasets.CONFIGS[args.workload]
    x = datasets.make(args.workload, seed=args.seed)
    n, grid = x.shape[0], cfg["grid"]
    t = g.Triangulator(grid_size=grid, adjacency=True)
    d_xyz, _ = t.upload(x)
    for _ in range(args.warmup):
        t.triangulate_device(d_xyz, n)
    step_ms, pba_ms, stages = [], [], []
    nt = 0
    t_begin = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        nt = t.triangulate_device(d_xyz, n)  # returns after the device work completed
        step_ms.append((time.perf_counter() - s0) * 1e3)
        st = t.stats()
        pba_ms.append(sum(st["t_pba_pass_ms"]))
        stages.append(st)
    total_s = time.perf_counter() - t_begin
    value = n * args.steps / total_s
    G3 = grid ** 3
    pba_bytes = 24 * G3
    pba_avg_ms = statistics.mean(pba_ms)
    achieved = pba_bytes / (pba_avg_ms * 1e-3) / 1e9
    traffic = None
    pmc = os.path.join(ROOT, "profiles", "pba_pass_pmc.json")
    if os.path.exists(pmc):
        with open(pmc) as f:
            rec = json.load(f)
        if rec.get("grid") == grid:
            traffic = rec.get("hbm_bytes_per_triangulation")
    last = stages[-1]
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_s * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": args.workload, "points": n, "grid": grid, "seed": args.seed,
                   "distribution": cfg["gen"], "output": "tetrahedra + face adjacency (HBM)"},
        "roofline": {"bound": "hbm", "kernel": "k_pba_pass (z, y, x passes)",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": pba_bytes, "avg_ms": round(pba_avg_ms, 4)},
        "stages_ms": {k: round(statistics.mean(s[k] for s in stages), 3) for k in
                      ("t_morton_ms", "t_pba_ms", "t_pairs_ms", "t_star_ms", "t_splay_ms", "t_out_ms",
                       "t_adj_ms", "t_total_ms")},
        "pba_pass_ms": [round(v, 4) for v in last["t_pba_pass_ms"]],
        "num_tets": int(nt),
        "splay_rounds": last["splay_rounds"],
        "exact_calls": last["exact_calls"],
        "sos_calls": last["sos_calls"],
        "step_ms": [round(v, 3) for v in step_ms],
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.seed)
    t.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
