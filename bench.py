"""Benchmark: points/s to the exact 3D Delaunay tetrahedralisation on one MI355X.

Workload (default): 4,194,304 FP64 points i.i.d. uniform in [0,1)^3 (SplitMix64 stream, seed 0),
PBA grid 512^3, tetrahedra with face adjacency.  One step = one full triangulation of the
HBM-resident points (Morton sort, PBA, pairs and worksets, star creation, splaying rounds,
lower-hull output, adjacency), result left in HBM.

Prints one JSON line.  Roofline: the dominant kernel, the 32-vertex LDS star kernel (k_star32,
FP64 vector work): algorithmic flops = 111 FP64 flops per filtered orient4 evaluation (the
reference's insphere filter) times the evaluations the kernel counts on the device, over its
launch time measured with HIP events on the library's stream; `executed_flops` is what the
cached-cofactor filter actually spends (15 per evaluation plus 64 per cofactor computation).
`roofline_pba`: the three PBA axis passes (HBM bound), 24 G^3 algorithmic bytes per
triangulation.  Traffic figures come from rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised
under profiles/ when present.  CPU baseline: the oracle's C restatement of the same pipeline,
single thread, on a 524,288-point sample at the same density (G = 256), timed on this host.
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
FP64_PEAK_TFLOPS = 78.6
FLOPS_PER_ORIENT4 = 111     # filtered orient4 (insphere filter, GD/GDelaunay/GDelShewchukDevice.h:1074-1136)
FLOPS_FILTER_EXECUTED = 15  # cached-cofactor filter: 4-term dot product and error bound
FLOPS_COFACTORS = 64        # cofactors and permanents of one link triangle


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


def _profile_value(name, workload, grid, key):
    """A figure from a rocprofv3 summary under profiles/ for this workload, or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("grid") != grid or rec.get("workload", workload) != workload:
        return None
    return rec.get(key)


def run_batch(args):
    """Batched workload: independent sets (sizes log-uniform in [1000, 200000], each uniform or
    one Gaussian), every set's grid from g4d_batch_grid_size.  One step = all sets triangulated
    by batched calls over consecutive sets of at most --chunk-points points; the per-set baseline
    triangulates each set with its own call (same grids, same results)."""
    import numpy as np
    import gstar4d_amd as g
    from gstar4d_amd import datasets

    t_gen = time.perf_counter()
    sets = datasets.batch_sets(args.sets, seed=args.seed)
    sizes = np.array([x.shape[0] for x in sets], np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    xyz = np.concatenate(sets)
    del sets
    grids = np.array([g.batch_grid_size(int(n)) for n in sizes], np.int32)
    chunks, a = [], 0
    while a < len(sizes):
        b = a + 1
        while b < len(sizes) and off[b + 1] - off[a] <= args.chunk_points:
            b += 1
        chunks.append((a, b))
        a = b
    print("batch: %d sets, %d points, %d chunks, generated in %.1f s" % (len(sizes), int(off[-1]), len(chunks),
                                                                         time.perf_counter() - t_gen),
          file=sys.stderr, flush=True)
    t = g.Triangulator(grid_size=64, adjacency=True)
    d_xyz, n = t.upload(xyz)

    def batch_pass():
        nt = 0
        for a, b in chunks:
            sub = off[a:b + 1] - off[a]
            nt += t.triangulate_batch_device(d_xyz + 24 * int(off[a]), sub, grids[a:b])
        return nt

    for _ in range(args.warmup):
        batch_pass()
    step_ms = []
    t_begin = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        nt = batch_pass()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        print("batch step %.1f ms" % step_ms[-1], file=sys.stderr, flush=True)
    total_s = time.perf_counter() - t_begin
    value = n * args.steps / total_s
    per_set = None
    if not args.no_per_set:
        s0 = time.perf_counter()
        nt1 = 0
        for k in range(len(sizes)):
            nt1 += t.triangulate_batch_device(d_xyz + 24 * int(off[k]), np.array([0, sizes[k]]), grids[k:k + 1])
            if k % 512 == 511:
                print("per-set %d / %d" % (k + 1, len(sizes)), file=sys.stderr, flush=True)
        per_s = time.perf_counter() - s0
        per_set = {"seconds": round(per_s, 3), "points_per_s": round(n / per_s, 1), "num_tets": int(nt1),
                   "same_tet_count": int(nt1) == int(nt)}
    t.close()
    out = {
        "metric": METRIC, "value": round(value, 1), "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(total_s * 1e3 / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": args.workload, "sets": int(len(sizes)), "points": int(n),
                   "set_sizes": [int(sizes.min()), int(sizes.max())], "grids": [int(grids.min()), int(grids.max())],
                   "calls_per_step": len(chunks), "chunk_points": args.chunk_points,
                   "distribution": "per set: uniform or one Gaussian (sigma 0.1)",
                   "output": "tetrahedra + face adjacency per set (HBM)"},
        "num_tets": int(nt), "step_ms": [round(v, 3) for v in step_ms],
        "per_set_baseline": per_set,
        "batch_speedup_vs_per_set": round(per_set["seconds"] / (total_s / args.steps), 3) if per_set else None,
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_uniform_4m_g512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sets", type=int, default=4096, help="batched workload: number of point sets")
    ap.add_argument("--chunk-points", type=int, default=24_000_000,
                    help="batched workload: points per batched call (consecutive sets)")
    ap.add_argument("--no-per-set", action="store_true", help="batched workload: skip the per-set baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X; runs on more than one GPU are not supported.", file=sys.stderr)
        sys.exit(2)
    if args.workload.startswith("c5_batch"):
        return run_batch(args)

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
    # Outside the timed region: one call through the host-buffer entry point (points copied in,
    # tetrahedra and adjacency copied out), the PCIe-inclusive rate.
    h0 = time.perf_counter()
    t.triangulate(x)
    host_ms = (time.perf_counter() - h0) * 1e3
    G3 = grid ** 3
    pba_bytes = 24 * G3
    pba_avg_ms = statistics.mean(pba_ms)
    achieved = pba_bytes / (pba_avg_ms * 1e-3) / 1e9
    traffic = _profile_value("pba_pass_pmc.json", args.workload, grid, "hbm_bytes_per_triangulation")
    last = stages[-1]
    # dominant kernel: k_star32
    k_ms = statistics.mean(s["t_star_t0_ms"] for s in stages)
    k_launch = last["star_t0_launches"]
    evals = statistics.mean(s["star_t0_filter_evals"] for s in stages)
    cofs = statistics.mean(s["star_t0_cofactors"] for s in stages)
    alg_flops = FLOPS_PER_ORIENT4 * evals
    exe_flops = FLOPS_FILTER_EXECUTED * evals + FLOPS_COFACTORS * cofs
    k_tflops = alg_flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    star_traffic = _profile_value("star_t0_pmc.json", args.workload, grid, "hbm_bytes_per_launch")
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
        "roofline": {"bound": "mfma", "compute": "FP64 vector ALU (no MFMA); latency-bound", "kernel": "k_star32",
                     "achieved": round(k_tflops, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(k_tflops / FP64_PEAK_TFLOPS, 5), "traffic": star_traffic,
                     "algorithmic_flops_per_launch": round(alg_flops / max(k_launch, 1)),
                     "executed_flops_per_launch": round(exe_flops / max(k_launch, 1)),
                     "orient4_evals_per_triangulation": int(evals), "launches": k_launch,
                     "avg_ms": round(k_ms / max(k_launch, 1), 4), "ms_per_triangulation": round(k_ms, 3)},
        "roofline_pba": {"bound": "hbm", "kernel": "k_pba_pass (z, y, x passes)",
                         "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                         "algorithmic_bytes": pba_bytes, "avg_ms": round(pba_avg_ms / 3, 4)},
        "stages_ms": {k: round(statistics.mean(s[k] for s in stages), 3) for k in
                      ("t_morton_ms", "t_pba_ms", "t_pairs_ms", "t_star_ms", "t_splay_ms", "t_out_ms",
                       "t_adj_ms", "t_total_ms")},
        "pba_pass_ms": [round(v, 4) for v in last["t_pba_pass_ms"]],
        "num_tets": int(nt),
        "splay_rounds": last["splay_rounds"],
        "exact_calls": last["exact_calls"],
        "sos_calls": last["sos_calls"],
        "step_ms": [round(v, 3) for v in step_ms],
        "host_buffers": {"ms": round(host_ms, 3), "points_per_s": round(n / (host_ms * 1e-3), 1),
                         "note": "g4d_triangulate: host points in, host tetrahedra + adjacency out (PCIe-inclusive)"},
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.seed)
    t.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
