"""Benchmark: points/s to the exact 3D Delaunay tetrahedralisation on one MI355X.

Headline workload (`value`): 4,194,304 FP64 points i.i.d. uniform in [0,1)^3 (SplitMix64 stream,
seed 0), PBA grid 512^3, tetrahedra with face adjacency.  One step = one full triangulation of
the HBM-resident points (Morton sort, PBA, pairs and worksets, star creation, splaying rounds,
lower-hull output, adjacency), result left in HBM.

Sub-records on the same line (each timed the same way, inputs resident in HBM):
  c4: 8,000,000 clustered points (70 % from 256 Gaussian clusters, sigma log-uniform in
      [0.002, 0.05], + 30 % uniform), G = 512: the plain pipeline and, when a trained proposal
      network is present, the learned-proposal pipeline; their ratio.
  c5: 4096 independent sets (sizes log-uniform in [1000, 200000], uniform or one Gaussian each,
      153.5 M points): batched calls over consecutive sets against one call per set.

Roofline: the dominant kernel, the 32-vertex LDS star kernel (k_star32, FP64 vector ALU work,
no MFMA): algorithmic flops = 111 FP64 flops per filtered orient4 evaluation (the reference's
insphere filter) times the evaluations the kernel counts on the device, over its launch time
measured with HIP events on the library's stream.  `roofline_pba`: the three PBA axis passes
(HBM bound), 24 G^3 algorithmic bytes per triangulation.  Traffic figures come from the
rocprofv3 FETCH_SIZE / WRITE_SIZE summaries under profiles/ when present.
CPU baseline: the oracle's C restatement of the same pipeline on the 100,000-point uniform
workload at G = 128, single-threaded and with OpenMP over this host's cores, median of 3.
"""
import argparse
import json
import os
import platform
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(seed=0, reps=3):
    """The oracle (C restatement of the reference's stages with its own host predicates) on the
    100,000-point uniform workload at G = 128, 1 thread and all host threads, median of reps."""
    from gstar4d_amd import datasets
    from oracle import orc
    x = datasets.make("c1_uniform_100k_g128", seed=seed)
    n = x.shape[0]
    res = {}
    for th in (1, host_threads()):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            tets, _, _ = orc.triangulate(x, 128, canonical=False, threads=th)
            ts.append(time.perf_counter() - t0)
        res[th] = (statistics.median(ts), int(tets.shape[0]))
        log("cpu baseline: %d thread(s) %.2f s" % (th, res[th][0]))
    nall = host_threads()
    return {"value": round(n / res[nall][0], 1), "unit": "points/s", "cores": nall, "kind": "port",
            "sample": "100000 uniform points, G=128 (the reference's CPU-runnable case); oracle C restatement, "
                      "OpenMP over %d threads, median of %d" % (nall, reps),
            "seconds": round(res[nall][0], 3),
            "single_thread": {"value": round(n / res[1][0], 1), "seconds": round(res[1][0], 3), "cores": 1},
            "cpu_model": cpu_model(), "num_tets": res[nall][1]}


def _traffic(workload, kernel):
    """HBM bytes per dispatch of the kernels whose names contain `kernel`, summed, from the
    rocprofv3 traffic summary of this workload under profiles/ (None when absent)."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    if rec.get("workload") != workload:
        return None
    vals = [v["hbm_bytes"] for k, v in rec["per_dispatch"].items() if kernel in k]
    return round(sum(vals)) if vals else None


def time_plain(t, d_xyz, n, steps, warmup, label):
    """Times `steps` triangulations of HBM-resident points; returns (seconds, step_ms, stats)."""
    for _ in range(warmup):
        t.triangulate_device(d_xyz, n)
    step_ms, stages = [], []
    nt = 0
    t_begin = time.perf_counter()
    for _ in range(steps):
        s0 = time.perf_counter()
        nt = t.triangulate_device(d_xyz, n)  # returns after the device work completed
        step_ms.append((time.perf_counter() - s0) * 1e3)
        stages.append(t.stats())
        log("%s step %.1f ms" % (label, step_ms[-1]))
    return time.perf_counter() - t_begin, step_ms, stages, nt


def run_c4(args):
    """Clustered 8M points at G = 512: plain pipeline, and the learned proposal when present."""
    import gstar4d_amd as g
    from gstar4d_amd import datasets
    x = datasets.make("c4_clustered_8m_g512", seed=args.seed)
    n = x.shape[0]
    out = {"workload": "c4_clustered_8m_g512", "points": n, "grid": 512}
    with g.Triangulator(grid_size=512, adjacency=True) as t:
        d_xyz, _ = t.upload(x)
        secs, step_ms, stages, nt = time_plain(t, d_xyz, n, args.c4_steps, 1, "c4 plain")
        last = stages[-1]
        out["plain"] = {"value": round(n * len(step_ms) / secs, 1), "ms_per_step": round(secs * 1e3 / len(step_ms), 3),
                        "splay_rounds": last["splay_rounds"], "store_repacks": last["store_repacks"],
                        "splay_insertions": last["splay_insertions"],
                        "num_tets": int(nt), "stages_ms": _stage_means(stages)}
        from gstar4d_amd import proposal
        if proposal.available():
            params, thr, rec = proposal.load_weights()
            out["proposal"] = {"weights": os.path.relpath(proposal.WEIGHTS_PATH, ROOT), "threshold": thr,
                               "trained_on": "%d sets of %d points, G=%d" % (rec["sets"], rec["points"], rec["grid"]),
                               "holdout_metrics": rec.get("holdout_metrics")}
            t.set_proposal(params, thr)
            secs2, step2, stages2, nt2 = time_plain(t, d_xyz, n, args.c4_steps, 1, "c4 learned")
            l2 = stages2[-1]
            out["learned"] = {"value": round(n * len(step2) / secs2, 1),
                              "ms_per_step": round(secs2 * 1e3 / len(step2), 3),
                              "splay_rounds": l2["splay_rounds"], "store_repacks": l2["store_repacks"],
                              "num_tets": int(nt2), "same_tet_count": int(nt2) == int(nt),
                              "prop_pairs": l2["prop_pairs"], "t_prop_ms": round(l2["t_prop_ms"], 3),
                              "splay_insertions": l2["splay_insertions"], "stages_ms": _stage_means(stages2)}
            out["learned_vs_plain"] = round(out["learned"]["value"] / out["plain"]["value"], 3)
            # ablation: the same candidates, every one proposed (the network's decisions removed)
            t.set_proposal(params, -1e30)
            secs3, step3, stages3, nt3 = time_plain(t, d_xyz, n, args.c4_steps, 1, "c4 all candidates")
            l3 = stages3[-1]
            out["all_candidates"] = {"value": round(n * len(step3) / secs3, 1),
                                     "ms_per_step": round(secs3 * 1e3 / len(step3), 3),
                                     "splay_rounds": l3["splay_rounds"], "prop_pairs": l3["prop_pairs"],
                                     "splay_insertions": l3["splay_insertions"], "num_tets": int(nt3),
                                     "stages_ms": _stage_means(stages3)}
            out["learned_vs_all_candidates"] = round(out["learned"]["value"] / out["all_candidates"]["value"], 3)
            t.set_proposal(enable=False)
        else:
            out["learned"] = None
            out["learned_vs_plain"] = None
    return out


def _stage_means(stages):
    return {k: round(statistics.mean(s[k] for s in stages), 3) for k in
            ("t_morton_ms", "t_pba_ms", "t_pairs_ms", "t_star_ms", "t_splay_ms", "t_out_ms", "t_adj_ms",
             "t_total_ms")}


def run_c5(args, steps, warmup, per_set=True):
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
    from gstar4d_amd import autotune
    t = g.Triangulator(grid_size=64, adjacency=True)
    tuned = autotune.apply(t)  # per-size grid table and star kernel from the autotuner, when present
    grids = (autotune.grids(t, sizes) if tuned else np.array([g.batch_grid_size(int(n)) for n in sizes], np.int32))
    chunks, a = [], 0
    while a < len(sizes):
        b = a + 1
        while b < len(sizes) and off[b + 1] - off[a] <= args.chunk_points:
            b += 1
        chunks.append((a, b))
        a = b
    log("batch: %d sets, %d points, %d chunks, generated in %.1f s" % (len(sizes), int(off[-1]), len(chunks),
                                                                     time.perf_counter() - t_gen))
    d_xyz, n = t.upload(xyz)

    def batch_pass():
        nt = 0
        for a, b in chunks:
            sub = off[a:b + 1] - off[a]
            nt += t.triangulate_batch_device(d_xyz + 24 * int(off[a]), sub, grids[a:b])
        return nt

    for _ in range(warmup):
        batch_pass()
    step_ms = []
    t_begin = time.perf_counter()
    nt = 0
    for _ in range(steps):
        s0 = time.perf_counter()
        nt = batch_pass()
        step_ms.append((time.perf_counter() - s0) * 1e3)
        log("batch step %.1f ms" % step_ms[-1])
    total_s = time.perf_counter() - t_begin
    value = n * steps / total_s
    base = None
    if per_set:
        s0 = time.perf_counter()
        nt1 = 0
        for k in range(len(sizes)):
            nt1 += t.triangulate_batch_device(d_xyz + 24 * int(off[k]), np.array([0, sizes[k]]), grids[k:k + 1])
            if k % 1024 == 1023:
                log("per-set %d / %d" % (k + 1, len(sizes)))
        per_s = time.perf_counter() - s0
        base = {"seconds": round(per_s, 3), "value": round(n / per_s, 1), "num_tets": int(nt1),
                "same_tet_count": int(nt1) == int(nt)}
    t.close()
    return {"workload": "c5_batch_4096", "value": round(value, 1), "ms_per_step": round(total_s * 1e3 / steps, 3),
            "steps": steps, "sets": int(len(sizes)), "points": int(n),
            "set_sizes": [int(sizes.min()), int(sizes.max())], "grids": [int(grids.min()), int(grids.max())],
            "calls_per_step": len(chunks), "chunk_points": args.chunk_points, "num_tets": int(nt),
            "autotuned": ({"grid_table": [[b["max_points"], b["voxels_per_point"]] for b in tuned["buckets"]],
                           "star_kernel": tuned.get("star_kernel")} if tuned else None),
            "step_ms": [round(v, 3) for v in step_ms], "per_set": base,
            "batched_vs_per_set": round(value / base["value"], 3) if base else None}


def run_headline(args):
    import gstar4d_amd as g
    from gstar4d_amd import datasets

    cfg = datasets.CONFIGS[args.workload]
    x = datasets.make(args.workload, seed=args.seed)
    n, grid = x.shape[0], cfg["grid"]
    t = g.Triangulator(grid_size=grid, adjacency=True)
    d_xyz, _ = t.upload(x)
    total_s, step_ms, stages, nt = time_plain(t, d_xyz, n, args.steps, args.warmup, args.workload)
    value = n * args.steps / total_s
    # Outside the timed region: the host-buffer entry point (points copied in, tetrahedra and
    # adjacency copied out into the context's page-locked buffers), the PCIe-inclusive rate; the
    # second call is timed (the first allocates and page-locks the result buffers).
    t.triangulate(x)
    h0 = time.perf_counter()
    t.triangulate(x)
    host_ms = (time.perf_counter() - h0) * 1e3
    t.close()
    G3 = grid ** 3
    pba_bytes = 24 * G3
    pba_avg_ms = statistics.mean(sum(s["t_pba_pass_ms"]) for s in stages)
    achieved = pba_bytes / (pba_avg_ms * 1e-3) / 1e9
    traffic = _traffic(args.workload, "k_pba_pass")
    last = stages[-1]
    # dominant kernel: k_star32
    k_ms = statistics.mean(s["t_star_t0_ms"] for s in stages)
    k_launch = last["star_t0_launches"]
    evals = statistics.mean(s["star_t0_filter_evals"] for s in stages)
    cofs = statistics.mean(s["star_t0_cofactors"] for s in stages)
    alg_flops = FLOPS_PER_ORIENT4 * evals
    exe_flops = FLOPS_FILTER_EXECUTED * evals + FLOPS_COFACTORS * cofs
    k_tflops = alg_flops / (k_ms * 1e-3) / 1e12 if k_ms > 0 else 0.0
    star_traffic = _traffic(args.workload, "k_star32")
    return {
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
        "roofline": {"bound": "valu", "compute": "FP64 vector ALU (no MFMA); latency-bound", "kernel": "k_star32",
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
        "stages_ms": _stage_means(stages),
        "pba_pass_ms": [round(v, 4) for v in last["t_pba_pass_ms"]],
        "num_tets": int(nt),
        "splay_rounds": last["splay_rounds"],
        "exact_calls": last["exact_calls"],
        "sos_calls": last["sos_calls"],
        "step_ms": [round(v, 3) for v in step_ms],
        "host_buffers": {"ms": round(host_ms, 3), "points_per_s": round(n / (host_ms * 1e-3), 1),
                         "note": "g4d_triangulate: host points in, host tetrahedra + adjacency out (PCIe-inclusive), "
                                  "second call"},
    }


# ---------------------------------------------------------------------------------------
# Profiling support: the workload driver that rocprofv3 runs, and the summarisers that turn
# its CSV output into the files under profiles/.
#
#   rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace -o run -- \
#       python3 bench.py --profile-run --workload c2_uniform_4m_g512 --steps 1
#   rocprofv3 --pmc FETCH_SIZE --output-format csv -d OUT/fetch -o run -- python3 bench.py --profile-run ...
#   rocprofv3 --pmc WRITE_SIZE --output-format csv -d OUT/write -o run -- python3 bench.py --profile-run ...
#   python bench.py --summarize-trace OUT/trace --out profiles/<kernel>_<workload>_kernel_stats.txt --cmd "..."
#   python bench.py --summarize-traffic OUT/fetch OUT/write --kernels "k_star32|k_pba_pass" --out ... --cmd "..."
#   python bench.py --summarize-counters OUT/c1 OUT/c2 ... --kernels "k_star32" --out ... --cmd "..."
# ---------------------------------------------------------------------------------------

def profile_run(args):
    """One plain copy kernel of known size (it calibrates the HBM traffic counters), then
    `steps` triangulations of the workload on HBM-resident points (learned proposal on when
    --learned and weights are present)."""
    import gstar4d_amd as g
    from gstar4d_amd import datasets
    cfg = datasets.CONFIGS[args.workload]
    x = datasets.make(args.workload, seed=args.seed)
    with g.Triangulator(grid_size=cfg["grid"], adjacency=True) as t:
        copy_bytes = args.copy_mib << 20
        copy_ms = t.diag_copy(copy_bytes)
        if args.learned:
            from gstar4d_amd import proposal
            params, thr, _ = proposal.load_weights()
            t.set_proposal(params, thr)
        d, n = t.upload(x)
        nt = 0
        for k in range(args.steps):
            nt = t.triangulate_device(d, n)
            log("profile step %d done" % k)
        st = t.stats()
    print(json.dumps({"workload": args.workload, "points": n, "num_tets": nt, "copy_bytes": copy_bytes,
                      "copy_ms": round(copy_ms, 4), "copy_GBps": round(2 * copy_bytes / copy_ms / 1e6, 1),
                      "t_pba_pass_ms": st["t_pba_pass_ms"], "t_total_ms": st["t_total_ms"]}), flush=True)


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(root, suffix):
    for dp, _, files in sorted(os.walk(root)):
        for fn in sorted(files):
            if fn.endswith(suffix):
                return os.path.join(dp, fn)
    raise FileNotFoundError("no *%s under %s" % (suffix, root))


def summarize_trace(prof_dir, command, top=30):
    """Text table of a rocprofv3 --kernel-trace --stats run (its kernel_stats.csv)."""
    path = _find(prof_dir, "kernel_stats.csv")
    rows = _csv_rows(path)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    out = ["command: " + command, "source: " + os.path.basename(path), "",
           "%-58s %7s %12s %12s %6s" % ("kernel", "calls", "avg_us", "total_ms", "pct")]
    for r in rows[:top]:
        name = r["Name"].split("(")[0][:58]
        out.append("%-58s %7d %12.2f %12.3f %6.2f" % (name, int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                                                      float(r["TotalDurationNs"]) / 1e6,
                                                      100.0 * float(r["TotalDurationNs"]) / total))
    out.append("%-58s %7s %12s %12.3f" % ("all kernels", "", "", total / 1e6))
    return "\n".join(out) + "\n"


def pmc_per_kernel(prof_dir):
    """{counter: {kernel: mean value per dispatch}} from a --pmc run's counter_collection.csv."""
    acc = {}
    for r in _csv_rows(_find(prof_dir, "counter_collection.csv")):
        name = r["Kernel_Name"].split("(")[0]
        acc.setdefault(r["Counter_Name"], {}).setdefault(name, []).append(float(r["Counter_Value"]))
    return {c: {k: sum(v) / len(v) for k, v in ks.items()} for c, ks in acc.items()}


def summarize_traffic(fetch_dir, write_dir, kernels, copy_bytes, workload, command):
    """HBM bytes per dispatch of the kernels matching `kernels` (a regex) from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the same counter reads for the plain
    copy kernel, which reads and writes copy_bytes once."""
    import re
    f = pmc_per_kernel(fetch_dir)["FETCH_SIZE"]
    w = pmc_per_kernel(write_dir)["WRITE_SIZE"]
    ck = [k for k in f if "k_diag_copy" in k][0]
    sf = copy_bytes / (f[ck] * 1024.0)
    sw = copy_bytes / (w[ck] * 1024.0)
    per = {}
    for k in f:
        if k != ck and re.search(kernels, k):
            per[k] = {"fetch_bytes": f[k] * 1024.0 * sf, "write_bytes": w.get(k, 0.0) * 1024.0 * sw}
            per[k]["hbm_bytes"] = per[k]["fetch_bytes"] + per[k]["write_bytes"]
    return {"workload": workload, "command": command, "copy_bytes": copy_bytes, "fetch_scale": sf,
            "write_scale": sw, "copy_fetch_kib": f[ck], "copy_write_kib": w[ck], "per_dispatch": per}


def summarize_counters(dirs, kernels, command):
    """Mean value per dispatch of every counter collected in the --pmc runs `dirs`, for the
    kernels matching `kernels`, with the derived lane utilisation of the vector ALU
    (SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)) when both are present."""
    import re
    out = {}
    for d in dirs:
        for c, ks in pmc_per_kernel(d).items():
            for k, v in ks.items():
                if re.search(kernels, k):
                    out.setdefault(k, {})[c] = v
    for k, cs in out.items():
        if cs.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in cs:
            cs["valu_lane_utilisation"] = cs["SQ_THREAD_CYCLES_VALU"] / (64.0 * cs["SQ_ACTIVE_INST_VALU"])
    return {"command": command, "per_dispatch": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2_uniform_4m_g512")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the clustered 8M sub-record")
    ap.add_argument("--no-c5", action="store_true", help="skip the 4096-set batched sub-record")
    ap.add_argument("--c4-steps", type=int, default=3)
    ap.add_argument("--c5-steps", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--sets", type=int, default=4096, help="batched workload: number of point sets")
    ap.add_argument("--chunk-points", type=int, default=24_000_000,
                    help="batched workload: points per batched call (consecutive sets)")
    ap.add_argument("--no-per-set", action="store_true", help="batched workload: skip the per-set baseline")
    ap.add_argument("--profile-run", action="store_true",
                    help="profiling driver: one copy kernel, then --steps triangulations of --workload")
    ap.add_argument("--learned", action="store_true", help="profiling driver: learned proposal on")
    ap.add_argument("--copy-mib", type=int, default=1024, help="profiling driver: size of the copy kernel")
    ap.add_argument("--summarize-trace", metavar="DIR", help="write a kernel table of a --kernel-trace --stats run")
    ap.add_argument("--summarize-traffic", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="HBM bytes per dispatch from FETCH_SIZE and WRITE_SIZE runs")
    ap.add_argument("--summarize-counters", nargs="+", metavar="DIR", help="counters per dispatch of --pmc runs")
    ap.add_argument("--kernels", default="k_star32|k_pba_pass", help="summaries: kernel name regex")
    ap.add_argument("--cmd", default="", help="summaries: the command lines that produced the runs")
    ap.add_argument("--out", help="summaries: output file")
    args = ap.parse_args()
    if args.profile_run:
        profile_run(args)
        return
    if args.summarize_trace or args.summarize_traffic or args.summarize_counters:
        if args.summarize_trace:
            text = summarize_trace(args.summarize_trace, args.cmd)
        elif args.summarize_traffic:
            text = json.dumps(summarize_traffic(args.summarize_traffic[0], args.summarize_traffic[1], args.kernels,
                                                args.copy_mib << 20, args.workload, args.cmd), indent=1) + "\n"
        else:
            text = json.dumps(summarize_counters(args.summarize_counters, args.kernels, args.cmd), indent=1) + "\n"
        if args.out:
            with open(args.out, "w") as f:
                f.write(text)
        else:
            sys.stdout.write(text)
        return
    if args.gpus != 1:
        print("bench.py measures one MI355X; runs on more than one GPU are not supported.", file=sys.stderr)
        sys.exit(2)
    if args.workload.startswith("c5_batch"):
        rec = run_c5(args, args.steps, args.warmup, per_set=not args.no_per_set)
        out = {"metric": METRIC, "value": rec["value"], "unit": "points/s", "n_gpus": 1, "steps": args.steps,
               "warmup": args.warmup, "ms_per_step": rec["ms_per_step"], "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": args.workload, "sets": rec["sets"], "points": rec["points"]}, "c5": rec}
        print(json.dumps(out), flush=True)
        return
    out = run_headline(args)
    if not args.no_c4 and args.workload != "c4_clustered_8m_g512":
        out["c4"] = run_c4(args)
    if not args.no_c5:
        out["c5"] = run_c5(args, args.c5_steps, 1, per_set=not args.no_per_set)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.seed)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
