"""Driver for rocprofv3 runs: one plain copy kernel of known size (to calibrate the HBM
traffic counters) followed by triangulations of a named workload on HBM-resident points.

    python gstar4d_amd/profile_run.py --workload c2_uniform_4m_g512 --copy-mib 1024 --steps 1
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import gstar4d_amd as g  # noqa: E402
from gstar4d_amd import datasets  # noqa: E402


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="c2_uniform_4m_g512")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    cfg = datasets.CONFIGS[a.workload]
    x = datasets.make(a.workload, seed=a.seed)
    with g.Triangulator(grid_size=cfg["grid"], adjacency=True) as t:
        copy_bytes = a.copy_mib << 20
        copy_ms = t.diag_copy(copy_bytes)
        d, n = t.upload(x)
        for _ in range(a.steps):
            nt = t.triangulate_device(d, n)
        st = t.stats()
    print(json.dumps({"workload": a.workload, "points": n, "num_tets": nt, "copy_bytes": copy_bytes,
                      "copy_ms": round(copy_ms, 4), "copy_GBps": round(2 * copy_bytes / copy_ms / 1e6, 1),
                      "t_pba_pass_ms": st["t_pba_pass_ms"], "t_total_ms": st["t_total_ms"]}), flush=True)
    return 0


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _find(root, suffix):
    for dp, _, files in os.walk(root):
        for fn in sorted(files):
            if fn.endswith(suffix):
                return os.path.join(dp, fn)
    return None


def summarize_trace(prof_dir, command, top=25):
    """Text table of rocprofv3 --kernel-trace --stats output (kernel_stats.csv)."""
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


def pmc_per_kernel(prof_dir, counter):
    """Mean counter value per dispatch, by kernel name, from a --pmc run."""
    path = _find(prof_dir, "counter_collection.csv")
    acc = {}
    for r in _csv_rows(path):
        if r.get("Counter_Name") != counter:
            continue
        name = r["Kernel_Name"].split("(")[0]
        acc.setdefault(name, []).append(float(r["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in acc.items()}


def pba_traffic(fetch_dir, write_dir, copy_bytes, grid):
    """HBM bytes per triangulation of the three k_pba_pass launches from FETCH_SIZE and
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for the plain copy
    kernel of copy_bytes (which reads and writes copy_bytes once)."""
    f = pmc_per_kernel(fetch_dir, "FETCH_SIZE")
    w = pmc_per_kernel(write_dir, "WRITE_SIZE")
    ck = [k for k in f if "k_diag_copy" in k][0]
    sf = copy_bytes / (f[ck] * 1024.0)
    sw = copy_bytes / (w[ck] * 1024.0)
    passes = {}
    for k in f:
        if "k_pba_pass" in k:
            passes[k] = {"fetch_bytes": f[k] * 1024.0 * sf, "write_bytes": w.get(k, 0.0) * 1024.0 * sw}
    total = sum(v["fetch_bytes"] + v["write_bytes"] for v in passes.values())
    return {"grid": grid, "copy_bytes": copy_bytes, "fetch_scale": sf, "write_scale": sw,
            "copy_fetch_kib": f[ck], "copy_write_kib": w[ck], "passes": passes,
            "hbm_bytes_per_triangulation": total, "algorithmic_bytes": 24 * grid ** 3}



def kernel_traffic(fetch_dir, write_dir, copy_bytes, grid, kernel, workload):
    """Mean HBM bytes per launch of `kernel` (a name substring) from FETCH_SIZE and WRITE_SIZE,
    scaled by the copy kernel as in pba_traffic."""
    f = pmc_per_kernel(fetch_dir, "FETCH_SIZE")
    w = pmc_per_kernel(write_dir, "WRITE_SIZE")
    ck = [k for k in f if "k_diag_copy" in k][0]
    sf = copy_bytes / (f[ck] * 1024.0)
    sw = copy_bytes / (w[ck] * 1024.0)
    ks = [k for k in f if kernel in k and "k_diag_copy" not in k]
    fetch = sum(f[k] for k in ks) * 1024.0 * sf
    write = sum(w.get(k, 0.0) for k in ks) * 1024.0 * sw
    return {"workload": workload, "grid": grid, "kernel": kernel, "copy_bytes": copy_bytes, "fetch_scale": sf,
            "write_scale": sw, "fetch_bytes_per_launch": fetch, "write_bytes_per_launch": write,
            "hbm_bytes_per_launch": fetch + write}


if __name__ == "__main__":
    sys.exit(main())
