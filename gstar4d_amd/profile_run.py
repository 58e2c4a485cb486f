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


if __name__ == "__main__":
    sys.exit(main())
