"""Benchmark of the MI355X clustering path.

Workload (one "step"): ParallelCorrelationClusterer::Cluster, sync,
NBHR_MOVE, resolution 0.05, refine=true, on a seeded RMAT graph of scale 22,
edge factor 16 (a=0.57, b=0.19, c=0.19), symmetrized, deduplicated, self-loops
removed, FP64 U[-1, 1] weights rounded to float32, seed 2.  The graph is
generated on the host and uploaded before the timed region; every step runs
all levels and the refinement on device-resident data.

Prints one JSON line.  `value` is best-move GTEPS: edges scanned by the
best-move kernels over all rounds of the timed steps divided by those kernels'
time (HIP events on the library's stream).  `roofline` describes the dominant
best-move kernel (the degree bin with the most time); `cpu_baseline` times the
oracle's first best-move round of level 0 on the same graph on the host cores.
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "parallel-correlation-clustering_amd"))

HBM_PEAK_GBS = 8000.0
BIN_NAMES = ["deg<=4 (4 lanes)", "deg<=8 (8 lanes)", "deg<=16 (16 lanes)", "deg<=32 (32 lanes)",
             "deg<=64 (64 lanes)", "one-wave LDS hash", "one-workgroup LDS hash",
             "global-memory hash"]
BIN_KERNELS = ["k_bm_small<4, false>", "k_bm_small<8, false>", "k_bm_small<16, false>",
               "k_bm_small<32, false>", "k_bm_small<64, false>", "k_bm_hash<64, 1024, false>",
               "k_bm_hash<512, 8192, false>", "k_bm_hash<512, 0, false>"]


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--edge-factor", type=float, default=16.0)
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads for the CPU baseline (0: all, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the clustering path is single-GPU in this build)", file=sys.stderr)
        return 2

    import numpy as np
    import pcc_amd

    t0 = time.time()
    csr = pcc_amd.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
    gen_s = time.time() - t0
    cfg = pcc_amd.make_config(resolution=0.05, move_method=pcc_amd.NBHR_MOVE, refine=1)
    t0 = time.time()
    g = pcc_amd.DeviceGraph.from_csr(csr)
    upload_s = time.time() - t0
    g.set_option(pcc_amd.OPT_TIMING, 1)
    print(f"# graph n={csr.n} m={csr.m} gen={gen_s:.1f}s upload={upload_s:.2f}s", file=sys.stderr,
          flush=True)

    for i in range(args.warmup):
        ids, st = g.cluster(cfg)
        print(f"# warmup {i}: {st.cluster_seconds*1e3:.1f} ms clusters={st.num_clusters}",
              file=sys.stderr, flush=True)

    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ids, st = g.cluster(cfg)
        stats.append(st)
        print(f"# step {i}: {st.cluster_seconds*1e3:.1f} ms rounds={st.rounds} "
              f"TE={st.edges_scanned} clusters={st.num_clusters}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0

    te = sum(s.edges_scanned for s in stats)
    bm_s = sum(s.bestmove_seconds for s in stats)
    cl_s = sum(s.cluster_seconds for s in stats)
    gteps = te / bm_s / 1e9 if bm_s > 0 else 0.0
    # Dominant best-move kernel: the bin with the most kernel time.
    bin_s = [sum(s.bin_seconds[b] for s in stats) for b in range(8)]
    bin_e = [sum(s.bin_edges[b] for s in stats) for b in range(8)]
    bin_v = [sum(s.bin_vertices[b] for s in stats) for b in range(8)]
    bin_l = [sum(s.bin_launches[b] for s in stats) for b in range(8)]
    dom = int(np.argmax(bin_s))
    dom_bytes = 12.0 * bin_e[dom] + 40.0 * bin_v[dom]
    dom_gbs = dom_bytes / bin_s[dom] / 1e9 if bin_s[dom] > 0 else 0.0
    all_gbs = sum(s.bestmove_bytes for s in stats) / bm_s / 1e9 if bm_s > 0 else 0.0

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(csr, cfg, args.cpu_threads)

    line = {
        "metric": "best-move GTEPS (sync NBHR_MOVE Louvain rounds, all levels + refinement)",
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / max(args.steps, 1) * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (move deltas) / f32 (edge weights)",
        "data": "synthetic (seeded RMAT generated on the host, resident in HBM before timing)",
        "config": {
            "workload": f"RMAT scale {args.scale} edge factor {args.edge_factor:g} "
                        "(a .57 b .19 c .19), U[-1,1] weights, seed 2; "
                        "ParallelCorrelationClusterer sync NBHR_MOVE resolution 0.05 refine",
            "n": csr.n,
            "m_directed": csr.m,
            "parallelism": "single GPU",
        },
        "end_to_end_ms": round(cl_s / max(args.steps, 1) * 1e3, 3),
        "end_to_end_gteps": round(te / cl_s / 1e9, 4) if cl_s > 0 else None,
        "edges_scanned_per_step": te // max(args.steps, 1),
        "rounds_per_step": stats[0].rounds if stats else 0,
        "levels": stats[0].levels if stats else 0,
        "num_clusters": stats[0].num_clusters if stats else 0,
        "objective": stats[0].objective if stats else None,
        "fallback_vertices_per_step": stats[0].fallback_vertices if stats else 0,
        "bestmove_hbm_gbs_all_bins": round(all_gbs, 2),
        "contract_ms_per_step": round(sum(s.contract_seconds for s in stats) / max(args.steps, 1) * 1e3, 3),
        "apply_frontier_ms_per_step": round(sum(s.apply_seconds for s in stats) / max(args.steps, 1) * 1e3, 3),
        "bins": [
            {"bin": BIN_NAMES[b], "kernel": BIN_KERNELS[b], "seconds": round(bin_s[b], 6),
             "edges": bin_e[b], "vertices": bin_v[b], "launches": bin_l[b],
             "avg_launch_us": round(bin_s[b] / bin_l[b] * 1e6, 2) if bin_l[b] else None,
             "gbs": round((12.0 * bin_e[b] + 40.0 * bin_v[b]) / bin_s[b] / 1e9, 2) if bin_s[b] else None}
            for b in range(8)
        ],
        "roofline": {
            "bound": "hbm",
            "kernel": BIN_KERNELS[dom],
            "achieved": round(dom_gbs, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(csr, cfg, threads):
    """Oracle best-move round 0 of level 0 (all vertices, singleton state)."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t = threads if threads > 0 else min(16, os.cpu_count() or 1)
    oracle.set_threads(t)
    g = oracle.G.from_csr(csr)
    s = oracle.S.singletons(csr.n)
    ocfg = oracle.config_from(cfg)
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.sync_decisions(g, s, ocfg)
        reps += 1
        el = time.perf_counter() - t0
        if el > 10.0 or reps >= 20:
            break
    gteps = reps * csr.m / el / 1e9
    return {
        "value": round(gteps, 4),
        "unit": "GTEPS",
        "cores": t,
        "kind": "port",
        "sample": f"{reps} x best-move round 0 of level 0 (all {csr.n} vertices, {csr.m} edges) "
                  f"of the same graph, {el:.1f} s",
    }


if __name__ == "__main__":
    sys.exit(main())
