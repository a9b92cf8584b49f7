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
best-move kernel (the degree bin with the most time).  The same line carries
the metric's second half, `batched`: graphs/s of one batched launch sequence
over a size-stratified sample of the 4,096-graph batched workload against
plain per-graph runs of the same graphs, with the count of identical results;
`parity_full_size`: the headline run's ids against the digest of the oracle's
ids on the same graph (tests/golden/fullsize_rmat22_nbhr_refine.json); and
`cpu_baseline`: the oracle's whole Cluster span on a bounded sample (the same
generator and configuration at RMAT scale 20) on the host cores, beside the
GPU's span on that sample.
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
             "partitioned LDS hash (hub vertices)"]
BIN_KERNELS = ["k_bm_small<4>", "k_bm_small<8>", "k_bm_small<16>", "k_bm_small<32>",
               "k_bm_small<64>", "k_bm_hash<64,256> + k_bm_hash<64,512> + k_bm_hash<128,1024>",
               "k_bm_hash<256,2048> + k_bm_hash<512,4096> + k_bm_hash<1024,8192>",
               "k_hub_scan + k_hub_gather + k_bm_hub_part + k_bm_hub_final"]


def apply_tuning(g, args, pcc_amd):
    """Degree-bin thresholds and hub partition size (defaults: 512 / 4096 / 2048)."""
    if args.wave_max:
        g.set_option(pcc_amd.OPT_WAVE_MAX_DEG, args.wave_max)
    if args.block_max:
        g.set_option(pcc_amd.OPT_BLOCK_MAX_DEG, args.block_max)
    if args.hub_part_edges:
        g.set_option(pcc_amd.OPT_HUB_PART_EDGES, args.hub_part_edges)
    if args.async_subrounds is not None:
        g.set_option(pcc_amd.OPT_ASYNC_SUBROUNDS, args.async_subrounds)
    if args.pull_fraction is not None:
        g.set_option(pcc_amd.OPT_PULL_FRACTION, args.pull_fraction)
    if args.hash_waves is not None:
        g.set_option(pcc_amd.OPT_HASH_WAVES, args.hash_waves)


def set_tuned(g, t, pcc_amd):
    """Applies an autotune result (pcc_autotune) to another handle."""
    g.set_option(pcc_amd.OPT_WAVE_MAX_DEG, int(t["wave_max_deg"]))
    g.set_option(pcc_amd.OPT_BLOCK_MAX_DEG, int(t["block_max_deg"]))
    g.set_option(pcc_amd.OPT_HUB_PART_EDGES, int(t["hub_part_edges"]))


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=None,
                    help="RMAT scale (default 22; modularity-async: 24)")
    ap.add_argument("--edge-factor", type=float, default=None,
                    help="RMAT edge factor (default 16; modularity-async: 32)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads for the CPU baseline (0: all, at most 16)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="rmat22: skip the batched sample of the headline line")
    ap.add_argument("--batched-sample", type=int, default=512,
                    help="rmat22: graphs of the size-stratified batched sample (of 4,096)")
    ap.add_argument("--cpu-scale", type=int, default=20,
                    help="rmat22: RMAT scale of the CPU baseline's sample")
    ap.add_argument("--no-gather-ceiling", action="store_true",
                    help="skip the random-gather ceiling (a child process using torch)")
    ap.add_argument("--workload", default="rmat22",
                    choices=["rmat22", "batched", "modularity-async", "serving", "copy", "gather",
                             "cluster-moves", "subclustering", "sbm-allmove"],
                    help="rmat22: the headline line; the others are secondary measurements")
    ap.add_argument("--graphs", type=int, default=4096, help="batched: number of SBM graphs")
    ap.add_argument("--batch-edges", type=int, default=1 << 30,
                    help="batched: maximum directed edges per device batch")
    ap.add_argument("--plain-graphs", type=int, default=0,
                    help="batched: per-graph runs timed for the comparison (0: all)")
    ap.add_argument("--requests", type=int, default=64, help="serving: requests in the stream")
    ap.add_argument("--qps", type=float, default=8.0, help="serving: Poisson arrival rate")
    ap.add_argument("--qps-sweep", default="4,8,16,32,64,0",
                    help="serving: arrival rates to sweep (0: every request at once); empty: --qps only")
    ap.add_argument("--wave-max", type=int, default=0, help="upper degree of the one-wave bin")
    ap.add_argument("--block-max", type=int, default=0, help="upper degree of the workgroup bin")
    ap.add_argument("--hub-part-edges", type=int, default=0, help="edges per hub partition")
    ap.add_argument("--async-subrounds", type=int, default=None, help="async sub-rounds (0: auto)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched: keep the default degree-bin thresholds instead of pcc_autotune")
    ap.add_argument("--autotune", action="store_true",
                    help="rmat22: run pcc_autotune on the workload first and report its choice")
    ap.add_argument("--hash-waves", type=int, default=None,
                    help="hash-bin workgroups per launch, in chip-fulls (0: one per vertex)")
    ap.add_argument("--subclustering-method", type=int, default=None,
                    help="subclustering / cluster-moves: 1 TRIANGLE, 2 CONNECTIVITY, "
                         "3 SLOW_CONNECTIVITY (defaults 1 / 0)")
    ap.add_argument("--pull-fraction", type=int, default=None,
                    help="frontier pull threshold in 1/1024 of m + n (1024: push only)")
    args = ap.parse_args()
    big = args.workload == "modularity-async"
    if args.scale is None:
        args.scale = 24 if big else 22
    if args.edge_factor is None:
        args.edge_factor = 32.0 if big else 16.0
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the clustering path is single-GPU in this build)", file=sys.stderr)
        return 2
    if args.workload == "batched":
        return bench_batched(args)
    if args.workload == "modularity-async":
        return bench_modularity_async(args)
    if args.workload == "serving":
        return bench_serving(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "gather":
        return bench_gather(args)
    if args.workload == "cluster-moves":
        return bench_cluster_moves(args)
    if args.workload == "subclustering":
        return bench_subclustering(args)
    if args.workload == "sbm-allmove":
        return bench_sbm_allmove(args)

    import numpy as np
    import pcc_amd

    t0 = time.time()
    csr = pcc_amd.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
    gen_s = time.time() - t0
    cfg = pcc_amd.make_config(resolution=0.05, move_method=pcc_amd.NBHR_MOVE, refine=1)
    t0 = time.time()
    g = pcc_amd.DeviceGraph.from_csr(csr)
    upload_s = time.time() - t0
    apply_tuning(g, args, pcc_amd)
    print(f"# graph n={csr.n} m={csr.m} gen={gen_s:.1f}s upload={upload_s:.2f}s", file=sys.stderr,
          flush=True)
    tune = None
    if args.autotune:
        tune = g.autotune(cfg, max_trials=8)
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    g.set_option(pcc_amd.OPT_TIMING, 1)

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

    parity = full_size_parity(ids, stats[-1], args)
    g.close()
    # torch and this library bring separate HIP runtimes, so the ceiling is
    # measured in a child process.
    gceil = None if args.no_gather_ceiling else gather_ceiling_child(args.scale)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(csr, cfg, args)
    n_v, m_e = csr.n, csr.m
    del csr
    batched = None if args.no_batched else batched_sample(args)

    line = {
        # The headline metric has two halves: `value` is best-move GTEPS over
        # the sync NBHR_MOVE rounds of all levels and the refinement, the
        # `batched` object the batched graphs/s against plain runs.
        "metric": "best-move GTEPS + batched graphs/s vs plain GPU path at reference-equal objective",
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
            "n": n_v,
            "m_directed": m_e,
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
            "traffic": measured_traffic(args, dom),
            "l2_hit_rate": measured_group(args, dom, "l2_rmat22.json", "l2_hit_rate"),
            "algorithmic_bytes_per_launch": round(dom_bytes / max(bin_l[dom], 1)),
        },
        # The dominant kernels wait on random lookups rather than on bytes:
        # one ids[u] gather per scanned edge (plus K[c] per distinct cluster,
        # not counted here), against the measured random-gather rate.
        "gather_roofline": None if not gceil else {
            "bound": "random 4-byte gathers",
            "achieved": round(bin_e[dom] / bin_s[dom] / 1e9, 2) if bin_s[dom] > 0 else 0.0,
            "peak": round(gceil / 1e9, 2),
            "unit": "G gathers/s",
            "frac": round(bin_e[dom] / bin_s[dom] / gceil, 4) if bin_s[dom] > 0 else 0.0,
            "peak_source": f"torch.index_select of 45M random indices into {n_v} u32 entries "
                           "(bench.py --workload gather)",
        },
        "cpu_baseline": cpu,
        "parity_full_size": parity,
        "batched": batched,
    }
    if tune:
        line["autotune"] = tune
    print(json.dumps(line), flush=True)
    return 0


def full_size_parity(ids, st, args):
    """The headline run's ids against the oracle's (sha256 digests from
    tests/golden/fullsize_rmat22_nbhr_refine.json, written by
    tests/golden/make_fullsize.py)."""
    import hashlib

    import numpy as np

    path = os.path.join(ROOT, "tests", "golden", "fullsize_rmat22_nbhr_refine.json")
    if args.scale != 22 or args.edge_factor != 16.0 or not os.path.exists(path):
        return None
    with open(path) as f:
        fx = json.load(f)
    digest = hashlib.sha256(np.ascontiguousarray(ids, "<u4").tobytes()).hexdigest()
    return {
        "reference": "oracle (CPU restatement), whole Cluster run on the same graph",
        "ids_equal": digest == fx["ids_sha256"],
        "num_clusters_equal": int(st.num_clusters) == fx["num_clusters"],
        "objective_equal": float(st.objective) == float.fromhex(fx["objective_hex"]),
        "rounds_equal": int(st.rounds) == fx["rounds"],
    }


def bench_subclustering(args):
    """Subclustering between LOUVAIN levels on the default workload's graph and
    configuration (sync NBHR_MOVE, resolution 0.05, refine).  value = edges
    scanned by the best-move rounds per second of the whole Cluster call,
    subcluster numbering and compression included.  The CPU figure is the
    oracle's whole run, same configuration, on an RMAT graph of scale 16."""
    import numpy as np
    import pcc_amd

    names = {1: "TRIANGLE", 2: "CONNECTIVITY", 3: "SLOW_CONNECTIVITY"}
    method = 1 if args.subclustering_method is None else args.subclustering_method
    csr = pcc_amd.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
    cfg = pcc_amd.make_config(resolution=0.05, move_method=pcc_amd.NBHR_MOVE, refine=1,
                              subclustering_method=method)
    g = pcc_amd.DeviceGraph.from_csr(csr)
    g.set_option(pcc_amd.OPT_TIMING, 1)
    for _ in range(args.warmup):
        g.cluster(cfg)
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ids, st = g.cluster(cfg)
        stats.append(st)
        print(f"# step {i}: {st.cluster_seconds*1e3:.1f} ms rounds={st.rounds} levels={st.levels} "
              f"contract+subclusters={st.contract_seconds*1e3:.1f} ms clusters={st.num_clusters}",
              file=sys.stderr, flush=True)
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    st = stats[-1]
    rate = st.edges_scanned / st.cluster_seconds / 1e9
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        small = pcc_amd.gen_rmat(16, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
        r = oracle.cluster(oracle.G.from_csr(small), oracle.config_from(cfg))
        oracle.set_threads(1)
        cpu = {"value": round(r.edges_scanned / r.seconds / 1e9, 4), "unit": "GTEPS",
               "cores": threads, "kind": "port",
               "sample": f"whole run on RMAT scale 16 edge factor {args.edge_factor:g} "
                         f"(m={small.m}), same configuration: {r.seconds:.2f} s; the numbering "
                         "is single-threaded in the oracle"}
    line = {
        "metric": f"Cluster edge rate with {names.get(method, method)} subclustering "
                  "(sync NBHR_MOVE LOUVAIN, all levels + refinement)",
        "value": round(rate, 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (move deltas) / f32 (edge weights)",
        "data": "synthetic (seeded RMAT generated on the host, resident in HBM before timing)",
        "config": {"workload": f"RMAT scale {args.scale} edge factor {args.edge_factor:g}, "
                               "U[-1,1] weights, seed 2; sync NBHR_MOVE resolution 0.05, refine, "
                               f"subclustering_method {names.get(method, method)}",
                   "n": csr.n, "m_directed": csr.m, "parallelism": "single GPU"},
        "cluster_ms": round(st.cluster_seconds * 1e3, 3),
        "contract_and_subcluster_ms": round(st.contract_seconds * 1e3, 3),
        "bestmove_ms": round(st.bestmove_seconds * 1e3, 3),
        "rounds": st.rounds, "levels": st.levels, "edges_scanned": st.edges_scanned,
        "objective": st.objective, "num_clusters": st.num_clusters,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_sbm_allmove(args):
    """The small planted-partition configuration: SBM with 10,000 vertices in
    100 communities (p_in 0.1, p_out 0.001, unit weights), sync ALL_MOVE,
    resolution 0.01, 10 x 10 iterations, no refinement.  Every level is a
    chain of short dependent launches, so this measures launch and read-back
    latency: value = ms per Cluster call (wall).  The CPU figure is the
    oracle's single-threaded run of the same call, ids compared."""
    import numpy as np
    import pcc_amd

    csr = pcc_amd.gen_sbm(10000, 100, 0.1, 0.001, seed=1)
    cfg = pcc_amd.parse_config("correlation_clusterer_config { resolution: 0.01 move_method: ALL_MOVE "
                               "louvain_config { num_iterations: 10 num_inner_iterations: 10 } }")
    g = pcc_amd.DeviceGraph.from_csr(csr)
    steps = max(args.steps, 1)
    for _ in range(max(args.warmup, 1)):
        ids, st = g.cluster(cfg)
    walls = []
    for _ in range(steps):
        t0 = time.perf_counter()
        ids, st = g.cluster(cfg)
        walls.append(time.perf_counter() - t0)
    g.close()
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        oracle.set_threads(1)
        r = oracle.cluster(oracle.G.from_csr(csr), oracle.config_from(cfg))
        cpu = {"value": round(r.seconds * 1e3, 3), "unit": "ms per Cluster call", "cores": 1,
               "kind": "port", "sample": "the same Cluster call", "ids_equal": bool(np.array_equal(r.ids, ids))}
    line = {
        "metric": "ms per Cluster call (ParallelCorrelationClusterer sync ALL_MOVE)",
        "value": round(float(np.median(walls)) * 1e3, 3), "unit": "ms", "n_gpus": 1,
        "steps": steps, "warmup": max(args.warmup, 1),
        "ms_per_step": round(float(np.median(walls)) * 1e3, 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (move deltas) / f32 (edge weights)",
        "data": "synthetic (seeded SBM)",
        "config": {"workload": "SBM n=10000, 100 blocks, p_in 0.1, p_out 0.001, unit weights; sync ALL_MOVE "
                               "resolution 0.01, 10 x 10 iterations, no refinement",
                   "n": csr.n, "m_directed": csr.m, "parallelism": "single GPU"},
        "min_ms": round(min(walls) * 1e3, 3), "device_span_ms": round(st.cluster_seconds * 1e3, 3),
        "rounds": st.rounds, "levels": st.levels, "num_clusters": st.num_clusters,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_cluster_moves(args):
    """Whole-cluster moves (DEFAULT_CLUSTER_MOVES) on the RMAT graph of the
    default workload: one level of up to 20 sync NBHR_MOVE rounds, each a
    vertex phase and a whole-cluster phase.  The CPU figure is the oracle's
    first round (both phases) from singletons on the same graph."""
    import numpy as np
    import pcc_amd

    csr = pcc_amd.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
    sub = args.subclustering_method or 0
    names = {0: "", 1: ", TRIANGLE subclustering", 2: ", CONNECTIVITY subclustering",
             3: ", SLOW_CONNECTIVITY subclustering"}
    cfg = pcc_amd.make_config(resolution=0.05, move_method=pcc_amd.NBHR_MOVE,
                              clustering_moves_method=pcc_amd.DEFAULT_CLUSTER_MOVES,
                              subclustering_method=sub)
    g = pcc_amd.DeviceGraph.from_csr(csr)
    g.set_option(pcc_amd.OPT_TIMING, 1)
    for _ in range(args.warmup):
        g.cluster(cfg)
    stats = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ids, st = g.cluster(cfg)
        stats.append(st)
        print(f"# step {i}: {st.cluster_seconds*1e3:.1f} ms rounds={st.rounds} "
              f"cluster moves={st.cluster_moves} clusters={st.num_clusters}", file=sys.stderr,
              flush=True)
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    st = stats[-1]
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        oracle.set_threads(threads)
        og = oracle.G.from_csr(csr)
        s = oracle.S.singletons(csr.n)
        t1 = time.perf_counter()
        oracle.round_(og, s, oracle.config_from(cfg), np.ones(csr.n, np.uint8))
        cpu_s = time.perf_counter() - t1
        oracle.set_threads(1)
        cpu = {"value": round(cpu_s * 1e3, 3), "unit": "ms per round", "cores": threads,
               "kind": "port",
               "sample": "round 0 (vertex and whole-cluster phases) from singletons, same graph; "
                         "the whole-cluster phase of the oracle is single-threaded"}
    line = {
        "metric": "whole-cluster-move Cluster time (DEFAULT_CLUSTER_MOVES, sync NBHR_MOVE"
                  f"{names.get(sub, sub)})",
        "value": round(st.cluster_seconds * 1e3, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (move deltas) / f32 (edge weights)",
        "data": "synthetic (seeded RMAT generated on the host, resident in HBM before timing)",
        "config": {"workload": f"RMAT scale {args.scale} edge factor {args.edge_factor:g}, "
                               "U[-1,1] weights, seed 2; sync NBHR_MOVE resolution 0.05, "
                               "DEFAULT_CLUSTER_MOVES (one level, up to 20 rounds)"
                               f"{names.get(sub, sub)}",
                   "n": csr.n, "m_directed": csr.m, "parallelism": "single GPU"},
        "rounds": st.rounds, "cluster_moves": st.cluster_moves,
        "cluster_phase_ms": round(st.cluster_move_seconds * 1e3, 3),
        "bestmove_ms": round(st.bestmove_seconds * 1e3, 3),
        "gpu_ms_per_round": round(st.cluster_seconds * 1e3 / max(st.rounds, 1), 3),
        "objective": st.objective, "num_clusters": st.num_clusters,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def measured_group(args, dom, name, key):
    """A per-group figure of the dominant kernel group recorded in profiles/
    (rocprofv3 counter runs of the default workload and tuning; null for
    other workloads)."""
    default = (args.scale == 22 and args.edge_factor == 16.0 and not args.wave_max
               and not args.block_max and not args.hub_part_edges)
    path = os.path.join(ROOT, "profiles", name)
    if not default or not os.path.exists(path):
        return None
    with open(path) as f:
        grp = json.load(f).get("groups", {}).get(str(dom))
    return grp.get(key) if grp else None


def measured_traffic(args, dom):
    """HBM bytes per launch of the dominant kernel group, from the calibrated
    FETCH_SIZE / WRITE_SIZE runs (profiles/traffic_rmat22.json)."""
    return measured_group(args, dom, "traffic_rmat22.json", "traffic_bytes_per_round")


def c4_graphs(pcc_amd, count, seed=4, select=None):
    """Batched workload: SBM graphs, n ~ logU[200, 2e5], mean degree ~ U[8, 64],
    power-law (exponent 2) community sizes, p_out / p_in = 0.01, weights
    U[0, 1] rounded to float32, resolution ~ U[0.001, 0.1].  The parameters
    of all `count` graphs are drawn in order from one seeded stream; the
    graphs listed in `select` (all by default) are generated, in that order,
    by a pool of host threads (the generator releases the interpreter lock)."""
    params, res = c4_params(count, seed)
    sel = list(range(count)) if select is None else list(select)
    return c4_generate(pcc_amd, [params[i] for i in sel]), res[sel]


def c4_params(count, seed=4):
    import numpy as np

    rng = np.random.default_rng(seed)
    params, res = [], []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(200), np.log(200000))))
        d = rng.uniform(8, 64)
        sizes = []
        left = n
        while left > 0:
            # discrete power law P(s) ~ s^-2 on [8, n], by inverse transform
            u = rng.random()
            s_ = int(8 / (1 - u * (1 - 8 / max(n, 9))))
            s_ = max(1, min(s_, left))
            sizes.append(s_)
            left -= s_
        sizes = np.array(sizes, np.int64)
        # p_in so that the expected degree is d, with p_out = 0.01 p_in.
        denom = float((sizes * (sizes - 1)).sum() + 0.01 * (sizes * (n - sizes)).sum()) / n
        p_in = min(1.0, d / max(denom, 1e-9))
        params.append((sizes, p_in, seed * 100003 + i))
        res.append(rng.uniform(0.001, 0.1))
    return params, np.array(res)


def c4_generate(pcc_amd, params):
    from concurrent.futures import ThreadPoolExecutor

    count = len(params)
    graphs = [None] * count
    step = max(1, count // 10)
    workers = min(16, os.cpu_count() or 1)

    def make(i):
        sizes, p_in, gseed = params[i]
        graphs[i] = pcc_amd.gen_sbm_blocks(sizes, p_in, 0.01 * p_in, 0.0, 1.0, seed=gseed)
        return i

    with ThreadPoolExecutor(workers) as ex:
        for done, _ in enumerate(ex.map(make, range(count)), 1):
            if count >= 512 and done % step == 0:
                print(f"# generated {done}/{count} graphs", file=sys.stderr, flush=True)
    return graphs


# The degree-bin thresholds and hub partition size pcc_autotune chose for the
# batched workload (on its first device batch): wave bound 512, block bound
# 4096, 1024-edge hub partitions.  They only pick kernel shapes.
C4_TUNED = {"wave_max_deg": 512, "block_max_deg": 4096, "hub_part_edges": 1024}


def batched_sample(args, steps=3):
    """The metric's second half on a size-stratified sample of the batched
    workload: the 4,096 graphs ranked by vertex count, every k-th of them
    (k = 4096 / sample), clustered in one batch (the saved autotune choice),
    against plain per-graph pcc_cluster runs of the same graphs; results must
    be identical."""
    import numpy as np
    import pcc_amd

    total = 4096
    k = max(1, total // max(1, args.batched_sample))
    params, res_all = c4_params(total)
    order = np.argsort([int(p[0].sum()) for p in params], kind="stable")
    sel = [int(i) for i in order[k // 2::k]]
    t0 = time.time()
    graphs = c4_generate(pcc_amd, [params[i] for i in sel])
    res = res_all[sel]
    nv, ne = sum(g.n for g in graphs), sum(g.m for g in graphs)
    print(f"# batched sample: {len(graphs)} of {total} graphs (every {k}th by size), {nv} vertices, "
          f"{ne} edges, generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    cfg = pcc_amd.make_config(move_method=pcc_amd.NBHR_MOVE, refine=1)
    b = pcc_amd.DeviceBatch.from_csrs(graphs)
    set_tuned(b, C4_TUNED, pcc_amd)
    b.cluster_batch(cfg, res)
    t0 = time.perf_counter()
    for _ in range(steps):
        ids_b, obj_b, _, st = b.cluster_batch(cfg, res)
    batch_s = (time.perf_counter() - t0) / steps
    b.close()
    print(f"# batched sample: {batch_s * 1e3:.1f} ms per batch call", file=sys.stderr, flush=True)
    plain_s, same = 0.0, 0
    for j, gr in enumerate(graphs):
        g = pcc_amd.DeviceGraph.from_csr(gr)
        set_tuned(g, C4_TUNED, pcc_amd)
        c1 = pcc_amd.make_config(resolution=float(res[j]), move_method=pcc_amd.NBHR_MOVE, refine=1)
        t1 = time.perf_counter()
        ids, sst = g.cluster(c1)
        plain_s += time.perf_counter() - t1
        same += int(np.array_equal(ids, ids_b[j]) and sst.objective == obj_b[j])
        g.close()
    print(f"# batched sample: plain runs {plain_s:.1f} s", file=sys.stderr, flush=True)
    bgps, pgps = len(graphs) / batch_s, len(graphs) / plain_s
    return {
        "metric": "batched graphs/s vs plain per-graph GPU runs (identical ids and objective)",
        "value": round(bgps, 3), "unit": "graphs/s",
        "plain_graphs_per_s": round(pgps, 3), "speedup_vs_plain": round(bgps / pgps, 2),
        "identical_results": f"{same}/{len(graphs)}",
        "batch_ms": round(batch_s * 1e3, 3), "plain_s": round(plain_s, 3),
        "sample": f"{len(graphs)} of the 4,096 seeded SBM graphs (n logU[200,2e5], mean degree "
                  f"U[8,64], power-law blocks, p_out/p_in 0.01, U[0,1] weights, res U[.001,.1]), "
                  f"every {k}th by vertex count: {nv} vertices, {ne} directed edges; sync NBHR_MOVE "
                  f"refine; plain runs time each pcc_cluster call on a resident graph",
        "rounds_per_batch_call": st.rounds, "edges_scanned_per_batch_call": st.edges_scanned,
        "tuning": C4_TUNED,
    }


def bench_batched(args):
    import numpy as np
    import pcc_amd

    t0 = time.time()
    graphs, res = c4_graphs(pcc_amd, args.graphs)
    print(f"# {len(graphs)} graphs, {sum(g.n for g in graphs)} vertices, "
          f"{sum(g.m for g in graphs)} edges, generated in {time.time() - t0:.1f}s",
          file=sys.stderr, flush=True)
    cfg = pcc_amd.make_config(move_method=pcc_amd.NBHR_MOVE, refine=1)
    # Consecutive graphs are grouped into batches of at most --batch-edges
    # directed edges, each uploaded once and kept resident.
    groups, cur, cur_m = [], [], 0
    for i, gr in enumerate(graphs):
        if cur and cur_m + gr.m > args.batch_edges:
            groups.append(cur)
            cur, cur_m = [], 0
        cur.append(i)
        cur_m += gr.m
    if cur:
        groups.append(cur)
    batches = [pcc_amd.DeviceBatch.from_csrs([graphs[i] for i in grp]) for grp in groups]
    print(f"# {len(batches)} batch(es) of up to {args.batch_edges} edges", file=sys.stderr, flush=True)
    ids_b = [None] * len(graphs)
    obj_b = [None] * len(graphs)
    rounds = 0
    te_b = 0
    tune = None
    if not args.no_autotune:
        # Degree-bin thresholds and hub partition size chosen on the first
        # batch by pcc_autotune (results do not depend on them), then used
        # for every batch and for the plain runs.
        t0 = time.perf_counter()
        tune = batches[0].autotune(cfg, resolutions=res[groups[0]], max_trials=8)
        for b in batches[1:]:
            set_tuned(b, tune, pcc_amd)
        tune["tuning_seconds"] = round(time.perf_counter() - t0, 3)
        print(f"# autotune: {tune}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        for grp, b in zip(groups, batches):
            b.cluster_batch(cfg, res[grp])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        rounds = te_b = 0
        for grp, b in zip(groups, batches):
            ids, obj, _, st = b.cluster_batch(cfg, res[grp])
            for j, i in enumerate(grp):
                ids_b[i] = ids[j]
                obj_b[i] = obj[j]
            rounds += st.rounds
            te_b += st.edges_scanned
    batch_s = (time.perf_counter() - t0) / max(args.steps, 1)
    print(f"# batched: {batch_s*1e3:.1f} ms for all graphs", file=sys.stderr, flush=True)
    del batches
    plain_n = args.plain_graphs if args.plain_graphs > 0 else len(graphs)
    plain_s = 0.0
    same = 0
    for i in range(plain_n):
        g = pcc_amd.DeviceGraph.from_csr(graphs[i])
        if tune:
            set_tuned(g, tune, pcc_amd)
        c1 = pcc_amd.make_config(resolution=float(res[i]), move_method=pcc_amd.NBHR_MOVE, refine=1)
        t1 = time.perf_counter()
        ids, sst = g.cluster(c1)
        plain_s += time.perf_counter() - t1
        same += int(np.array_equal(ids, ids_b[i]) and sst.objective == obj_b[i])
        g.close()
        if (i + 1) % 64 == 0 or i + 1 == plain_n:
            print(f"# plain runs: {i + 1}/{plain_n}, {plain_s:.1f} s", file=sys.stderr, flush=True)
    plain_gps = plain_n / plain_s
    batch_gps = len(graphs) / batch_s
    line = {
        "metric": "batched graphs/s vs plain per-graph GPU runs (identical ids and objective)",
        "value": round(batch_gps, 3), "unit": "graphs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(batch_s * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64 (move deltas) / f32 (edge weights)",
        "data": "synthetic (seeded SBM batch)",
        "config": {"workload": f"{len(graphs)} SBM graphs (n logU[200,2e5], mean degree U[8,64], "
                               "power-law blocks, p_out/p_in 0.01, U[0,1] weights, res U[.001,.1]), "
                               "sync NBHR_MOVE refine", "parallelism": "single GPU"},
        "plain_graphs_per_s": round(plain_gps, 3), "plain_graphs_timed": plain_n,
        "speedup_vs_plain": round(batch_gps / plain_gps, 2),
        "identical_results": f"{same}/{plain_n}",
        "batches": len(groups), "batch_rounds": rounds, "batch_edges_scanned": te_b,
        "autotune": tune,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_modularity_async(args):
    import pcc_amd

    t0 = time.time()
    csr = pcc_amd.gen_rmat(args.scale, args.edge_factor, 0.57, 0.19, 0.19, 1.0, 1.0, seed=3)
    print(f"# generated in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    cfg = pcc_amd.make_config(resolution=1.0, move_method=pcc_amd.NBHR_CLUSTER_MOVE, refine=1,
                              permute=1)
    cfg.async_ = 1
    t0 = time.time()
    g = pcc_amd.DeviceGraph.from_csr(csr)
    print(f"# graph n={csr.n} m={csr.m} upload={time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    apply_tuning(g, args, pcc_amd)
    for i in range(args.warmup):
        _, st = g.cluster(cfg, pcc_amd.OBJECTIVE_MODULARITY)
        print(f"# warmup {i}: {st.cluster_seconds*1e3:.1f} ms modularity={st.modularity:.5f}",
              file=sys.stderr, flush=True)
    # The timed steps run as a caller would (no per-launch events, so bins
    # 0-4 and small hashed sub-rounds take one launch each); one more step
    # with per-bin launches and HIP events gives the best-move time and the
    # per-bin breakdown.
    t0 = time.perf_counter()
    stats = []
    for i in range(args.steps):
        ids, st = g.cluster(cfg, pcc_amd.OBJECTIVE_MODULARITY)
        stats.append(st)
        print(f"# step {i}: {st.cluster_seconds*1e3:.1f} ms rounds={st.rounds} "
              f"modularity={st.modularity:.5f} clusters={st.num_clusters}", file=sys.stderr, flush=True)
    wall = (time.perf_counter() - t0) / max(args.steps, 1)
    g.set_option(pcc_amd.OPT_TIMING, 1)
    _, ist = g.cluster(cfg, pcc_amd.OBJECTIVE_MODULARITY)
    print(f"# instrumented step: {ist.cluster_seconds*1e3:.1f} ms rounds={ist.rounds} "
          f"best-move {ist.bestmove_seconds*1e3:.1f} ms modularity={ist.modularity:.5f}",
          file=sys.stderr, flush=True)
    te = sum(s.edges_scanned for s in stats)
    line = {
        "metric": "best-move GTEPS (ParallelModularityClusterer async NBHR_CLUSTER_MOVE refine permute)",
        "value": round(ist.edges_scanned / ist.bestmove_seconds / 1e9, 4) if ist.bestmove_seconds else 0.0,
        "value_from": "one instrumented step (per-bin launches timed with HIP events) after the timed steps",
        "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (move deltas)", "data": "synthetic (seeded RMAT, unit weights)",
        "config": {"workload": f"RMAT scale {args.scale} edge factor {args.edge_factor:g}, unit weights, "
                               "gamma 1.0", "n": csr.n, "m_directed": csr.m, "parallelism": "single GPU"},
        "modularity": stats[0].modularity, "num_clusters": stats[0].num_clusters,
        "end_to_end_ms": round(sum(s.cluster_seconds for s in stats) / len(stats) * 1e3, 3),
        "rounds_per_step": stats[0].rounds, "levels": stats[0].levels,
        "edges_scanned_per_step": te // max(len(stats), 1),
        "instrumented_step": {"cluster_ms": round(ist.cluster_seconds * 1e3, 3), "rounds": ist.rounds,
                              "bestmove_ms": round(ist.bestmove_seconds * 1e3, 3),
                              "contract_ms": round(ist.contract_seconds * 1e3, 3),
                              "apply_frontier_ms": round(ist.apply_seconds * 1e3, 3),
                              "modularity": ist.modularity},
        "bins": [{"bin": BIN_NAMES[b], "seconds": round(ist.bin_seconds[b], 6),
                  "edges": ist.bin_edges[b], "launches": ist.bin_launches[b]} for b in range(8)],
    }
    print(json.dumps(line), flush=True)
    return 0


def serving_requests(pcc_amd, count, seed=5):
    """The serving stream's requests: a seeded 50/50 mix of SBM and RMAT
    graphs, n ~ logU[1e3, 2e6], mean degree 16, resolution ~ U[0.001, 0.1]."""
    import numpy as np

    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e6))))
        if i % 2 == 0:
            g = pcc_amd.gen_sbm(n, max(1, n // 50), min(1.0, 16 / 50.0 * 0.9), 16 * 0.1 / n, 0.0, 1.0,
                                seed=5000 + i)
        else:
            scale = max(10, int(round(np.log2(n))))
            g = pcc_amd.gen_rmat(scale, 8.0, 0.57, 0.19, 0.19, 0.0, 1.0, seed=5000 + i)
        reqs.append((g, float(rng.uniform(0.001, 0.1))))
    return reqs


def serve_stream(pcc_amd, cfg, reqs, qps, seed):
    """Submits every request with Poisson gaps at `qps` (0: all at once) and
    waits for all; returns the server's statistics and the wall time."""
    import numpy as np

    rng = np.random.default_rng(seed)
    srv = pcc_amd.Server(cfg, workers=2, max_batch_graphs=64)
    gaps = rng.exponential(1.0 / qps, size=len(reqs)) if qps > 0 else np.zeros(len(reqs))
    t0 = time.perf_counter()
    ids = []
    for (g, r), gap in zip(reqs, gaps):
        if gap > 0:
            time.sleep(gap)
        ids.append(srv.submit(g, r))
    for rid in ids:
        srv.wait(rid)
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    return st, wall


def bench_serving(args):
    """Continuous batching under a Poisson stream.  The arrival rate is swept
    (--qps-sweep) until throughput stops growing; capacity is the largest
    graphs/s reached, and one more stream at 80% of it gives the latency the
    server holds below saturation.  The headline value is that capacity."""
    import pcc_amd

    reqs = serving_requests(pcc_amd, args.requests)
    cfg = pcc_amd.make_config(move_method=pcc_amd.NBHR_MOVE, refine=1)
    serve_stream(pcc_amd, cfg, reqs[:8], 0, 1)  # warm-up
    levels = [float(x) for x in args.qps_sweep.split(",")] if args.qps_sweep else [args.qps]
    sweep = []
    for qps in levels:
        st, wall = serve_stream(pcc_amd, cfg, reqs, qps, 11)
        sweep.append({"offered_qps": qps if qps > 0 else "all at once",
                      "graphs_per_s": round(st.graphs_per_second, 3),
                      "p50_latency_s": round(st.p50_latency_seconds, 4),
                      "p99_latency_s": round(st.p99_latency_seconds, 4),
                      "batches": st.batches, "mean_batch_graphs": round(st.mean_batch_graphs, 2),
                      "failed": st.failed})
        print(f"# serving {sweep[-1]}", file=sys.stderr, flush=True)
    capacity = max(x["graphs_per_s"] for x in sweep)
    at80 = None
    if len(levels) > 1:
        st, wall = serve_stream(pcc_amd, cfg, reqs, 0.8 * capacity, 12)
        at80 = {"offered_qps": round(0.8 * capacity, 3), "graphs_per_s": round(st.graphs_per_second, 3),
                "p50_latency_s": round(st.p50_latency_seconds, 4),
                "p99_latency_s": round(st.p99_latency_seconds, 4), "failed": st.failed}
        print(f"# serving at 80% of capacity: {at80}", file=sys.stderr, flush=True)
    nv = sum(g.n for g, _ in reqs)
    line = {
        "metric": "serving capacity graphs/s (continuous batching, Poisson arrivals swept to saturation)",
        "value": capacity, "unit": "graphs/s", "n_gpus": 1,
        "steps": len(reqs) * len(levels), "warmup": 8, "ms_per_step": None,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 (move deltas) / f32 (edge weights)", "data": "synthetic (seeded SBM/RMAT stream)",
        "config": {"workload": f"{len(reqs)} requests per stream ({nv} vertices in all), 50/50 SBM/RMAT, "
                               "n logU[1e3, 2e6], mean degree 16, res U[.001,.1], sync NBHR_MOVE refine",
                   "parallelism": "single GPU, 2 workers, batches of up to 64 graphs"},
        "sweep": sweep,
        "at_80_percent_capacity": at80,
    }
    print(json.dumps(line), flush=True)
    return 0


def bench_copy(args):
    """Plain device-to-device copy of a known size: the calibration of the
    FETCH_SIZE / WRITE_SIZE counters used for roofline.traffic."""
    import torch

    n = 1 << 28  # 1 GiB of float32
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(args.steps, 1)
    line = {"metric": "device copy bandwidth (counter calibration)", "value": round(8 * n / dt / 1e9, 2),
            "unit": "GB/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
            "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": f"copy of {4 * n} bytes (read {4 * n} + write {4 * n} per step)"},
            "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n}
    print(json.dumps(line), flush=True)
    return 0


def gather_ceiling(n_ids, gathers=45_000_000, reps=10):
    """Random 4-byte gathers per second from an array of n_ids entries
    (torch.index_select with uniform random indices): the access pattern of
    the best-move kernels' ids[u] lookups, measured on this device."""
    import torch

    g = torch.Generator(device="cuda")
    g.manual_seed(7)
    src = torch.randint(0, 1 << 30, (n_ids,), device="cuda", dtype=torch.int32, generator=g)
    idx = torch.randint(0, n_ids, (gathers,), device="cuda", dtype=torch.int64, generator=g)
    out = torch.empty(gathers, device="cuda", dtype=torch.int32)
    torch.index_select(src, 0, idx, out=out)
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(reps):
        torch.index_select(src, 0, idx, out=out)
    t1.record()
    torch.cuda.synchronize()
    return gathers * reps / (t0.elapsed_time(t1) * 1e-3)


def gather_ceiling_child(scale):
    """gather_ceiling in a child process (bench.py --workload gather)."""
    import subprocess

    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--workload", "gather",
                        "--scale", str(scale)], capture_output=True, text=True, timeout=300)
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    if r.returncode != 0 or not lines:
        print(f"# gather ceiling unavailable (rc={r.returncode})", file=sys.stderr, flush=True)
        return None
    return json.loads(lines[-1])["value"] * 1e9


def bench_gather(args):
    """Random-gather ceiling for the level-0 id array of the headline workload."""
    n_ids = 1 << args.scale
    rate = gather_ceiling(n_ids)
    line = {"metric": "random 4-byte gathers (ceiling of the ids[u] lookups)", "value": round(rate / 1e9, 2),
            "unit": "G gathers/s", "n_gpus": 1, "steps": 10, "warmup": 1, "ms_per_step": None,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic", "config": {"workload": f"45M uniform random gathers from {n_ids} u32 entries"}}
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline(csr, cfg, args):
    """The oracle (the CPU restatement; the reference cannot be built here)
    on the host cores: its whole Cluster span (all levels and the refinement,
    the reference's "Cluster Time:" span) on a bounded sample, the same RMAT
    generator and configuration at scale --cpu-scale, beside the GPU's span on
    that sample; plus its best-move round 0 on the full graph and one round of
    the sequential clusterer (SEQ-CC)."""
    import numpy as np
    import pcc_amd

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
    ocfg = oracle.config_from(cfg)
    small = pcc_amd.gen_rmat(args.cpu_scale, args.edge_factor, 0.57, 0.19, 0.19, -1.0, 1.0, seed=2)
    gs = pcc_amd.DeviceGraph.from_csr(small)
    gs.cluster(cfg)
    gpu_ids, gst = gs.cluster(cfg)
    gs.close()
    oracle.set_threads(t)
    oracle.phase_seconds(reset=True)
    r = oracle.cluster(oracle.G.from_csr(small), ocfg)
    phases = {k: round(v, 3) for k, v in oracle.phase_seconds(reset=True).items()}
    same = bool(np.array_equal(r.ids, gpu_ids))
    print(f"# cpu baseline: scale {args.cpu_scale} Cluster span {r.seconds:.1f} s on {t} threads "
          f"(GPU {gst.cluster_seconds * 1e3:.1f} ms), ids equal {same}", file=sys.stderr, flush=True)
    # Round 0 of level 0 on the full graph (all vertices, singleton state).
    g = oracle.G.from_csr(csr)
    s = oracle.S.singletons(csr.n)
    reps = 0
    t0 = time.perf_counter()
    while True:
        oracle.sync_decisions(g, s, ocfg)
        reps += 1
        el = time.perf_counter() - t0
        if el > 4.0 or reps >= 10:
            break
    round0 = reps * csr.m / el / 1e9
    # The paper's sequential baseline (SEQ-CC): one sequential round over all
    # vertices from singletons, single-threaded.
    oracle.set_threads(1)
    scfg = oracle.config_from(cfg)
    scfg.sequential = 1
    s = oracle.S.singletons(csr.n)
    t1 = time.perf_counter()
    oracle.round_(g, s, scfg, np.ones(csr.n, np.uint8))
    seq_s = time.perf_counter() - t1
    return {
        "value": round(r.edges_scanned / r.seconds / 1e9, 4),
        "unit": "GTEPS",
        "cores": t,
        "kind": "port",
        "sample": f"whole Cluster span (all {r.levels} levels, {r.rounds} rounds, refinement) on "
                  f"RMAT scale {args.cpu_scale} edge factor {args.edge_factor:g} with the same "
                  f"generator and configuration (n={small.n}, m={small.m}): {r.edges_scanned} edges "
                  f"scanned in {r.seconds:.2f} s",
        "end_to_end_s": round(r.seconds, 3),
        "phase_seconds": phases,
        "gpu_same_sample": {
            "end_to_end_ms": round(gst.cluster_seconds * 1e3, 3),
            "end_to_end_gteps": round(gst.edges_scanned / gst.cluster_seconds / 1e9, 4),
            "ids_equal_to_cpu": same,
            "speedup_end_to_end": round(r.seconds / gst.cluster_seconds, 1),
        },
        "round0": {
            "value": round(round0, 4), "unit": "GTEPS", "cores": t, "kind": "port",
            "sample": f"{reps} x best-move round 0 of level 0 (all {csr.n} vertices, {csr.m} edges) "
                      f"of the headline graph, {el:.1f} s",
        },
        "sequential": {
            "value": round(csr.m / seq_s / 1e9, 4), "unit": "GTEPS", "cores": 1, "kind": "port",
            "sample": f"round 0 of level 0 of the sequential clusterer (SEQ-CC, vertices moving "
                      f"one after another) on the headline graph, {seq_s:.1f} s",
        },
    }


if __name__ == "__main__":
    sys.exit(main())
