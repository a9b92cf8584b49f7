"""Writes the golden vectors in this directory from the CPU oracle.

The reference ships no fixtures (its tests were not published), so these
vectors pin the HIP path and future versions of the oracle to the oracle as
validated by tests/test_oracle.py.  Each .npz holds the input CSR (as the
seeded generators produce it), the config, and the oracle's final cluster ids,
objective and modularity.

Run: python tests/golden/make_golden.py [case ...]   (all cases by default)
"""

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "parallel-correlation-clustering_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
import pcc_amd  # noqa: E402

CASES = {
    # SBM planted partition, sync ALL_MOVE, res 0.01, no refine, 10 x 10 (the
    # small-graph configuration, at 1/5 of its size).
    "sbm2000_allmove": (dict(kind="sbm", n=2000, blocks=20, p_in=0.1, p_out=0.001, seed=1),
                        "correlation_clusterer_config { resolution: 0.01 move_method: ALL_MOVE "
                        "louvain_config { num_iterations: 10 num_inner_iterations: 10 } }", 0),
    # RMAT, signed FP weights, sync NBHR_MOVE, res 0.05, refine.
    "rmat10_nbhr_refine": (dict(kind="rmat", scale=10, ef=16, lo=-1.0, hi=1.0, seed=2),
                           "correlation_clusterer_config { resolution: 0.05 move_method: NBHR_MOVE "
                           "refine: true }", 0),
    # RMAT, unit weights, modularity, sync NBHR_CLUSTER_MOVE, refine.
    "rmat11_modularity": (dict(kind="rmat", scale=11, ef=8, lo=1.0, hi=1.0, seed=3),
                          "correlation_clusterer_config { resolution: 1.0 refine: true }", 1),
    # RMAT, signed FP weights, whole-cluster moves (DEFAULT_CLUSTER_MOVES),
    # sync NBHR_MOVE, res 0.05, 12 rounds.
    "rmat10_cluster_moves": (dict(kind="rmat", scale=10, ef=16, lo=-1.0, hi=1.0, seed=4),
                             "correlation_clusterer_config { resolution: 0.05 move_method: NBHR_MOVE "
                             "clustering_moves_method: DEFAULT_CLUSTER_MOVES num_iterations: 12 }", 0),
    # RMAT, signed FP weights, TRIANGLE subclusters as the next levels'
    # vertices, sync NBHR_MOVE, res 0.05, refine.
    "rmat10_triangle_subclusters": (dict(kind="rmat", scale=10, ef=16, lo=-1.0, hi=1.0, seed=5),
                                    "correlation_clusterer_config { resolution: 0.05 "
                                    "move_method: NBHR_MOVE refine: true "
                                    "subclustering_method: TRIANGLE_SUBCLUSTERING }", 0),
    # SBM, CONNECTIVITY subclusters between levels, modularity, refine.
    "sbm1500_connectivity_subclusters": (dict(kind="sbm", n=1500, blocks=15, p_in=0.1, p_out=0.005,
                                              seed=7),
                                         "correlation_clusterer_config { resolution: 1.0 "
                                         "move_method: NBHR_MOVE refine: true "
                                         "subclustering_method: CONNECTIVITY_SUBCLUSTERING }", 1),
    # RMAT, signed FP weights, TRIANGLE subcluster moves inside
    # DEFAULT_CLUSTER_MOVES rounds, NBHR_MOVE, 10 rounds.
    "rmat10_triangle_subcluster_moves": (dict(kind="rmat", scale=10, ef=16, lo=-1.0, hi=1.0, seed=8),
                                         "correlation_clusterer_config { resolution: 0.05 "
                                         "move_method: NBHR_MOVE num_iterations: 10 "
                                         "clustering_moves_method: DEFAULT_CLUSTER_MOVES "
                                         "subclustering_method: TRIANGLE_SUBCLUSTERING }", 0),
    # SBM, SLOW_CONNECTIVITY subcluster moves inside DEFAULT_CLUSTER_MOVES
    # rounds, sync, res 0.02.
    "sbm1000_slow_subcluster_moves": (dict(kind="sbm", n=1000, blocks=10, p_in=0.2, p_out=0.01,
                                           seed=6),
                                      "correlation_clusterer_config { resolution: 0.02 "
                                      "clustering_moves_method: DEFAULT_CLUSTER_MOVES "
                                      "subclustering_method: SLOW_CONNECTIVITY_SUBCLUSTERING }", 0),
}


def make_graph(spec):
    if spec["kind"] == "sbm":
        return pcc_amd.gen_sbm(spec["n"], spec["blocks"], spec["p_in"], spec["p_out"], seed=spec["seed"])
    return pcc_amd.gen_rmat(spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["lo"], spec["hi"],
                            seed=spec["seed"])


def main(names):
    for name, (spec, text, modularity) in CASES.items():
        if names and name not in names:
            continue
        csr = make_graph(spec)
        cfg = pcc_amd.parse_config(text)
        r = oracle.cluster(oracle.G.from_csr(csr), oracle.config_from(cfg), bool(modularity))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), offsets=csr.offsets, nbrs=csr.nbrs,
                            weights=csr.weights, ids=r.ids, objective=r.objective,
                            modularity=r.modularity, spec=json.dumps(spec), config=text,
                            modularity_objective=modularity)
        print(name, csr.n, csr.m, r.num_clusters, r.objective)


if __name__ == "__main__":
    main(sys.argv[1:])
