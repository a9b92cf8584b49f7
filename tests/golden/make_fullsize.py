"""Writes fingerprints of the oracle's results on the full-size workloads.

The workloads are too large to store as vectors (the RMAT scale-22 graph has
128M directed edges), so each fixture holds sha256 digests of the seeded
generator's CSR and of the oracle's final cluster ids (little-endian uint32,
the reference's cluster_ids vector before OutputIndicesById), with the
cluster count, the objective (as a hex float) and the round counts.  The GPU
tests regenerate the graph, check its digest, run the HIP path and compare the
digest of its ids.

Run: python tests/golden/make_fullsize.py [case ...]   (all cases by default;
the RMAT scale-22 case takes a few minutes on 8 host threads)
"""

import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "parallel-correlation-clustering_amd"), os.path.join(ROOT, "oracle")]

import oracle  # noqa: E402
import pcc_amd  # noqa: E402

CASES = {
    # Planted-partition SBM, 10,000 vertices in 100 blocks, p_in 0.1, p_out
    # 0.001, unit weights; sync ALL_MOVE, resolution 0.01, no refinement,
    # 10 x 10 iterations.
    "sbm10000_allmove": (dict(kind="sbm", n=10000, blocks=100, p_in=0.1, p_out=0.001, seed=1),
                         "correlation_clusterer_config { resolution: 0.01 move_method: ALL_MOVE "
                         "louvain_config { num_iterations: 10 num_inner_iterations: 10 } }", 0),
    # RMAT scale 22, edge factor 16 (a .57 b .19 c .19), U[-1, 1] weights;
    # sync NBHR_MOVE, resolution 0.05, refinement.
    "rmat22_nbhr_refine": (dict(kind="rmat", scale=22, ef=16, lo=-1.0, hi=1.0, seed=2),
                           "correlation_clusterer_config { resolution: 0.05 move_method: NBHR_MOVE "
                           "refine: true }", 0),
}


def make_graph(spec):
    if spec["kind"] == "sbm":
        return pcc_amd.gen_sbm(spec["n"], spec["blocks"], spec["p_in"], spec["p_out"], seed=spec["seed"])
    return pcc_amd.gen_rmat(spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["lo"], spec["hi"],
                            seed=spec["seed"])


def csr_digest(csr) -> str:
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(csr.offsets, np.int64), np.ascontiguousarray(csr.nbrs, np.uint32),
              np.ascontiguousarray(csr.weights, np.float32).view(np.uint32)):
        h.update(a.astype(a.dtype.newbyteorder("<"), copy=False).tobytes())
    return h.hexdigest()


def ids_digest(ids) -> str:
    return hashlib.sha256(np.ascontiguousarray(ids, "<u4").tobytes()).hexdigest()


def main(names):
    oracle.set_threads(min(16, os.cpu_count() or 1))
    for name, (spec, text, modularity) in CASES.items():
        if names and name not in names:
            continue
        csr = make_graph(spec)
        cfg = pcc_amd.parse_config(text)
        t0 = time.time()
        r = oracle.cluster(oracle.G.from_csr(csr), oracle.config_from(cfg), bool(modularity))
        out = {
            "spec": spec, "config": text, "modularity_objective": modularity,
            "n": int(csr.n), "m": int(csr.m), "csr_sha256": csr_digest(csr),
            "ids_sha256": ids_digest(r.ids), "num_clusters": int(r.num_clusters),
            "objective_hex": float(r.objective).hex(), "objective": float(r.objective),
            "modularity": float(r.modularity), "rounds": int(r.rounds), "levels": int(r.levels),
            "edges_scanned": int(r.edges_scanned),
        }
        with open(os.path.join(HERE, f"fullsize_{name}.json"), "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")
        print(f"{name}: n={csr.n} m={csr.m} clusters={r.num_clusters} objective={r.objective!r} "
              f"oracle {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
