"""The reference-shaped API on the GPU: Import/FinishImport, Cluster,
RefineClusters, the CLI clone, and the error statuses the reference returns."""

import os
import subprocess
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "parallel-correlation-clustering_amd", "cluster-in-memory")


def test_import_path_matches_csr_path(pcc, orc):
    csr = pcc.gen_sbm(500, 5, 0.1, 0.01, 0.0, 1.0, seed=21)
    print("import path", flush=True)
    cl = pcc.ParallelCorrelationClusterer()
    g = cl.MutableGraph()
    # Concurrent Import calls, as the interface allows (in-memory-clusterer.h:60-65).
    def worker(vs):
        for v in vs:
            s, e = csr.offsets[v], csr.offsets[v + 1]
            g.Import(pcc.AdjacencyList(int(v), 1.0, list(zip(csr.nbrs[s:e].tolist(),
                                                             csr.weights[s:e].tolist()))))
    ts = [threading.Thread(target=worker, args=(range(i, csr.n, 4),)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    g.FinishImport()
    back = g.device.to_csr()
    assert np.array_equal(back.offsets, csr.offsets)
    assert np.array_equal(back.nbrs, csr.nbrs)
    assert np.array_equal(back.weights, csr.weights)
    text = "correlation_clusterer_config { resolution: 0.05 refine: true move_method: NBHR_MOVE }"
    clustering = cl.Cluster(text)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(pcc.parse_config(text)))
    assert clustering == pcc.output_indices_by_id(r.ids)


def test_refine_clusters_from_initial_clustering(pcc, orc):
    csr = pcc.gen_sbm(600, 6, 0.15, 0.01, seed=22)
    print("refine", flush=True)
    rng = np.random.default_rng(1)
    init_ids = rng.integers(0, 40, size=csr.n)
    # Number clusters 0..k-1 by first appearance in a Clustering vector.
    clustering = pcc.output_indices_by_id(init_ids)
    cl = pcc.ParallelCorrelationClusterer()
    cl.MutableGraph().set_csr(csr)
    cfg = pcc.make_config(resolution=0.02, move_method=0)
    ids0 = pcc.clustering_to_ids(clustering, csr.n)
    cl.RefineClusters(cfg, clustering)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), initial_ids=ids0)
    assert clustering == pcc.output_indices_by_id(r.ids)
    # node weights overload
    nw = rng.integers(1, 4, size=csr.n).astype(np.float64)
    c2 = pcc.output_indices_by_id(init_ids)
    cl.RefineClusters(cfg, c2, node_weights=nw)
    r2 = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), node_weights=nw,
                     initial_ids=pcc.clustering_to_ids(pcc.output_indices_by_id(init_ids), csr.n))
    assert c2 == pcc.output_indices_by_id(r2.ids)


def test_modularity_clusterer(pcc, orc):
    csr = pcc.gen_rmat(12, 8, seed=23)
    print("modularity", flush=True)
    cl = pcc.ParallelModularityClusterer()
    cl.MutableGraph().set_csr(csr)
    text = "correlation_clusterer_config { resolution: 1.0 move_method: NBHR_CLUSTER_MOVE }"
    clustering = cl.Cluster(text)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(pcc.parse_config(text)),
                    modularity_objective=True)
    assert clustering == pcc.output_indices_by_id(r.ids)
    assert cl.last_stats.modularity == pytest.approx(r.modularity, abs=1e-12)


def test_error_statuses(pcc):
    csr = pcc.gen_sbm(50, 2, 0.3, 0.0, seed=24)
    g = pcc.DeviceGraph.from_csr(csr)
    with pytest.raises(pcc.StatusError) as e:
        g.cluster(pcc.make_config(clustering_moves_method=5))  # PCC.cc:501-505
    assert e.value.code == pcc.UNIMPLEMENTED
    with pytest.raises(pcc.StatusError) as e:
        g.cluster(pcc.make_config(subclustering_method=2, clustering_moves_method=0, async_=1))
    assert e.value.code == pcc.UNIMPLEMENTED
    with pytest.raises(pcc.StatusError) as e:
        g.cluster(pcc.make_config(subclustering_method=7))
    assert e.value.code == pcc.INVALID_ARGUMENT
    bad = pcc.Csr(csr.n, csr.offsets, csr.nbrs.copy(), csr.weights)
    bad.nbrs[3] = csr.n + 7
    with pytest.raises(pcc.StatusError) as e:
        pcc.DeviceGraph.from_csr(bad)
    assert e.value.code == pcc.INVALID_ARGUMENT
    with pytest.raises(pcc.StatusError) as e:
        g.refine(pcc.make_config(), np.full(csr.n, csr.n + 1, np.uint32))
    assert e.value.code == pcc.INVALID_ARGUMENT


def test_edge_cases(pcc, orc):
    print("edge cases", flush=True)
    # No edges at all, a single vertex, isolated vertices mixed in.
    for csr in [pcc.Csr(5, np.zeros(6, np.int64), np.zeros(0, np.uint32), np.zeros(0, np.float32)),
                pcc.Csr(1, np.zeros(2, np.int64), np.zeros(0, np.uint32), np.zeros(0, np.float32)),
                pcc.gen_sbm_blocks([1, 1, 30, 1, 20], 0.5, 0.0, seed=25)]:
        for kw in [dict(move_method=1), dict(move_method=2, refine=1), dict(move_method=0)]:
            cfg = pcc.make_config(resolution=0.01, **kw)
            ids, st = pcc.DeviceGraph.from_csr(csr).cluster(cfg)
            r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
            assert np.array_equal(ids, r.ids)


def test_cli_clone(pcc, orc, tmp_path):
    assert os.path.exists(CLI), "build the CLI first"
    csr = pcc.gen_sbm(300, 3, 0.2, 0.01, 0.0, 1.0, seed=26)
    path = str(tmp_path / "g.adj")
    pcc.write_adjacency_graph(path, csr, float_weighted=True)
    out = str(tmp_path / "clusters.tsv")
    comm = tmp_path / "comm.tsv"
    comm.write_text("\n".join("\t".join(str(v) for v in range(b * 100, b * 100 + 100))
                              for b in range(3)) + "\n")
    text = "correlation_clusterer_config { resolution: 0.01 move_method: ALL_MOVE }"
    p = subprocess.run([CLI, "--clusterer_name=ParallelCorrelationClusterer",
                        f"--clusterer_config={text}", f"--input_graph={path}", "--float_weighted",
                        f"--output_clustering={out}", f"--input_communities={comm}"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "Objective:" in p.stdout and "Number of Clusters:" in p.stdout
    assert "Avg precision:" in p.stdout
    lines = open(out).read().splitlines()
    got = [[int(x) for x in l.split("\t") if x] for l in lines]
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(pcc.parse_config(text)))
    assert got == pcc.output_indices_by_id(r.ids)
    bad = subprocess.run([CLI, "--clusterer_name=Nope", "--clusterer_config=", f"--input_graph={path}"],
                         capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "UNIMPLEMENTED" in bad.stderr
    # Subclustering through the text config (enum names as the proto spells them).
    for method in ("TRIANGLE_SUBCLUSTERING", "CONNECTIVITY_SUBCLUSTERING",
                   "SLOW_CONNECTIVITY_SUBCLUSTERING"):
        text = ("correlation_clusterer_config { resolution: 0.01 move_method: NBHR_MOVE refine: true "
                f"subclustering_method: {method} }}")
        p = subprocess.run([CLI, "--clusterer_name=ParallelCorrelationClusterer",
                            f"--clusterer_config={text}", f"--input_graph={path}", "--float_weighted",
                            f"--output_clustering={out}"], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stderr
        got = [[int(x) for x in l.split("\t") if x] for l in open(out).read().splitlines()]
        r = orc.cluster(orc.G.from_csr(csr), orc.config_from(pcc.parse_config(text)))
        assert got == pcc.output_indices_by_id(r.ids), method


def test_autotune_keeps_results(pcc):
    csr = pcc.gen_rmat(14, 16, w_lo=-1.0, w_hi=1.0, seed=77)
    cfg = pcc.make_config(resolution=0.05, move_method=2, refine=1)
    g = pcc.DeviceGraph.from_csr(csr)
    ref, _ = g.cluster(cfg)
    t = g.autotune(cfg, max_trials=6)
    assert 2 <= t["trials"] <= 6
    assert t["wave_max_deg"] in (128, 256, 512)
    assert t["block_max_deg"] in (1024, 2048, 4096)
    assert t["hub_part_edges"] in (1024, 2048, 4096)
    assert 0 < t["best_seconds"] <= t["initial_seconds"]
    ids, _ = g.cluster(cfg)
    assert np.array_equal(ids, ref)


def test_autotune_batch(pcc):
    graphs = [pcc.gen_rmat(s, 8, seed=90 + s) for s in (8, 11, 9, 12)]
    res = np.array([0.01, 0.05, 0.02, 0.08])
    cfg = pcc.make_config(move_method=0)
    b = pcc.DeviceBatch.from_csrs(graphs)
    ref, _, _, _ = b.cluster_batch(cfg, res)
    t = b.autotune(cfg, resolutions=res)
    assert t["trials"] >= 2
    ids, _, _, _ = b.cluster_batch(cfg, res)
    for a, r in zip(ids, ref):
        assert np.array_equal(a, r)
