"""Subclustering against the oracle, bit-exact:
  * between LOUVAIN levels (CompressSubclusters,
    clustering/clusterers/parallel-correlation-clusterer.cc:436-474, 564-567):
    the subcluster numbering (union-find components, triangle scores), the
    compression by subcluster and the next level started from the parent
    clusters, with and without refinement;
  * inside sync DEFAULT_CLUSTER_MOVES rounds (PCC.cc:175-297): subcluster set
    moves, then set moves of the clusters as they were before them."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GRAPHS = {
    "sbm_dense": lambda pcc: pcc.gen_sbm(600, 12, 0.7, 0.01, seed=31),
    "sbm_sparse": lambda pcc: pcc.gen_sbm(2000, 20, 0.08, 0.002, seed=32),
    "rmat_unit": lambda pcc: pcc.gen_rmat(12, 8, seed=33),
    "rmat_signed": lambda pcc: pcc.gen_rmat(12, 8, w_lo=-1.0, w_hi=1.0, seed=34),
    "rmat_hubs": lambda pcc: pcc.gen_rmat(15, 16, seed=35),
}

CASES = []
for name in GRAPHS:
    for method in (1, 2, 3):
        if (name, method) != ("rmat_hubs", 1):  # covered below at resolution 1
            CASES.append((name, method, dict(resolution=0.05), False))
CASES += [
    ("sbm_dense", 1, dict(resolution=0.05, refine=1), False),
    ("sbm_sparse", 2, dict(resolution=0.01, refine=1), False),
    ("rmat_hubs", 2, dict(resolution=1.0, refine=1), True),
    ("rmat_hubs", 1, dict(resolution=1.0), True),
    ("rmat_signed", 2, dict(resolution=0.1, edge_weight_offset=0.25, move_method=2), False),
    ("sbm_sparse", 1, dict(resolution=0.01, move_method=0, louvain_num_iterations=3), False),
]
# DEFAULT_CLUSTER_MOVES rounds (one level).
for name, method, kw, modularity in [
        ("sbm_dense", 1, dict(resolution=0.05), False),
        ("sbm_dense", 2, dict(resolution=0.05, move_method=1), False),
        ("sbm_sparse", 2, dict(resolution=0.01), False),
        ("sbm_sparse", 1, dict(resolution=0.01, move_method=2), False),
        ("rmat_unit", 2, dict(resolution=0.05), False),
        ("rmat_signed", 1, dict(resolution=0.1, edge_weight_offset=0.25), False),
        ("rmat_signed", 2, dict(resolution=0.1, move_method=1, num_iterations=5), False),
        ("rmat_hubs", 2, dict(resolution=1.0), True),
        ("rmat_hubs", 1, dict(resolution=1.0, move_method=2), True),
        ("sbm_dense", 3, dict(resolution=0.05), False),
        ("rmat_signed", 3, dict(resolution=0.1, edge_weight_offset=0.25), False),
        ("rmat_hubs", 3, dict(resolution=1.0), True)]:
    CASES.append((name, method, dict(clustering_moves_method=0, **kw), modularity))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_subclustering_bit_exact(pcc, orc, case):
    name, method, kw, modularity = CASES[case]
    csr = GRAPHS[name](pcc)
    print(f"subclustering {name} method={method} {kw} modularity={modularity}", flush=True)
    cfg = pcc.make_config(subclustering_method=method, **kw)
    dg = pcc.DeviceGraph.from_csr(csr)
    ids, st = dg.cluster(cfg, pcc.OBJECTIVE_MODULARITY if modularity else pcc.OBJECTIVE_CORRELATION)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), modularity)
    assert np.array_equal(ids, r.ids)
    assert st.levels == r.levels
    assert st.num_clusters == r.num_clusters
    assert st.objective == pytest.approx(r.objective, rel=1e-9, abs=1e-6)


def test_subclustering_changes_the_levels(pcc, orc):
    """The subcluster path is not the plain one: on dense blocks merged at
    level 0, TRIANGLE splits the clusters and the runs differ."""
    csr = GRAPHS["sbm_dense"](pcc)
    dg = pcc.DeviceGraph.from_csr(csr)
    plain, _ = dg.cluster(pcc.make_config(resolution=0.05))
    tri, _ = dg.cluster(pcc.make_config(resolution=0.05, subclustering_method=1))
    assert not np.array_equal(plain, tri)


def test_triangle_needs_ascending_rows(pcc):
    csr = GRAPHS["sbm_dense"](pcc)
    nbrs = csr.nbrs.copy()
    a, b = int(csr.offsets[5]), int(csr.offsets[6])
    nbrs[a:b] = nbrs[a:b][::-1].copy()  # one row descending
    bad = pcc.Csr(csr.n, csr.offsets, nbrs, csr.weights)
    dg = pcc.DeviceGraph.from_csr(bad)
    with pytest.raises(pcc.StatusError) as e:
        dg.cluster(pcc.make_config(resolution=0.05, subclustering_method=1))
    assert e.value.code == pcc.INVALID_ARGUMENT
    ids, _ = dg.cluster(pcc.make_config(resolution=0.05, subclustering_method=2))
    assert ids.shape[0] == csr.n


def test_async_subcluster_moves_unimplemented(pcc):
    dg = pcc.DeviceGraph.from_csr(GRAPHS["sbm_dense"](pcc))
    with pytest.raises(pcc.StatusError) as e:
        dg.cluster(pcc.make_config(resolution=0.05, clustering_moves_method=0, async_=1,
                                   subclustering_method=2))
    assert e.value.code == pcc.UNIMPLEMENTED


def test_subclustering_batches_run_graph_by_graph(pcc):
    graphs = [pcc.gen_sbm(200 + 100 * s, 4, 0.3, 0.02, seed=s) for s in (1, 2, 3)]
    graphs.append(pcc.gen_rmat(9, 8, w_lo=-1.0, w_hi=1.0, seed=4))
    res = np.array([0.05, 0.02, 0.1, 0.05])
    for method in (1, 2, 3):
        cfg = pcc.make_config(subclustering_method=method, move_method=2, refine=1)
        ids, obj, _, st = pcc.DeviceBatch.from_csrs(graphs).cluster_batch(cfg, res)
        for k, csr in enumerate(graphs):
            single, sst = pcc.DeviceGraph.from_csr(csr).cluster(
                pcc.make_config(subclustering_method=method, move_method=2, refine=1,
                                resolution=float(res[k])))
            assert np.array_equal(ids[k], single), (method, k)
            assert obj[k] == pytest.approx(sst.objective, rel=1e-12, abs=1e-9)


def _csr_from_edges(pcc, n, edges, w=1.0):
    adj = [[] for _ in range(n)]
    for u, v in edges:
        adj[u].append(v)
        if u != v:
            adj[v].append(u)
    off = np.zeros(n + 1, np.int64)
    nbrs = []
    for u in range(n):
        row = sorted(adj[u])
        off[u + 1] = off[u] + len(row)
        nbrs += row
    nb = np.array(nbrs, np.uint32)
    return pcc.Csr(n, off, nb, np.full(nb.shape[0], w, np.float32))


EDGE_CASES = {
    "one_vertex": (1, []),
    "isolated": (6, [(1, 2)]),
    "self_loops_only": (4, [(0, 0), (1, 1), (3, 3)]),
    "two_triangles_bridge": (7, [(0, 1), (1, 2), (0, 2), (3, 4), (4, 5), (3, 5), (2, 3), (6, 6)]),
    "star": (9, [(0, k) for k in range(1, 9)]),
    "clique8": (8, [(a, b) for a in range(8) for b in range(a + 1, 8)]),
}


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("name", sorted(EDGE_CASES))
@pytest.mark.parametrize("moves", [0, 1])
def test_subclustering_edge_cases(pcc, orc, name, method, moves):
    n, edges = EDGE_CASES[name]
    print(f"subclustering edge case {name} method={method} moves={moves}", flush=True)
    csr = _csr_from_edges(pcc, n, edges)
    # Without refinement at 0.1: with it, a run whose last moving level is
    # level 0 returns singletons (PCC.cc:526-530, 589), as every
    # DEFAULT_CLUSTER_MOVES run would.
    for res, refine in ((0.1, 0), (0.6, 1)):
        cfg = pcc.make_config(resolution=res, subclustering_method=method,
                              clustering_moves_method=moves, refine=refine)
        ids, st = pcc.DeviceGraph.from_csr(csr).cluster(cfg)
        r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
        assert np.array_equal(ids, r.ids), (name, method, moves, res)


# Async LOUVAIN rounds with subclustering between levels: the statistical
# contract of test_gpu_async.py (the oracle's async runs over eight
# processing orders).
ASYNC_STRICT = [
    (dict(kind="sbm", n=4000, blocks=40, seed=32),
     dict(resolution=0.01, move_method=2, refine=0, subclustering_method=2), False),
    (dict(kind="rmat", scale=15, ef=16, w_lo=-1.0, w_hi=1.0, seed=35),
     dict(resolution=0.05, move_method=2, refine=1, subclustering_method=3), False),
]
ASYNC_RANGE = [
    (dict(kind="rmat", scale=14, ef=16, seed=31),
     dict(resolution=1.0, move_method=0, refine=1, permute=1, subclustering_method=2), True),
]


@pytest.mark.parametrize("case", range(len(ASYNC_STRICT)))
def test_async_subclustering_within_half_percent(pcc, orc, case):
    from test_gpu_async import _run

    spec, kw, modularity = ASYNC_STRICT[case]
    gpu, ref = _run(pcc, orc, spec, kw, modularity)
    assert abs(gpu - ref.mean()) <= 0.005 * abs(ref.mean())


@pytest.mark.parametrize("case", range(len(ASYNC_RANGE)))
def test_async_subclustering_within_band_of_noisy_orders(pcc, orc, case):
    from test_gpu_async import _run, async_band

    spec, kw, modularity = ASYNC_RANGE[case]
    gpu, ref = _run(pcc, orc, spec, kw, modularity)
    assert abs(gpu - ref.mean()) <= async_band(ref)


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("moves", [0, 1])
def test_refine_from_clustering_with_subclustering(pcc, orc, method, moves):
    """RefineClusters from a given clustering (pcc_refine_clusters) with node
    weights: the first level starts from the clustering, whose clusters are
    split into subclusters before compression (LOUVAIN) or inside the rounds
    (DEFAULT_CLUSTER_MOVES)."""
    csr = GRAPHS["sbm_dense"](pcc)
    rng = np.random.default_rng(60 + method)
    init = (np.arange(csr.n) // 90).astype(np.uint32)  # blocks of 90 across the planted ones
    nw = rng.integers(1, 4, size=csr.n).astype(np.float64)
    cfg = pcc.make_config(resolution=0.05, move_method=2, subclustering_method=method,
                          clustering_moves_method=moves)
    out, st = pcc.DeviceGraph.from_csr(csr).refine(cfg, init, node_weights=nw)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), False, nw, init)
    assert np.array_equal(out, r.ids)
