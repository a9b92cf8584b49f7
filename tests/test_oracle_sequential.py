"""Pins the oracle's sequential clusterer (SEQ-CC / SEQ-MOD, the paper's own
CPU baselines): clustering/clusterers/correlation-clusterer.cc and
correlation-clusterer-internal.cc.

Checks:
  * one sequential round equals a brute-force simulation written from the
    mathematics: vertices in ascending order, each moving to the candidate
    (its neighbours' clusters, or a new cluster = the smallest empty id from
    0) whose dense pair-sum objective change is largest, ties to "new" and
    then to the smaller id, iff that change is positive.  Integer weights and
    a dyadic resolution make every value exact, so ids must match exactly;
  * a converged one-level run (all_iter) is a local optimum: no single vertex
    move raises the dense objective;
  * modularity runs report the modularity of their own ids.
"""

import numpy as np
import pytest

from test_oracle import cc_pairs, dense_graph, to_csr


def seq_round_bruteforce(A, W, ids, k, res, off):
    n = A.shape[0]
    ids = ids.copy()
    sizes = np.bincount(ids, minlength=n)
    for v in range(n):
        base = cc_pairs(A, W, ids, k, res, off)
        best_c, best_d = None, None
        # "new cluster": the reference picks the first empty id from 0
        # (correlation-clusterer-internal.cc:102-112); its value is the change
        # of isolating v.
        alone = ids.copy()
        alone[v] = n + 1000
        new_d = cc_pairs(A, W, alone, k, res, off) - base
        best_d = new_d
        for c in sorted(set(int(ids[u]) for u in np.flatnonzero(A[v])) - {int(ids[v])}):
            moved = ids.copy()
            moved[v] = c
            d = cc_pairs(A, W, moved, k, res, off) - base
            if d > best_d:
                best_c, best_d = c, d
        if not best_d > 0:
            continue
        cur = int(ids[v])
        sizes[cur] -= 1
        if best_c is None:
            best_c = int(np.flatnonzero(sizes == 0)[0])
        sizes[best_c] += 1
        ids[v] = best_c
    return ids


@pytest.mark.parametrize("seed", range(10))
def test_sequential_round_equals_bruteforce(orc, seed):
    rng = np.random.default_rng(300 + seed)
    n = 11
    A, W = dense_graph(rng, n, 0.45, integer=True)
    g = orc.G(*to_csr(A, W))
    ids = rng.integers(0, 5, size=n).astype(np.uint32)
    k = rng.integers(1, 3, size=n).astype(np.float64)
    res, off = (0.25, 0.5) if seed % 2 else (0.125, 0.0)
    want = seq_round_bruteforce(A, W, ids.astype(np.int64), k, res, off)
    s = orc.S.from_ids(ids, k)
    cfg = orc.config(resolution=res, edge_weight_offset=off, sequential=True, move_method=1)
    orc.round_(g, s, cfg, np.ones(n, np.uint8))
    assert np.array_equal(s.ids, want.astype(np.uint32))
    assert np.array_equal(s.sizes, np.bincount(s.ids, minlength=n).astype(np.uint32))
    K = np.zeros(n)
    np.add.at(K, s.ids, k)
    assert np.array_equal(s.K, K)


def test_new_cluster_takes_smallest_empty_id(orc):
    # Vertex 2 leaves cluster {1, 2} (edge weight below the resolution
    # penalty); ids 0..? : cluster 0 = {0}, 1 = {1, 2}, 3 = {3}; empty: 2.
    n = 4
    A = np.zeros((n, n), bool)
    A[1, 2] = A[2, 1] = True
    W = A.astype(np.float32) * np.float32(0.25)
    g = orc.G(*to_csr(A, W))
    ids = np.array([0, 1, 1, 3], np.uint32)
    s = orc.S.from_ids(ids, np.ones(n))
    cfg = orc.config(resolution=1.0, sequential=True, move_method=1)
    orc.round_(g, s, cfg, np.ones(n, np.uint8))
    # Vertex 1 moves first and takes id 2 (the parallel clusterer's probe
    # would start at the vertex's own id), then vertex 2 stays alone in 1.
    assert list(s.ids) == [0, 2, 1, 3]


@pytest.mark.parametrize("seed", range(4))
def test_converged_level_is_local_optimum(orc, seed):
    rng = np.random.default_rng(400 + seed)
    n = 14
    A, W = dense_graph(rng, n, 0.35, integer=True)
    g = orc.G(*to_csr(A, W))
    res = 0.25
    # One level run until nothing moves, as all_iter does
    # (correlation-clusterer.cc:173-183); NBHR_MOVE so that the frontier
    # empties (with ALL_MOVE and all_iter the reference never stops).
    cfg = orc.config(resolution=res, sequential=True, all_iter=True, move_method=2)
    s = orc.S.singletons(n)
    fr = np.ones(n, np.uint8)
    for _ in range(1000):
        fr, size = orc.round_(g, s, cfg, fr)
        if size == 0:
            break
    assert size == 0
    ids = s.ids.astype(np.int64)
    k = np.ones(n)
    base = cc_pairs(A, W, ids, k, res, 0.0)
    assert base > 0
    for v in range(n):
        for c in set(int(ids[u]) for u in np.flatnonzero(A[v])) | {n + 7}:
            moved = ids.copy()
            moved[v] = c
            assert cc_pairs(A, W, moved, k, res, 0.0) <= base + 1e-9


def test_sequential_pipeline(orc, pcc):
    csr = pcc.gen_sbm(2000, 20, 0.1, 0.002, seed=1)
    g = orc.G.from_csr(csr)
    r = orc.cluster(g, orc.config(resolution=0.01, sequential=True))
    assert r.num_clusters == 20 and r.levels >= 1
    blocks = np.arange(csr.n) // 100
    agree = sum(np.bincount(r.ids[blocks == b]).max() for b in range(20))
    assert agree >= 0.99 * csr.n  # the planted blocks, up to a few noisy vertices
    m = orc.cluster(orc.G.from_csr(pcc.gen_rmat(12, 8, seed=3)),
                    orc.config(resolution=1.0, sequential=True), True)
    assert m.modularity > 0.1
