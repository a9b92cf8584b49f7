"""Pins the oracle's whole-cluster moves (DEFAULT_CLUSTER_MOVES).

Reference: BestMove(graph, moving_nodes)
(parallel-correlation-clusterer-internal.cc:644-707) with BestMoveFromStats
(correlation-clusterer-util.h:70-113), AsyncMove(graph, moving_nodes)
(parallel-correlation-clusterer-internal.cc:405-445) and the cluster phase of
BestMovesForVertexSubset (parallel-correlation-clusterer.cc:167-297).

Checks:
  * brute force: for random vertex sets (whole clusters, parts of clusters and
    sets spanning several clusters) every candidate value equals the change of
    the correlation objective, written as a dense sum over vertex pairs, when
    the whole set moves to that cluster, and "new cluster" equals the change
    when the set forms a cluster of its own;
  * hand cases: two cliques that no single vertex wants to leave are merged by
    a cluster move, into the larger id in sync mode
    (parallel-correlation-clusterer.cc:268-276); the one-level refine quirk
    (parallel-correlation-clusterer.cc:488-492, 550-556, 589).
"""

import numpy as np
import pytest

from test_oracle import cc_pairs, dense_graph, to_csr


def brute_best(A, W, ids, k, res, off, members):
    n = A.shape[0]
    base = cc_pairs(A, W, ids, k, res, off)
    mset = set(int(x) for x in members)
    cands = {}
    for u in members:
        for v in np.flatnonzero(A[u]):
            if int(v) in mset:
                continue
            c = int(ids[v])
            if c in cands:
                continue
            moved = ids.copy()
            moved[list(mset)] = c
            cands[c] = cc_pairs(A, W, moved, k, res, off) - base
    alone = ids.copy()
    alone[list(mset)] = 1000 + n
    new_delta = cc_pairs(A, W, alone, k, res, off) - base
    best_c, best_d = None, new_delta
    for c in sorted(cands):
        if cands[c] > best_d + 1e-12:
            best_c, best_d = c, cands[c]
    near = [x for x in list(cands.values()) + [new_delta] if abs(x - best_d) < 1e-5]
    return best_c, best_d, len(near) == 1


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("res,off", [(0.0, 0.0), (0.3, 0.0), (0.15, 0.2)])
def test_set_move_deltas_are_objective_changes(orc, seed, res, off):
    rng = np.random.default_rng(100 + seed)
    n = 10
    A, W = dense_graph(rng, n, 0.5, signed=(seed % 2 == 0))
    g = orc.G(*to_csr(A, W))
    ids = rng.integers(0, 4, size=n).astype(np.uint32)
    k = rng.integers(1, 4, size=n).astype(np.float64)
    s = orc.S.from_ids(ids, k)
    cfg = orc.config(resolution=res, edge_weight_offset=off, clustering_moves_method=0)
    sets = [np.flatnonzero(ids == c) for c in np.unique(ids)]           # whole clusters
    sets += [np.sort(rng.choice(n, size=sz, replace=False)) for sz in (1, 2, 3, 5)]
    # The reference's set lambda takes the weight as float, so w - offset is
    # rounded to float (parallel-correlation-clusterer-internal.cc:668-669):
    # with an offset the values differ from the exact change by that rounding.
    tol = 1e-9 if off == 0.0 else 1e-6
    checked = 0
    for members in sets:
        if len(members) == 0 or len(members) == n:
            continue
        t, d = orc.best_move_set(g, s, cfg, members)
        best_c, best_d, unique = brute_best(A, W, ids, k, res, off, members)
        assert d == pytest.approx(best_d, abs=tol)
        if unique:
            assert t == (n if best_c is None else best_c)
        checked += 1
    assert checked >= 5


def two_stuck_cliques():
    """Two 4-cliques (ids 0-3, 4-7); every vertex also has two edges into the
    other clique.  At resolution 0.1 a single vertex loses by switching
    (-2.7 + 1.6), while the whole clique gains 8 - 0.1 * 16 = 6.4."""
    n = 8
    A = np.zeros((n, n), bool)
    for base in (0, 4):
        for i in range(4):
            for j in range(4):
                if i != j:
                    A[base + i, base + j] = True
    for i in range(4):
        for j in (4 + i, 4 + (i + 1) % 4):
            A[i, j] = A[j, i] = True
    return A, A.astype(np.float32)


def test_whole_cluster_move_delta_hand_case(orc):
    A, W = two_stuck_cliques()
    g = orc.G(*to_csr(A, W))
    ids = np.array([0] * 4 + [4] * 4, np.uint32)
    s = orc.S.from_ids(ids, np.ones(8))
    cfg = orc.config(resolution=0.1, clustering_moves_method=0)
    assert orc.best_move_set(g, s, cfg, np.arange(4)) == (4, pytest.approx(6.4, abs=1e-12))
    assert orc.best_move_set(g, s, cfg, np.arange(4, 8)) == (0, pytest.approx(6.4, abs=1e-12))
    # No single vertex moves.
    for v in range(8):
        t, d = orc.best_move(g, s, cfg, v)
        assert d < 0


@pytest.mark.parametrize("async_", [False, True])
def test_cluster_moves_merge_stuck_cliques(orc, async_):
    A, W = two_stuck_cliques()
    g = orc.G(*to_csr(A, W))
    init = np.array([0] * 4 + [4] * 4, np.uint32)
    louvain = orc.cluster(g, orc.config(resolution=0.1, louvain_num_iterations=1,
                                        async_=async_), initial_ids=init)
    # Vertex moves alone are stuck, so nothing moves and the reference returns
    # its identity cluster_ids (parallel-correlation-clusterer.cc:543-547).
    assert louvain.rounds == 1 and np.array_equal(louvain.ids, np.arange(8))
    # Async runs ALL_MOVE: see test_async_cluster_move_marks_vertex_ids.
    r = orc.cluster(g, orc.config(resolution=0.1, clustering_moves_method=0, async_=async_,
                                  move_method=1 if async_ else 0, num_iterations=3),
                    initial_ids=init)
    # Sync: cluster 4 may not move into the smaller id 0, cluster 0 moves
    # into 4.  Async (clusters in ascending order): 0 moves first, then
    # {4..7} gains nothing by leaving.
    assert np.array_equal(r.ids, np.full(8, 4, np.uint32))
    base = orc.objective(g, init, np.ones(8), 0.1)
    assert r.objective == pytest.approx(base + 6.4, abs=1e-9)


def test_async_cluster_move_marks_vertex_ids(orc):
    # The async cluster phase marks moved_clusters at the members' vertex ids
    # (parallel-correlation-clusterer.cc:257-259).  Here cluster 0 = {0..3}
    # moves into 4, entries 0..3 are marked, but every vertex is now in
    # cluster 4, so the NBHR_CLUSTER_MOVE frontier is empty, the level reports
    # no move and the identity comes back (:543-547).
    A, W = two_stuck_cliques()
    g = orc.G(*to_csr(A, W))
    init = np.array([0] * 4 + [4] * 4, np.uint32)
    r = orc.cluster(g, orc.config(resolution=0.1, clustering_moves_method=0, async_=True),
                    initial_ids=init)
    assert r.rounds == 1 and np.array_equal(r.ids, np.arange(8))
    s = orc.S.from_ids(init, np.ones(8))
    fr, size = orc.round_(g, s, orc.config(resolution=0.1, clustering_moves_method=0,
                                           async_=True), np.ones(8, np.uint8))
    assert size == 0 and np.array_equal(s.ids, np.full(8, 4, np.uint32))
    assert list(s.sizes) == [0, 0, 0, 0, 8, 0, 0, 0] and s.K[4] == 8.0


def test_default_cluster_moves_refine_quirk(orc, pcc):
    # One level only: with refine the reference keeps cluster_ids at the
    # identity (parallel-correlation-clusterer.cc:488-492, 550-556, 589).
    csr = pcc.gen_sbm(300, 6, 0.3, 0.01, seed=4)
    g = orc.G.from_csr(csr)
    r = orc.cluster(g, orc.config(resolution=0.01, clustering_moves_method=0, refine=True))
    assert np.array_equal(r.ids, np.arange(csr.n)) and r.levels == 1
    r2 = orc.cluster(g, orc.config(resolution=0.01, clustering_moves_method=0))
    assert r2.num_clusters < 60 and r2.levels == 1


@pytest.mark.parametrize("move_method", [0, 1, 2])
def test_default_cluster_moves_rounds_and_threads(orc, pcc, move_method):
    csr = pcc.gen_rmat(10, 8, w_lo=-1, w_hi=1, seed=11)
    g = orc.G.from_csr(csr)
    cfg = orc.config(resolution=0.05, clustering_moves_method=0, move_method=move_method,
                     num_iterations=6)
    orc.set_threads(1)
    a = orc.cluster(g, cfg)
    orc.set_threads(4)
    b = orc.cluster(g, cfg)
    orc.set_threads(1)
    assert np.array_equal(a.ids, b.ids)
    assert a.rounds <= 6 and (move_method != 1 or a.rounds == 6)
    single = orc.objective(g, np.arange(csr.n), np.ones(csr.n), 0.05)
    assert a.objective > single
