"""Pins the CPU oracle to what the mathematics fixes.

The reference ships no tests or fixtures, so the oracle is checked against:
  * brute force on tiny graphs: every best-move delta equals the change of the
    correlation-clustering objective written independently as a sum over
    unordered vertex pairs (dense numpy), except the own-cluster candidate,
    whose coded value is -res * k^2
    (parallel-correlation-clusterer-internal.cc:483-504);
  * hand-built cases for the tie, new-cluster and singleton rules
    (parallel-correlation-clusterer-internal.cc:490-503,
    parallel-correlation-clusterer.cc:145-156) and for empty-id reuse
    (parallel-correlation-clusterer-internal.cc:339-366);
  * networkx: frontier sets, quotient-graph weights, modularity;
  * numpy closed forms for flatten and rank relabel.
"""

import itertools

import networkx as nx
import numpy as np
import pytest

from conftest import random_state


def dense_graph(rng, n, p, signed=False, integer=False):
    A = np.triu(rng.random((n, n)) < p, 1)
    if integer:
        W = rng.integers(1, 4, size=(n, n)).astype(np.float32)
    elif signed:
        W = (rng.random((n, n)) * 2 - 1).astype(np.float32)
    else:
        W = np.ones((n, n), np.float32)
    W = np.triu(W, 1)
    A = A | A.T
    W = (W + W.T) * A
    return A, W


def to_csr(A, W):
    n = A.shape[0]
    off = np.zeros(n + 1, np.int64)
    nb, ws = [], []
    for u in range(n):
        vs = np.flatnonzero(A[u])
        off[u + 1] = off[u] + len(vs)
        nb.extend(vs.tolist())
        ws.extend(W[u, vs].tolist())
    return off, np.array(nb, np.uint32), np.array(ws, np.float32)


def cc_pairs(A, W, ids, k, res, off):
    """Correlation-clustering objective as a sum over unordered pairs u < v in
    one cluster: (w_uv - offset) for edges, minus res * k_u * k_v."""
    n = A.shape[0]
    total = 0.0
    for u in range(n):
        for v in range(u + 1, n):
            if ids[u] != ids[v]:
                continue
            if A[u, v]:
                total += float(W[u, v]) - off
            total -= res * k[u] * k[v]
    return total


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("res,off", [(0.0, 0.0), (0.3, 0.0), (0.15, 0.2)])
def test_best_move_deltas_are_objective_changes(orc, seed, res, off):
    rng = np.random.default_rng(seed)
    n = 9
    A, W = dense_graph(rng, n, 0.5, signed=(seed % 2 == 0))
    g = orc.G(*to_csr(A, W))
    ids = rng.integers(0, 4, size=n).astype(np.uint32)
    k = rng.integers(1, 4, size=n).astype(np.float64)
    s = orc.S.from_ids(ids, k)
    cfg = orc.config(resolution=res, edge_weight_offset=off)
    base = cc_pairs(A, W, ids, k, res, off)
    for v in range(n):
        t, d = orc.best_move(g, s, cfg, v)
        cands = {}
        for u in np.flatnonzero(A[v]):
            c = int(ids[u])
            if c == ids[v]:
                cands[c] = -res * k[v] * k[v]  # coded own-cluster value
            else:
                moved = ids.copy()
                moved[v] = c
                cands[c] = cc_pairs(A, W, moved, k, res, off) - base
        alone = ids.copy()
        alone[v] = 1000 + v
        new_delta = cc_pairs(A, W, alone, k, res, off) - base
        best_c, best_d = None, new_delta
        for c in sorted(cands):
            if cands[c] > best_d + 1e-12:
                best_c, best_d = c, cands[c]
        near = [x for x in list(cands.values()) + [new_delta] if abs(x - best_d) < 1e-9]
        assert d == pytest.approx(best_d, abs=1e-9)
        if len(near) == 1:
            assert t == (n if best_c is None else best_c)


def path_graph(n, w=1.0):
    A = np.zeros((n, n), bool)
    for i in range(n - 1):
        A[i, i + 1] = A[i + 1, i] = True
    return A, A * np.float32(w)


def test_tie_prefers_smaller_cluster_id(orc):
    # Vertex 2 sees clusters 7 and 5 with identical weights: smaller id wins.
    A = np.zeros((8, 8), bool)
    A[2, 0] = A[0, 2] = A[2, 1] = A[1, 2] = True
    W = A.astype(np.float32)
    g = orc.G(*to_csr(A, W))
    ids = np.array([7, 5, 2, 3, 4, 5, 6, 7], np.uint32)
    s = orc.S.from_ids(ids, np.ones(8))
    t, d = orc.best_move(g, s, orc.config(resolution=0.0), 2)
    assert (t, d) == (5, 1.0)


def test_new_cluster_wins_ties(orc):
    # change == delta_B == 0 at res 0.5: "new cluster" (id n) is kept.
    n = 4
    A = np.zeros((n, n), bool)
    A[0, 1] = A[1, 0] = A[0, 2] = A[2, 0] = True
    W = np.zeros((n, n), np.float32)
    W[0, 1] = W[1, 0] = 0.5
    W[0, 2] = W[2, 0] = 0.5
    g = orc.G(*to_csr(A, W))
    ids = np.array([0, 0, 2, 3], np.uint32)  # cluster 0 = {0, 1}; 2 = {2}
    s = orc.S.from_ids(ids, np.ones(n))
    t, d = orc.best_move(g, s, orc.config(resolution=0.5), 0)
    # change = -(0.5 - 0.5 * 1 * (2 - 1)) = 0; delta_2 = 0 + 0.5 - 0.5 = 0.
    assert t == n and d == 0.0


def test_singleton_rule(orc):
    A, W = path_graph(3)
    g = orc.G(*to_csr(A, W))
    s = orc.S.singletons(3)
    moves, moved, tgt, dlt = orc.sync_decisions(g, s, orc.config(resolution=0.0))
    # 0 -> 1 (0 < 1 so it may move); 1 -> 0 blocked (1 >= 0); 2 -> 1 blocked.
    assert list(tgt) == [1, 0, 1]
    assert list(moves) == [1, orc.NO_MOVE, orc.NO_MOVE]
    assert list(moved) == [1, 0, 0]


def test_apply_reuses_empty_ids_in_order(orc):
    n = 6
    ids = np.array([0, 0, 0, 3, 3, 5], np.uint32)
    s = orc.S.from_ids(ids, np.ones(n))
    moves = np.full(n, orc.NO_MOVE, np.uint32)
    moves[2] = n      # new cluster
    moves[4] = n      # new cluster
    moves[5] = 0      # joins cluster 0
    mod = orc.apply_moves(s, moves)
    # After the moves: cluster 0 = {0,1,5}, 3 = {3}, 5 empty; empty ids
    # ascending = 1, 2, 4, 5 -> vertex 2 gets 1, vertex 4 gets 2.
    assert list(s.ids) == [0, 0, 1, 3, 2, 0]
    assert list(s.sizes[:6]) == [3, 1, 1, 1, 0, 0]
    assert list(s.K[:6]) == [3.0, 1.0, 1.0, 1.0, 0.0, 0.0]
    assert list(np.flatnonzero(mod)) == [0, 1, 2, 3, 5]


@pytest.mark.parametrize("seed", range(4))
def test_apply_invariants(orc, seed):
    rng = np.random.default_rng(seed)
    n = 300
    ids, sizes, nw, K = random_state(rng, n, 40, nw=rng.integers(1, 5, n).astype(float))
    s = orc.S(ids, sizes, nw, K)
    moves = np.full(n, orc.NO_MOVE, np.uint32)
    sel = rng.random(n) < 0.4
    moves[sel] = rng.integers(0, n + 1, size=sel.sum())
    moves[sel & (rng.random(n) < 0.3)] = n
    orc.apply_moves(s, moves)
    assert int(s.sizes.sum()) == n
    assert np.array_equal(s.sizes, np.bincount(s.ids, minlength=n)[:n])
    Kx = np.zeros(n)
    np.add.at(Kx, s.ids, nw)
    assert np.array_equal(s.K, Kx)


@pytest.mark.parametrize("move_method", [0, 2])
def test_frontier_is_neighbourhood_of_seeds(orc, move_method):
    rng = np.random.default_rng(4)
    A, W = dense_graph(rng, 60, 0.08)
    g = orc.G(*to_csr(A, W))
    s = orc.S.singletons(60)
    before = s.ids.copy()
    fr, size = orc.round_(g, s, orc.config(resolution=0.02, move_method=move_method),
                          np.ones(60, np.uint8))
    G = nx.from_numpy_array(A.astype(int))
    if move_method == 2:  # NBHR_MOVE: neighbours of vertices that moved
        seeds = {v for v in range(60) if s.ids[v] != before[v]}
        # vertices with delta > 0 but no actual change also seed; recompute
        moves, moved, _, _ = orc.sync_decisions(g, orc.S.singletons(60),
                                                orc.config(resolution=0.02))
        seeds = set(np.flatnonzero(moved))
    else:  # NBHR_CLUSTER_MOVE: neighbours of vertices in modified clusters
        changed = set(int(before[v]) for v in range(60) if s.ids[v] != before[v])
        changed |= set(int(s.ids[v]) for v in range(60) if s.ids[v] != before[v])
        seeds = {v for v in range(60) if int(s.ids[v]) in changed}
    expect = set()
    for v in seeds:
        expect |= set(G.neighbors(v))
    assert set(np.flatnonzero(fr)) == expect
    assert size == len(expect)


@pytest.mark.parametrize("seed", range(3))
def test_compress_invariants(orc, seed):
    rng = np.random.default_rng(seed)
    n = 80
    A, W = dense_graph(rng, n, 0.1, integer=True)
    g = orc.G(*to_csr(A, W))
    ids = rng.integers(0, 12, size=n).astype(np.uint32)
    nw = rng.integers(1, 4, size=n).astype(np.float64)
    cg, nw2 = orc.compress(g, ids, nw)
    assert cg.n == int(ids.max()) + 1
    assert nw2.sum() == nw.sum()
    for c in range(cg.n):
        assert nw2[c] == nw[ids == c].sum()
    # Quotient weights: inter-cluster weight appears in both directions, the
    # intra-cluster weight once as a self-loop.
    Q = {}
    for u in range(n):
        for v in range(n):
            if A[u, v] and (v <= u or ids[u] != ids[v]):
                Q[(ids[u], ids[v])] = Q.get((ids[u], ids[v]), 0.0) + float(W[u, v])
    got = {}
    for c in range(cg.n):
        for p in range(cg.offsets[c], cg.offsets[c + 1]):
            got[(c, int(cg.nbrs[p]))] = float(cg.weights[p])
        nbrs = cg.nbrs[cg.offsets[c]:cg.offsets[c + 1]]
        assert np.all(np.diff(nbrs.astype(np.int64)) > 0)  # sorted, deduplicated
    assert got == pytest.approx(Q)
    G = nx.Graph()
    G.add_nodes_from(range(n))
    for u in range(n):
        for v in np.flatnonzero(A[u]):
            G.add_edge(u, int(v), weight=float(W[u, v]))
    parts = [set(np.flatnonzero(ids == c)) for c in range(cg.n) if (ids == c).any()]
    qg = nx.quotient_graph(G, parts, relabel=False,
                           edge_data=lambda b, c: {"weight": sum(
                               G[x][y]["weight"] for x in b for y in c if G.has_edge(x, y))})
    for b, c, d in qg.edges(data=True):
        cb, cc_ = int(ids[next(iter(b))]), int(ids[next(iter(c))])
        assert got[(cb, cc_)] == pytest.approx(d["weight"])
        assert got[(cc_, cb)] == pytest.approx(d["weight"])
    # Objective invariance: for two clusterings X, Y of the contracted graph,
    # the objective difference equals that of their flattenings on the input
    # graph (self-loop halves and within-cluster pair terms cancel).
    res = 0.05
    X = rng.integers(0, 4, size=cg.n).astype(np.uint32)
    Y = rng.integers(0, 4, size=cg.n).astype(np.uint32)
    dc = orc.objective(cg, X, nw2, res) - orc.objective(cg, Y, nw2, res)
    do = orc.objective(g, X[ids], nw, res) - orc.objective(g, Y[ids], nw, res)
    assert dc == pytest.approx(do, rel=1e-9, abs=1e-9)


def test_compress_long_run_float_sum(orc):
    # 3000 parallel contributions to one pair; with small integers every
    # summation order is exact, so the run sum must be the exact total.
    n = 3002
    A = np.zeros((n, n), bool)
    A[0, 2:] = A[2:, 0] = True
    A[1, 2:] = A[2:, 1] = True
    W = A.astype(np.float32) * 3
    g = orc.G(*to_csr(A, W))
    ids = np.zeros(n, np.uint32)
    ids[1] = 1
    ids[2:] = 1
    cg, _ = orc.compress(g, ids, np.ones(n))
    c0 = dict(zip(cg.nbrs[cg.offsets[0]:cg.offsets[1]].tolist(),
                  cg.weights[cg.offsets[0]:cg.offsets[1]].tolist()))
    assert c0[1] == 3.0 * (n - 2)


def test_objective_closed_form(orc):
    rng = np.random.default_rng(8)
    n = 12
    A, W = dense_graph(rng, n, 0.4, signed=True)
    g = orc.G(*to_csr(A, W))
    for _ in range(5):
        ids = rng.integers(0, 4, size=n).astype(np.uint32)
        k = rng.integers(1, 3, size=n).astype(np.float64)
        assert orc.objective(g, ids, k, 0.2, 0.1) == pytest.approx(
            cc_pairs(A, W, ids, k, 0.2, 0.1), abs=1e-9)


def test_modularity_matches_networkx(orc):
    G = nx.karate_club_graph()
    n = G.number_of_nodes()
    A = nx.to_numpy_array(G, weight=None).astype(bool)
    g = orc.G(*to_csr(A, A.astype(np.float32)))
    rng = np.random.default_rng(1)
    for gamma in (1.0, 0.5):
        for _ in range(3):
            ids = rng.integers(0, 5, size=n).astype(np.uint32)
            comms = [set(np.flatnonzero(ids == c)) for c in np.unique(ids)]
            ref = nx.algorithms.community.modularity(G, comms, weight=None, resolution=gamma)
            assert orc.modularity(g, ids, gamma) == pytest.approx(ref, abs=1e-12)


def test_flatten_and_relabel_closed_forms(orc):
    rng = np.random.default_rng(3)
    ids = rng.integers(0, 50, size=200).astype(np.uint32)
    comp = rng.integers(0, 30, size=50).astype(np.uint32)
    assert np.array_equal(orc.flatten(ids, comp), comp[ids])
    _, inv = np.unique(ids, return_inverse=True)
    assert np.array_equal(orc.rank_relabel(ids), inv.astype(np.uint32))


def test_refine_quirk_single_level_is_singletons(orc, pcc):
    # refine with one level keeps cluster_ids at the identity
    # (parallel-correlation-clusterer.cc:526-530, 589).
    csr = pcc.gen_sbm(300, 6, 0.3, 0.01, seed=4)
    g = orc.G.from_csr(csr)
    r = orc.cluster(g, orc.config(resolution=0.01, refine=True, louvain_num_iterations=1,
                                  move_method=1))
    assert np.array_equal(r.ids, np.arange(csr.n))
    r2 = orc.cluster(g, orc.config(resolution=0.01, refine=False, louvain_num_iterations=1,
                                   move_method=1))
    assert r2.num_clusters < csr.n


def test_all_move_runs_every_level(orc, pcc):
    csr = pcc.gen_sbm(500, 5, 0.2, 0.005, seed=5)
    r = orc.cluster(orc.G.from_csr(csr), orc.config(resolution=0.01, move_method=1,
                                                    louvain_num_iterations=4,
                                                    louvain_num_inner_iterations=3))
    assert r.levels == 4 and r.rounds == 12


def test_planted_partition_recovered(orc, pcc):
    csr = pcc.gen_sbm(400, 4, 0.5, 0.0, seed=6)
    r = orc.cluster(orc.G.from_csr(csr), orc.config(resolution=0.01, move_method=1))
    blocks = np.arange(400) // 100
    for b in range(4):
        assert len(set(r.ids[blocks == b].tolist())) == 1
    assert r.num_clusters == 4


@pytest.mark.parametrize("kw", [dict(move_method=0, refine=True), dict(move_method=2),
                                dict(move_method=1, refine=True)])
def test_sync_result_independent_of_threads(orc, pcc, kw):
    csr = pcc.gen_rmat(11, 8, w_lo=-1, w_hi=1, seed=9)
    g = orc.G.from_csr(csr)
    cfg = orc.config(resolution=0.05, **kw)
    orc.set_threads(1)
    a = orc.cluster(g, cfg)
    orc.set_threads(4)
    b = orc.cluster(g, cfg)
    orc.set_threads(1)
    assert np.array_equal(a.ids, b.ids)


def test_modularity_pipeline_improves_modularity(orc, pcc):
    csr = pcc.gen_sbm(600, 6, 0.2, 0.01, seed=7)
    g = orc.G.from_csr(csr)
    r = orc.cluster(g, orc.config(resolution=1.0), modularity_objective=True)
    single = orc.modularity(g, np.arange(csr.n), 1.0)
    assert r.modularity > single + 0.3
    assert r.modularity == pytest.approx(orc.modularity(g, r.ids, 1.0), abs=1e-12)
    # The blocks are found on level 0, level 1 makes no move, and with
    # refinement the reference then returns its identity cluster_ids
    # (parallel-correlation-clusterer.cc:543-547, 589).
    q = orc.cluster(g, orc.config(resolution=1.0, refine=True), modularity_objective=True)
    assert q.levels == 1 and np.array_equal(q.ids, np.arange(csr.n))


def test_async_oracle_objective_close_to_sync(orc, pcc):
    csr = pcc.gen_sbm(800, 8, 0.15, 0.01, seed=8)
    g = orc.G.from_csr(csr)
    s = orc.cluster(g, orc.config(resolution=0.01))
    a = orc.cluster(g, orc.config(resolution=0.01, async_=True))
    assert a.objective == pytest.approx(s.objective, rel=0.05)
