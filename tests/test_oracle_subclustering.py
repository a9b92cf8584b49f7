"""Pins the oracle's subclustering (CompressSubclusters between LOUVAIN
levels): clustering/clusterers/parallel-correlation-clusterer.cc:436-474,
564-567 and parallel-correlation-clusterer-internal.cc:786-1120.

Checks:
  * the per-edge triangle score equals (A @ A)[u, v] / (deg u + deg v) from a
    dense adjacency matrix (self-loops included as diagonal entries);
  * the subcluster numbering equals a breadth-first search over each
    cluster's members written here from the definition: components of the
    cluster's own edges (TRIANGLE: edges scoring above 0.25), labelled by the
    position of their smallest member, spans laid out in cluster order;
  * a two-level run (with and without refinement) equals the composition of
    the oracle's own primitives: level-0 rounds, the numbering, compression by
    subcluster, the next level started from the parent clusters, and the
    result read through the subcluster ids.
"""

from collections import deque

import numpy as np
import pytest

from test_oracle import to_csr


def random_graph(rng, n, p, self_loops=False):
    A = np.triu(rng.random((n, n)) < p, 1)
    A = A | A.T
    if self_loops:
        A[np.diag_indices(n)] = rng.random(n) < 0.3
    W = np.where(A, rng.integers(1, 4, size=(n, n)), 0).astype(np.float32)
    W = np.triu(W) + np.triu(W, 1).T
    return A, W


def dense_scores(A):
    Ai = A.astype(np.int64)
    common = Ai @ Ai
    deg = Ai.sum(1)
    with np.errstate(invalid="ignore"):  # 0 / 0 between isolated vertices
        return common / (deg[:, None] + deg[None, :]).astype(np.float64)


def bfs_subclusters(A, ids, method):
    n = A.shape[0]
    S = dense_scores(A) if method == 1 else None
    sub = np.zeros(n, np.int64)
    parent = []
    nxt = 0
    for r, c in enumerate(sorted(set(ids.tolist()))):
        mem = [v for v in range(n) if ids[v] == c]
        where = {v: i for i, v in enumerate(mem)}
        label = {}
        for v in mem:
            if v in label:
                continue
            comp, q = [v], deque([v])
            label[v] = -1
            while q:
                u = q.popleft()
                for x in np.flatnonzero(A[u]):
                    x = int(x)
                    if x not in where or x in label:
                        continue
                    if method == 1 and not S[u, x] > 0.25:
                        continue
                    label[x] = -1
                    comp.append(x)
                    q.append(x)
            lo = min(where[x] for x in comp)
            for x in comp:
                label[x] = lo
        for v in mem:
            sub[v] = nxt + label[v]
        span = len(mem) if method == 2 else max(label.values()) + 1
        parent += [r] * span
        nxt += span
    return sub, np.array(parent, np.int64), nxt


@pytest.mark.parametrize("seed", range(4))
def test_triangle_score_matches_dense(orc, seed):
    rng = np.random.default_rng(900 + seed)
    n = 24
    A, W = random_graph(rng, n, 0.3, self_loops=True)
    g = orc.G(*to_csr(A, W))
    S = dense_scores(A)
    for u in range(n):
        for v in np.flatnonzero(A[u]):
            assert orc.triangle_score(g, u, int(v)) == S[u, v]


@pytest.mark.parametrize("method", [1, 2])
@pytest.mark.parametrize("seed", range(6))
def test_subclusters_match_bfs(orc, method, seed):
    rng = np.random.default_rng(1000 + seed)
    n = 30
    A, W = random_graph(rng, n, 0.12 + 0.05 * (seed % 3), self_loops=seed % 2 == 1)
    g = orc.G(*to_csr(A, W))
    # Clusters with unused ids in between, so spans and ranks differ from ids.
    ids = (rng.integers(0, 6, size=n) * 3 + 1).astype(np.uint32)
    sub, parent, total = orc.subclusters(g, ids, method)
    want_sub, want_parent, want_total = bfs_subclusters(A, ids.astype(np.int64), method)
    assert total == want_total
    assert np.array_equal(sub, want_sub)
    k = min(total, n)
    assert np.array_equal(parent[:k], want_parent[:k])


def test_subclusters_split_disconnected_cluster(orc):
    # Cluster 0 = {0, 1, 2, 3}: edges 0-1 and 2-3 only -> two subclusters.
    n = 5
    A = np.zeros((n, n), bool)
    for u, v in [(0, 1), (2, 3), (3, 4)]:
        A[u, v] = A[v, u] = True
    W = A.astype(np.float32)
    g = orc.G(*to_csr(A, W))
    ids = np.array([0, 0, 0, 0, 4], np.uint32)
    sub, parent, total = orc.subclusters(g, ids, 2)
    assert list(sub) == [0, 0, 2, 2, 4] and total == 5
    assert list(parent) == [0, 0, 0, 0, 1]
    # Triangle: no edge lies on a triangle -> every member alone.
    sub, parent, total = orc.subclusters(g, ids, 1)
    assert list(sub) == [0, 1, 2, 3, 4] and total == 5


@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("refine", [False, True])
def test_two_levels_compose(orc, pcc, method, refine):
    # Dense blocks for TRIANGLE (edges need common neighbours), so that the
    # level-0 clusters (several blocks each) split into subclusters.
    csr = pcc.gen_sbm(300, 10, 0.7 if method == 1 else 0.3, 0.02, seed=5 + method)
    g = orc.G.from_csr(csr)
    n = csr.n
    res = 0.05 if method == 1 else 0.02
    cfg = orc.config(resolution=res, louvain_num_iterations=2, refine=refine,
                     subclustering_method=method)
    got = orc.cluster(g, cfg)
    # Level 0 from singletons.
    plain = orc.config(resolution=res, louvain_num_iterations=2, refine=refine)
    s0 = orc.S.singletons(n)
    assert orc.iterate_best_moves(g, s0, plain, 10, plain.permute_seed) == 1
    sub, parent, total = orc.subclusters(g, s0.ids, method, 0.0, res)
    assert len(set(sub.tolist())) > len(set(s0.ids.tolist()))  # some cluster splits
    g1, nw1 = orc.compress(g, sub, np.ones(n))
    n1 = g1.n
    assert n1 == int(sub.max()) + 1
    s1 = orc.S.from_ids(parent[:n1], nw1)
    assert np.array_equal(s1.sizes[: int(parent[:n1].max()) + 1],
                          np.bincount(parent[:n1]))
    moved1 = orc.iterate_best_moves(g1, s1, plain, 10, plain.permute_seed + 1)
    if not refine:
        want = s1.ids[sub]
    elif moved1:
        # Refinement of level 0 from the level-1 clusters read through the
        # subcluster ids.
        s0r = orc.S.from_ids(orc.rank_relabel(s1.ids[sub]), np.ones(n))
        orc.iterate_best_moves(g, s0r, plain, 10, plain.permute_seed + 7919)
        want = s0r.ids
    else:
        want = np.arange(n)  # refinement with one level keeps the identity
    assert np.array_equal(orc.rank_relabel(got.ids), orc.rank_relabel(want))


def _group_moves(orc, g, s, cfg, key):
    """Sync set moves of the groups of key, written from PCC.cc:220-241 and
    263-285 with the oracle's own set move (pinned by brute force in
    test_oracle_cluster_moves.py)."""
    n = g.n
    moves = np.full(n, 0xFFFFFFFF, np.uint32)
    moved = 0
    for kv in np.unique(key):
        members = np.flatnonzero(key == kv)
        t, d = orc.best_move_set(g, s, cfg, members)
        cur = int(s.ids[members[0]])
        if t < n and cur >= t:
            d = 0.0
        if d > 0:
            moves[members] = t
            moved += 1
    orc.apply_moves(s, moves)
    return moved


def _round_case(orc, method, seed):
    """One DEFAULT_CLUSTER_MOVES round with subclustering (PCC.cc:121-297):
    vertex moves, then the subclusters of the resulting clusters move as sets,
    then the clusters as they were before the subcluster moves.  Returns the
    number of subclusters that moved."""
    rng = np.random.default_rng(1200 + seed)
    n = 36
    A, W = random_graph(rng, n, {1: 0.3, 2: 0.12, 3: 0.5}[method])
    g = orc.G(*to_csr(A, W))
    ids = rng.integers(0, 5, size=n).astype(np.uint32) * 7
    k = rng.integers(1, 3, size=n).astype(np.float64)
    cfg = orc.config(resolution=0.08, clustering_moves_method=0, move_method=1,
                     subclustering_method=method)
    got = orc.S.from_ids(ids, k)
    orc.round_(g, got, cfg, np.ones(n, np.uint8))
    want = orc.S.from_ids(ids, k)
    vmoves, _, _, _ = orc.sync_decisions(g, want, cfg)
    orc.apply_moves(want, vmoves)
    before = want.ids.copy()
    sub, _, _ = orc.subclusters(g, want.ids, method, 0.0, 0.08)
    sub_moves = _group_moves(orc, g, want, cfg, sub)
    _group_moves(orc, g, want, cfg, before)
    assert np.array_equal(got.ids, want.ids)
    assert np.array_equal(got.sizes, want.sizes)
    assert np.array_equal(got.K, want.K)
    return sub_moves


@pytest.mark.parametrize("method", [1, 2, 3])
def test_cluster_move_round_with_subclusters(orc, method):
    moved = sum(_round_case(orc, method, seed) for seed in range(5))
    assert moved > 0  # the subcluster phase is exercised


def slow_labels(A, cl, offset, res):
    """ComputeSubclusterConnectivitySlow (PCCI.cc:897-1013) for one cluster,
    written here from the reference's text with sets instead of marks."""
    size = len(cl)
    inside = set(cl)
    tmp = list(range(size))
    single = [False] * size
    nb = [set(int(x) for x in np.flatnonzero(A[v])) for v in range(A.shape[0])]
    for i in range(size):
        if single[i]:
            continue
        v = cl[i]
        if len(nb[v] - inside) < size - 1:
            continue
        best, pick = 0.0, None
        for t in sorted(set(tmp)):
            grp = [j for j in range(size) if tmp[j] == t]
            if grp[0] == i:
                continue
            leaving = sum(len(nb[cl[j]] - inside) for j in grp)
            if leaving < len(grp) * (size - len(grp)):
                continue
            weight = float(len(nb[v] & set(cl[j] for j in grp)))
            # The test's offsets and resolutions are dyadic: the reference's
            # fused multiply-add and this sum agree exactly.
            off = offset * len(grp) + res * len(grp)
            if weight < off:
                continue
            if weight - off >= best:
                best, pick = weight - off, grp
        if pick is None:
            continue
        tmp[i] = tmp[pick[0]]
        single[i] = True
        single[pick[0]] = True
    return tmp


@pytest.mark.parametrize("seed", range(8))
def test_slow_connectivity_matches_restatement(orc, seed):
    rng = np.random.default_rng(1400 + seed)
    n = 32
    A, W = random_graph(rng, n, 0.5, self_loops=seed % 2 == 1)
    g = orc.G(*to_csr(A, W))
    ids = (rng.integers(0, 10, size=n) * 2).astype(np.uint32)  # clusters of ~3
    res, off = [(0.375, 0.0), (0.5, 0.25), (1.0, 0.0), (0.0625, 0.0)][seed % 4]
    sub, parent, total = orc.subclusters(g, ids, 3, off, res)
    want = np.zeros(n, np.int64)
    nxt = 0
    merged = 0
    for r, c in enumerate(sorted(set(ids.tolist()))):
        cl = [v for v in range(n) if ids[v] == c]
        tmp = slow_labels(A, cl, off, res)
        merged += len(cl) - len(set(tmp))
        for j, v in enumerate(cl):
            want[v] = nxt + tmp[j]
        span = max(tmp) + 1
        assert np.all(parent[nxt:nxt + span] == r)
        nxt += span
    assert total == nxt
    assert np.array_equal(sub, want)
    if seed % 4 != 2:
        assert merged > 0  # members do join groups


def test_slow_connectivity_hand_case(orc):
    # Cluster {0, 1, 2}; vertices 3..8 outside.  0-1 adjacent, 2 alone; each
    # of 0, 1, 2 has three edges leaving the cluster (>= size - 1 = 2).
    # i = 0: groups {1} (leaving 3 >= 1 * 2, weight 1 >= res 0.5) and {2}
    # (weight 0 < 0.5): 0 joins 1's group, label 1.  i = 1 is no longer alone,
    # i = 2: group {0, 1} needs 6 >= 2 * 1 leaving edges and weight 0 >= 1.0:
    # no.  Labels 1, 1, 2.
    n = 9
    A = np.zeros((n, n), bool)
    edges = [(0, 1), (0, 3), (0, 4), (0, 5), (1, 6), (1, 7), (1, 8), (2, 3), (2, 6), (2, 8)]
    for u, v in edges:
        A[u, v] = A[v, u] = True
    g = orc.G(*to_csr(A, A.astype(np.float32)))
    ids = np.array([0, 0, 0, 3, 4, 5, 6, 7, 8], np.uint32)
    sub, parent, total = orc.subclusters(g, ids, 3, 0.0, 0.5)
    assert list(sub[:3]) == [1, 1, 2]
    assert list(sub[3:]) == [3, 4, 5, 6, 7, 8] and total == 9


def test_subclustering_unsupported_combinations(orc, pcc):
    g = orc.G.from_csr(pcc.gen_sbm(100, 4, 0.3, 0.02, seed=1))
    for cfg in [orc.config(resolution=0.05, subclustering_method=4),
                orc.config(resolution=0.05, subclustering_method=2, clustering_moves_method=0,
                           async_=True)]:
        with pytest.raises(RuntimeError):
            orc.cluster(g, cfg)
