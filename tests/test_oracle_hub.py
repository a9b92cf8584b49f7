"""Pins the oracle where its hub branch and async apply differ from the
small-degree code (parallel-correlation-clusterer-internal.cc:516-582 for
degree > 1000, :371-403 for AsyncMove) and where the modularity objective is
mapped onto the correlation one (parallel-graph-utils.cc:141-153,
parallel-modularity-clusterer.cc:68-96).

What pins each part, independently of the oracle's code:
  * every candidate delta of a hub vertex (degree 1001-3000, random clusters,
    signed weights, self-loops, edge_weight_offset != 0) equals the change of
    the correlation objective written as an exactly rounded sum over edges and
    cluster weights, with the hub's edge weights taken as float32(w - offset),
    the float subtraction of :528-529;
  * a known answer for the pbbs-blocked own-cluster total (blocks of
    max(1024, 4 ceil(sqrt(deg))) folded left, then the block sums): a 2^60
    edge followed by 2,999 unit edges, where the blocked sum and an
    adjacency-order fold differ by 2048;
  * the max-reduce identity (0, 0.0) of a hub with no candidate and a
    negative change (:565-582), against the (n, change) of the small-degree
    path on the same vertex with one self-loop fewer;
  * the async new-cluster probe v, v+1, ... mod n (:380-401) on hand cases;
  * the modularity mapping (k = degree, res = gamma / sum deg, offset 0):
    CC - m Q is the same constant for every clustering, with Q from networkx,
    and one sync round of the oracle's modularity pipeline equals the moves
    derived from networkx-verified modularity gains in exact rational
    arithmetic.
"""

import math
from fractions import Fraction

import networkx as nx
import numpy as np
import pytest


def hub_csr(rng, deg, extra, signed, self_loops):
    """Vertex 0 adjacent to 1..deg (plus self-loops), extra random edges among
    1..deg; symmetric, float32 weights."""
    n = deg + 1
    pairs = {(0, u): None for u in range(1, n)}
    while len(pairs) < deg + extra:
        a, b = sorted(rng.integers(1, n, size=2).tolist())
        if a != b:
            pairs[(a, b)] = None
    rows = [[] for _ in range(n)]
    for (a, b) in pairs:
        w = np.float32(rng.uniform(-1, 1) if signed else 1.0)
        rows[a].append((b, w))
        rows[b].append((a, w))
    for _ in range(self_loops):
        rows[0].append((0, np.float32(rng.uniform(-1, 1) if signed else 1.0)))
    for r in rows:
        rng.shuffle(r)
    off = np.zeros(n + 1, np.int64)
    for v in range(n):
        off[v + 1] = off[v] + len(rows[v])
    nb = np.array([u for r in rows for (u, _) in r], np.uint32)
    ws = np.array([w for r in rows for (_, w) in r], np.float32)
    return off, nb, ws


def objective_exact(off, nb, ws, ids, k, res, hub_w):
    """Correlation objective: sum over undirected non-loop edges inside a
    cluster of their weight, minus res * sum over unordered same-cluster pairs
    of k_u k_x.  hub_w(u, x, w) gives an edge's weight; fsum makes each total
    exactly rounded."""
    n = len(off) - 1
    terms = []
    for u in range(n):
        for p in range(off[u], off[u + 1]):
            x = int(nb[p])
            if x > u and ids[x] == ids[u]:
                terms.append(hub_w(u, x, float(ws[p])))
    edge_part = math.fsum(terms)
    pair_terms = []
    K = {}
    for v in range(n):
        K[int(ids[v])] = K.get(int(ids[v]), 0.0) + float(k[v])
    for c, kc in K.items():
        pair_terms.append(kc * kc)
    pair_terms.extend(-float(x) * float(x) for x in k)
    return edge_part - res * math.fsum(pair_terms) / 2


@pytest.mark.parametrize("seed", range(6))
def test_hub_candidate_deltas_are_objective_changes(orc, seed):
    rng = np.random.default_rng(100 + seed)
    deg = int(rng.integers(1001, 3001))
    off, nb, ws = hub_csr(rng, deg, extra=2 * deg, signed=seed % 3 != 2, self_loops=seed % 3)
    n = len(off) - 1
    res = [0.0, 0.02, 0.003][seed % 3]
    offset = [0.0, 0.3, -0.15][seed % 3] if seed < 3 else 0.1 * (seed - 2)
    g = orc.G(off, nb, ws)
    ncl = int(rng.integers(5, 40))
    ids = rng.integers(0, ncl, size=n).astype(np.uint32)
    k = rng.integers(1, 4, size=n).astype(np.float64)
    s = orc.S.from_ids(ids, k)
    cfg = orc.config(resolution=res, edge_weight_offset=offset)

    def w_of(u, x, w):
        # The hub path subtracts the offset in float (PCCI.cc:528-529); only
        # pairs containing vertex 0 enter the delta of vertex 0.
        return float(np.float32(float(w) - offset)) if 0 in (u, x) else w - offset

    base = objective_exact(off, nb, ws, ids, k, res, w_of)
    cands, change = orc.best_move_candidates(g, s, cfg, 0)
    nbr_clusters = {int(ids[u]) for u in nb[off[0]:off[1]] if u != 0}
    assert set(cands) == nbr_clusters
    for c, d in cands.items():
        if c == int(ids[0]):
            # The coded own-cluster value: change + T - res k K_own = -res k^2.
            assert d == pytest.approx(-res * k[0] * k[0], abs=1e-9)
            continue
        moved = ids.copy()
        moved[0] = c
        assert d == pytest.approx(objective_exact(off, nb, ws, moved, k, res, w_of) - base,
                                  abs=1e-9), c
    alone = ids.copy()
    alone[0] = n + 5
    assert change == pytest.approx(objective_exact(off, nb, ws, alone, k, res, w_of) - base,
                                   abs=1e-9)
    # The selection rule: the largest delta, the smaller id on ties, "new
    # cluster" (id n) when change is at least as large.
    t, d = orc.best_move(g, s, cfg, 0)
    best_c, best_d = min(cands.items(), key=lambda kv: (-kv[1], kv[0]))
    if change >= best_d:
        assert (t, d) == (n, change)
    else:
        assert (t, d) == (best_c, best_d)


def test_hub_offset_is_subtracted_in_float(orc):
    """With an offset whose float32 subtraction rounds, a hub's candidate sums
    are sums of float32(w - offset), not of the double w - offset that the
    small-degree path uses: vertex 0 alone in its cluster at resolution 0 has
    change 0, so each candidate delta is that cluster's edge sum."""
    rng = np.random.default_rng(7)
    deg = 1500
    off, nb, ws = hub_csr(rng, deg, extra=0, signed=True, self_loops=0)
    n = len(off) - 1
    offset = 0.1  # not a float32 value, so float32(w - 0.1) rounds
    ids = np.zeros(n, np.uint32)
    ids[1:] = rng.integers(1, 4, size=n - 1)
    s = orc.S.from_ids(ids, np.ones(n))
    g = orc.G(off, nb, ws)
    cands, change = orc.best_move_candidates(g, s, orc.config(edge_weight_offset=offset), 0)
    assert change == 0.0 and set(cands) == {1, 2, 3}
    row_ids, row_w = ids[nb[off[0]:off[1]]], ws[off[0]:off[1]]
    for c, d in cands.items():
        wf = row_w[row_ids == c]
        as_float = math.fsum(float(np.float32(float(w) - offset)) for w in wf)
        as_double = math.fsum(float(w) - offset for w in wf)
        assert abs(as_float - as_double) > 1e-7  # the two readings are far apart here
        assert d == pytest.approx(as_float, abs=1e-10)


def test_hub_own_cluster_total_uses_pbbs_blocks(orc):
    """Known answer: v has 3,000 neighbours in its own cluster, the first edge
    2^60 and the others 1.  The adjacency-order sum of the own cluster's map
    entry stays 2^60 (each +1 is below half an ulp); the own-cluster total T
    is reduced in pbbs blocks of max(1024, 4*ceil(sqrt(3000))) = 1024: blocks
    2^60, 1024, 952, whose left fold rounds to 2^60 + 2048.  With resolution
    0 the own-cluster candidate is S - T = -2048, exactly."""
    deg = 3000
    n = deg + 1
    off = np.zeros(n + 1, np.int64)
    off[1] = deg
    off[2:] = deg + np.arange(1, n)
    nb = np.concatenate([np.arange(1, n), np.zeros(deg)]).astype(np.uint32)
    w = np.ones(deg, np.float32)
    w[0] = np.float32(2.0 ** 60)
    ws = np.concatenate([w, w]).astype(np.float32)
    g = orc.G(off, nb, ws)
    s = orc.S.from_ids(np.zeros(n, np.uint32), np.ones(n))
    t, d = orc.best_move(g, s, orc.config(resolution=0.0), 0)
    assert (t, d) == (0, -2048.0)
    # An adjacency-order T would give 0 here, 4*ceil(sqrt(n)) blocks -3328.


def self_loop_csr(loops):
    """Vertex 0 with `loops` self-loops, vertices 1 and 2 isolated."""
    off = np.array([0, loops, loops, loops], np.int64)
    return np.zeros(loops, np.uint32), np.ones(loops, np.float32), off


def test_hub_without_candidates_returns_max_identity(orc):
    # change = -(0 - res k (K - k)) = -0.5 < 0 with res = -0.5, K = 2, k = 1.
    ids = np.array([1, 1, 2], np.uint32)
    cfg = orc.config(resolution=-0.5)
    nb, w, off = self_loop_csr(1001)
    g = orc.G(off, nb, w)
    s = orc.S.from_ids(ids, np.ones(3))
    cands, change = orc.best_move_candidates(g, s, cfg, 0)
    assert cands == {} and change == -0.5
    assert orc.best_move(g, s, cfg, 0) == (0, 0.0)  # PCCI.cc:574-582
    # The small-degree path has no such identity: (n, change).
    nb, w, off = self_loop_csr(1000)
    g = orc.G(off, nb, w)
    assert orc.best_move(g, s, cfg, 0) == (3, -0.5)


def probe_case(orc, ids, v=4):
    """Vertex v = 4 of 6 shares cluster ids[v] with vertex 0 through one edge
    of weight -1, so its best move (resolution 0) is a new cluster with delta
    1; it is the only frontier vertex of one async round."""
    n = 6
    off = np.zeros(n + 1, np.int64)
    off[1:] = [1, 1, 1, 1, 2, 2]
    nb = np.array([4, 0], np.uint32)
    w = np.array([-1.0, -1.0], np.float32)
    g = orc.G(off, nb, w)
    s = orc.S.from_ids(np.asarray(ids, np.uint32), np.ones(n))
    assert orc.best_move(g, s, orc.config(), v) == (n, 1.0)
    fr = np.zeros(n, np.uint8)
    fr[v] = 1
    orc.round_(g, s, orc.config(async_=True, move_method=2), fr)
    return s


def test_async_new_cluster_probe_starts_at_the_vertex(orc):
    # Empty ids are 1 and 5; the probe starts at v = 4 (occupied by vertex 1)
    # and takes 5, where a smallest-empty-id rule would take 1.
    s = probe_case(orc, [0, 4, 2, 3, 0, 3])
    assert s.ids[4] == 5 and s.sizes[5] == 1 and s.K[5] == 1.0
    assert s.sizes[0] == 1 and s.K[0] == 1.0


def test_async_new_cluster_probe_wraps_around(orc):
    # Only id 1 is empty: 4, 5, then 0 (still holding vertex 0) are skipped.
    s = probe_case(orc, [0, 4, 2, 3, 0, 5])
    assert s.ids[4] == 1 and s.sizes[1] == 1 and s.K[1] == 1.0


def test_async_new_cluster_probe_reuses_own_id(orc):
    # v = 4 sits in cluster 0 and its own id 4 is empty: the probe stops there.
    s = probe_case(orc, [0, 1, 2, 3, 0, 5])
    assert s.ids[4] == 4 and s.sizes[4] == 1 and s.K[4] == 1.0


def small_graph(rng, n, p):
    G = nx.gnp_random_graph(n, p, seed=int(rng.integers(1 << 30)))
    G.remove_nodes_from([v for v in list(G) if G.degree(v) == 0])
    G = nx.convert_node_labels_to_integers(G)
    n = G.number_of_nodes()
    off = np.zeros(n + 1, np.int64)
    nb = []
    for v in range(n):
        vs = sorted(G.neighbors(v))
        off[v + 1] = off[v] + len(vs)
        nb.extend(vs)
    return G, off, np.array(nb, np.uint32), np.ones(len(nb), np.float32)


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("gamma", [1.0, 0.5, 2.0])
def test_modularity_mapping_is_cc_up_to_a_constant(orc, seed, gamma):
    """CC(k = deg, res = gamma / 2m) - m Q_gamma = gamma / (4m) sum deg^2 for
    every clustering (networkx's modularity)."""
    rng = np.random.default_rng(seed)
    G, off, nb, ws = small_graph(rng, 40, 0.12)
    g = orc.G(off, nb, ws)
    m = G.number_of_edges()
    deg = np.diff(off).astype(np.float64)
    const = gamma / (4 * m) * float((deg * deg).sum())
    for _ in range(5):
        ids = rng.integers(0, 6, size=len(deg)).astype(np.uint32)
        cc = orc.objective(g, ids, deg, gamma / deg.sum())
        comms = [set(np.flatnonzero(ids == c).tolist()) for c in np.unique(ids)]
        q = nx.algorithms.community.modularity(G, comms, resolution=gamma)
        assert cc - m * q == pytest.approx(const, abs=1e-9)
        assert orc.modularity(g, ids, gamma) == pytest.approx(q, abs=1e-12)


def exact_q(G, ids, gamma):
    m = G.number_of_edges()
    L, D = {}, {}
    for u, x in G.edges():
        if ids[u] == ids[x]:
            L[ids[u]] = L.get(ids[u], 0) + 1
    for v in G:
        D[ids[v]] = D.get(ids[v], 0) + G.degree(v)
    return sum(Fraction(L.get(c, 0), m) - gamma * Fraction(D[c], 2 * m) ** 2 for c in D)


@pytest.mark.parametrize("seed,adaptive", [(0, False), (1, False), (2, True), (3, True),
                                           (4, True), (5, True)])
def test_modularity_pipeline_round_follows_exact_gains(orc, seed, adaptive):
    """One sync ALL_MOVE round of the oracle's ParallelModularityClusterer
    from singletons equals the moves derived from exact modularity gains: v
    joins the neighbour cluster of largest gain (smaller id on ties) when the
    gain is positive, except that a singleton does not join a singleton of
    smaller id (the singleton rule, PCC.cc:145-156).  The adaptive cases pick
    gamma so that about a third of the gains are positive, which pins the
    resolution's scale too."""
    rng = np.random.default_rng(40 + seed)
    G, off, nb, ws = small_graph(rng, 30, 0.15)
    n = G.number_of_nodes()
    m = G.number_of_edges()
    gamma = Fraction(1)
    if adaptive:
        prods = sorted(G.degree(u) * G.degree(x) for u, x in G.edges())
        gamma = Fraction(2 * m, prods[len(prods) // 3]) + Fraction(1, 7)
    if any(gamma * G.degree(u) * G.degree(x) == 2 * m for u, x in G.edges()):
        pytest.skip("a gain is exactly zero; its float sign is rounding")
    gam = gamma
    single = list(range(n))
    assert float(exact_q(G, single, gam)) == pytest.approx(
        nx.algorithms.community.modularity(G, [{v} for v in range(n)], resolution=float(gamma)),
        abs=1e-12)
    q0 = exact_q(G, single, gam)
    expect = list(range(n))
    signs = set()
    for v in range(n):
        best, best_gain = None, Fraction(0)
        for t in sorted(G.neighbors(v)):
            moved = list(single)
            moved[v] = t
            gain = exact_q(G, moved, gam) - q0
            signs.add(gain > 0)
            if gain > best_gain:
                best, best_gain = t, gain
        if best is not None and best > v:
            expect[v] = best
    cfg = orc.config(resolution=float(gamma), move_method=1, louvain_num_iterations=1,
                     louvain_num_inner_iterations=1)
    r = orc.cluster(orc.G(off, nb, ws), cfg, modularity_objective=True)
    assert list(r.ids) == expect
    if adaptive:
        assert signs == {True, False}
