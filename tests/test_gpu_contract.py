"""Contraction parity: the cluster-grouped path (default) against the oracle and
against the all-edge radix-sort path (PCC_OPT_CONTRACT_METHOD 1), bit for bit.

The grouped path sorts a cluster's terms within one wave when its capacity
(the degree sum of its members) is at most 1024, across up to eight waves up
to 8192 and by the radix sort beyond; the cases below put clusters in every
class and on the class boundaries, runs of more than 1024 terms inside one
workgroup's sort (the blocked pbbs sum), clusters whose members are mostly
isolated vertices, and self-loops.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _same(c, og):
    assert c.n == og.n and c.m == og.m
    assert np.array_equal(c.offsets, og.offsets)
    assert np.array_equal(c.nbrs, og.nbrs)
    assert np.array_equal(c.weights.view(np.uint32), og.weights.view(np.uint32))


def _check(pcc, orc, csr, ids, nw, label):
    print(f"contract {label} n={csr.n} m={csr.m}", flush=True)
    og, onw = orc.compress(orc.G.from_csr(csr), ids, nw)
    for method in (0, 1):
        dg = pcc.DeviceGraph.from_csr(csr)
        dg.set_option(pcc.OPT_CONTRACT_METHOD, method)
        cg, gnw = dg.compress(ids, nw)
        _same(cg.to_csr(), og)
        assert np.array_equal(gnw.view(np.uint64), onw.view(np.uint64)), f"method {method}"


def _sym(n, pairs, w):
    """Symmetric CSR from undirected (u, v) pairs (u == v: one self-loop)."""
    u = np.concatenate([pairs[:, 0], pairs[pairs[:, 0] != pairs[:, 1], 1]])
    v = np.concatenate([pairs[:, 1], pairs[pairs[:, 0] != pairs[:, 1], 0]])
    ww = np.concatenate([w, w[pairs[:, 0] != pairs[:, 1]]]).astype(np.float32)
    order = np.lexsort((v, u))
    u, v, ww = u[order], v[order], ww[order]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, u + 1, 1)
    return np.cumsum(off), v.astype(np.uint32), ww


@pytest.mark.parametrize("seed", [1, 2])
def test_power_law_clusters_every_class(pcc, orc, seed):
    """RMAT graph, cluster sizes from 1 to thousands of vertices: capacities
    fall in every class."""
    csr = pcc.gen_rmat(13, 12, w_lo=-1.0, w_hi=1.0, seed=20 + seed)
    rng = np.random.default_rng(seed)
    sizes = np.minimum((rng.pareto(1.2, size=csr.n) + 1).astype(np.int64), 3000)
    ids = np.repeat(np.arange(len(sizes)), sizes)[: csr.n]
    ids = rng.permutation(ids).astype(np.uint32)
    nw = rng.uniform(0.0, 1.0, size=csr.n)
    _check(pcc, orc, csr, ids, nw, f"power-law seed {seed}")


def test_long_runs_inside_one_workgroup(pcc, orc):
    """Complete bipartite blocks between pairs of clusters: 40 x 40 (1600 terms
    in one run, capacity 1600), 64 x 64 (4096), 65 x 64 (4160), 90 x 91 (a run
    of 8190 terms, eight pbbs blocks, in the eight-wave sort) and 130 x 64
    (8320: the radix path), signed weights."""
    rng = np.random.default_rng(5)
    blocks = [(40, 40), (64, 64), (65, 64), (90, 91), (130, 64), (3, 5)]
    pairs, ids, base, cid = [], [], 0, 0
    for a, b in blocks:
        left = np.arange(base, base + a)
        right = np.arange(base + a, base + a + b)
        pairs.append(np.stack(np.meshgrid(left, right, indexing="ij"), -1).reshape(-1, 2))
        ids += [cid] * a + [cid + 1] * b
        base += a + b
        cid += 2
    pairs = np.concatenate(pairs)
    n = base
    off, nbrs, w = _sym(n, pairs, rng.uniform(-1.0, 1.0, size=len(pairs)))
    csr = pcc.Csr(n, off, nbrs, w)
    ids = np.array(ids, np.uint32)
    _check(pcc, orc, csr, ids, np.ones(n), "bipartite blocks")
    # The same graph with the cluster ids scattered (ids need not be members).
    perm = rng.permutation(n).astype(np.uint32)
    _check(pcc, orc, csr, perm[ids], rng.integers(1, 9, size=n).astype(np.float64),
           "bipartite blocks, scattered ids")


def test_isolated_members_and_self_loops(pcc, orc):
    """Clusters of up to 300 members of which only a few have edges (more
    members than lanes), self-loops, and unused ids."""
    rng = np.random.default_rng(9)
    n = 5000
    m_und = 6000
    pairs = rng.integers(0, 600, size=(m_und, 2))  # vertices >= 600 are isolated
    loops = np.repeat(rng.integers(0, 600, size=(50, 1)), 2, axis=1)
    pairs = np.unique(np.concatenate([pairs, loops]), axis=0)
    pairs = pairs[pairs[:, 0] <= pairs[:, 1]]
    off, nbrs, w = _sym(n, pairs, rng.uniform(-1.0, 1.0, size=len(pairs)))
    csr = pcc.Csr(n, off, nbrs, w)
    for k in (17, 40, 300):
        ids = (rng.integers(0, k, size=n) * 3).astype(np.uint32)  # ids 3j: most ids unused
        _check(pcc, orc, csr, ids, rng.uniform(0.5, 2.0, size=n), f"isolated members k={k}")


def test_cluster_runs_equal_between_methods(pcc):
    """Whole Cluster runs (all levels, refinement) give the same ids and
    objective with either contraction."""
    csr = pcc.gen_rmat(14, 16, w_lo=-1.0, w_hi=1.0, seed=31)
    cfg = pcc.make_config(resolution=0.05, move_method=pcc.NBHR_MOVE, refine=1)
    out = []
    for method in (0, 1):
        dg = pcc.DeviceGraph.from_csr(csr)
        dg.set_option(pcc.OPT_CONTRACT_METHOD, method)
        out.append(dg.cluster(cfg))
    (i0, s0), (i1, s1) = out
    assert np.array_equal(i0, i1)
    assert s0.objective == s1.objective and s0.rounds == s1.rounds


def test_radix_path_in_chunks(pcc, orc):
    """Clusters beyond 8192 terms go through the radix sort in chunks of at
    most PCC_OPT_CONTRACT_CHUNK terms (a cluster larger than a chunk alone);
    every chunk size gives the oracle's graph."""
    csr = pcc.gen_rmat(15, 16, w_lo=-1.0, w_hi=1.0, seed=33)
    rng = np.random.default_rng(12)
    ids = rng.integers(0, 40, size=csr.n).astype(np.uint32)  # ~40 clusters of ~13K terms
    nw = rng.integers(1, 4, size=csr.n).astype(np.float64)
    print(f"contract chunks n={csr.n} m={csr.m}", flush=True)
    og, onw = orc.compress(orc.G.from_csr(csr), ids, nw)
    for chunk in (1, 9000, 40000, 1 << 31):
        dg = pcc.DeviceGraph.from_csr(csr)
        dg.set_option(pcc.OPT_CONTRACT_CHUNK, chunk)
        cg, gnw = dg.compress(ids, nw)
        _same(cg.to_csr(), og)
        assert np.array_equal(gnw.view(np.uint64), onw.view(np.uint64)), f"chunk {chunk}"


def test_contract_options_checked(pcc):
    dg = pcc.DeviceGraph.from_csr(pcc.gen_sbm(100, 4, 0.2, 0.01, seed=3))
    for key, bad in ((pcc.OPT_CONTRACT_METHOD, 2), (pcc.OPT_CONTRACT_METHOD, -1),
                     (pcc.OPT_CONTRACT_CHUNK, 0), (pcc.OPT_CONTRACT_CHUNK, (1 << 31) + 1)):
        with pytest.raises(pcc.StatusError):
            dg.set_option(key, bad)
