"""HIP path vs the CPU oracle, phase by phase and end to end.

Sync mode must be bit-exact: per-vertex best-move targets and FP64 deltas,
cluster state after a round, the next frontier, the contracted CSR (float
weights compared bit for bit) and the final cluster ids.
"""

import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu


def _graphs(pcc):
    """Seeded inputs covering every degree bin, signed weights and self-loops."""
    out = {
        "sbm_unit": pcc.gen_sbm(600, 12, 0.15, 0.01, seed=11),
        "rmat_signed": pcc.gen_rmat(12, 8, w_lo=-1.0, w_hi=1.0, seed=12),
        "rmat_hubs": pcc.gen_rmat(15, 16, w_lo=-1.0, w_hi=1.0, seed=13),
        "rmat_unit_hubs": pcc.gen_rmat(15, 16, seed=14),
    }
    return out


def _dev(pcc, csr):
    return pcc.DeviceGraph.from_csr(csr)


def _compressed(pcc, orc, csr, rng):
    """A contracted graph (self-loops, aggregated weights) from the oracle."""
    g = orc.G.from_csr(csr)
    ids = rng.integers(0, csr.n // 3 + 1, size=csr.n).astype(np.uint32)
    cg, nw2 = orc.compress(g, ids, np.ones(csr.n))
    return pcc.Csr(cg.n, cg.offsets, cg.nbrs, cg.weights), nw2


@pytest.mark.parametrize("name", ["sbm_unit", "rmat_signed", "rmat_hubs", "rmat_unit_hubs"])
@pytest.mark.parametrize("state_kind", ["singletons", "random"])
def test_best_moves_bit_exact(pcc, orc, name, state_kind):
    csr = _graphs(pcc)[name]
    print(f"best_moves {name} {state_kind} n={csr.n} m={csr.m}", flush=True)
    rng = np.random.default_rng(7)
    n = csr.n
    if state_kind == "singletons":
        ids, sizes, nw, K = random_state(rng, n, n)
        ids = np.arange(n, dtype=np.uint32)
        sizes = np.ones(n, np.uint32)
        K = nw.copy()
    else:
        ids, sizes, nw, K = random_state(rng, n, max(2, n // 5))
    for res, off in [(0.01, 0.0), (0.05, 0.0), (0.3, 0.25)]:
        cfg = pcc.make_config(resolution=res, edge_weight_offset=off)
        ocfg = orc.config_from(cfg)
        dg = _dev(pcc, csr)
        g_moves, g_moved, g_t, g_d = dg.best_moves(cfg, ids, sizes, nw, K)
        o_moves, o_moved, o_t, o_d = orc.sync_decisions(orc.G.from_csr(csr),
                                                        orc.S(ids, sizes, nw, K), ocfg)
        assert np.array_equal(g_t, o_t), f"target mismatch res={res}"
        assert np.array_equal(g_d.view(np.uint64), o_d.view(np.uint64)), f"delta mismatch res={res}"
        assert np.array_equal(g_moves, o_moves)
        assert np.array_equal(g_moved, o_moved)


@pytest.mark.parametrize("name", ["rmat_signed", "rmat_hubs"])
def test_ordered_path_matches(pcc, orc, name):
    """The ordered fallback kernel alone reproduces the oracle (forced)."""
    csr = _graphs(pcc)[name]
    print(f"ordered {name} n={csr.n}", flush=True)
    rng = np.random.default_rng(3)
    ids, sizes, nw, K = random_state(rng, csr.n, max(2, csr.n // 7))
    cfg = pcc.make_config(resolution=0.05)
    dg = _dev(pcc, csr)
    dg.set_option(pcc.OPT_FORCE_ORDERED, 1)
    g = dg.best_moves(cfg, ids, sizes, nw, K)
    o = orc.sync_decisions(orc.G.from_csr(csr), orc.S(ids, sizes, nw, K), orc.config_from(cfg))
    for a, b in zip(g, o):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def test_best_moves_compressed_graph(pcc, orc):
    rng = np.random.default_rng(5)
    csr, nw2 = _compressed(pcc, orc, _graphs(pcc)["rmat_signed"], rng)
    print(f"compressed best_moves n={csr.n} m={csr.m}", flush=True)
    n = csr.n
    ids = np.arange(n, dtype=np.uint32)
    sizes = np.ones(n, np.uint32)
    cfg = pcc.make_config(resolution=0.05)
    g = _dev(pcc, csr).best_moves(cfg, ids, sizes, nw2, nw2.copy())
    o = orc.sync_decisions(orc.G.from_csr(csr), orc.S(ids, sizes, nw2, nw2.copy()),
                           orc.config_from(cfg))
    for a, b in zip(g, o):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.mark.parametrize("pull", [512, 0, 1024])
@pytest.mark.parametrize("move_method", [0, 1, 2])
@pytest.mark.parametrize("name", ["sbm_unit", "rmat_signed", "rmat_hubs"])
def test_round_bit_exact(pcc, orc, name, move_method, pull):
    """pull: the frontier expansion's pull threshold (0: pull whenever any
    vertex is a seed, 1024: push only, 512: the default switch)."""
    csr = _graphs(pcc)[name]
    print(f"round {name} move_method={move_method} pull={pull}", flush=True)
    rng = np.random.default_rng(9)
    n = csr.n
    ids = np.arange(n, dtype=np.uint32)
    sizes = np.ones(n, np.uint32)
    nw = np.ones(n)
    K = nw.copy()
    frontier = np.ones(n, np.uint8)
    cfg = pcc.make_config(resolution=0.05, move_method=move_method)
    ocfg = orc.config_from(cfg)
    dg = _dev(pcc, csr)
    dg.set_option(pcc.OPT_PULL_FRACTION, pull)
    og = orc.G.from_csr(csr)
    ostate = orc.S(ids, sizes, nw, K)
    gstate = (ids, sizes, nw, K)
    ofr = frontier.copy()
    gfr = frontier.copy()
    for r in range(4):
        gi, gs, gk, gfr, gsize = dg.round(cfg, gstate[0], gstate[1], nw, gstate[3], gfr)
        ofr, osize = orc.round_(og, ostate, ocfg, ofr)
        assert np.array_equal(gi, ostate.ids), f"ids differ after round {r}"
        assert np.array_equal(gs, ostate.sizes), f"sizes differ after round {r}"
        assert np.array_equal(gk, ostate.K), f"cluster weights differ after round {r}"
        assert np.array_equal(gfr, ofr), f"frontier differs after round {r}"
        assert gsize == osize
        gstate = (gi, gs, nw, gk)
        if osize == 0:
            break


def _asymmetric(pcc, csr):
    """Drops every third edge u -> v with u < v (by a hash of the pair), so
    some edges have no reverse."""
    deg = np.diff(csr.offsets)
    src = np.repeat(np.arange(csr.n, dtype=np.int64), deg)
    dst = csr.nbrs.astype(np.int64)
    h = ((src * 1000003 + dst) * 0x9E3779B1) & 0xFFFFFFFF
    keep = ~((src < dst) & ((h >> 5) % 3 == 0))
    off = np.zeros(csr.n + 1, np.int64)
    np.cumsum(np.bincount(src[keep], minlength=csr.n), out=off[1:])
    return pcc.Csr(csr.n, off, csr.nbrs[keep].copy(), csr.weights[keep].copy())


@pytest.mark.parametrize("move_method", [1, 2])
def test_asymmetric_graph_expands_by_push(pcc, orc, move_method):
    """On a graph whose edge set is not symmetric, pull and push frontiers
    differ; the pull side must then stay off, whatever the threshold."""
    csr = _asymmetric(pcc, _graphs(pcc)["rmat_signed"])
    print(f"asymmetric round move_method={move_method} n={csr.n} m={csr.m}", flush=True)
    n = csr.n
    ids = np.arange(n, dtype=np.uint32)
    sizes = np.ones(n, np.uint32)
    nw = np.ones(n)
    cfg = pcc.make_config(resolution=0.05, move_method=move_method)
    dg = _dev(pcc, csr)
    dg.set_option(pcc.OPT_PULL_FRACTION, 0)
    og = orc.G.from_csr(csr)
    ostate = orc.S(ids, sizes, nw, nw.copy())
    gstate = (ids, sizes, nw.copy())
    ofr = np.ones(n, np.uint8)
    gfr = ofr.copy()
    for r in range(3):
        gi, gs, gk, gfr, gsize = dg.round(cfg, gstate[0], gstate[1], nw, gstate[2], gfr)
        ofr, osize = orc.round_(og, ostate, orc.config_from(cfg), ofr)
        assert np.array_equal(gi, ostate.ids), f"ids differ after round {r}"
        assert np.array_equal(gfr, ofr), f"frontier differs after round {r}"
        gstate = (gi, gs, gk)
        if osize == 0:
            break


@pytest.mark.parametrize("pull", [0, 1024])
@pytest.mark.parametrize("move_method", [1, 2])
def test_cluster_pull_and_push_agree(pcc, orc, move_method, pull):
    csr = _graphs(pcc)["rmat_unit_hubs"]
    print(f"cluster pull={pull} move_method={move_method}", flush=True)
    cfg = pcc.make_config(resolution=1.0, move_method=move_method, refine=1)
    dg = _dev(pcc, csr)
    dg.set_option(pcc.OPT_PULL_FRACTION, pull)
    ids, st = dg.cluster(cfg, pcc.OBJECTIVE_MODULARITY)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), True)
    assert np.array_equal(ids, r.ids)


def test_compress_long_runs(pcc, orc):
    """Two clusters joined by 4200 x 4200 edges with signed float weights: the
    run of 17.6M terms has more than 1024 level-1 blocks, so the blocked sum
    recurses; order-sensitive weights force ordered folds in every block."""
    a = 4200
    n = 2 * a
    rng = np.random.default_rng(41)
    off = np.arange(n + 1, dtype=np.int64) * a
    left = np.arange(a, dtype=np.uint32)
    nbrs = np.concatenate([np.tile(left + a, a), np.tile(left, a)])
    w = rng.uniform(-1.0, 1.0, size=2 * a * a).astype(np.float32)
    csr = pcc.Csr(n, off, nbrs, w)
    ids = np.repeat(np.array([0, 1], np.uint32), a)
    nw = rng.uniform(0.5, 2.0, size=n)
    print(f"compress long runs n={n} m={csr.m}", flush=True)
    cg, gnw = _dev(pcc, csr).compress(ids, nw)
    c = cg.to_csr()
    og, onw = orc.compress(orc.G.from_csr(csr), ids, nw)
    assert c.n == og.n and c.m == og.m
    assert np.array_equal(c.offsets, og.offsets)
    assert np.array_equal(c.nbrs, og.nbrs)
    assert np.array_equal(c.weights.view(np.uint32), og.weights.view(np.uint32))
    assert np.array_equal(gnw.view(np.uint64), onw.view(np.uint64))


@pytest.mark.parametrize("name", ["sbm_unit", "rmat_signed", "rmat_hubs"])
def test_compress_bit_exact(pcc, orc, name):
    csr = _graphs(pcc)[name]
    print(f"compress {name}", flush=True)
    rng = np.random.default_rng(2)
    for k, fine in ((1, False), (7, False), (csr.n // 10 + 1, False), (csr.n, False), (1, True),
                    (7, True)):
        ids = rng.integers(0, k, size=csr.n).astype(np.uint32)
        # Integer node weights sum exactly in any order; uniform ones do not.
        nw = rng.uniform(0.0, 1.0, size=csr.n) if fine else \
            rng.integers(1, 5, size=csr.n).astype(np.float64)
        dg = _dev(pcc, csr)
        cg, gnw = dg.compress(ids, nw)
        c = cg.to_csr()
        og, onw = orc.compress(orc.G.from_csr(csr), ids, nw)
        assert c.n == og.n and c.m == og.m
        assert np.array_equal(c.offsets, og.offsets)
        assert np.array_equal(c.nbrs, og.nbrs)
        assert np.array_equal(c.weights.view(np.uint32), og.weights.view(np.uint32))
        assert np.array_equal(gnw.view(np.uint64), onw.view(np.uint64))


CLUSTER_CASES = [
    # (graph, config kwargs, modularity)
    ("sbm_unit", dict(resolution=0.01, move_method=1, louvain_num_iterations=10,
                      louvain_num_inner_iterations=10), False),
    ("rmat_signed", dict(resolution=0.05, move_method=2, refine=1), False),
    ("rmat_signed", dict(resolution=0.05, move_method=0, refine=0), False),
    ("rmat_hubs", dict(resolution=0.05, move_method=2, refine=1), False),
    ("rmat_unit_hubs", dict(resolution=1.0, move_method=0, refine=1), True),
    ("sbm_unit", dict(resolution=0.01, move_method=1, refine=1, louvain_num_iterations=1), False),
]


@pytest.mark.parametrize("case", range(len(CLUSTER_CASES)))
def test_cluster_bit_exact(pcc, orc, case):
    name, kw, modularity = CLUSTER_CASES[case]
    csr = _graphs(pcc)[name]
    print(f"cluster {name} {kw} modularity={modularity}", flush=True)
    cfg = pcc.make_config(**kw)
    dg = _dev(pcc, csr)
    ids, st = dg.cluster(cfg, pcc.OBJECTIVE_MODULARITY if modularity else pcc.OBJECTIVE_CORRELATION)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), modularity)
    assert np.array_equal(ids, r.ids)
    assert st.num_clusters == r.num_clusters
    assert st.objective == pytest.approx(r.objective, rel=1e-9, abs=1e-6)
    if modularity:
        assert st.modularity == pytest.approx(r.modularity, rel=1e-9, abs=1e-12)


# ---------------------------------------------------------------------------
# Hub partitions and the order-free test of the hashed bins.

HUB_OPTS = [(64, 64, 16), (64, 128, 64), (512, 4096, 4096)]


def _hub_dev(pcc, csr, opts):
    wave_max, block_max, part = opts
    dg = _dev(pcc, csr)
    dg.set_option(pcc.OPT_WAVE_MAX_DEG, wave_max)
    dg.set_option(pcc.OPT_BLOCK_MAX_DEG, block_max)
    dg.set_option(pcc.OPT_HUB_PART_EDGES, part)
    return dg


@pytest.mark.parametrize("opts", HUB_OPTS)
@pytest.mark.parametrize("name", ["rmat_signed", "rmat_hubs"])
def test_best_moves_hub_partitions(pcc, orc, name, opts):
    """Every vertex above the block bin split into many hash partitions."""
    csr = _graphs(pcc)[name]
    print(f"hub partitions {name} opts={opts}", flush=True)
    rng = np.random.default_rng(17)
    for kclusters in (csr.n, max(2, csr.n // 5), 3):
        ids, sizes, nw, K = random_state(rng, csr.n, kclusters)
        for res, off in [(0.05, 0.0), (0.3, 0.25)]:
            cfg = pcc.make_config(resolution=res, edge_weight_offset=off)
            g = _hub_dev(pcc, csr, opts).best_moves(cfg, ids, sizes, nw, K)
            o = orc.sync_decisions(orc.G.from_csr(csr), orc.S(ids, sizes, nw, K),
                                   orc.config_from(cfg))
            for a, b in zip(g, o):
                assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8)), \
                    f"k={kclusters} res={res}"


def _scaled_weights(csr, every, scale):
    """Scales the weight of every `every`-th undirected edge (by a hash of the
    endpoint pair, so both directions agree) by `scale`."""
    deg = np.diff(csr.offsets)
    src = np.repeat(np.arange(csr.n, dtype=np.int64), deg)
    a = np.minimum(src, csr.nbrs.astype(np.int64))
    b = np.maximum(src, csr.nbrs.astype(np.int64))
    h = ((a * 1000003 + b) * 0x9E3779B1) & 0xFFFFFFFF
    w = csr.weights.copy()
    sel = (h >> 7) % every == 0
    w[sel] = (w[sel].astype(np.float64) * scale).astype(np.float32)
    return sel, w


def _with_weights(pcc, csr, w):
    return pcc.Csr(csr.n, csr.offsets, csr.nbrs, w)


@pytest.mark.parametrize("opts", [(512, 4096, 4096), (64, 64, 16)])
def test_fine_weights_take_ordered_path(pcc, orc, opts):
    """Weights of 2^-40 next to weights near 1 make multi-term cluster sums
    order-sensitive; those vertices must still match the sequential order."""
    base = _graphs(pcc)["rmat_hubs"]
    _, w = _scaled_weights(base, 16, 2.0 ** -40)
    csr = _with_weights(pcc, base, w)
    print(f"fine weights opts={opts} n={csr.n} m={csr.m}", flush=True)
    rng = np.random.default_rng(23)
    ids, sizes, nw, K = random_state(rng, csr.n, max(2, csr.n // 50))
    cfg = pcc.make_config(resolution=0.05)
    g = _hub_dev(pcc, csr, opts).best_moves(cfg, ids, sizes, nw, K)
    o = orc.sync_decisions(orc.G.from_csr(csr), orc.S(ids, sizes, nw, K), orc.config_from(cfg))
    for a, b in zip(g, o):
        assert np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))
    # End to end, with the number of ordered-path vertices reported.
    cfg = pcc.make_config(resolution=0.05, move_method=1, refine=1)
    ids2, st = _hub_dev(pcc, csr, opts).cluster(cfg, pcc.OBJECTIVE_CORRELATION)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), False)
    print(f"  ordered-path vertices {st.fallback_vertices}", flush=True)
    assert np.array_equal(ids2, r.ids)
    assert st.fallback_vertices > 0


def test_dyadic_weights_stay_on_hash_path(pcc, orc):
    """Weights that are multiples of 2^-10 sum exactly in any order: no vertex
    needs the ordered path, and the hashed result is still bit-exact."""
    base = _graphs(pcc)["rmat_hubs"]
    w = (np.round(base.weights.astype(np.float64) * 1024.0) / 1024.0).astype(np.float32)
    csr = _with_weights(pcc, base, w)
    cfg = pcc.make_config(resolution=0.05, move_method=1, refine=1)
    dg = _dev(pcc, csr)
    ids, st = dg.cluster(cfg, pcc.OBJECTIVE_CORRELATION)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), False)
    assert np.array_equal(ids, r.ids)
    assert st.fallback_vertices == 0


HUB_CLUSTER_CASES = [
    ("rmat_signed", dict(resolution=0.05, move_method=1, refine=1), False),
    ("rmat_hubs", dict(resolution=0.05, move_method=2, refine=1), False),
    ("rmat_hubs", dict(resolution=0.05, move_method=0, refine=0), False),
    ("rmat_unit_hubs", dict(resolution=1.0, move_method=2, refine=1), True),
]


@pytest.mark.parametrize("case", range(len(HUB_CLUSTER_CASES)))
def test_cluster_hub_partitions(pcc, orc, case):
    """End-to-end parity with every vertex of degree > 64 in the partitioned
    bin (partition expansion of the frontier included)."""
    name, kw, modularity = HUB_CLUSTER_CASES[case]
    csr = _graphs(pcc)[name]
    print(f"cluster hubs {name} {kw}", flush=True)
    cfg = pcc.make_config(**kw)
    dg = _hub_dev(pcc, csr, (64, 64, 16))
    ids, st = dg.cluster(cfg, pcc.OBJECTIVE_MODULARITY if modularity else pcc.OBJECTIVE_CORRELATION)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), modularity)
    assert np.array_equal(ids, r.ids)


@pytest.mark.parametrize("weights", ["unit", "signed"])
def test_objective_long_rows(pcc, orc, weights):
    """A vertex of degree 40000 (above the split-row threshold) next to a
    ring: the objective and modularity of its long row are summed over many
    workgroups (order-free sum or the ordered fold) and must match the oracle."""
    n = 40001
    rng = np.random.default_rng(43)
    hub = np.zeros(n - 1, np.int64)
    leaves = np.arange(1, n, dtype=np.int64)
    ring_a = leaves
    ring_b = np.roll(leaves, 1)
    src = np.concatenate([hub, leaves, ring_a, ring_b])
    dst = np.concatenate([leaves, hub, ring_b, ring_a])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    off = np.zeros(n + 1, np.int64)
    np.cumsum(np.bincount(src, minlength=n), out=off[1:])
    if weights == "unit":
        w = np.ones(len(src), np.float32)
    else:
        a, b = np.minimum(src, dst), np.maximum(src, dst)
        w = np.sin((a * 7919 + b * 104729) % 100003).astype(np.float32)  # symmetric, signed
    csr = pcc.Csr(n, off, dst.astype(np.uint32), w)
    ids = rng.integers(0, 5, size=n).astype(np.uint32)
    nw = rng.integers(1, 4, size=n).astype(np.float64)
    print(f"objective long rows {weights} n={n} m={csr.m}", flush=True)
    dg = _dev(pcc, csr)
    g = orc.G.from_csr(csr)
    for res, offv in [(0.01, 0.0), (0.3, 0.25)]:
        cfg = pcc.make_config(resolution=res, edge_weight_offset=offv)
        got = dg.objective(cfg, ids, nw)
        ref = orc.objective(g, ids, nw, res, offv)
        assert got == pytest.approx(ref, rel=1e-12, abs=1e-9)
    for gamma in (0.5, 1.0):
        assert dg.modularity(gamma, ids) == pytest.approx(orc.modularity(g, ids, gamma), rel=1e-12,
                                                         abs=1e-12)


def _hub_graph(pcc, fine_weight):
    """An SBM with three vertices joined to 20,000 others each (rows longer
    than the 16384 entries of the per-wave bound pass); unit weights, except
    one hub edge of weight fine_weight (both directions) when given."""
    base = pcc.gen_sbm(60000, 600, 0.08, 0.0002, seed=51)
    rng = np.random.default_rng(52)
    n = base.n
    src = np.repeat(np.arange(n, dtype=np.int64), np.diff(base.offsets))
    us, vs, ws = [src], [base.nbrs.astype(np.int64)], [base.weights]
    for h in (7, 20011, 41017):
        nb = rng.choice(n, size=20000, replace=False)
        nb = nb[nb != h]
        us += [np.full(nb.shape[0], h), nb]
        vs += [nb, np.full(nb.shape[0], h)]
        w = np.ones(nb.shape[0], np.float32)
        if fine_weight is not None and h == 20011:
            w[5] = np.float32(fine_weight)
        ws += [w, w]
    u = np.concatenate(us)
    v = np.concatenate(vs)
    w = np.concatenate(ws)
    key = u * n + v
    order = np.argsort(key, kind="stable")
    key, u, v, w = key[order], u[order], v[order], w[order]
    keep = np.concatenate([[True], key[1:] != key[:-1]])  # drop duplicate pairs
    u, v, w = u[keep], v[keep], w[keep]
    off = np.zeros(n + 1, np.int64)
    np.add.at(off, u + 1, 1)
    return pcc.Csr(n, np.cumsum(off), v.astype(np.uint32), w.astype(np.float32))


@pytest.mark.parametrize("fine", [None, 2.0 ** -40])
def test_long_rows_coarse_and_fine(pcc, orc, fine):
    """Rows longer than 16384 entries take the chunked bound pass; with unit
    weights they are tagged coarse (no per-slot records), with one 2^-40
    weight the hub's slots are tracked again.  Both bit-exact."""
    csr = _hub_graph(pcc, fine)
    assert int(np.diff(csr.offsets).max()) > 16384
    print(f"long rows fine={fine} n={csr.n} m={csr.m}", flush=True)
    cfg = pcc.make_config(resolution=0.02, move_method=2, refine=1)
    ids, st = _dev(pcc, csr).cluster(cfg)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
    assert np.array_equal(ids, r.ids)
    assert st.objective == pytest.approx(r.objective, rel=1e-9, abs=1e-6)
