"""Batched form: every graph of a batch gets exactly the result of its own
pcc_cluster run (ids, objective, modularity), whatever else is in the batch."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(pcc, seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(20), np.log(3000))))
        if i % 3 == 2:
            scale = max(4, int(np.log2(n)))
            out.append(pcc.gen_rmat(scale, float(rng.uniform(2, 10)), w_lo=0.0, w_hi=1.0,
                                    seed=seed * 1000 + i))
        else:
            blocks = max(1, n // int(rng.integers(10, 60)))
            out.append(pcc.gen_sbm(n, blocks, float(rng.uniform(0.05, 0.4)), 0.002, 0.0, 1.0,
                                   seed=seed * 1000 + i))
    return out


BATCH_CASES = [
    dict(move_method=0, refine=0),
    dict(move_method=2, refine=1),
    dict(move_method=1, refine=0, louvain_num_iterations=4, louvain_num_inner_iterations=4),
    dict(move_method=1, refine=1, louvain_num_iterations=3),
]


@pytest.mark.parametrize("case", range(len(BATCH_CASES)))
def test_batch_equals_single_runs(pcc, orc, case):
    kw = BATCH_CASES[case]
    graphs = _graphs(pcc, 40 + case, 24)
    print(f"batch case {kw}: {len(graphs)} graphs, {sum(g.m for g in graphs)} edges", flush=True)
    rng = np.random.default_rng(case)
    res = rng.uniform(0.001, 0.1, size=len(graphs))
    cfg = pcc.make_config(**kw)
    b = pcc.DeviceBatch.from_csrs(graphs)
    ids, obj, _, st = b.cluster_batch(cfg, res)
    for i, g in enumerate(graphs):
        c1 = pcc.make_config(resolution=float(res[i]), **kw)
        single, sst = pcc.DeviceGraph.from_csr(g).cluster(c1)
        assert np.array_equal(ids[i], single), f"graph {i} differs from its single run"
        # The per-graph objective sums follow the graph's own chunking, so a
        # graph in a batch gets the same bits as the graph alone.
        assert obj[i] == sst.objective
        r = orc.cluster(orc.G.from_csr(g), orc.config_from(c1))
        assert np.array_equal(ids[i], r.ids), f"graph {i} differs from the oracle"


def test_batch_modularity(pcc, orc):
    graphs = [pcc.gen_rmat(s, 8, seed=50 + s) for s in (6, 9, 7, 10, 5)]
    cfg = pcc.make_config(move_method=0, refine=1)
    gam = np.array([1.0, 0.5, 1.0, 2.0, 1.0])
    b = pcc.DeviceBatch.from_csrs(graphs)
    ids, obj, mod, st = b.cluster_batch(cfg, gam, pcc.OBJECTIVE_MODULARITY)
    for i, g in enumerate(graphs):
        c1 = pcc.make_config(resolution=float(gam[i]), move_method=0, refine=1)
        r = orc.cluster(orc.G.from_csr(g), orc.config_from(c1), modularity_objective=True)
        assert np.array_equal(ids[i], r.ids)
        assert mod[i] == pytest.approx(r.modularity, abs=1e-12)


def test_batch_result_independent_of_mix(pcc):
    graphs = _graphs(pcc, 60, 12)
    cfg = pcc.make_config(move_method=2, refine=1, resolution=0.02)
    full, _, _, _ = pcc.DeviceBatch.from_csrs(graphs).cluster_batch(cfg)
    part, _, _, _ = pcc.DeviceBatch.from_csrs(graphs[3:9][::-1]).cluster_batch(cfg)
    for j, i in enumerate(range(8, 2, -1)):
        assert np.array_equal(part[j], full[i])


def _power_law_graphs(pcc, seed, count):
    """The batched workload's generator at test sizes: n ~ logU[200, 4000],
    mean degree ~ U[8, 64], community sizes from a discrete power law with
    exponent 2 on [8, n], p_out / p_in = 0.01, U[0, 1] weights."""
    rng = np.random.default_rng(seed)
    graphs, res = [], []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(200), np.log(4000))))
        d = rng.uniform(8, 64)
        sizes, left = [], n
        while left > 0:
            u = rng.random()
            s_ = max(1, min(int(8 / (1 - u * (1 - 8 / max(n, 9)))), left))
            sizes.append(s_)
            left -= s_
        sizes = np.array(sizes, np.int64)
        denom = float((sizes * (sizes - 1)).sum() + 0.01 * (sizes * (n - sizes)).sum()) / n
        p_in = min(1.0, d / max(denom, 1e-9))
        graphs.append(pcc.gen_sbm_blocks(sizes, p_in, 0.01 * p_in, 0.0, 1.0, seed=seed * 997 + i))
        res.append(rng.uniform(0.001, 0.1))
    return graphs, np.array(res)


def test_power_law_batch_with_tuned_bins(pcc, orc):
    """Power-law community graphs of the batched workload, with the degree-bin
    bounds and hub partition size its autotuning picked (512, 4096, 1024
    edges): every graph equals its own run and the oracle."""
    graphs, res = _power_law_graphs(pcc, 71, 16)
    print(f"power-law batch: {len(graphs)} graphs, {sum(g.m for g in graphs)} edges", flush=True)
    cfg = pcc.make_config(move_method=2, refine=1)
    b = pcc.DeviceBatch.from_csrs(graphs)
    for key, val in ((pcc.OPT_WAVE_MAX_DEG, 512), (pcc.OPT_BLOCK_MAX_DEG, 4096),
                     (pcc.OPT_HUB_PART_EDGES, 1024)):
        b.set_option(key, val)
    ids, obj, _, st = b.cluster_batch(cfg, res)
    for i, g in enumerate(graphs):
        c1 = pcc.make_config(resolution=float(res[i]), move_method=2, refine=1)
        single, sst = pcc.DeviceGraph.from_csr(g).cluster(c1)
        assert np.array_equal(ids[i], single), f"graph {i} differs from its single run"
        assert obj[i] == sst.objective
        r = orc.cluster(orc.G.from_csr(g), orc.config_from(c1))
        assert np.array_equal(ids[i], r.ids), f"graph {i} differs from the oracle"
