"""Node weights on which FP64 atomics are not exact, and the config checks of
the one-phase entry points.

RefineClusters accepts arbitrary node weights
(parallel-correlation-clusterer.h:58-91).  With fractional weights (or
integers whose magnitudes sum to 2^53 or more) the order in which cluster
weights are updated changes their last bits, so the library then updates
them in the reference's order (MoveNodesToCluster's per-cluster loops,
parallel-correlation-clusterer-internal.cc:283-337, and SetClustering's,
:68-75) and the sync results must still equal the oracle's bit for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _graphs(pcc):
    return [
        ("sbm", pcc.gen_sbm(3000, 30, 0.08, 0.004, -0.2, 1.0, seed=61)),
        ("rmat", pcc.gen_rmat(13, 16, w_lo=-1.0, w_hi=1.0, seed=62)),
    ]


@pytest.mark.parametrize("kind", ["fractional", "huge"])
@pytest.mark.parametrize("which", [0, 1])
def test_refine_with_inexact_node_weights_bit_exact(pcc, orc, kind, which):
    name, csr = _graphs(pcc)[which]
    rng = np.random.default_rng(70 + which)
    if kind == "fractional":
        nw = rng.uniform(0.1, 3.0, size=csr.n)
        res = 0.03
    else:  # integers, but their sum passes 2^53: partial sums round
        nw = rng.integers(1, 1 << 12, size=csr.n).astype(np.float64) * 2.0 ** 40 + 1.0
        res = 0.03 / 2.0 ** 80
    init = rng.integers(0, 50, size=csr.n).astype(np.uint32)
    print(f"inexact weights {kind} {name} n={csr.n} m={csr.m}", flush=True)
    for mm in (0, 2):
        cfg = pcc.make_config(resolution=res, move_method=mm, refine=1)
        g = pcc.DeviceGraph.from_csr(csr)
        ids, st = g.refine(cfg, init, node_weights=nw)
        r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), node_weights=nw,
                        initial_ids=init)
        assert np.array_equal(ids, r.ids), mm
        assert st.objective == pytest.approx(r.objective, rel=1e-9)


def test_round_with_fractional_weights_state_bit_exact(pcc, orc):
    csr = pcc.gen_rmat(12, 16, w_lo=-1.0, w_hi=1.0, seed=63)
    rng = np.random.default_rng(5)
    nw = rng.uniform(0.25, 2.0, size=csr.n)
    print(f"fractional round n={csr.n}", flush=True)
    g = pcc.DeviceGraph.from_csr(csr)
    og = orc.G.from_csr(csr)
    s = orc.S.from_ids(rng.integers(0, 200, size=csr.n), nw)
    ids, sizes, K = s.ids.copy(), s.sizes.copy(), s.K.copy()
    fr = np.ones(csr.n, np.uint8)
    for mm in (2, 0, 2):
        cfg = pcc.make_config(resolution=0.02, move_method=mm)
        ids, sizes, K, fr_gpu, _ = g.round(cfg, ids, sizes, nw, K, fr)
        fr_cpu, _ = orc.round_(og, s, orc.config_from(cfg), fr)
        assert np.array_equal(ids, s.ids)
        assert np.array_equal(sizes, s.sizes)
        assert np.array_equal(K.view(np.uint64), s.K.view(np.uint64))
        assert np.array_equal(fr_gpu, fr_cpu)
        fr = fr_cpu


def test_cluster_moves_reject_inexact_node_weights(pcc):
    csr = pcc.gen_sbm(500, 5, 0.1, 0.01, seed=64)
    g = pcc.DeviceGraph.from_csr(csr)
    cfg = pcc.make_config(resolution=0.02, clustering_moves_method=pcc.DEFAULT_CLUSTER_MOVES)
    nw = np.full(csr.n, 1.5)
    with pytest.raises(pcc.StatusError) as e:
        g.refine(cfg, np.arange(csr.n, dtype=np.uint32), node_weights=nw)
    assert e.value.code == pcc.INVALID_ARGUMENT
    ids = np.arange(csr.n, dtype=np.uint32)
    with pytest.raises(pcc.StatusError) as e:
        g.round(cfg, ids, np.ones(csr.n, np.uint32), nw, nw.copy(), np.ones(csr.n, np.uint8))
    assert e.value.code == pcc.INVALID_ARGUMENT
    # Integer weights still run.
    out, _ = g.refine(cfg, ids, node_weights=np.full(csr.n, 2.0))
    assert out.shape == (csr.n,)


def test_one_phase_entry_points_check_the_config(pcc):
    csr = pcc.gen_sbm(300, 3, 0.1, 0.01, seed=65)
    g = pcc.DeviceGraph.from_csr(csr)
    n = csr.n
    ids = np.arange(n, dtype=np.uint32)
    ones = np.ones(n)
    args = (ids, np.ones(n, np.uint32), ones, ones.copy())

    def code(fn, **kw):
        with pytest.raises(pcc.StatusError) as e:
            fn(pcc.make_config(resolution=0.02, **kw))
        return e.value.code

    rnd = lambda c: g.round(c, *args, np.ones(n, np.uint8))  # noqa: E731
    bm = lambda c: g.best_moves(c, *args)  # noqa: E731
    # Async subcluster moves inside DEFAULT_CLUSTER_MOVES rounds are not on
    # this path (the reference indexes past its arrays there).
    assert code(rnd, clustering_moves_method=pcc.DEFAULT_CLUSTER_MOVES, subclustering_method=1,
                **{"async": 1}) == pcc.UNIMPLEMENTED
    for fn in (rnd, bm):
        assert code(fn, subclustering_method=7) == pcc.INVALID_ARGUMENT
        assert code(fn, move_method=5) == pcc.INVALID_ARGUMENT
        assert code(fn, clustering_moves_method=4) == pcc.UNIMPLEMENTED
    # TRIANGLE needs ascending rows: reverse every row.
    nb = csr.nbrs.copy()
    for v in range(n):
        s, e = csr.offsets[v], csr.offsets[v + 1]
        nb[s:e] = nb[s:e][::-1]
    rev = pcc.DeviceGraph.from_csr(pcc.Csr(n, csr.offsets, nb, csr.weights))
    with pytest.raises(pcc.StatusError) as e:
        rev.round(pcc.make_config(resolution=0.02, subclustering_method=1,
                                  clustering_moves_method=pcc.DEFAULT_CLUSTER_MOVES),
                  *args, np.ones(n, np.uint8))
    assert e.value.code == pcc.INVALID_ARGUMENT
