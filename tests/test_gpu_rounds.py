"""Device-driven rounds and their HIP-graph replay.

With PCC_OPT_DEVICE_ROUNDS, sync LOUVAIN levels run every best-move round
with the list counts on device: without per-kernel timing one round is
captured as a HIP graph and replayed for each round of the level; with
timing (or option value 2) the same rounds are launched one by one.  Both must give the oracle's ids and round counts
exactly (IterateBestMoves, parallel-correlation-clusterer.cc:340-365), for a
single graph and for batches, on every move method and expansion direction.
Graphs are replayed many times per call, with the library's own eager work
(other levels, contraction, the objective) and synchronizations between
calls; every call is checked.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cases(pcc):
    return [
        ("sbm", pcc.gen_sbm(2000, 20, 0.1, 0.002, seed=91)),
        ("rmat_signed", pcc.gen_rmat(13, 16, w_lo=-1.0, w_hi=1.0, seed=92)),
        ("rmat_hubs", pcc.gen_rmat(15, 16, seed=93)),
    ]


CONFIGS = [
    dict(resolution=0.01, move_method=1, louvain_num_iterations=6, louvain_num_inner_iterations=7),
    dict(resolution=0.05, move_method=2, refine=1),
    dict(resolution=0.05, move_method=0, refine=1),
    dict(resolution=0.02, move_method=2, louvain_num_inner_iterations=3),
]


@pytest.mark.parametrize("which", [0, 1, 2])
@pytest.mark.parametrize("ci", range(len(CONFIGS)))
def test_device_rounds_replay_equals_oracle(pcc, orc, which, ci):
    name, csr = _cases(pcc)[which]
    cfg = pcc.make_config(**CONFIGS[ci])
    print(f"device rounds {name} {CONFIGS[ci]}", flush=True)
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
    g = pcc.DeviceGraph.from_csr(csr)
    # (device rounds, timing): graph replay, device rounds with timers, device
    # rounds launched one by one, host-driven rounds, graph replay again.
    for dev, timing in ((1, 0), (1, 1), (2, 0), (0, 0), (1, 0)):
        g.set_option(pcc.OPT_DEVICE_ROUNDS, dev)
        g.set_option(pcc.OPT_TIMING, timing)
        for _ in range(2):
            ids, st = g.cluster(cfg)
            assert np.array_equal(ids, r.ids), (dev, timing, name)
            assert (st.rounds, st.levels, st.edges_scanned) == (r.rounds, r.levels, r.edges_scanned)
            # Eager work on the same handle and stream between the calls.
            assert g.objective(cfg, ids) == pytest.approx(st.objective, rel=1e-12)
    g.close()


@pytest.mark.parametrize("pull", [0, 1024])
def test_device_rounds_both_expansion_directions(pcc, orc, pull):
    csr = pcc.gen_rmat(14, 16, w_lo=-1.0, w_hi=1.0, seed=94)
    for mm in (0, 2):
        cfg = pcc.make_config(resolution=0.05, move_method=mm, refine=1)
        g = pcc.DeviceGraph.from_csr(csr)
        g.set_option(pcc.OPT_DEVICE_ROUNDS, 1)
        g.set_option(pcc.OPT_PULL_FRACTION, pull)  # 0: pull whenever possible; 1024: push only
        ids, st = g.cluster(cfg)
        r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
        assert np.array_equal(ids, r.ids)
        assert st.rounds == r.rounds
        g.close()


def test_device_rounds_in_batches(pcc, orc):
    graphs = [c for _, c in _cases(pcc)] + [pcc.gen_sbm(300, 3, 0.2, 0.01, seed=95)]
    res = [0.01, 0.05, 0.05, 0.02]
    b = pcc.DeviceBatch.from_csrs(graphs)
    b.set_option(pcc.OPT_DEVICE_ROUNDS, 1)
    for mm in (0, 1, 2):
        cfg = pcc.make_config(move_method=mm, refine=1, louvain_num_inner_iterations=6)
        rounds = []
        for timing in (0, 1):
            b.set_option(pcc.OPT_TIMING, timing)
            ids, obj, _, st = b.cluster_batch(cfg, res)
            most = 0
            for j, gr in enumerate(graphs):
                c1 = pcc.make_config(resolution=res[j], move_method=mm, refine=1,
                                     louvain_num_inner_iterations=6)
                r = orc.cluster(orc.G.from_csr(gr), orc.config_from(c1))
                assert np.array_equal(ids[j], r.ids), (mm, timing, j)
                most = max(most, r.rounds)
            assert st.rounds >= most
            rounds.append(st.rounds)
        assert rounds[0] == rounds[1]
    b.close()
