"""Serving loop: requests submitted while others run are batched as workers
and HBM free up, and each result equals the request's own pcc_cluster run."""

import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(pcc, count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(np.exp(rng.uniform(np.log(100), np.log(20000))))
        if i % 2:
            g = pcc.gen_rmat(max(5, int(np.log2(n))), 8, w_lo=-1.0, w_hi=1.0, seed=seed * 100 + i)
        else:
            g = pcc.gen_sbm(n, max(1, n // 40), 0.3, 0.001, -1.0, 1.0, seed=seed * 100 + i)
        out.append((g, float(rng.uniform(0.001, 0.1))))
    return out


def test_server_results_match_single_runs(pcc):
    cfg = pcc.make_config(move_method=2, refine=1)
    reqs = _requests(pcc, 24, 70)
    print(f"serving {len(reqs)} requests", flush=True)
    srv = pcc.Server(cfg, workers=2, max_batch_graphs=8)
    ids = {}
    # Submit from two threads while workers are already running batches.
    def submit(part):
        for j in part:
            ids[j] = srv.submit(reqs[j][0], reqs[j][1])
            time.sleep(0.002)
    ts = [threading.Thread(target=submit, args=(range(k, len(reqs), 2),)) for k in range(2)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    results = {j: srv.wait(ids[j]) for j in range(len(reqs))}
    st = srv.stats()
    srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    assert st.batches >= 2
    assert st.p99_latency_seconds >= st.p50_latency_seconds > 0
    for j, (g, res) in enumerate(reqs):
        single, sst = pcc.DeviceGraph.from_csr(g).cluster(pcc.make_config(resolution=res, move_method=2,
                                                                          refine=1))
        got, obj, lat = results[j]
        assert np.array_equal(got, single), f"request {j} differs from its single run"
        assert obj == pytest.approx(sst.objective, rel=1e-12, abs=1e-9)


def test_server_with_subclustering_runs_requests_alone(pcc):
    cfg = pcc.make_config(move_method=2, refine=1, subclustering_method=2)
    reqs = _requests(pcc, 6, 71)
    srv = pcc.Server(cfg, workers=2, max_batch_graphs=8)
    ids = [srv.submit(g, res) for g, res in reqs]
    results = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    assert st.mean_batch_graphs == 1.0
    for (g, res), (got, obj, lat) in zip(reqs, results):
        single, _ = pcc.DeviceGraph.from_csr(g).cluster(
            pcc.make_config(resolution=res, move_method=2, refine=1, subclustering_method=2))
        assert np.array_equal(got, single)


def test_server_rejects_request_over_budget(pcc):
    cfg = pcc.make_config(move_method=0)
    srv = pcc.Server(cfg, workers=1, hbm_budget_bytes=4e6)
    small = pcc.gen_sbm(200, 4, 0.2, 0.01, seed=1)
    big = pcc.gen_rmat(14, 16, seed=2)
    a = srv.submit(small, 0.01)
    b = srv.submit(big, 0.01)
    srv.wait(a)
    with pytest.raises(pcc.StatusError) as e:
        srv.wait(b)
    assert e.value.code == pcc.RESOURCE_EXHAUSTED
    srv.close()


def test_server_mixed_sizes_up_to_200k_vertices(pcc):
    """Requests from the serving workload's size range (1,000 up to 200,000
    vertices, SBM and RMAT, mean degree 16, resolution U[0.001, 0.1]),
    submitted in a Poisson stream while batches of other sizes run: every
    result equals the request's own pcc_cluster run."""
    rng = np.random.default_rng(72)
    reqs = []
    for i in range(16):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        if i % 2:
            g = pcc.gen_rmat(max(10, int(round(np.log2(n)))), 8.0, w_lo=0.0, w_hi=1.0, seed=7200 + i)
        else:
            g = pcc.gen_sbm(n, max(1, n // 50), min(1.0, 16 / 50.0 * 0.9), 16 * 0.1 / n, 0.0, 1.0,
                            seed=7200 + i)
        reqs.append((g, float(rng.uniform(0.001, 0.1))))
    print(f"serving mixed sizes, max n={max(g.n for g, _ in reqs)}", flush=True)
    cfg = pcc.make_config(move_method=2, refine=1)
    srv = pcc.Server(cfg, workers=2, max_batch_graphs=16)
    ids = []
    for g, res in reqs:
        ids.append(srv.submit(g, res))
        time.sleep(float(rng.exponential(0.01)))
    results = [srv.wait(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    for (g, res), (got, obj, _) in zip(reqs, results):
        single, sst = pcc.DeviceGraph.from_csr(g).cluster(
            pcc.make_config(resolution=res, move_method=2, refine=1))
        assert np.array_equal(got, single)
        assert obj == pytest.approx(sst.objective, rel=1e-12, abs=1e-9)
