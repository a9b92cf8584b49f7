"""Parity at the full sizes of the measured workloads.

* The 10,000-vertex planted-partition SBM (100 blocks, p_in 0.1, p_out
  0.001, sync ALL_MOVE, resolution 0.01, 10 x 10 iterations): the HIP ids
  equal the oracle's, which runs here in well under a second.
* The RMAT scale-22 graph (4,194,304 vertices, 128,312,844 directed edges,
  U[-1, 1] weights, sync NBHR_MOVE, resolution 0.05, refinement): the oracle
  needs minutes for it, so tests/golden/fullsize_rmat22_nbhr_refine.json
  (written by tests/golden/make_fullsize.py) holds the sha256 of the
  oracle's ids; the HIP ids must hash to the same digest, after the
  regenerated graph hashes to the recorded CSR digest.
* The RMAT scale-24, edge-factor-32 async modularity workload (about 10^9
  directed edges): async results are not reproducible run to run, so the
  check is a property: the modularity and objective the library reports
  equal the oracle's evaluation of the returned ids, and the ids are a valid
  clustering.
"""

import hashlib
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fixture(name):
    with open(os.path.join(GOLDEN, f"fullsize_{name}.json")) as f:
        return json.load(f)


def _csr_digest(csr):
    h = hashlib.sha256()
    for a in (np.ascontiguousarray(csr.offsets, "<i8"), np.ascontiguousarray(csr.nbrs, "<u4"),
              np.ascontiguousarray(csr.weights, np.float32).view("<u4")):
        h.update(a.tobytes())
    return h.hexdigest()


def _ids_digest(ids):
    return hashlib.sha256(np.ascontiguousarray(ids, "<u4").tobytes()).hexdigest()


def _graph(pcc, spec):
    if spec["kind"] == "sbm":
        return pcc.gen_sbm(spec["n"], spec["blocks"], spec["p_in"], spec["p_out"], seed=spec["seed"])
    return pcc.gen_rmat(spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["lo"], spec["hi"],
                        seed=spec["seed"])


def test_sbm10000_allmove_bit_exact(pcc, orc):
    fx = _fixture("sbm10000_allmove")
    csr = _graph(pcc, fx["spec"])
    print(f"fullsize sbm10000 n={csr.n} m={csr.m}", flush=True)
    assert (csr.n, csr.m) == (fx["n"], fx["m"]) and _csr_digest(csr) == fx["csr_sha256"]
    cfg = pcc.parse_config(fx["config"])
    ids, st = pcc.DeviceGraph.from_csr(csr).cluster(cfg)
    ref = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg))
    assert np.array_equal(ids, ref.ids)
    assert _ids_digest(ids) == fx["ids_sha256"]
    assert st.num_clusters == fx["num_clusters"] and st.rounds == fx["rounds"]
    assert st.objective == pytest.approx(fx["objective"], rel=1e-12)


@pytest.mark.timeout(600)
def test_rmat22_nbhr_refine_matches_oracle_digest(pcc):
    fx = _fixture("rmat22_nbhr_refine")
    csr = _graph(pcc, fx["spec"])
    print(f"fullsize rmat22 n={csr.n} m={csr.m}", flush=True)
    assert (csr.n, csr.m) == (fx["n"], fx["m"]) and _csr_digest(csr) == fx["csr_sha256"]
    cfg = pcc.parse_config(fx["config"])
    ids, st = pcc.DeviceGraph.from_csr(csr).cluster(cfg)
    assert _ids_digest(ids) == fx["ids_sha256"]
    assert st.num_clusters == fx["num_clusters"]
    assert (st.rounds, st.levels, st.edges_scanned) == (fx["rounds"], fx["levels"],
                                                         fx["edges_scanned"])
    assert st.objective == pytest.approx(fx["objective"], rel=1e-12)


@pytest.mark.timeout(900)
def test_rmat24_async_modularity_reports_its_ids(pcc, orc):
    csr = pcc.gen_rmat(24, 32, 0.57, 0.19, 0.19, 1.0, 1.0, seed=3)
    print(f"fullsize rmat24 n={csr.n} m={csr.m}", flush=True)
    cfg = pcc.make_config(resolution=1.0, move_method=pcc.NBHR_CLUSTER_MOVE, refine=1, permute=1,
                          **{"async": 1})
    g = pcc.DeviceGraph.from_csr(csr)
    ids, st = g.cluster(cfg, pcc.OBJECTIVE_MODULARITY)
    g.close()
    assert ids.shape == (csr.n,) and int(ids.max()) < csr.n
    og = orc.G.from_csr(csr)
    q = orc.modularity(og, ids, 1.0)
    print(f"  modularity gpu {st.modularity:.6f} oracle {q:.6f} clusters {st.num_clusters}",
          flush=True)
    assert st.modularity == pytest.approx(q, rel=1e-9, abs=1e-12)
    deg = np.diff(csr.offsets).astype(np.float64)
    obj = orc.objective(og, ids, deg, 1.0 / float(csr.m))
    assert st.objective == pytest.approx(obj, rel=1e-9)
    assert st.num_clusters == len(np.unique(ids))
    # The same workload's quality: modularity 0.0481-0.0488 has been measured
    # on it; a clustering that merges or scatters everything is far outside.
    assert 0.03 < st.modularity < 0.1
