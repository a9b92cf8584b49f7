"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the seeded generators still produce the stored graphs and the oracle
still produces the stored clusterings.  GPU: the HIP path reproduces the
stored cluster ids exactly."""

import glob
import json
import os

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FILES = sorted(glob.glob(os.path.join(HERE, "*.npz")))


def _load(path):
    z = np.load(path, allow_pickle=False)
    return {k: z[k] for k in z.files}


def _graph(pcc, spec):
    if spec["kind"] == "sbm":
        return pcc.gen_sbm(spec["n"], spec["blocks"], spec["p_in"], spec["p_out"], seed=spec["seed"])
    return pcc.gen_rmat(spec["scale"], spec["ef"], 0.57, 0.19, 0.19, spec["lo"], spec["hi"],
                        seed=spec["seed"])


def test_fixtures_present():
    assert len(FILES) >= 4


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_generators_and_oracle_reproduce_golden(pcc, orc, path):
    z = _load(path)
    csr = _graph(pcc, json.loads(str(z["spec"])))
    assert np.array_equal(csr.offsets, z["offsets"]) and np.array_equal(csr.nbrs, z["nbrs"])
    assert np.array_equal(csr.weights.view(np.uint32), z["weights"].view(np.uint32))
    cfg = pcc.parse_config(str(z["config"]))
    r = orc.cluster(orc.G.from_csr(csr), orc.config_from(cfg), bool(int(z["modularity_objective"])))
    assert np.array_equal(r.ids, z["ids"])
    assert r.objective == float(z["objective"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(f) for f in FILES])
def test_hip_reproduces_golden(pcc, path):
    z = _load(path)
    print(f"golden {os.path.basename(path)}", flush=True)
    csr = pcc.Csr(int(z["offsets"].shape[0] - 1), z["offsets"], z["nbrs"], z["weights"])
    cfg = pcc.parse_config(str(z["config"]))
    obj = pcc.OBJECTIVE_MODULARITY if int(z["modularity_objective"]) else pcc.OBJECTIVE_CORRELATION
    ids, st = pcc.DeviceGraph.from_csr(csr).cluster(cfg, obj)
    assert np.array_equal(ids, z["ids"])
    assert st.objective == pytest.approx(float(z["objective"]), rel=1e-9, abs=1e-6)
