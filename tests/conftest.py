"""Shared pytest set-up: the `gpu` marker, import paths and helpers.

`-m "not gpu"` tests run on the CPU only (oracle checks, host logic, the C ABI
surface).  `-m gpu` tests call the HIP path through the C ABI and compare it
with the oracle.
"""

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "parallel-correlation-clustering_amd")
for p in (PKG, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


@pytest.fixture(scope="session")
def pcc():
    import pcc_amd
    pcc_amd.lib()
    return pcc_amd


@pytest.fixture(scope="session")
def orc():
    import oracle
    oracle.lib()
    return oracle


def random_state(rng, n, num_clusters, nw=None):
    """A random ClusteringHelper state with consistent sizes / weights."""
    ids = rng.integers(0, max(1, num_clusters), size=n).astype(np.uint32)
    nw = np.ones(n) if nw is None else np.asarray(nw, np.float64)
    sizes = np.bincount(ids, minlength=n)[:n].astype(np.uint32)
    K = np.zeros(n)
    np.add.at(K, ids, nw)
    return ids, sizes, nw, K
