"""Async mode on the GPU.

The reference's async mode applies moves in place while other threads read
the state (parallel-correlation-clusterer-internal.cc:371-403), so its result
depends on thread timing and processing order (PAPER §4, footnote 3).  Even
the sequential oracle moves by 0.2-0.7% between processing orders on small
RMAT graphs.  The contract is therefore statistical, against the oracle's
async runs over eight seeded processing orders (permute seeds 1..8), and
two-sided:

* on graphs whose order-to-order spread is small (std <= 0.25% of the mean),
  the GPU objective lies within 0.5% of the oracle's mean;
* on the noisier small graphs a single reference run is itself only within
  about 3 standard deviations of its mean, so the band is
  |gpu - mean| <= 0.5% |mean| + 3 std.
The objective the GPU reports is checked separately against the oracle's own
evaluation of the GPU's cluster ids, so an inflated value cannot pass.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ORDERS = range(1, 9)


def _make(pcc, spec):
    kind = spec["kind"]
    if kind == "rmat":
        return pcc.gen_rmat(spec["scale"], spec["ef"], w_lo=spec.get("w_lo", 1.0),
                            w_hi=spec.get("w_hi", 1.0), seed=spec["seed"])
    if kind == "sbm":
        return pcc.gen_sbm(spec["n"], spec["blocks"], 0.1, 0.002, seed=spec["seed"])
    sizes = np.maximum(20, (2000 / (np.arange(1, 60) ** 1.0))).astype(int)
    return pcc.gen_sbm_blocks(list(sizes), 0.2, 0.002, seed=spec["seed"])


STRICT_CASES = [
    # (graph, config kwargs, modularity)
    (dict(kind="rmat", scale=16, ef=16, seed=31),
     dict(resolution=1.0, move_method=0, refine=1, permute=1), True),
    (dict(kind="sbm", n=4000, blocks=40, seed=32),
     dict(resolution=0.01, move_method=2, refine=0), False),
    (dict(kind="sbm_powerlaw", seed=34),
     dict(resolution=0.01, move_method=2, refine=1), False),
    (dict(kind="rmat", scale=15, ef=16, w_lo=-1.0, w_hi=1.0, seed=35),
     dict(resolution=0.05, move_method=2, refine=1), False),
]

RANGE_CASES = [
    (dict(kind="rmat", scale=14, ef=16, seed=31),
     dict(resolution=1.0, move_method=0, refine=1, permute=1), True),
    (dict(kind="rmat", scale=13, ef=8, seed=31),
     dict(resolution=1.0, move_method=0, refine=0), True),
    (dict(kind="rmat", scale=15, ef=16, seed=33),
     dict(resolution=1.0, move_method=2, refine=0), True),
]


def _value(st_or_r, modularity):
    return st_or_r.modularity if modularity else st_or_r.objective


def _run(pcc, orc, spec, kw, modularity, opts=()):
    csr = _make(pcc, spec)
    print(f"async {spec} {kw} opts={list(opts)} n={csr.n} m={csr.m}", flush=True)
    obj = pcc.OBJECTIVE_MODULARITY if modularity else pcc.OBJECTIVE_CORRELATION
    cfg = pcc.make_config(**{**kw, "async": 1})
    dg = pcc.DeviceGraph.from_csr(csr)
    for opt, value in opts:
        dg.set_option(opt, value)
    ids, st = dg.cluster(cfg, obj)
    g = orc.G.from_csr(csr)
    ref = []
    for seed in ORDERS:
        ocfg = pcc.make_config(**{**kw, "async": 1, "permute": 1, "permute_seed": seed})
        ref.append(_value(orc.cluster(g, orc.config_from(ocfg), modularity), modularity))
    ref = np.array(ref)
    gpu = _value(st, modularity)
    if modularity:
        again = orc.modularity(g, ids, kw["resolution"])
    else:
        again = orc.objective(g, ids, np.ones(csr.n), kw["resolution"],
                              kw.get("edge_weight_offset", 0.0))
    assert gpu == pytest.approx(again, rel=1e-9, abs=1e-9)
    print(f"  gpu {gpu:.6f} clusters {st.num_clusters}; oracle mean {ref.mean():.6f} "
          f"std {ref.std() / abs(ref.mean()) * 100:.3f}% range [{ref.min():.6f}, {ref.max():.6f}]",
          flush=True)
    return gpu, ref


def async_band(ref):
    """Half-width of the accepted band around the oracle's mean: 0.5% of the
    mean, plus 3 standard deviations of the order-to-order spread when that
    spread exceeds 0.25% of the mean."""
    mean, std = ref.mean(), ref.std()
    extra = 3 * std if std > 0.0025 * abs(mean) else 0.0
    return 0.005 * abs(mean) + extra


@pytest.mark.parametrize("case", range(len(STRICT_CASES)))
def test_async_within_half_percent_of_oracle_mean(pcc, orc, case):
    spec, kw, modularity = STRICT_CASES[case]
    gpu, ref = _run(pcc, orc, spec, kw, modularity)
    mean = ref.mean()
    assert ref.std() <= 0.0025 * abs(mean)  # the case is a strict one
    assert abs(gpu - mean) <= 0.005 * abs(mean)


@pytest.mark.parametrize("case", range(len(RANGE_CASES)))
def test_async_within_band_of_noisy_oracle_orders(pcc, orc, case):
    spec, kw, modularity = RANGE_CASES[case]
    gpu, ref = _run(pcc, orc, spec, kw, modularity)
    assert abs(gpu - ref.mean()) <= async_band(ref)


def test_async_small_hub_partitions(pcc, orc):
    """Async rounds with every vertex of degree > 64 in the partitioned bin and
    16 edges per partition: the buckets hold 8 * 16 + 64 pairs, so the heavy
    clusters of the hubs (unit-weight modularity gathers a hub's neighbours
    into few clusters) overflow them, and those partitions stream the
    adjacency-order cluster copy of the gather pass instead.  The scale-14
    graph's processing orders spread by about 0.5%, so the band of the noisy
    cases applies."""
    spec, kw, modularity = RANGE_CASES[0]
    opts = ((pcc.OPT_WAVE_MAX_DEG, 64), (pcc.OPT_BLOCK_MAX_DEG, 64), (pcc.OPT_HUB_PART_EDGES, 16))
    gpu, ref = _run(pcc, orc, spec, kw, modularity, opts)
    assert abs(gpu - ref.mean()) <= async_band(ref)
