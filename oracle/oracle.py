"""ctypes loader of the CPU oracle (oracle/build/libpcc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product library.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# PCC_ORACLE_LIB loads another build of the same source (a build with a
# planted mistake, to check that the oracle's tests catch it).
LIB_PATH = os.environ.get("PCC_ORACLE_LIB") or os.path.join(_HERE, "build", "libpcc_oracle.so")
NO_MOVE = 0xFFFFFFFF


class OConfig(ctypes.Structure):
    _fields_ = [
        ("resolution", ctypes.c_double),
        ("edge_weight_offset", ctypes.c_double),
        ("num_iterations", ctypes.c_int64),
        ("louvain_num_iterations", ctypes.c_int64),
        ("louvain_num_inner_iterations", ctypes.c_int64),
        ("clustering_moves_method", ctypes.c_int32),
        ("move_method", ctypes.c_int32),
        ("refine", ctypes.c_int32),
        ("async_", ctypes.c_int32),
        ("permute", ctypes.c_int32),
        ("permute_seed", ctypes.c_uint64),
        ("sequential", ctypes.c_int32),
        ("all_iter", ctypes.c_int32),
        ("subclustering_method", ctypes.c_int32),
    ]


class OGraph(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("m", ctypes.c_int64),
        ("offsets", ctypes.c_void_p),
        ("nbrs", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
    ]


class OState(ctypes.Structure):
    _fields_ = [
        ("cluster_ids", ctypes.c_void_p),
        ("cluster_sizes", ctypes.c_void_p),
        ("node_weights", ctypes.c_void_p),
        ("cluster_weights", ctypes.c_void_p),
    ]


class OResult(ctypes.Structure):
    _fields_ = [
        ("objective", ctypes.c_double),
        ("modularity", ctypes.c_double),
        ("levels", ctypes.c_int64),
        ("rounds", ctypes.c_int64),
        ("edges_scanned", ctypes.c_int64),
        ("num_clusters", ctypes.c_int64),
        ("seconds", ctypes.c_double),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        l.oracle_set_threads.argtypes = [ctypes.c_int]
        l.oracle_get_threads.restype = ctypes.c_int
        l.oracle_phase_seconds.argtypes = [P, ctypes.c_int]
        l.oracle_best_move.argtypes = [P, P, P, ctypes.c_uint32, P, P]
        l.oracle_best_move_candidates.argtypes = [P, P, P, ctypes.c_uint32, P, P, ctypes.c_int64, P]
        l.oracle_best_move_candidates.restype = ctypes.c_int64
        l.oracle_best_move_set.argtypes = [P, P, P, P, ctypes.c_int64, P, P]
        l.oracle_sync_decisions.argtypes = [P, P, P, P, P, P, P, P]
        l.oracle_apply_moves.argtypes = [ctypes.c_int64, P, P, P]
        l.oracle_round.argtypes = [P, P, P, P, ctypes.c_uint64]
        l.oracle_round.restype = ctypes.c_int64
        l.oracle_iterate_best_moves.argtypes = [P, P, P, ctypes.c_int64, ctypes.c_uint64, P]
        l.oracle_iterate_best_moves.restype = ctypes.c_int
        l.oracle_compress.argtypes = [P, P, P]
        l.oracle_compress.restype = P
        l.oracle_csr_sizes.argtypes = [P, P, P]
        l.oracle_csr_copy.argtypes = [P, P, P, P, P]
        l.oracle_csr_free.argtypes = [P]
        l.oracle_objective.argtypes = [P, P, P, ctypes.c_double, ctypes.c_double]
        l.oracle_objective.restype = ctypes.c_double
        l.oracle_modularity.argtypes = [P, P, ctypes.c_double]
        l.oracle_modularity.restype = ctypes.c_double
        l.oracle_cluster.argtypes = [P, P, ctypes.c_int, P, P, P, P]
        l.oracle_cluster.restype = ctypes.c_int
        l.oracle_flatten.argtypes = [ctypes.c_int64, P, P, P]
        l.oracle_subclusters.argtypes = [P, P, ctypes.c_int, P, P, ctypes.c_double,
                                         ctypes.c_double]
        l.oracle_subclusters.restype = ctypes.c_int64
        l.oracle_triangle_score.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32]
        l.oracle_triangle_score.restype = ctypes.c_double
        l.oracle_rank_relabel.argtypes = [ctypes.c_int64, P, P]
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def set_threads(t: int) -> None:
    lib().oracle_set_threads(int(t))


PHASES = ("best_move", "apply", "frontier", "contraction", "refine_relabel")


def phase_seconds(reset: bool = True) -> dict:
    """Wall seconds per phase of the oracle's rounds and pipeline runs."""
    out = np.zeros(5)
    lib().oracle_phase_seconds(_p(out), int(reset))
    return dict(zip(PHASES, out.tolist()))


class G:
    """Keeps the numpy arrays alive behind an OGraph view."""

    def __init__(self, offsets, nbrs, weights):
        self.offsets = np.ascontiguousarray(offsets, dtype=np.int64)
        self.nbrs = np.ascontiguousarray(nbrs, dtype=np.uint32)
        self.weights = np.ascontiguousarray(weights, dtype=np.float32)
        self.n = int(self.offsets.shape[0] - 1)
        self.m = int(self.nbrs.shape[0])
        self.view = OGraph(self.n, self.m, _p(self.offsets), _p(self.nbrs), _p(self.weights))

    @classmethod
    def from_csr(cls, csr) -> "G":
        return cls(csr.offsets, csr.nbrs, csr.weights)


class S:
    """ClusteringHelper state arrays."""

    def __init__(self, ids, sizes, nw, K):
        self.ids = np.ascontiguousarray(ids, dtype=np.uint32).copy()
        self.sizes = np.ascontiguousarray(sizes, dtype=np.uint32).copy()
        self.nw = np.ascontiguousarray(nw, dtype=np.float64).copy()
        self.K = np.ascontiguousarray(K, dtype=np.float64).copy()
        self.view = OState(_p(self.ids), _p(self.sizes), _p(self.nw), _p(self.K))

    @classmethod
    def singletons(cls, n: int, nw=None) -> "S":
        nw = np.ones(n) if nw is None else np.asarray(nw, np.float64)
        return cls(np.arange(n), np.ones(n), nw, nw.copy())

    @classmethod
    def from_ids(cls, ids, nw) -> "S":
        ids = np.asarray(ids, np.int64)
        n = ids.shape[0]
        nw = np.asarray(nw, np.float64)
        sizes = np.bincount(ids, minlength=n)[:n]
        K = np.zeros(n)
        np.add.at(K, ids, nw)
        return cls(ids, sizes, nw, K)


def config(resolution=0.0, edge_weight_offset=0.0, louvain_num_iterations=0,
           louvain_num_inner_iterations=0, clustering_moves_method=1, move_method=0, refine=False,
           async_=False, permute=False, permute_seed=0x5EED, num_iterations=0,
           sequential=False, all_iter=False, subclustering_method=0) -> OConfig:
    return OConfig(resolution, edge_weight_offset, num_iterations, louvain_num_iterations,
                   louvain_num_inner_iterations, clustering_moves_method, move_method, int(refine),
                   int(async_), int(permute), permute_seed, int(sequential), int(all_iter),
                   subclustering_method)


def config_from(pcfg) -> OConfig:
    """From the product's Config struct (field for field)."""
    return config(pcfg.resolution, pcfg.edge_weight_offset, pcfg.louvain_num_iterations,
                  pcfg.louvain_num_inner_iterations, pcfg.clustering_moves_method,
                  pcfg.move_method, bool(pcfg.refine), bool(pcfg.async_), bool(pcfg.permute),
                  pcfg.permute_seed, pcfg.num_iterations,
                  subclustering_method=pcfg.subclustering_method)


def best_move(g: G, s: S, cfg: OConfig, v: int) -> Tuple[int, float]:
    t = ctypes.c_uint32()
    d = ctypes.c_double()
    lib().oracle_best_move(ctypes.byref(g.view), ctypes.byref(s.view), ctypes.byref(cfg), v,
                           ctypes.byref(t), ctypes.byref(d))
    return int(t.value), float(d.value)


def best_move_candidates(g: G, s: S, cfg: OConfig, v: int):
    """Every candidate EfficientBestMove evaluates for v: ({cluster: delta},
    change), change being the "new cluster" value."""
    cap = int(g.offsets[v + 1] - g.offsets[v]) + 1
    cl = np.zeros(cap, np.uint32)
    dl = np.zeros(cap, np.float64)
    ch = ctypes.c_double()
    cnt = lib().oracle_best_move_candidates(ctypes.byref(g.view), ctypes.byref(s.view),
                                            ctypes.byref(cfg), v, _p(cl), _p(dl), cap,
                                            ctypes.byref(ch))
    assert cnt <= cap
    return {int(c): float(d) for c, d in zip(cl[:cnt], dl[:cnt])}, float(ch.value)


def best_move_set(g: G, s: S, cfg: OConfig, members) -> Tuple[int, float]:
    mem = np.ascontiguousarray(members, dtype=np.uint32)
    t = ctypes.c_uint32()
    d = ctypes.c_double()
    lib().oracle_best_move_set(ctypes.byref(g.view), ctypes.byref(s.view), ctypes.byref(cfg),
                               _p(mem), mem.shape[0], ctypes.byref(t), ctypes.byref(d))
    return int(t.value), float(d.value)


def sync_decisions(g: G, s: S, cfg: OConfig, frontier=None):
    n = g.n
    moves = np.zeros(n, np.uint32)
    moved = np.zeros(n, np.uint8)
    tgt = np.zeros(n, np.uint32)
    dlt = np.zeros(n, np.float64)
    fr = None if frontier is None else np.ascontiguousarray(frontier, dtype=np.uint8)
    lib().oracle_sync_decisions(ctypes.byref(g.view), ctypes.byref(s.view), ctypes.byref(cfg),
                                _p(fr), _p(moves), _p(moved), _p(tgt), _p(dlt))
    return moves, moved, tgt, dlt


def apply_moves(s: S, moves) -> np.ndarray:
    moves = np.ascontiguousarray(moves, dtype=np.uint32)
    n = moves.shape[0]
    mod = np.zeros(n, np.uint8)
    lib().oracle_apply_moves(n, ctypes.byref(s.view), _p(moves), _p(mod))
    return mod


def round_(g: G, s: S, cfg: OConfig, frontier, seed: int = 0):
    fr = np.ascontiguousarray(frontier, dtype=np.uint8).copy()
    size = lib().oracle_round(ctypes.byref(g.view), ctypes.byref(s.view), ctypes.byref(cfg),
                              _p(fr), seed)
    return fr, int(size)


def iterate_best_moves(g: G, s: S, cfg: OConfig, num_inner: int, seed: int = 0) -> int:
    """IterateBestMoves on one level; 1 if anything moved."""
    rounds = ctypes.c_int64(0)
    return int(lib().oracle_iterate_best_moves(ctypes.byref(g.view), ctypes.byref(s.view),
                                               ctypes.byref(cfg), num_inner, seed,
                                               ctypes.byref(rounds)))


def compress(g: G, ids, nw):
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    nw = np.ascontiguousarray(nw, dtype=np.float64)
    h = lib().oracle_compress(ctypes.byref(g.view), _p(ids), _p(nw))
    n = ctypes.c_int64()
    m = ctypes.c_int64()
    lib().oracle_csr_sizes(h, ctypes.byref(n), ctypes.byref(m))
    off = np.zeros(n.value + 1, np.int64)
    nb = np.zeros(max(m.value, 1), np.uint32)
    w = np.zeros(max(m.value, 1), np.float32)
    nw2 = np.zeros(max(n.value, 1), np.float64)
    lib().oracle_csr_copy(h, _p(off), _p(nb), _p(w), _p(nw2))
    lib().oracle_csr_free(h)
    return G(off, nb[: m.value], w[: m.value]), nw2[: n.value]


def objective(g: G, ids, nw, resolution, offset=0.0) -> float:
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    nw = np.ascontiguousarray(nw, dtype=np.float64)
    return float(lib().oracle_objective(ctypes.byref(g.view), _p(ids), _p(nw), resolution, offset))


def modularity(g: G, ids, resolution) -> float:
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    return float(lib().oracle_modularity(ctypes.byref(g.view), _p(ids), resolution))


@dataclass
class Result:
    ids: np.ndarray
    objective: float
    modularity: float
    levels: int
    rounds: int
    edges_scanned: int
    num_clusters: int
    seconds: float


def cluster(g: G, cfg: OConfig, modularity_objective: bool = False, node_weights=None,
            initial_ids=None) -> Result:
    out = np.zeros(max(g.n, 1), np.uint32)
    r = OResult()
    nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
    init = None if initial_ids is None else np.ascontiguousarray(initial_ids, dtype=np.uint32)
    rc = lib().oracle_cluster(ctypes.byref(g.view), ctypes.byref(cfg), int(modularity_objective),
                              _p(nw), _p(init), _p(out), ctypes.byref(r))
    if rc != 0:
        raise RuntimeError(f"oracle_cluster returned {rc}")
    return Result(out[: g.n], r.objective, r.modularity, r.levels, r.rounds, r.edges_scanned,
                  r.num_clusters, r.seconds)


def subclusters(g: G, ids, method: int, offset: float = 0.0,
                resolution: float = 0.0) -> Tuple[np.ndarray, np.ndarray, int]:
    """CompressSubclusters' numbering: (sub ids, parent rank per id, total)."""
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    sub = np.zeros(g.n, np.uint32)
    parent = np.zeros(g.n, np.uint32)
    total = lib().oracle_subclusters(ctypes.byref(g.view), _p(ids), method, _p(sub), _p(parent),
                                     offset, resolution)
    if total < 0:
        raise ValueError(f"subclustering method {method} is not restated")
    return sub, parent, int(total)


def triangle_score(g: G, u: int, v: int) -> float:
    return float(lib().oracle_triangle_score(ctypes.byref(g.view), u, v))


def flatten(ids, compressed) -> np.ndarray:
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    compressed = np.ascontiguousarray(compressed, dtype=np.uint32)
    out = np.zeros_like(ids)
    lib().oracle_flatten(ids.shape[0], _p(ids), _p(compressed), _p(out))
    return out


def rank_relabel(ids) -> np.ndarray:
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    out = np.zeros_like(ids)
    lib().oracle_rank_relabel(ids.shape[0], _p(ids), _p(out))
    return out
