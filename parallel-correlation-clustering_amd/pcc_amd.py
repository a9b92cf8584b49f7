"""Python binding of libpcc_amd.so, the MI355X ParallelCorrelationClusterer /
ParallelModularityClusterer hot path.

The classes at the bottom mirror the reference's ``InMemoryClusterer`` API
(clustering/in-memory-clusterer.h:34-121): ``MutableGraph().Import(...)``,
``FinishImport()``, ``Cluster(config)`` returning a ``Clustering`` (list of
clusters, ascending ids, members ascending, as OutputIndicesById produces it,
parallel/parallel-sequence-ops.h:183-209) and ``RefineClusters``.  Errors
surface as ``StatusError`` carrying the absl status code.

The shared library must have been built (``make`` in this directory); there is
no fallback implementation.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcc_amd.so")

OK = 0
INVALID_ARGUMENT = 3
NOT_FOUND = 5
RESOURCE_EXHAUSTED = 8
UNIMPLEMENTED = 12
INTERNAL = 13

NO_MOVE = 0xFFFFFFFF

DEFAULT_CLUSTER_MOVES = 0
LOUVAIN = 1
NBHR_CLUSTER_MOVE = 0
ALL_MOVE = 1
NBHR_MOVE = 2

OBJECTIVE_CORRELATION = 0
OBJECTIVE_MODULARITY = 1

OPT_TIMING = 1
OPT_FORCE_ORDERED = 2
OPT_WAVE_MAX_DEG = 3
OPT_BLOCK_MAX_DEG = 4
OPT_ASYNC_SUBROUNDS = 5
OPT_ASYNC_SINGLETON_RULE = 6
OPT_HUB_PART_EDGES = 7
OPT_PULL_FRACTION = 8
OPT_CLUSTER_SUBROUNDS = 9
OPT_HASH_WAVES = 10
OPT_DEVICE_ROUNDS = 11
OPT_CONTRACT_METHOD = 12
OPT_CONTRACT_CHUNK = 13


class StatusError(RuntimeError):
    """A non-OK absl-style status returned by the library."""

    def __init__(self, code: int, message: str):
        super().__init__(f"status {code}: {message}")
        self.code = code
        self.message = message


class Config(ctypes.Structure):
    """CorrelationClustererConfig (clustering/config.proto:94-183)."""

    _fields_ = [
        ("resolution", ctypes.c_double),
        ("edge_weight_offset", ctypes.c_double),
        ("random_seed", ctypes.c_uint32),
        ("num_iterations", ctypes.c_int64),
        ("louvain_num_iterations", ctypes.c_int64),
        ("louvain_num_inner_iterations", ctypes.c_int64),
        ("clustering_moves_method", ctypes.c_int32),
        ("subclustering_method", ctypes.c_int32),
        ("preclustering_method", ctypes.c_int32),
        ("move_method", ctypes.c_int32),
        ("refine", ctypes.c_uint8),
        ("async_", ctypes.c_uint8),
        ("connect_stats", ctypes.c_uint8),
        ("all_iter", ctypes.c_uint8),
        ("permute", ctypes.c_uint8),
        ("_pad", ctypes.c_uint8 * 3),
        ("permute_seed", ctypes.c_uint64),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("levels", ctypes.c_int64),
        ("rounds", ctypes.c_int64),
        ("edges_scanned", ctypes.c_int64),
        ("frontier_vertices", ctypes.c_int64),
        ("num_clusters", ctypes.c_int64),
        ("fallback_vertices", ctypes.c_int64),
        ("objective", ctypes.c_double),
        ("modularity", ctypes.c_double),
        ("cluster_seconds", ctypes.c_double),
        ("bestmove_seconds", ctypes.c_double),
        ("bestmove_bytes", ctypes.c_double),
        ("contract_seconds", ctypes.c_double),
        ("apply_seconds", ctypes.c_double),
        ("bestmove_launches", ctypes.c_int64),
        ("ordered_seconds", ctypes.c_double),
        ("bin_seconds", ctypes.c_double * 8),
        ("bin_edges", ctypes.c_int64 * 8),
        ("bin_vertices", ctypes.c_int64 * 8),
        ("bin_launches", ctypes.c_int64 * 8),
        ("cluster_move_seconds", ctypes.c_double),
        ("cluster_moves", ctypes.c_int64),
    ]

    def as_dict(self) -> dict:
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class TuneResult(ctypes.Structure):
    _fields_ = [
        ("wave_max_deg", ctypes.c_int64),
        ("block_max_deg", ctypes.c_int64),
        ("hub_part_edges", ctypes.c_int64),
        ("trials", ctypes.c_int64),
        ("initial_seconds", ctypes.c_double),
        ("best_seconds", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerOptions(ctypes.Structure):
    _fields_ = [
        ("workers", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
        ("max_batch_graphs", ctypes.c_int64),
        ("max_batch_edges", ctypes.c_int64),
        ("hbm_budget_bytes", ctypes.c_double),
        ("batch_window_seconds", ctypes.c_double),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("completed", ctypes.c_int64),
        ("failed", ctypes.c_int64),
        ("batches", ctypes.c_int64),
        ("max_graphs_in_flight", ctypes.c_int64),
        ("p50_latency_seconds", ctypes.c_double),
        ("p99_latency_seconds", ctypes.c_double),
        ("mean_batch_graphs", ctypes.c_double),
        ("graphs_per_second", ctypes.c_double),
    ]


class HostCsr(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("m", ctypes.c_int64),
        ("offsets", ctypes.POINTER(ctypes.c_int64)),
        ("nbrs", ctypes.POINTER(ctypes.c_uint32)),
        ("weights", ctypes.POINTER(ctypes.c_float)),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_D = ctypes.c_double
_U64 = ctypes.c_uint64

_SIGS = {
    "pcc_config_default": (None, [ctypes.POINTER(Config)]),
    "pcc_config_parse_text": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(Config)]),
    "pcc_graph_new": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "pcc_graph_import": (ctypes.c_int, [_P, ctypes.c_uint32, _D, _P, _P, _U64]),
    "pcc_graph_finish_import": (ctypes.c_int, [_P]),
    "pcc_graph_from_csr": (ctypes.c_int, [_I64, _I64, _P, _P, _P, ctypes.POINTER(_P)]),
    "pcc_graph_num_nodes": (_I64, [_P]),
    "pcc_graph_num_edges": (_I64, [_P]),
    "pcc_graph_copy_csr": (ctypes.c_int, [_P, _P, _P, _P]),
    "pcc_graph_destroy": (None, [_P]),
    "pcc_cluster": (ctypes.c_int, [_P, ctypes.POINTER(Config), ctypes.c_int, _P, ctypes.POINTER(Stats)]),
    "pcc_refine_clusters": (ctypes.c_int, [_P, ctypes.POINTER(Config), ctypes.c_int, _P, _P, ctypes.POINTER(Stats)]),
    "pcc_compute_objective": (ctypes.c_int, [_P, ctypes.POINTER(Config), _P, _P, ctypes.POINTER(_D)]),
    "pcc_compute_objective_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Config), _P, _P, _P, _P]),
    "pcc_compute_modularity": (ctypes.c_int, [_P, _D, _P, ctypes.POINTER(_D)]),
    "pcc_best_moves": (ctypes.c_int, [_P, ctypes.POINTER(Config), _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "pcc_round": (ctypes.c_int, [_P, ctypes.POINTER(Config), _P, _P, _P, _P, _P, ctypes.POINTER(_I64)]),
    "pcc_compress": (ctypes.c_int, [_P, _P, _P, ctypes.POINTER(_P), _P]),
    "pcc_set_option": (ctypes.c_int, [_P, ctypes.c_int, _I64]),
    "pcc_autotune": (ctypes.c_int, [_P, ctypes.POINTER(Config), ctypes.c_int, _P, ctypes.c_int,
                                    ctypes.POINTER(TuneResult)]),
    "pcc_batch_from_csr": (ctypes.c_int, [_I64, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "pcc_batch_num_graphs": (_I64, [_P]),
    "pcc_batch_cluster": (ctypes.c_int, [_P, ctypes.POINTER(Config), ctypes.c_int, _P, _P, _P, _P,
                                         ctypes.POINTER(Stats)]),
    "pcc_server_options_default": (None, [ctypes.POINTER(ServerOptions)]),
    "pcc_server_create": (ctypes.c_int, [ctypes.POINTER(Config), ctypes.c_int,
                                         ctypes.POINTER(ServerOptions), ctypes.POINTER(_P)]),
    "pcc_server_submit": (ctypes.c_int, [_P, _I64, _I64, _P, _P, _P, _D, ctypes.POINTER(_U64)]),
    "pcc_server_wait": (ctypes.c_int, [_P, _U64, _P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "pcc_server_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    "pcc_server_destroy": (None, [_P]),
    "pcc_host_csr_free": (None, [ctypes.POINTER(HostCsr)]),
    "pcc_gen_sbm": (ctypes.c_int, [_I64, _I64, _D, _D, _D, _D, _U64, ctypes.POINTER(HostCsr)]),
    "pcc_gen_sbm_blocks": (ctypes.c_int, [_I64, _P, _D, _D, _D, _D, _U64, ctypes.POINTER(HostCsr)]),
    "pcc_gen_rmat": (ctypes.c_int, [ctypes.c_int, _D, _D, _D, _D, _D, _D, _U64, ctypes.POINTER(HostCsr)]),
    "pcc_read_adjacency_graph": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(HostCsr)]),
    "pcc_write_adjacency_graph": (ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(HostCsr), ctypes.c_int]),
    "pcc_last_error": (ctypes.c_char_p, []),
    "pcc_version": (ctypes.c_char_p, []),
}

_lib = None


def lib() -> ctypes.CDLL:
    """Loads libpcc_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `make -C {_HERE}` "
                "(or __graft_entry__.build()); there is no CPU fallback")
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def exported_symbols() -> List[str]:
    return list(_SIGS)


def _check(rc: int) -> None:
    if rc != OK:
        raise StatusError(rc, lib().pcc_last_error().decode(errors="replace"))


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ------------------------------------------------------------------ config --

def default_config() -> Config:
    c = Config()
    lib().pcc_config_default(ctypes.byref(c))
    return c


def parse_config(text: str) -> Config:
    """Text-format ClustererConfig -> Config (the CLI's --clusterer_config)."""
    c = Config()
    _check(lib().pcc_config_parse_text(text.encode(), ctypes.byref(c)))
    return c


def make_config(**kw) -> Config:
    c = default_config()
    for k, v in kw.items():
        setattr(c, "async_" if k == "async" else k, v)
    return c


# -------------------------------------------------------------- host CSR --

@dataclass
class Csr:
    """Host CSR: offsets int64[n+1], nbrs uint32[m], weights float32[m]."""

    n: int
    offsets: np.ndarray
    nbrs: np.ndarray
    weights: np.ndarray

    @property
    def m(self) -> int:
        return int(self.nbrs.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.offsets)


def _take(h: HostCsr) -> Csr:
    n, m = h.n, h.m
    off = np.ctypeslib.as_array(h.offsets, shape=(n + 1,)).copy()
    nb = np.ctypeslib.as_array(h.nbrs, shape=(max(m, 1),))[:m].copy()
    w = np.ctypeslib.as_array(h.weights, shape=(max(m, 1),))[:m].copy()
    lib().pcc_host_csr_free(ctypes.byref(h))
    return Csr(int(n), off, nb, w)


def gen_sbm(n: int, num_blocks: int, p_in: float, p_out: float, w_lo: float = 1.0,
            w_hi: float = 1.0, seed: int = 1) -> Csr:
    h = HostCsr()
    _check(lib().pcc_gen_sbm(n, num_blocks, p_in, p_out, w_lo, w_hi, seed, ctypes.byref(h)))
    return _take(h)


def gen_sbm_blocks(block_sizes: Sequence[int], p_in: float, p_out: float, w_lo: float = 1.0,
                   w_hi: float = 1.0, seed: int = 1) -> Csr:
    bs = np.ascontiguousarray(block_sizes, dtype=np.int64)
    h = HostCsr()
    _check(lib().pcc_gen_sbm_blocks(len(bs), _ptr(bs), p_in, p_out, w_lo, w_hi, seed,
                                    ctypes.byref(h)))
    return _take(h)


def gen_rmat(scale: int, edge_factor: float, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             w_lo: float = 1.0, w_hi: float = 1.0, seed: int = 2) -> Csr:
    h = HostCsr()
    _check(lib().pcc_gen_rmat(scale, edge_factor, a, b, c, w_lo, w_hi, seed, ctypes.byref(h)))
    return _take(h)


def read_adjacency_graph(path: str, float_weighted: bool = False) -> Csr:
    h = HostCsr()
    _check(lib().pcc_read_adjacency_graph(path.encode(), int(float_weighted), ctypes.byref(h)))
    return _take(h)


def write_adjacency_graph(path: str, csr: Csr, float_weighted: bool = False) -> None:
    off = np.ascontiguousarray(csr.offsets, dtype=np.int64)
    nb = np.ascontiguousarray(csr.nbrs, dtype=np.uint32)
    w = np.ascontiguousarray(csr.weights, dtype=np.float32)
    h = HostCsr(csr.n, csr.m, off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                nb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                w.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    _check(lib().pcc_write_adjacency_graph(path.encode(), ctypes.byref(h), int(float_weighted)))


# ------------------------------------------------------------ device graph --

class DeviceGraph:
    """A pcc_graph handle: CSR resident in HBM plus a HIP stream."""

    def __init__(self, handle):
        self._h = handle

    @classmethod
    def from_csr(cls, csr: Csr) -> "DeviceGraph":
        off = np.ascontiguousarray(csr.offsets, dtype=np.int64)
        nb = np.ascontiguousarray(csr.nbrs, dtype=np.uint32)
        w = np.ascontiguousarray(csr.weights, dtype=np.float32)
        h = ctypes.c_void_p()
        _check(lib().pcc_graph_from_csr(csr.n, csr.m, _ptr(off), _ptr(nb), _ptr(w), ctypes.byref(h)))
        return cls(h)

    def close(self) -> None:
        if self._h:
            lib().pcc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def n(self) -> int:
        return int(lib().pcc_graph_num_nodes(self._h))

    @property
    def m(self) -> int:
        return int(lib().pcc_graph_num_edges(self._h))

    def set_option(self, key: int, value: int) -> None:
        _check(lib().pcc_set_option(self._h, key, value))

    def autotune(self, cfg: Config, objective: int = OBJECTIVE_CORRELATION, resolutions=None,
                 max_trials: int = 0) -> dict:
        """pcc_autotune: picks the degree-bin thresholds and hub partition
        size for this workload and leaves them set on the handle."""
        res = None if resolutions is None else np.ascontiguousarray(resolutions, dtype=np.float64)
        out = TuneResult()
        _check(lib().pcc_autotune(self._h, ctypes.byref(cfg), objective, _ptr(res), max_trials,
                                  ctypes.byref(out)))
        return out.as_dict()

    def to_csr(self) -> Csr:
        n, m = self.n, self.m
        off = np.zeros(n + 1, np.int64)
        nb = np.zeros(max(m, 1), np.uint32)
        w = np.zeros(max(m, 1), np.float32)
        _check(lib().pcc_graph_copy_csr(self._h, _ptr(off), _ptr(nb), _ptr(w)))
        return Csr(n, off, nb[:m], w[:m])

    def cluster(self, cfg: Config, objective: int = OBJECTIVE_CORRELATION) -> Tuple[np.ndarray, Stats]:
        ids = np.zeros(max(self.n, 1), np.uint32)
        st = Stats()
        _check(lib().pcc_cluster(self._h, ctypes.byref(cfg), objective, _ptr(ids), ctypes.byref(st)))
        return ids[: self.n], st

    def refine(self, cfg: Config, ids: np.ndarray, objective: int = OBJECTIVE_CORRELATION,
               node_weights: Optional[np.ndarray] = None) -> Tuple[np.ndarray, Stats]:
        out = np.ascontiguousarray(ids, dtype=np.uint32).copy()
        nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
        st = Stats()
        _check(lib().pcc_refine_clusters(self._h, ctypes.byref(cfg), objective, _ptr(nw), _ptr(out),
                                         ctypes.byref(st)))
        return out, st

    def objective(self, cfg: Config, ids: np.ndarray, node_weights: Optional[np.ndarray] = None) -> float:
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        nw = None if node_weights is None else np.ascontiguousarray(node_weights, dtype=np.float64)
        out = ctypes.c_double()
        _check(lib().pcc_compute_objective(self._h, ctypes.byref(cfg), _ptr(ids), _ptr(nw), ctypes.byref(out)))
        return out.value

    def objective_noalloc(self, cfg: Config, ids_dev: int, nw_dev: Optional[int], out_dev: int,
                          stream: Optional[int] = None) -> None:
        """pcc_compute_objective_noalloc: device pointers (integers) in and out;
        asynchronous on `stream` (a hipStream_t, None: the handle's stream)."""
        _check(lib().pcc_compute_objective_noalloc(self._h, ctypes.byref(cfg), ids_dev, nw_dev,
                                                   out_dev, stream))

    def modularity(self, resolution: float, ids: np.ndarray) -> float:
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        out = ctypes.c_double()
        _check(lib().pcc_compute_modularity(self._h, resolution, _ptr(ids), ctypes.byref(out)))
        return out.value

    def best_moves(self, cfg: Config, ids, sizes, nw, K, frontier=None):
        n = self.n
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
        nw = np.ascontiguousarray(nw, dtype=np.float64)
        K = np.ascontiguousarray(K, dtype=np.float64)
        fr = None if frontier is None else np.ascontiguousarray(frontier, dtype=np.uint8)
        moves = np.zeros(n, np.uint32)
        moved = np.zeros(n, np.uint8)
        tgt = np.zeros(n, np.uint32)
        dlt = np.zeros(n, np.float64)
        _check(lib().pcc_best_moves(self._h, ctypes.byref(cfg), _ptr(ids), _ptr(sizes), _ptr(nw), _ptr(K),
                                    _ptr(fr), _ptr(moves), _ptr(moved), _ptr(tgt), _ptr(dlt)))
        return moves, moved, tgt, dlt

    def round(self, cfg: Config, ids, sizes, nw, K, frontier):
        ids = np.ascontiguousarray(ids, dtype=np.uint32).copy()
        sizes = np.ascontiguousarray(sizes, dtype=np.uint32).copy()
        nw = np.ascontiguousarray(nw, dtype=np.float64)
        K = np.ascontiguousarray(K, dtype=np.float64).copy()
        fr = np.ascontiguousarray(frontier, dtype=np.uint8).copy()
        fs = ctypes.c_int64()
        _check(lib().pcc_round(self._h, ctypes.byref(cfg), _ptr(ids), _ptr(sizes), _ptr(nw), _ptr(K),
                               _ptr(fr), ctypes.byref(fs)))
        return ids, sizes, K, fr, fs.value

    def compress(self, ids, nw) -> Tuple["DeviceGraph", np.ndarray]:
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        nw = np.ascontiguousarray(nw, dtype=np.float64)
        out_nw = np.zeros(max(self.n, 1), np.float64)
        h = ctypes.c_void_p()
        _check(lib().pcc_compress(self._h, _ptr(ids), _ptr(nw), ctypes.byref(h), _ptr(out_nw)))
        g = DeviceGraph(h)
        return g, out_nw[: g.n]


class DeviceBatch(DeviceGraph):
    """Many independent graphs in one device structure (pcc_batch_*)."""

    @classmethod
    def from_csrs(cls, csrs: Sequence[Csr]) -> "DeviceBatch":
        counts = np.array([c.n for c in csrs], np.int64)
        off = [np.zeros(1, np.int64)]
        run = 0
        for c in csrs:
            off.append(c.offsets[1:] + run)
            run += c.m
        offsets = np.ascontiguousarray(np.concatenate(off), dtype=np.int64)
        nbrs = np.ascontiguousarray(np.concatenate([c.nbrs for c in csrs]) if run else
                                    np.zeros(0, np.uint32), dtype=np.uint32)
        w = np.ascontiguousarray(np.concatenate([c.weights for c in csrs]) if run else
                                 np.zeros(0, np.float32), dtype=np.float32)
        h = ctypes.c_void_p()
        _check(lib().pcc_batch_from_csr(len(csrs), _ptr(counts), _ptr(offsets), _ptr(nbrs), _ptr(w),
                                        ctypes.byref(h)))
        b = cls(h)
        b.counts = counts
        return b

    @property
    def num_graphs(self) -> int:
        return int(lib().pcc_batch_num_graphs(self._h))

    def cluster_batch(self, cfg: Config, resolutions=None, objective: int = OBJECTIVE_CORRELATION):
        """Returns (list of per-graph local id arrays, objectives, modularities, stats)."""
        G = self.num_graphs
        ids = np.zeros(max(self.n, 1), np.uint32)
        obj = np.zeros(G)
        mod = np.zeros(G)
        res = None if resolutions is None else np.ascontiguousarray(resolutions, dtype=np.float64)
        st = Stats()
        _check(lib().pcc_batch_cluster(self._h, ctypes.byref(cfg), objective, _ptr(res), _ptr(ids),
                                       _ptr(obj), _ptr(mod), ctypes.byref(st)))
        bounds = np.concatenate([[0], np.cumsum(self.counts)])
        return [ids[bounds[i]:bounds[i + 1]] for i in range(G)], obj, mod, st


class Server:
    """Continuous-batching request loop (pcc_server_*)."""

    def __init__(self, cfg: Config, objective: int = OBJECTIVE_CORRELATION, workers: int = 2,
                 max_batch_graphs: int = 512, hbm_budget_bytes: float = 0.0,
                 batch_window_seconds: float = 0.0):
        o = ServerOptions()
        lib().pcc_server_options_default(ctypes.byref(o))
        o.workers = workers
        o.max_batch_graphs = max_batch_graphs
        o.hbm_budget_bytes = hbm_budget_bytes
        o.batch_window_seconds = batch_window_seconds
        h = ctypes.c_void_p()
        _check(lib().pcc_server_create(ctypes.byref(cfg), objective, ctypes.byref(o), ctypes.byref(h)))
        self._h = h
        self._n = {}

    def submit(self, csr: Csr, resolution: float) -> int:
        off = np.ascontiguousarray(csr.offsets, dtype=np.int64)
        nb = np.ascontiguousarray(csr.nbrs, dtype=np.uint32)
        w = np.ascontiguousarray(csr.weights, dtype=np.float32)
        rid = ctypes.c_uint64()
        _check(lib().pcc_server_submit(self._h, csr.n, csr.m, _ptr(off), _ptr(nb), _ptr(w),
                                       resolution, ctypes.byref(rid)))
        self._n[rid.value] = csr.n
        return rid.value

    def wait(self, rid: int):
        ids = np.zeros(max(self._n.pop(rid), 1), np.uint32)
        obj = ctypes.c_double()
        lat = ctypes.c_double()
        _check(lib().pcc_server_wait(self._h, rid, _ptr(ids), ctypes.byref(obj), ctypes.byref(lat)))
        return ids, obj.value, lat.value

    def stats(self) -> ServerStats:
        st = ServerStats()
        _check(lib().pcc_server_get_stats(self._h, ctypes.byref(st)))
        return st

    def close(self) -> None:
        if self._h:
            lib().pcc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------- InMemoryClusterer mirror API --

def output_indices_by_id(ids: np.ndarray) -> List[List[int]]:
    """OutputIndicesById (parallel/parallel-sequence-ops.h:183-209)."""
    ids = np.asarray(ids)
    order = np.argsort(ids, kind="stable")
    sorted_ids = ids[order]
    bounds = np.flatnonzero(np.diff(sorted_ids)) + 1
    return [list(map(int, g)) for g in np.split(order, bounds)] if ids.size else []


def clustering_to_ids(clustering: Sequence[Sequence[int]], n: int) -> np.ndarray:
    ids = np.full(n, NO_MOVE, np.uint32)
    for i, cl in enumerate(clustering):
        for v in cl:
            ids[v] = i
    if (ids == NO_MOVE).any():
        raise StatusError(INVALID_ARGUMENT, "clustering must cover every node exactly once")
    return ids


@dataclass
class AdjacencyList:
    """InMemoryClusterer::Graph::AdjacencyList (in-memory-clusterer.h:45-49)."""

    id: int = -1
    weight: float = 1.0
    outgoing_edges: List[Tuple[int, float]] = field(default_factory=list)


class Graph:
    """GbbsGraph equivalent (clustering/gbbs-graph.h:39-56) feeding the device."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(lib().pcc_graph_new(ctypes.byref(h)))
        self._dev = DeviceGraph(h)
        self._finished = False

    def Import(self, adjacency_list: AdjacencyList) -> None:
        e = adjacency_list.outgoing_edges
        nb = np.ascontiguousarray([x[0] for x in e], dtype=np.uint32)
        w = np.ascontiguousarray([x[1] for x in e], dtype=np.float64)
        _check(lib().pcc_graph_import(self._dev.handle, adjacency_list.id, adjacency_list.weight,
                                      _ptr(nb), _ptr(w), len(e)))

    def FinishImport(self) -> None:
        _check(lib().pcc_graph_finish_import(self._dev.handle))
        self._finished = True

    def set_csr(self, csr: Csr) -> None:
        """The CLI path: graph_ assigned from a GBBS adjacency file (MAIN.cc:306)."""
        self._dev = DeviceGraph.from_csr(csr)
        self._finished = True

    @property
    def device(self) -> DeviceGraph:
        return self._dev


def _as_config(config) -> Config:
    if isinstance(config, Config):
        return config
    return parse_config(config)


class ParallelCorrelationClusterer:
    """ParallelCorrelationClusterer (clusterers/parallel-correlation-clusterer.h:44-95)."""

    objective = OBJECTIVE_CORRELATION

    def __init__(self):
        self._graph = Graph()
        self.last_stats: Optional[Stats] = None

    def MutableGraph(self) -> Graph:
        return self._graph

    def Cluster(self, config) -> List[List[int]]:
        ids, st = self._graph.device.cluster(_as_config(config), self.objective)
        self.last_stats = st
        return output_indices_by_id(ids)

    def RefineClusters(self, config, clustering: List[List[int]],
                       node_weights: Optional[Sequence[float]] = None) -> List[List[int]]:
        n = self._graph.device.n
        ids = clustering_to_ids(clustering, n)
        nw = None if node_weights is None else np.asarray(node_weights, np.float64)
        out, st = self._graph.device.refine(_as_config(config), ids, self.objective, nw)
        self.last_stats = st
        clustering[:] = output_indices_by_id(out)
        return clustering


class ParallelModularityClusterer(ParallelCorrelationClusterer):
    """ParallelModularityClusterer (clusterers/parallel-modularity-clusterer.h:45-56)."""

    objective = OBJECTIVE_MODULARITY
