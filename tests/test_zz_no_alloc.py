"""pcc_compute_objective_noalloc against pcc_compute_objective.

The allocation-free entry point takes device arrays and writes a device
double on a caller stream; its work memory is a fixed per-device array.  The
test allocates the device arrays itself through the HIP runtime the library
links (libamdhip64.so.7, the one already loaded with libpcc_amd.so), calls
the entry point twice in a row (the second time on a smaller graph, so the
array keeps the first call's contents), and compares each value with the
allocating entry point's bit for bit.
"""

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_H2D, _D2H = 1, 2


def _hip():
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return h


class DeviceArrays:
    """Device copies of host arrays (the test's own allocations)."""

    def __init__(self, hip):
        self.hip = hip
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        self.ptrs.append(p)
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, _H2D) == 0
        return p.value

    def get_f64(self, ptr):
        out = np.zeros(1, np.float64)
        assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ptr, 8, _D2H) == 0
        return out[0]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def noalloc_objective(pcc, hip, g, cfg, ids, nw):
    d = DeviceArrays(hip)
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        ids_dev = d.put(np.asarray(ids, np.uint32))
        nw_dev = None if nw is None else d.put(np.asarray(nw, np.float64))
        out_dev = d.put(np.full(1, 12345.0))
        g.objective_noalloc(cfg, ids_dev, nw_dev, out_dev, stream.value)
        assert hip.hipStreamSynchronize(stream) == 0
        return d.get_f64(out_dev)
    finally:
        hip.hipStreamDestroy(stream)
        d.free()


def test_noalloc_objective_equals_allocating_one(pcc):
    hip = _hip()
    # First a graph with rows longer than 16384 entries (the long-row
    # kernels), then a smaller one: the second call runs on the contents the
    # first one left in the shared array.
    big = pcc.gen_rmat(19, 16, w_lo=-1.0, w_hi=1.0, seed=81)
    small = pcc.gen_sbm(3000, 30, 0.08, 0.004, -0.5, 1.0, seed=82)
    assert np.diff(big.offsets).max() > 16384
    rng = np.random.default_rng(9)
    for name, csr in (("rmat19", big), ("sbm3000", small)):
        print(f"noalloc {name} n={csr.n} m={csr.m}", flush=True)
        g = pcc.DeviceGraph.from_csr(csr)
        ids = rng.integers(0, max(2, csr.n // 50), size=csr.n).astype(np.uint32)
        # Quarter-integer weights: every cluster weight is an exact sum, so
        # the order of the atomic adds that form it cannot change a bit (with
        # arbitrary fractional weights two calls of either entry point may
        # differ in the last bit).
        nw = rng.integers(2, 9, size=csr.n) / 4.0
        for weights in (nw, None):
            cfg = pcc.make_config(resolution=0.01, edge_weight_offset=0.125)
            want = g.objective(cfg, ids, weights)
            got = noalloc_objective(pcc, hip, g, cfg, ids, weights)
            assert np.float64(got).view(np.uint64) == np.float64(want).view(np.uint64), (name, got,
                                                                                        want)
        g.close()


def test_noalloc_objective_rejects_sizes_beyond_its_array(pcc):
    hip = _hip()
    n = (1 << 21) + 1
    csr = pcc.Csr(n, np.zeros(n + 1, np.int64), np.zeros(0, np.uint32), np.zeros(0, np.float32))
    g = pcc.DeviceGraph.from_csr(csr)
    d = DeviceArrays(hip)
    try:
        ids_dev = d.put(np.zeros(n, np.uint32))
        out_dev = d.put(np.zeros(1))
        with pytest.raises(pcc.StatusError) as e:
            g.objective_noalloc(pcc.make_config(resolution=0.01), ids_dev, None, out_dev)
        assert e.value.code == pcc.INVALID_ARGUMENT and "2^21" in e.value.message
    finally:
        d.free()
        g.close()


def test_noalloc_objective_out_of_range_ids_give_nan(pcc):
    hip = _hip()
    csr = pcc.gen_sbm(500, 5, 0.1, 0.01, seed=83)
    g = pcc.DeviceGraph.from_csr(csr)
    ids = np.arange(csr.n, dtype=np.uint32)
    ids[7] = csr.n + 3
    got = noalloc_objective(pcc, hip, g, pcc.make_config(resolution=0.01), ids, None)
    assert np.isnan(got)
    # The next call with valid ids is unaffected by the flag left behind.
    ids[7] = 0
    got = noalloc_objective(pcc, hip, g, pcc.make_config(resolution=0.01), ids, None)
    assert got == g.objective(pcc.make_config(resolution=0.01), ids)
    g.close()
