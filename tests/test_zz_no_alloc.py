"""scd_count_triangles_noalloc: the triangle statistics of scd_count_triangles
with outputs in device memory and every temporary in the library's static
device array -- no allocation -- equal to the allocating entry point's
results (and to the graph files' own tri / tri_deg, which the reference
convert wrote) on two calls in a row, the second on a smaller graph; a graph
too large for the array is rejected with SCD_ELIMIT."""
import ctypes

import numpy as np
import pytest

import scd_amd
from scd_amd import _lib
from golden_graphs import graph

pytestmark = pytest.mark.gpu

ELIMIT = -5
ESTATE = -6


class DevBufs:
    def __init__(self):
        self.hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.ptrs = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 4))) == 0
        self.ptrs.append(p)
        return p

    def fill(self, p, nbytes, byte):
        assert self.hip.hipMemset(p, byte, ctypes.c_size_t(nbytes)) == 0

    def get(self, p, count, dtype):
        out = np.empty(count, dtype)
        assert self.hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), p, ctypes.c_size_t(out.nbytes), 2) == 0
        return out

    def free(self):
        self.hip.hipDeviceSynchronize()
        for p in self.ptrs:
            self.hip.hipFree(p)


def triangles_noalloc(eng, bufs, n, with_deg=True):
    d_tri = bufs.alloc(4 * n)
    d_td = bufs.alloc(4 * n) if with_deg else None
    bufs.fill(d_tri, 4 * n, 0xAB)  # garbage the call must overwrite
    if with_deg:
        bufs.fill(d_td, 4 * n, 0xCD)
    rc = _lib.hip().scd_count_triangles_noalloc(eng._ctx, d_tri, d_td)
    assert rc == 0, _lib.hip().scd_last_error().decode()
    return bufs.get(d_tri, n, np.int32), (bufs.get(d_td, n, np.int32) if with_deg else None)


def test_noalloc_equals_allocating_entry_two_calls():
    bufs = DevBufs()
    eng = scd_amd.Engine(device=0)
    try:
        for name in ("rmat16", "pp8x40_s5"):  # the second call on a smaller graph
            g = graph(name)
            print(f"[gpu] noalloc triangles {name} n={g.n} m={g.m}", flush=True)
            eng.transfer_graph(g)
            want_tri, want_td = eng.count_triangles()
            tri, td = triangles_noalloc(eng, bufs, g.n)
            assert np.array_equal(tri, want_tri) and np.array_equal(td, want_td)
            assert np.array_equal(tri, g.tri) and np.array_equal(td, g.tri_deg)
            tri2, _ = triangles_noalloc(eng, bufs, g.n, with_deg=False)
            assert np.array_equal(tri2, want_tri)
    finally:
        eng.close()
        bufs.free()


def test_noalloc_rejects_graph_beyond_scratch():
    # 2^21 disjoint triangles: 8n + 2m + 4n bytes exceed the 64 MiB array
    k = 1 << 21
    n = 3 * k
    base = np.arange(k, dtype=np.uint32) * 3
    adj = np.stack([base + 1, base + 2, base, base + 2, base, base + 1], axis=1).reshape(-1)
    indices = np.arange(n + 1, dtype=np.uint32) * 2
    tri = np.full(n, 2, np.int32)
    g = scd_amd.Graph(n, 3 * k, indices, adj, tri, np.full(n, 2, np.int32))
    bufs = DevBufs()
    eng = scd_amd.Engine(device=0)
    try:
        eng.transfer_graph(g)
        d = bufs.alloc(4 * n)
        rc = _lib.hip().scd_count_triangles_noalloc(eng._ctx, d, None)
        assert rc == ELIMIT, _lib.hip().scd_last_error().decode()
        eng.set_profiling(True)
        small = graph("k9")
        eng.transfer_graph(small)
        assert _lib.hip().scd_count_triangles_noalloc(eng._ctx, d, None) == ESTATE
        eng.set_profiling(False)
        tri, td = triangles_noalloc(eng, bufs, small.n)
        assert np.array_equal(tri, small.tri) and np.array_equal(td, small.tri_deg)
    finally:
        eng.close()
        bufs.free()
