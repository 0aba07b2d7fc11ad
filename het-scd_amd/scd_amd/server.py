"""Serving loop (include/scd_serve.h): graphs arrive over time, are admitted
within an HBM budget, refined together with continuous batching, and each
comes back with the result the single-graph path gives it alone."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check_serve, ptr, u32p, i32p


class Server:
    """Requests are queued with submit(graph, arrival_ms) and served by
    run(), which replays the arrival times in wall-clock time."""

    def __init__(self, device=0, hbm_budget=0, max_graphs=0, join_wait_ms=5.0, max_iters=1000, graph_mode=False,
                 intake_graphs=0, intake_bytes=0):
        o = _lib.ScdServerOptsC(device, hbm_budget, max_graphs, join_wait_ms, max_iters, 1 if graph_mode else 0,
                                intake_graphs, intake_bytes)
        self._s = ctypes.c_void_p()
        check_serve(_lib.serve().scd_server_create(ctypes.byref(o), ctypes.byref(self._s)))
        self._keep = []  # host arrays must outlive run()
        self._n = []

    def close(self):
        if self._s:
            _lib.serve().scd_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, g, arrival_ms=0.0, clustering_coef=None):
        cc = g.clustering_coef if clustering_coef is None else clustering_coef
        r = _lib.ScdRequestC(g.n, g.m, ptr(g.indices, u32p), ptr(g.adj, u32p), ptr(g.tri, i32p),
                             ptr(g.tri_deg, i32p), cc, float(arrival_ms))
        rid = ctypes.c_uint64()
        check_serve(_lib.serve().scd_server_submit(self._s, ctypes.byref(r), ctypes.byref(rid)))
        self._keep.append(g)
        self._n.append(g.n)
        return rid.value

    def run(self):
        check_serve(_lib.serve().scd_server_run(self._s))

    def result(self, rid):
        res = _lib.ScdResultC()
        labels = np.empty(self._n[rid], np.uint32)
        check_serve(_lib.serve().scd_server_result(self._s, rid, ptr(labels, u32p), ctypes.byref(res)))
        out = {k: getattr(res, k) for k, _ in _lib.ScdResultC._fields_}
        out["labels"] = labels if res.status == 0 else None
        return out

    def stats(self):
        st = _lib.ScdServerStatsC()
        check_serve(_lib.serve().scd_server_get_stats(self._s, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.ScdServerStatsC._fields_}

    @staticmethod
    def footprint(n, m):
        return int(_lib.serve().scd_server_footprint(n, m))
