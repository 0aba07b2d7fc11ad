"""Python mirror of the reference's single-GPU interface (src/cuda.hpp:23-28)
over the C ABI of libscd_hip.so (include/scd_hip.h).

    eng = Engine(device=0)                   # cuda_init
    eng.transfer_graph(g)                    # cuda_transfer_graph
    eng.labels_to_device(labels)             # cuda_labels_to_device
    wcc = eng.improve_labels()               # cuda_improve_labels
    eng.store_labels()                       # cuda_store_labels
    best = eng.labels_from_device()          # cuda_labels_from_device

Errors raise ScdError instead of exiting the process.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check_hip, f64p, i32p, ptr, u32p, u64p
from .graph import Graph


def device_count():
    return check_hip(_lib.hip().scd_device_count())


def _get_tuning(ctx):
    t = _lib.ScdTuningC()
    check_hip(_lib.hip().scd_get_tuning(ctx, ctypes.byref(t)))
    return {"lds_tile": t.lds_tile, "merge_chunk": t.merge_chunk, "wave_max": t.wave_max, "hub_table": t.hub_table,
            "lds_hash": t.lds_hash, "lanes": t.lanes, "dense_wcc": t.dense_wcc}


def _set_tuning(ctx, lds_tile=None, merge_chunk=None, wave_max=None, hub_table=None, lds_hash=None, lanes=None,
                dense_wcc=None):
    t = _lib.ScdTuningC()
    check_hip(_lib.hip().scd_get_tuning(ctx, ctypes.byref(t)))
    if lds_tile is not None:
        t.lds_tile = lds_tile
    if merge_chunk is not None:
        t.merge_chunk = merge_chunk
    if wave_max is not None:
        t.wave_max = wave_max
    if hub_table is not None:
        t.hub_table = hub_table
    if lds_hash is not None:
        t.lds_hash = lds_hash
    if lanes is not None:
        t.lanes = lanes
    if dense_wcc is not None:
        t.dense_wcc = dense_wcc
    check_hip(_lib.hip().scd_set_tuning(ctx, ctypes.byref(t)))


def _autotune(ctx, cache_path, reps):
    t = _lib.ScdTuningC()
    ms = ctypes.c_double()
    path = None if cache_path is None else str(cache_path).encode()
    check_hip(_lib.hip().scd_autotune(ctx, path, reps, ctypes.byref(t), ctypes.byref(ms)))
    return {"lds_tile": t.lds_tile, "merge_chunk": t.merge_chunk, "wave_max": t.wave_max,
            "hub_table": t.hub_table, "lds_hash": t.lds_hash, "lanes": t.lanes,
            "dense_wcc": t.dense_wcc}, ms.value


class Engine:
    def __init__(self, device=0, work_group_size=256, profile=False):
        self._ctx = ctypes.c_void_p()
        check_hip(_lib.hip().scd_init(device, work_group_size, 1 if profile else 0, ctypes.byref(self._ctx)))
        self.n = 0
        self.m = 0
        self.graph = None

    def close(self):
        if self._ctx:
            _lib.hip().scd_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- cuda.hpp mirror -------------------------------------------------
    def transfer_graph(self, g, clustering_coef=None):
        cc = g.clustering_coef if clustering_coef is None else clustering_coef
        check_hip(_lib.hip().scd_upload(self._ctx, g.n, g.m, ptr(g.indices, u32p), ptr(g.adj, u32p),
                                        ptr(g.tri, i32p), ptr(g.tri_deg, i32p), cc))
        self.n = g.n
        self.m = g.m
        self.graph = g

    def labels_to_device(self, labels):
        labels = np.ascontiguousarray(labels, dtype=np.uint32)
        if labels.shape != (self.n,):
            raise ValueError(f"labels must hold n = {self.n} entries")
        check_hip(_lib.hip().scd_labels_to_device(self._ctx, ptr(labels, u32p)))

    def set_partition(self, labels, num_comms=None):
        """scd_set_partition without caller statistics: any labels below
        num_comms (default: the largest valid label + 1) or INVALID become
        current, with their statistics computed on the device."""
        labels = np.ascontiguousarray(labels, dtype=np.uint32)
        if labels.shape != (self.n,):
            raise ValueError(f"labels must hold n = {self.n} entries")
        if num_comms is None:
            valid = labels[labels != 0xFFFFFFFF]
            num_comms = int(valid.max()) + 1 if valid.size else 0
        check_hip(_lib.hip().scd_set_partition(self._ctx, ptr(labels, u32p), num_comms, None, None, None))

    def improve_labels(self, return_labels=False):
        wcc = ctypes.c_double()
        out = np.empty(self.n, np.uint32) if return_labels else None
        check_hip(_lib.hip().scd_improve(self._ctx, ctypes.byref(wcc), ptr(out, u32p)))
        return (wcc.value, out) if return_labels else wcc.value

    def store_labels(self):
        check_hip(_lib.hip().scd_store_best(self._ctx))

    def labels_from_device(self):
        out = np.empty(self.n, np.uint32)
        check_hip(_lib.hip().scd_labels_from_device(self._ctx, ptr(out, u32p)))
        return out

    # -- the rest of the path ------------------------------------------------
    def current_labels(self):
        out = np.empty(self.n, np.uint32)
        check_hip(_lib.hip().scd_current_labels(self._ctx, ptr(out, u32p)))
        return out

    def count_triangles(self, replace=False, outputs=True):
        """t / vt of the uploaded graph (src/convert.cpp:271-340); with
        outputs=False nothing is copied back (returns None)."""
        tri = np.empty(self.n, np.int32) if outputs else None
        td = np.empty(self.n, np.int32) if outputs else None
        check_hip(_lib.hip().scd_count_triangles(self._ctx, ptr(tri, i32p), ptr(td, i32p), 1 if replace else 0))
        return (tri, td) if outputs else None

    def set_triangle_split(self, tc_heavy=128, wcc_heavy=0):
        """scd_set_triangle_split: out-list length above which an edge is
        walked from its longer out-list's side (0: never), for graph triangle
        counting / convert and for the WCC's internal triangles."""
        check_hip(_lib.hip().scd_set_triangle_split(self._ctx, tc_heavy, wcc_heavy))

    def convert(self, edges, download=True):
        """The convert tool (src/convert.cpp:466-537) on the device for an
        Edges list: the result replaces the engine's graph, as
        transfer_graph would.  Returns the converted Graph (downloaded, equal
        byte for byte to edges.convert()) or, with download=False, its
        (n, m, clustering_coef)."""
        return self._convert(len(edges), edges.u.ctypes.data, edges.v.ctypes.data, 0, download)

    def convert_device(self, count, u_ptr, v_ptr, download=False):
        """As convert, for an edge list already in device memory: u_ptr and
        v_ptr address count uint32 endpoints each on this engine's device."""
        return self._convert(count, u_ptr, v_ptr, 1, download)

    def _convert(self, count, u_ptr, v_ptr, on_device, download):
        n, m, cc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        check_hip(_lib.hip().scd_convert_edges(self._ctx, count, ctypes.c_void_p(u_ptr), ctypes.c_void_p(v_ptr),
                                               on_device, ctypes.byref(n), ctypes.byref(m), ctypes.byref(cc)))
        self.n = n.value
        self.m = m.value
        self.graph = None
        if not download:
            return n.value, m.value, cc.value
        g = self.download_graph() if n.value else Graph(0, 0, np.zeros(1, np.uint32), np.zeros(0, np.uint32),
                                                          np.zeros(0, np.int32), np.zeros(0, np.int32))
        g._cc = cc.value
        self.graph = g if n.value else None
        return g

    def convert_info(self):
        """Counts and per-step device times of the last convert, as a dict."""
        info = _lib.ScdConvertInfoC()
        check_hip(_lib.hip().scd_convert_last_info(self._ctx, ctypes.byref(info)))
        return {f: getattr(info, f) for f, _ in info._fields_}

    def download_graph(self):
        """The engine's graph copied back to the host as a Graph."""
        n, m, cc = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_double()
        check_hip(_lib.hip().scd_graph_shape(self._ctx, ctypes.byref(n), ctypes.byref(m), ctypes.byref(cc)))
        n, m = n.value, m.value
        idx = np.empty(n + 1, np.uint32)
        adj = np.empty(2 * m, np.uint32)
        tri = np.empty(n, np.int32)
        td = np.empty(n, np.int32)
        check_hip(_lib.hip().scd_download_graph(self._ctx, ptr(idx, u32p), ptr(adj, u32p), ptr(tri, i32p),
                                                ptr(td, i32p)))
        g = Graph(n, m, idx, adj, tri, td)
        g._cc = cc.value
        return g

    def initial_partition(self, want_labels=True):
        out = np.empty(self.n, np.uint32) if want_labels else None
        k = ctypes.c_uint32()
        check_hip(_lib.hip().scd_initial_partition(self._ctx, ptr(out, u32p), ctypes.byref(k)))
        return out, k.value

    def compute_wcc(self, details=False):
        wcc = ctypes.c_double()
        if not details:
            check_hip(_lib.hip().scd_compute_wcc(self._ctx, ctypes.byref(wcc), None, None, None))
            return wcc.value
        terms = np.empty(self.n, np.float64)
        it = np.empty(self.n, np.int32)
        itd = np.empty(self.n, np.int32)
        check_hip(_lib.hip().scd_compute_wcc(self._ctx, ctypes.byref(wcc), ptr(terms, f64p), ptr(it, i32p),
                                             ptr(itd, i32p)))
        return wcc.value, terms, it, itd

    def stats(self):
        nc = ctypes.c_uint32()
        s = np.empty(self.n, np.int32)
        e = np.empty(self.n, np.int32)
        i = np.empty(self.n, np.int32)
        check_hip(_lib.hip().scd_get_stats(self._ctx, ctypes.byref(nc), ptr(s, i32p), ptr(e, i32p), ptr(i, i32p)))
        k = nc.value
        return s[:k].copy(), e[:k].copy(), i[:k].copy()

    def neighbor_label_pairs(self):
        """(sorted[2m], run_offsets[n+1], run_label[P], run_count[P])."""
        E = 2 * self.m
        srt = np.empty(max(E, 1), np.uint32)
        off = np.empty(self.n + 1, np.uint32)
        rl = np.empty(max(E, 1), np.uint32)
        rc = np.empty(max(E, 1), np.uint32)
        P = ctypes.c_uint64()
        check_hip(_lib.hip().scd_neighbor_label_pairs(self._ctx, ptr(srt, u32p), ptr(off, u32p), ptr(rl, u32p),
                                                      ptr(rc, u32p), ctypes.byref(P)))
        P = P.value
        return srt[:E], off, rl[:P].copy(), rc[:P].copy()

    def refine(self, initial_wcc, max_iters=1000):
        """The loop of src/main-gpu.cpp:71-96.  Returns (best_wcc, iterations,
        per-iteration WCC)."""
        best = ctypes.c_double()
        iters = ctypes.c_uint32()
        hist = np.zeros(max_iters, np.float64)
        check_hip(_lib.hip().scd_refine(self._ctx, initial_wcc, max_iters, ctypes.byref(best),
                                        ctypes.byref(iters), ptr(hist, f64p)))
        return best.value, iters.value, hist[:iters.value].copy()

    def get_tuning(self):
        return _get_tuning(self._ctx)

    def set_tuning(self, lds_tile=None, merge_chunk=None, wave_max=None, hub_table=None, lds_hash=None,
                   lanes=None, dense_wcc=None):
        _set_tuning(self._ctx, lds_tile, merge_chunk, wave_max, hub_table, lds_hash, lanes, dense_wcc)

    def autotune(self, cache_path=None, reps=3):
        """scd_autotune: fastest tuning for this graph's size class (cached);
        returns (tuning, iteration ms)."""
        return _autotune(self._ctx, cache_path, reps)

    def last_counters(self):
        out = np.zeros(4, np.uint64)
        check_hip(_lib.hip().scd_last_counters(self._ctx, ptr(out, u64p)))
        return {"pairs": int(out[0]), "internal_entries": int(out[1]), "internal_deg_sq": int(out[2]),
                "hubs": int(out[3])}

    def dense_wcc_info(self):
        """scd_dense_wcc_info: (row units requested, vertices counted by bitmap
        rows) of the last internal triangle count."""
        out = np.zeros(2, np.uint32)
        check_hip(_lib.hip().scd_dense_wcc_info(self._ctx, ptr(out, u32p)))
        return int(out[0]), int(out[1])

    def profile(self):
        """{kernel name: (launches, total ms)} from the HIP events."""
        lib = _lib.hip()
        res = {}
        for i in range(lib.scd_prof_count(self._ctx)):
            name = ctypes.c_char_p()
            la = ctypes.c_uint64()
            ms = ctypes.c_double()
            check_hip(lib.scd_prof_get(self._ctx, i, ctypes.byref(name), ctypes.byref(la), ctypes.byref(ms)))
            res[name.value.decode()] = (la.value, ms.value)
        return res

    def profile_reset(self):
        check_hip(_lib.hip().scd_prof_reset(self._ctx))

    def bin_info(self):
        """Per degree bin: name -> (vertices, adjacency entries, pairs
        evaluated in the last move step)."""
        names = (ctypes.c_char * 16 * 16)()
        v = np.zeros(16, np.uint64)
        e = np.zeros(16, np.uint64)
        p = np.zeros(16, np.uint64)
        k = check_hip(_lib.hip().scd_bin_info(self._ctx, 16, ctypes.cast(names, ctypes.c_void_p), ptr(v, u64p),
                                              ptr(e, u64p), ptr(p, u64p)))
        return {names[i].value.decode(): (int(v[i]), int(e[i]), int(p[i])) for i in range(k)}

    def set_profiling(self, on=True):
        check_hip(_lib.hip().scd_set_profiling(self._ctx, 1 if on else 0))

    def capture_iterations(self):
        """scd_capture_iterations: record the next iterations' hipGraphs now."""
        check_hip(_lib.hip().scd_capture_iterations(self._ctx))

    def set_phase_timing(self, on=True):
        """scd_set_phase_timing: device clock stamps at the phase boundaries
        of every iteration (captured iterations included)."""
        check_hip(_lib.hip().scd_set_phase_timing(self._ctx, 1 if on else 0))

    def last_phase_ms(self):
        """(move, statistics, removal of external edges, WCC) of the last
        iteration in ms (src/cuda.cu:347-360's split)."""
        ms = np.zeros(4, np.float64)
        check_hip(_lib.hip().scd_last_phase_ms(self._ctx, ptr(ms, f64p)))
        return tuple(float(x) for x in ms)

    def set_graph_mode(self, on=True):
        check_hip(_lib.hip().scd_set_graph_mode(self._ctx, 1 if on else 0))

    def set_counting(self, on=True):
        check_hip(_lib.hip().scd_set_counting(self._ctx, 1 if on else 0))

    def snapshot(self):
        check_hip(_lib.hip().scd_snapshot(self._ctx))

    def restore(self):
        check_hip(_lib.hip().scd_restore(self._ctx))

    def calibration_copy(self, nbytes, reps=10):
        ms = ctypes.c_double()
        check_hip(_lib.hip().scd_calibration_copy(self._ctx, nbytes, reps, ctypes.byref(ms)))
        return ms.value

    def device_bytes(self):
        return int(_lib.hip().scd_device_bytes(self._ctx))

    def eval_improvement(self, r, d_in, d_out, c_out, p_in, p_ext):
        """calculate_improvement (src/scd.cpp:173-219, alpha 1.0) per row on
        the device, by the move kernels' own device function."""
        a = [np.ascontiguousarray(x, np.int32) for x in (r, d_in, d_out, c_out)]
        f = [np.ascontiguousarray(x, np.float64) for x in (p_in, p_ext)]
        out = np.empty(len(a[0]), np.float64)
        check_hip(_lib.hip().scd_eval_improvement(self._ctx, len(out), *(ptr(x, i32p) for x in a),
                                                  *(ptr(x, f64p) for x in f), ptr(out, f64p)))
        return out

    def eval_candidate(self, size, ext_edges, int_edges, label, freq, degree, my_label, clustering_coef):
        """The candidate set-up of src/scd.cpp:280-307 plus the improvement,
        per row on the device (scd_eval_candidate)."""
        s = [np.ascontiguousarray(x, np.int32) for x in (size, ext_edges, int_edges)]
        u = [np.ascontiguousarray(x, np.uint32) for x in (label, freq, degree, my_label)]
        out = np.empty(len(s[0]), np.float64)
        check_hip(_lib.hip().scd_eval_candidate(self._ctx, len(out), *(ptr(x, i32p) for x in s),
                                                *(ptr(x, u32p) for x in u), clustering_coef, ptr(out, f64p)))
        return out


class BatchEngine:
    """Batched refinement of an Arena (scd_batch_*): every graph follows the
    reference loop with its own stopping rule; labels are local per graph."""

    def __init__(self, device=0, profile=False):
        self.eng = Engine(device=device, profile=profile)
        self.arena = None

    @property
    def _ctx(self):
        return self.eng._ctx

    def close(self):
        self.eng.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def upload(self, arena):
        g = arena.graph
        check_hip(_lib.hip().scd_batch_upload(self._ctx, arena.count, ptr(arena.vtx_off, u64p), ptr(g.indices, u32p),
                                              ptr(g.adj, u32p), ptr(g.tri, i32p), ptr(g.tri_deg, i32p),
                                              ptr(arena.cc, f64p)))
        self.arena = arena
        self.eng.n = g.n
        self.eng.graph = g

    def initial_partition(self, want_labels=True):
        lab = np.empty(self.arena.graph.n, np.uint32) if want_labels else None
        k = np.empty(self.arena.count, np.uint32)
        check_hip(_lib.hip().scd_batch_initial_partition(self._ctx, ptr(lab, u32p), ptr(k, u32p)))
        return lab, k

    def compute_wcc(self):
        w = np.empty(self.arena.count, np.float64)
        check_hip(_lib.hip().scd_batch_compute_wcc(self._ctx, ptr(w, f64p)))
        return w

    def refine(self, initial_wcc, max_iters=1000, history=False):
        G = self.arena.count
        w0 = np.ascontiguousarray(initial_wcc, np.float64)
        best = np.empty(G, np.float64)
        iters = np.empty(G, np.uint32)
        done = np.empty(G, np.float64)
        hist = np.zeros((G, max_iters), np.float64) if history else None
        check_hip(_lib.hip().scd_batch_refine(self._ctx, ptr(w0, f64p), max_iters, ptr(best, f64p), ptr(iters, u32p),
                                              ptr(done, f64p), ptr(hist, f64p)))
        return (best, iters, done, hist) if history else (best, iters, done)

    def labels(self):
        lab = np.empty(self.arena.graph.n, np.uint32)
        check_hip(_lib.hip().scd_batch_labels(self._ctx, ptr(lab, u32p)))
        return lab

    def autotune(self, cache_path=None, reps=3):
        """scd_autotune on the arena (after the initial partition)."""
        return _autotune(self._ctx, cache_path, reps)

    def get_tuning(self):
        return _get_tuning(self._ctx)

    def set_tuning(self, lds_tile=None, merge_chunk=None, wave_max=None, hub_table=None, lds_hash=None,
                   lanes=None, dense_wcc=None):
        _set_tuning(self._ctx, lds_tile, merge_chunk, wave_max, hub_table, lds_hash, lanes, dense_wcc)

    def current_labels(self):
        lab = np.empty(self.arena.graph.n, np.uint32)
        check_hip(_lib.hip().scd_batch_current_labels(self._ctx, ptr(lab, u32p)))
        return lab

    def graph_labels(self, graphs, best=True):
        """Local best (or current) labels of the listed graphs, one array
        each, through scd_batch_graphs_labels (one download for all)."""
        gs = np.ascontiguousarray(graphs, np.uint32)
        off = self.arena.vtx_off
        outs = [np.empty(int(off[g + 1] - off[g]) if g < self.arena.count else 1, np.uint32) for g in gs]
        arr = (u32p * max(len(outs), 1))(*[ptr(o, u32p) for o in outs])
        check_hip(_lib.hip().scd_batch_graphs_labels(self._ctx, len(gs), ptr(gs, u32p), 1 if best else 0, arr))
        return outs
