"""ctypes binding of the CPU oracle (oracle/liboracle.so) — test
infrastructure only: the checker the HIP engine is compared against."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

u32p = ctypes.POINTER(ctypes.c_uint32)
i32p = ctypes.POINTER(ctypes.c_int32)
f64p = ctypes.POINTER(ctypes.c_double)
u8pp = ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))
INVALID = 0xFFFFFFFF


class OrcRun(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("iterations", ctypes.c_uint32), ("num_communities", ctypes.c_uint32),
                ("cc", ctypes.c_double), ("initial_wcc", ctypes.c_double), ("final_wcc", ctypes.c_double),
                ("wcc", f64p), ("labels", u32p), ("best_labels", u32p), ("refine_ms", ctypes.c_double),
                ("init_ms", ctypes.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = ctypes.CDLL(path)
        sig = {
            "orc_clustering_coef": (ctypes.c_double, [ctypes.c_uint32, u32p, i32p]),
            "orc_collect_stats": (None, [ctypes.c_uint32, u32p, u32p, u32p, ctypes.c_uint32, i32p, i32p, i32p]),
            "orc_initial_partition": (ctypes.c_uint32, [ctypes.c_uint32, u32p, u32p, i32p, u32p]),
            "orc_improvement": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "orc_candidate_improvement": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                            ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                            ctypes.c_uint32, ctypes.c_double, ctypes.c_double]),
            "orc_improve": (None, [ctypes.c_uint32, u32p, u32p, u32p, i32p, i32p, i32p, ctypes.c_double,
                                   ctypes.c_double, u32p]),
            "orc_compute_wcc": (ctypes.c_double, [ctypes.c_uint32, u32p, u32p, i32p, i32p, u32p, i32p,
                                                  ctypes.c_double, f64p, i32p, i32p]),
            "orc_neighbor_label_pairs": (ctypes.c_uint64, [ctypes.c_uint32, u32p, u32p, u32p, u32p, u32p, u32p,
                                                           u32p]),
            "orc_triangles": (None, [ctypes.c_uint32, u32p, u32p, i32p, i32p]),
            "orc_convert": (ctypes.c_int, [ctypes.c_uint64, u32p, u32p, u8pp, ctypes.POINTER(ctypes.c_uint64)]),
            "orc_free": (None, [ctypes.c_void_p]),
            "orc_compress_count": (ctypes.c_uint32, [ctypes.c_uint32, u32p]),
            "orc_run": (ctypes.c_int, [ctypes.c_uint32, u32p, u32p, i32p, i32p, ctypes.c_uint32,
                                       ctypes.POINTER(OrcRun)]),
            "orc_run_free": (None, [ctypes.POINTER(OrcRun)]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def P(a, t):
    return None if a is None else a.ctypes.data_as(t)


def clustering_coef(g):
    return lib().orc_clustering_coef(g.n, P(g.indices, u32p), P(g.tri, i32p))


def initial_partition(g):
    labels = np.empty(g.n, np.uint32)
    k = lib().orc_initial_partition(g.n, P(g.indices, u32p), P(g.adj, u32p), P(g.tri, i32p), P(labels, u32p))
    return labels, k


def collect_stats(g, labels, num_comms):
    s = np.zeros(max(num_comms, 1), np.int32)
    e = np.zeros_like(s)
    i = np.zeros_like(s)
    labels = np.ascontiguousarray(labels, np.uint32)
    lib().orc_collect_stats(g.n, P(g.indices, u32p), P(g.adj, u32p), P(labels, u32p), num_comms,
                            P(s, i32p), P(e, i32p), P(i, i32p))
    return s[:num_comms], e[:num_comms], i[:num_comms]


def improvement(r, d_in, d_out, c_out, p_in, p_ext, alpha=1.0):
    return lib().orc_improvement(r, d_in, d_out, c_out, p_in, p_ext, alpha)


def candidate_improvement(size, ext_edges, int_edges, label, freq, degree, my_label, cc, alpha=1.0):
    return lib().orc_candidate_improvement(size, ext_edges, int_edges, label, freq, degree, my_label, cc, alpha)


def improve(g, labels, sizes, ext, inte, cc):
    labels = np.ascontiguousarray(labels, np.uint32)
    out = np.empty(g.n, np.uint32)
    lib().orc_improve(g.n, P(g.indices, u32p), P(g.adj, u32p), P(labels, u32p),
                      P(np.ascontiguousarray(sizes, np.int32), i32p), P(np.ascontiguousarray(ext, np.int32), i32p),
                      P(np.ascontiguousarray(inte, np.int32), i32p), cc, 1.0, P(out, u32p))
    return out


def compute_wcc(g, labels, sizes):
    labels = np.ascontiguousarray(labels, np.uint32)
    sizes = np.ascontiguousarray(sizes, np.int32)
    terms = np.empty(g.n, np.float64)
    it = np.empty(g.n, np.int32)
    itd = np.empty(g.n, np.int32)
    w = lib().orc_compute_wcc(g.n, P(g.indices, u32p), P(g.adj, u32p), P(g.tri, i32p), P(g.tri_deg, i32p),
                              P(labels, u32p), P(sizes, i32p), 1.0, P(terms, f64p), P(it, i32p), P(itd, i32p))
    return w, terms, it, itd


def neighbor_label_pairs(g, labels):
    E = 2 * g.m
    labels = np.ascontiguousarray(labels, np.uint32)
    srt = np.empty(max(E, 1), np.uint32)
    off = np.empty(g.n + 1, np.uint32)
    rl = np.empty(max(E, 1), np.uint32)
    rc = np.empty(max(E, 1), np.uint32)
    tot = lib().orc_neighbor_label_pairs(g.n, P(g.indices, u32p), P(g.adj, u32p), P(labels, u32p), P(srt, u32p),
                                         P(off, u32p), P(rl, u32p), P(rc, u32p))
    return srt[:E], off, rl[:tot].copy(), rc[:tot].copy()


def triangles(g):
    tri = np.empty(g.n, np.int32)
    td = np.empty(g.n, np.int32)
    lib().orc_triangles(g.n, P(g.indices, u32p), P(g.adj, u32p), P(tri, i32p), P(td, i32p))
    return tri, td


def convert_bytes(edges):
    out = ctypes.POINTER(ctypes.c_uint8)()
    ln = ctypes.c_uint64()
    rc = lib().orc_convert(len(edges.u), P(edges.u, u32p), P(edges.v, u32p), ctypes.byref(out), ctypes.byref(ln))
    if rc != 0:
        raise MemoryError("orc_convert")
    b = ctypes.string_at(out, ln.value)
    lib().orc_free(out)
    return b


def compress_count(labels):
    labels = np.ascontiguousarray(labels, np.uint32)
    return lib().orc_compress_count(len(labels), P(labels, u32p))


def run(g, max_iters=1000):
    """The driver of src/main.cpp:35-93: dict with cc, initial_wcc, wcc per
    iteration, labels per iteration ((iters+1) x n), best labels, final WCC,
    printed community count and timings."""
    r = OrcRun()
    rc = lib().orc_run(g.n, P(g.indices, u32p), P(g.adj, u32p), P(g.tri, i32p), P(g.tri_deg, i32p), max_iters,
                       ctypes.byref(r))
    if rc != 0:
        raise MemoryError("orc_run")
    it = r.iterations
    res = {
        "n": r.n, "iterations": it, "num_communities": r.num_communities, "cc": r.cc,
        "initial_wcc": r.initial_wcc, "final_wcc": r.final_wcc,
        "wcc": np.ctypeslib.as_array(r.wcc, (it,)).copy() if it else np.zeros(0),
        "labels": np.ctypeslib.as_array(r.labels, ((it + 1) * r.n,)).reshape(it + 1, r.n).copy(),
        "best_labels": np.ctypeslib.as_array(r.best_labels, (r.n,)).copy(),
        "refine_ms": r.refine_ms, "init_ms": r.init_ms,
    }
    lib().orc_run_free(ctypes.byref(r))
    return res
