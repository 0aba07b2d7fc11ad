"""Host-side graph model: the reference's binary CSR (class Graph,
src/graph.hpp:23-62) as numpy arrays, plus the deterministic generators and
the convert pipeline of libscd_host.so.
"""
import ctypes
import hashlib

import numpy as np

from . import _lib
from ._lib import check_host, i32p, ptr, u32p


class Graph:
    """Preprocessed undirected graph: n, m, indices[n+1], adj[2m] (sorted
    lists), tri[n] (2 x triangles per vertex), tri_deg[n] (edges in a
    triangle) and clustering_coef (Graph::clustering_coef,
    src/graph.cpp:57-68, summed in vertex order)."""

    def __init__(self, n, m, indices, adj, tri, tri_deg):
        self.n = int(n)
        self.m = int(m)
        self.indices = np.ascontiguousarray(indices, dtype=np.uint32)
        self.adj = np.ascontiguousarray(adj, dtype=np.uint32)
        self.tri = np.ascontiguousarray(tri, dtype=np.int32)
        self.tri_deg = np.ascontiguousarray(tri_deg, dtype=np.int32)
        if self.indices.shape != (self.n + 1,) or self.adj.shape != (2 * self.m,):
            raise ValueError("indices must hold n+1 and adj 2m entries")
        if self.tri.shape != (self.n,) or self.tri_deg.shape != (self.n,):
            raise ValueError("tri and tri_deg must hold n entries")
        self._cc = None

    @property
    def clustering_coef(self):
        if self._cc is None:
            g = self._as_c()
            self._cc = _lib.host().scd_graph_clustering_coef(ctypes.byref(g)) if self.n else float("nan")
        return self._cc

    def degree(self, v=None):
        d = np.diff(self.indices.astype(np.int64))
        return d if v is None else int(d[v])

    def _as_c(self):
        g = _lib.ScdGraphC()
        g.n, g.m = self.n, self.m
        g.indices = ptr(self.indices, u32p)
        g.adj = ptr(self.adj, u32p)
        g.tri = ptr(self.tri, i32p)
        g.tri_deg = ptr(self.tri_deg, i32p)
        return g

    @staticmethod
    def _from_c(gp):
        g = gp.contents
        n, m = g.n, g.m
        G = Graph(n, m,
                  np.ctypeslib.as_array(g.indices, (n + 1,)).copy(),
                  np.ctypeslib.as_array(g.adj, (2 * m,)).copy() if m else np.zeros(0, np.uint32),
                  np.ctypeslib.as_array(g.tri, (n,)).copy() if n else np.zeros(0, np.int32),
                  np.ctypeslib.as_array(g.tri_deg, (n,)).copy() if n else np.zeros(0, np.int32))
        _lib.host().scd_graph_free(gp)
        return G

    @staticmethod
    def read(path):
        """Graph::read (src/graph.cpp:86-134): raises ScdError on a missing,
        short or over-long file."""
        gp = ctypes.POINTER(_lib.ScdGraphC)()
        check_host(_lib.host().scd_graph_read(str(path).encode(), ctypes.byref(gp)))
        return Graph._from_c(gp)

    def write(self, path):
        g = self._as_c()
        check_host(_lib.host().scd_graph_write(ctypes.byref(g), str(path).encode()))

    def to_bytes(self):
        return (np.array([self.n, self.m], np.uint32).tobytes() + self.indices.tobytes() +
                self.adj.tobytes() + self.tri.tobytes() + self.tri_deg.tobytes())

    def sha256(self):
        return hashlib.sha256(self.to_bytes()).hexdigest()


class Edges:
    """An edge list (u[i], v[i]) in generation order."""

    def __init__(self, u, v):
        self.u = np.ascontiguousarray(u, dtype=np.uint32)
        self.v = np.ascontiguousarray(v, dtype=np.uint32)

    def __len__(self):
        return len(self.u)

    @staticmethod
    def _from_c(ep):
        e = ep.contents
        cnt = e.count
        E = Edges(np.ctypeslib.as_array(e.u, (cnt,)).copy() if cnt else np.zeros(0, np.uint32),
                  np.ctypeslib.as_array(e.v, (cnt,)).copy() if cnt else np.zeros(0, np.uint32))
        _lib.host().scd_edges_free(ep)
        return E

    def _as_c(self):
        ep = ctypes.POINTER(_lib.ScdEdgesC)()
        check_host(_lib.host().scd_edges_from_arrays(len(self.u), ptr(self.u, u32p), ptr(self.v, u32p),
                                                     ctypes.byref(ep)))
        return ep

    def convert(self):
        """The reference convert pipeline (src/convert.cpp:466-537) on this
        edge sequence: reindex, dedup, drop triangle-free edges, relabel by
        degree.  Returns the preprocessed Graph."""
        ep = self._as_c()
        gp = ctypes.POINTER(_lib.ScdGraphC)()
        try:
            check_host(_lib.host().scd_convert(ep, ctypes.byref(gp)))
        finally:
            _lib.host().scd_edges_free(ep)
        return Graph._from_c(gp)

    @staticmethod
    def read_text(path):
        """A text edge list read with the reference convert's rules
        (src/convert.cpp:74-136); raises ScdError("line <i> is corrupt")
        like the reference's reader reports it."""
        ep = ctypes.POINTER(_lib.ScdEdgesC)()
        check_host(_lib.host().scd_edges_read_text(str(path).encode(), ctypes.byref(ep)))
        return Edges._from_c(ep)

    def write_text(self, path):
        ep = self._as_c()
        try:
            check_host(_lib.host().scd_edges_write_text(ep, str(path).encode()))
        finally:
            _lib.host().scd_edges_free(ep)


def planted_partition(sizes, p_in, p_out, seed):
    sizes = np.ascontiguousarray(sizes, dtype=np.uint32)
    ep = ctypes.POINTER(_lib.ScdEdgesC)()
    check_host(_lib.host().scd_gen_planted_partition(len(sizes), ptr(sizes, u32p), p_in, p_out, seed,
                                                     ctypes.byref(ep)))
    return Edges._from_c(ep)


def planted_random(seed, n_min, n_max, k_min, k_max, p_in, p_out):
    """Planted partition with n log-uniform in [n_min, n_max], k uniform in
    [k_min, k_max], Dirichlet(1) community sizes.  Returns (edges, n, k)."""
    ep = ctypes.POINTER(_lib.ScdEdgesC)()
    n = ctypes.c_uint32()
    k = ctypes.c_uint32()
    check_host(_lib.host().scd_gen_planted_random(seed, n_min, n_max, k_min, k_max, p_in, p_out,
                                                  ctypes.byref(n), ctypes.byref(k), ctypes.byref(ep)))
    return Edges._from_c(ep), n.value, k.value


def rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1):
    ep = ctypes.POINTER(_lib.ScdEdgesC)()
    check_host(_lib.host().scd_gen_rmat(scale, edge_factor, a, b, c, seed, ctypes.byref(ep)))
    return Edges._from_c(ep)


def from_edge_pairs(pairs):
    """Edges from a sequence of (u, v) pairs."""
    a = np.asarray(pairs, dtype=np.uint32).reshape(-1, 2)
    return Edges(a[:, 0], a[:, 1])


def compress_count(labels):
    """The community count the reference prints (compress_labels,
    src/scd.cpp:32-55, whose inverted map test makes it n minus the number
    of distinct valid labels); scd_labels_printed_count."""
    labels = np.ascontiguousarray(labels, dtype=np.uint32)
    return int(_lib.host().scd_labels_printed_count(len(labels), ptr(labels, u32p)))


def compact_labels(labels):
    """The compaction compress_labels intends (scd_labels_compact): distinct
    valid labels numbered 0.. in order of first appearance, every INVALID
    vertex a community of its own.  Returns (labels, number of communities)."""
    labels = np.ascontiguousarray(labels, dtype=np.uint32)
    out = np.empty_like(labels)
    k = ctypes.c_uint32()
    check_host(_lib.host().scd_labels_compact(len(labels), ptr(labels, u32p), ptr(out, u32p), ctypes.byref(k)))
    return out, k.value


def write_labels(path, labels):
    """Labels file [u32 n][u32 communities][u32 compact labels n]."""
    labels = np.ascontiguousarray(labels, dtype=np.uint32)
    check_host(_lib.host().scd_labels_write(str(path).encode(), len(labels), ptr(labels, u32p)))


def read_labels(path):
    """(labels, number of communities) from a labels file."""
    n = ctypes.c_uint32()
    k = ctypes.c_uint32()
    lp = u32p()
    check_host(_lib.host().scd_labels_read(str(path).encode(), ctypes.byref(n), ctypes.byref(k), ctypes.byref(lp)))
    try:
        out = np.ctypeslib.as_array(lp, (n.value,)).copy() if n.value else np.zeros(0, np.uint32)
    finally:
        _lib.host().scd_labels_free(lp)
    return out, k.value


class Arena:
    """Several graphs as one disjoint union (scd_graph_concat): `graph` is
    the union, graph i owns vertices [vtx_off[i], vtx_off[i+1]), cc[i] is its
    clustering coefficient."""

    def __init__(self, graph, vtx_off, cc, seeds=None):
        self.graph = graph
        self.vtx_off = np.ascontiguousarray(vtx_off, dtype=np.uint64)
        self.cc = np.ascontiguousarray(cc, dtype=np.float64)
        self.seeds = seeds

    @property
    def count(self):
        return len(self.cc)

    def member(self, i):
        """Graph i on its own (ids 0-based)."""
        a, b = int(self.vtx_off[i]), int(self.vtx_off[i + 1])
        g = self.graph
        e0, e1 = int(g.indices[a]), int(g.indices[b])
        return Graph(b - a, (e1 - e0) // 2, g.indices[a:b + 1] - g.indices[a], g.adj[e0:e1] - np.uint32(a),
                     g.tri[a:b], g.tri_deg[a:b])

    @staticmethod
    def concat(graphs):
        arr = (_lib.ScdGraphC * len(graphs))()
        for i, g in enumerate(graphs):
            arr[i] = g._as_c()
        ptrs = (ctypes.POINTER(_lib.ScdGraphC) * len(graphs))(*[ctypes.pointer(arr[i]) for i in range(len(graphs))])
        gp = ctypes.POINTER(_lib.ScdGraphC)()
        off = np.zeros(len(graphs) + 1, np.uint64)
        check_host(_lib.host().scd_graph_concat(len(graphs), ptrs, ctypes.byref(gp), ptr(off, _lib.u64p)))
        return Arena(Graph._from_c(gp), off, [g.clustering_coef for g in graphs])


def planted_batch(count, seed_base=1, n_min=64, n_max=20000, k_min=4, k_max=64, p_in=0.15, p_out=0.002):
    """`count` planted-partition graphs (n log-uniform in [n_min, n_max], k
    uniform in [k_min, k_max], Dirichlet(1) sizes), each converted, drawn in
    parallel and concatenated into an Arena."""
    gp = ctypes.POINTER(_lib.ScdGraphC)()
    off = np.zeros(count + 1, np.uint64)
    cc = np.zeros(count, np.float64)
    seeds = np.zeros(count, np.uint64)
    check_host(_lib.host().scd_gen_planted_batch(seed_base, count, n_min, n_max, k_min, k_max, p_in, p_out,
                                                 ctypes.byref(gp), ptr(off, _lib.u64p), ptr(cc, _lib.f64p),
                                                 ptr(seeds, _lib.u64p)))
    return Arena(Graph._from_c(gp), off, cc, seeds)
