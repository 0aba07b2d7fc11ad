"""Binary format and convert pipeline (Graph::read src/graph.cpp:86-134,
write_edges src/convert.cpp:417-463, convert src/convert.cpp:466-537)."""
import hashlib
import struct

import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from golden_graphs import GRAPHS, MEDIUM, SMALL, graph, make_edges


def test_bowtie_bytes():
    # edges 0-1,1-2,2-0,2-3,3-4,4-2: vertex 2 has degree 4 and becomes id 0
    g = scd_amd.from_edge_pairs([(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 2)]).convert()
    words = [5, 6, 0, 4, 6, 8, 10, 12, 1, 2, 3, 4, 0, 2, 0, 1, 0, 4, 0, 3, 4, 2, 2, 2, 2, 4, 2, 2, 2, 2]
    assert g.to_bytes() == struct.pack("<30I", *words)
    assert len(g.to_bytes()) == 120


def test_triangle_free_graph_becomes_empty():
    # a 4-cycle has no triangle: every edge and vertex is dropped
    g = scd_amd.from_edge_pairs([(0, 1), (1, 2), (2, 3), (3, 0)]).convert()
    assert (g.n, g.m) == (0, 0)
    assert g.to_bytes() == struct.pack("<3I", 0, 0, 0)


def test_read_write_roundtrip(tmp_path):
    g = graph("pp4x16_s4")
    p = tmp_path / "g.bin"
    g.write(p)
    assert p.read_bytes() == g.to_bytes()
    h = scd_amd.Graph.read(p)
    assert h.to_bytes() == g.to_bytes()


def test_reader_rejects_short_and_long_files(tmp_path):
    g = graph("bowtie")
    b = g.to_bytes()
    for bad in (b[:-1], b + b"\0", b[:7], b""):
        p = tmp_path / "bad.bin"
        p.write_bytes(bad)
        with pytest.raises(scd_amd.ScdError) as ei:
            scd_amd.Graph.read(p)
        assert ei.value.code == -2
    with pytest.raises(scd_amd.ScdError):
        scd_amd.Graph.read(tmp_path / "missing.bin")


@pytest.mark.parametrize("name", SMALL + MEDIUM)
def test_convert_matches_reference_bytes(name, manifest):
    """scd_convert == the reference convert tool's output (SHA-256 recorded
    from oracle/_ref/convert) == the oracle's step-by-step restatement."""
    edges = make_edges(GRAPHS[name])
    b = edges.convert().to_bytes()
    assert hashlib.sha256(b).hexdigest() == manifest["graphs"][name]["convert_sha256"]
    if len(edges) <= 300000:
        assert orc.convert_bytes(edges) == b


def test_generators_are_deterministic():
    a = scd_amd.rmat(10, 16, 0.57, 0.19, 0.19, 7)
    b = scd_amd.rmat(10, 16, 0.57, 0.19, 0.19, 7)
    c = scd_amd.rmat(10, 16, 0.57, 0.19, 0.19, 8)
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v)
    assert not np.array_equal(a.u, c.u)
    assert len(a) == 16 << 10
    e, n, k = scd_amd.planted_random(11, 64, 20000, 4, 64, 0.15, 0.002)
    assert 64 <= n <= 20000 and 4 <= k <= 64
    assert int(max(e.u.max(), e.v.max())) < n


def test_convert_invariants():
    """convert's output: sorted duplicate-free lists without loops, every
    edge in a triangle, tri_deg == degree, degree-descending ids."""
    for name in ("rmat12", "pp8x40_s5"):
        g = graph(name)
        deg = g.degree()
        assert np.array_equal(g.tri_deg, deg)
        assert np.all(deg[:-1] >= deg[1:])
        for v in range(g.n):
            lst = g.adj[g.indices[v]:g.indices[v + 1]]
            assert np.all(np.diff(lst.astype(np.int64)) > 0) and v not in lst
        assert int(g.tri.astype(np.int64).sum()) % 6 == 0
