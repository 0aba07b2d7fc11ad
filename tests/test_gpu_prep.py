"""HIP engine parity (MI355X) for the per-graph set-up: triangle statistics
(src/convert.cpp:271-340) and the initial partition (src/scd.cpp:113-171),
bit-exact against the graph files' own arrays (pinned by the reference
convert), the reference fixtures and the oracle."""
import hashlib

import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from golden_graphs import LARGE, MEDIUM, SMALL, graph

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def raw_graph(edges):
    """A simple undirected CSR straight from an edge list (loops and
    duplicates dropped, ids kept, triangle-free edges kept) — input the
    convert invariants do not hold for.  tri/tri_deg are zero."""
    u = edges.u.astype(np.int64)
    v = edges.v.astype(np.int64)
    keep = u != v
    a = np.minimum(u[keep], v[keep])
    b = np.maximum(u[keep], v[keep])
    n = int(max(a.max(), b.max())) + 1
    key = np.unique(a * n + b)
    a, b = key // n, key % n
    src = np.concatenate([a, b])
    dst = np.concatenate([b, a])
    order = np.lexsort((dst, src))
    src, dst = src[order], dst[order]
    indices = np.zeros(n + 1, np.int64)
    np.add.at(indices, src + 1, 1)
    indices = np.cumsum(indices).astype(np.uint32)
    z = np.zeros(n, np.int32)
    return scd_amd.Graph(n, len(key), indices, dst.astype(np.uint32), z, z)


@pytest.mark.parametrize("name", SMALL + MEDIUM + LARGE)
def test_triangles_match_file(engine, name):
    g = graph(name)
    print(f"[gpu] triangles {name} n={g.n} m={g.m}", flush=True)
    engine.transfer_graph(g)
    tri, td = engine.count_triangles()
    assert np.array_equal(tri, g.tri) and np.array_equal(td, g.tri_deg)


@pytest.mark.parametrize("scale,seed", [(9, 1), (12, 2), (14, 3)])
def test_triangles_raw_graph(engine, scale, seed):
    """Graphs with triangle-free edges and unsorted degrees (no convert):
    t/vt against the oracle (pinned by brute force on small graphs)."""
    g = raw_graph(scd_amd.rmat(scale, 8, 0.45, 0.25, 0.15, seed))
    engine.transfer_graph(g, clustering_coef=0.0)
    tri, td = engine.count_triangles()
    otri, otd = orc.triangles(g)
    assert np.array_equal(tri, otri) and np.array_equal(td, otd)
    assert (td < g.degree()).any()  # some edges lie in no triangle


@pytest.mark.parametrize("name", SMALL + MEDIUM + LARGE)
def test_initial_partition_matches_reference(engine, manifest, name):
    ent = manifest["graphs"][name]
    g = graph(name)
    print(f"[gpu] initial partition {name}", flush=True)
    engine.transfer_graph(g)
    lab, k = engine.initial_partition()
    assert sha(lab) == ent["labels_sha256"][0]
    assert k == ent["num_comms"]
    s, e, i = engine.stats()
    assert sha(np.stack([s, e, i])) == ent["stats_sha256"][0]
    assert np.array_equal(engine.current_labels(), lab)


def test_initial_partition_rejects_degree_one(engine):
    # a pendant vertex: deg * (deg - 1.0) = 0 makes the reference's key NaN
    g = raw_graph(scd_amd.from_edge_pairs([(0, 1), (1, 2), (2, 0), (2, 3)]))
    engine.transfer_graph(g, clustering_coef=0.0)
    with pytest.raises(scd_amd.ScdError) as ei:
        engine.initial_partition()
    assert ei.value.code == -1


@pytest.mark.parametrize("name", ["pp4x16_s4", "c1", "rmat14", "c2"])
def test_whole_device_path(engine, manifest, name):
    """Everything on the device: t/vt recomputed into the device copy,
    initial partition, initial WCC and the refinement loop; result equals
    the reference's run."""
    ent = manifest["graphs"][name]
    g = graph(name)
    print(f"[gpu] whole path {name}", flush=True)
    zero = np.zeros(g.n, np.int32)
    blank = scd_amd.Graph(g.n, g.m, g.indices, g.adj, zero, zero)
    engine.transfer_graph(blank, clustering_coef=g.clustering_coef)
    engine.count_triangles(replace=True)
    lab, k = engine.initial_partition()
    assert sha(lab) == ent["labels_sha256"][0]
    w0 = engine.compute_wcc()
    best, iters, hist = engine.refine(w0)
    assert iters == ent["iterations"]
    assert sha(engine.current_labels()) == ent["labels_sha256"][iters]
    import struct
    final = struct.unpack("<d", bytes.fromhex(ent["final_wcc_hex"]))[0]
    assert best == pytest.approx(final, rel=1e-9, abs=0)


@pytest.mark.parametrize("name", ["pp8x40_s5", "c1", "rmat14", "rmat16", "c2"])
def test_triangle_split_invariance(engine, name):
    """Walking an edge from either end's out-list (scd_set_triangle_split)
    changes the work, never t / vt."""
    g = graph(name)
    engine.transfer_graph(g)
    try:
        for h in [0, 64, 100, 128, 512, 2047]:
            print(f"[gpu] triangle split {name} tc_heavy={h}", flush=True)
            engine.set_triangle_split(tc_heavy=h)
            tri, td = engine.count_triangles()
            assert np.array_equal(tri, g.tri) and np.array_equal(td, g.tri_deg), h
    finally:
        engine.set_triangle_split()


@pytest.mark.parametrize("scale,seed", [(12, 2), (15, 4)])
def test_triangle_split_raw_graph(engine, scale, seed):
    g = raw_graph(scd_amd.rmat(scale, 16, 0.57, 0.19, 0.19, seed))
    engine.transfer_graph(g, clustering_coef=0.0)
    otri, otd = orc.triangles(g)
    try:
        for h in [0, 64, 256]:
            engine.set_triangle_split(tc_heavy=h)
            tri, td = engine.count_triangles()
            assert np.array_equal(tri, otri) and np.array_equal(td, otd), h
    finally:
        engine.set_triangle_split()


def test_triangle_split_limits(engine):
    for bad in [1, 63, 2048, 5000]:
        with pytest.raises(scd_amd.ScdError):
            engine.set_triangle_split(tc_heavy=bad)
    engine.set_triangle_split()


@pytest.mark.parametrize("name,k", [("rmat14", 4), ("rmat16", 16), ("c2", 64)])
def test_wcc_triangle_split_invariance(engine, name, k):
    """The WCC's internal triangle counts (coarse random partitions, so that
    internal out-lists are long) do not depend on the walk side, and equal
    the oracle's."""
    g = graph(name)
    rng = np.random.default_rng(7)
    labels = rng.integers(0, k, g.n).astype(np.uint32)
    engine.transfer_graph(g)
    outs = []
    try:
        for h in [0, 64, 256]:
            print(f"[gpu] WCC triangle split {name} wcc_heavy={h}", flush=True)
            engine.set_triangle_split(wcc_heavy=h)
            engine.labels_to_device(labels)
            outs.append(engine.compute_wcc(details=True))
    finally:
        engine.set_triangle_split()
    for w, terms, it, itd in outs[1:]:
        assert w == outs[0][0]
        assert np.array_equal(terms, outs[0][1])
        assert np.array_equal(it, outs[0][2]) and np.array_equal(itd, outs[0][3])
    if g.n <= 100000:
        _, _, rit, ritd = orc.compute_wcc(g, labels, np.bincount(labels, minlength=k))
        assert np.array_equal(outs[0][2], rit) and np.array_equal(outs[0][3], ritd)
