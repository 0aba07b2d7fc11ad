"""The convert tool on the device (scd_convert_edges): byte-identical to the
reference's convert output (the manifest's convert_sha256, written from the
reference's own convert program) and to the host restatement scd_convert,
including the edge cases of src/convert.cpp:466-537 (loops, duplicates in
both directions, triangle-free edges, sparse ids, empty results)."""
import hashlib
import struct

import numpy as np
import pytest

import scd_amd
from scd_amd import _lib
from golden_graphs import GRAPHS, LARGE, MEDIUM, SMALL, make_edges

pytestmark = pytest.mark.gpu


def sha(b):
    return hashlib.sha256(b).hexdigest()


def cc_hex(x):
    return struct.pack("<d", x).hex()


def same_graph(dev, host):
    assert dev.n == host.n and dev.m == host.m
    assert dev.to_bytes() == host.to_bytes()
    if host.n:
        assert cc_hex(dev.clustering_coef) == cc_hex(host.clustering_coef)
    else:
        assert np.isnan(dev.clustering_coef)


@pytest.mark.parametrize("name", SMALL + MEDIUM + LARGE)
def test_convert_matches_reference(engine, manifest, name):
    ent = manifest["graphs"][name]
    print(f"[gpu] convert {name}", flush=True)
    e = make_edges(GRAPHS[name])
    g = engine.convert(e)
    assert (g.n, g.m) == (ent["n"], ent["m"])
    assert sha(g.to_bytes()) == ent["convert_sha256"]
    assert cc_hex(g.clustering_coef) == ent["cc_hex"]


@pytest.mark.parametrize("pairs", [
    [],                                              # no edges
    [(3, 3), (7, 7)],                                # loops only
    [(0, 1), (1, 2), (2, 3), (3, 0)],                # a square: no triangle
    [(5, 1), (1, 5), (5, 1), (9, 9)],                # one edge, repeated both ways
    [(0, 1), (1, 2), (2, 0)],                        # one triangle
    [(2, 0), (0, 1), (1, 2), (2, 2), (1, 0), (3, 4), (4, 5), (5, 3), (0, 3)],
    [(200000000, 7), (7, 123456), (123456, 200000000), (7, 7), (8, 9)],  # sparse ids
    [(0, 0), (1, 2), (2, 3), (3, 1), (1, 4), (4, 2), (10, 11)],
])
def test_convert_edge_cases(engine, pairs):
    print(f"[gpu] convert edge case {pairs[:3]}", flush=True)
    if pairs:
        u = np.array([p[0] for p in pairs], np.uint32)
        v = np.array([p[1] for p in pairs], np.uint32)
    else:
        u = v = np.zeros(0, np.uint32)
    e = scd_amd.Edges(u, v)
    same_graph(engine.convert(e), e.convert())


@pytest.mark.parametrize("scale,ef,seed", [(8, 4, 11), (13, 8, 12), (15, 16, 13), (17, 24, 14)])
def test_convert_rmat_matches_host(engine, scale, ef, seed):
    """R-MAT edge lists of other shapes than the fixtures, against the host
    restatement."""
    print(f"[gpu] convert rmat s{scale} ef{ef} seed{seed}", flush=True)
    e = scd_amd.rmat(scale, ef, 0.57, 0.19, 0.19, seed)
    same_graph(engine.convert(e), e.convert())


def test_convert_shuffled_edge_order(engine):
    """Ids follow first appearance, so a shuffled edge list gives another
    (equally valid) numbering: host and device still agree."""
    e = make_edges(GRAPHS["rmat12"])
    rng = np.random.default_rng(5)
    p = rng.permutation(len(e))
    sw = rng.random(len(e)) < 0.5
    u = np.where(sw, e.v[p], e.u[p])
    v = np.where(sw, e.u[p], e.v[p])
    s = scd_amd.Edges(u, v)
    same_graph(engine.convert(s), s.convert())


def test_convert_device_edges_then_refine(engine, manifest):
    """Edges already in HBM; the converted graph stays uploaded and the
    device path runs on it with the reference's result."""
    import torch
    name = "c1"
    ent = manifest["graphs"][name]
    e = make_edges(GRAPHS[name])
    du = torch.from_numpy(e.u.view(np.int32)).to("cuda:0")
    dv = torch.from_numpy(e.v.view(np.int32)).to("cuda:0")
    torch.cuda.synchronize()
    n, m, cc = engine.convert_device(len(e), du.data_ptr(), dv.data_ptr())
    assert (n, m) == (ent["n"], ent["m"])
    assert cc_hex(cc) == ent["cc_hex"]
    g = engine.download_graph()
    assert sha(g.to_bytes()) == ent["convert_sha256"]
    lab, k = engine.initial_partition()
    assert hashlib.sha256(lab.tobytes()).hexdigest() == ent["labels_sha256"][0]
    w0 = engine.compute_wcc()
    best, iters, _ = engine.refine(w0)
    assert iters == ent["iterations"]
    assert hashlib.sha256(engine.current_labels().tobytes()).hexdigest() == ent["labels_sha256"][iters]
    final = struct.unpack("<d", bytes.fromhex(ent["final_wcc_hex"]))[0]
    assert best == pytest.approx(final, rel=1e-9, abs=0)


def test_convert_empty_leaves_no_graph(engine):
    e = scd_amd.Edges(np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))  # a path
    g = engine.convert(e)
    assert g.n == 0 and g.m == 0
    with pytest.raises(scd_amd.ScdError):
        engine.initial_partition()


def test_convert_limits(engine):
    with pytest.raises(scd_amd.ScdError) as ei:
        engine.convert_device(1 << 31, 0, 0)
    assert ei.value.code == _lib.SCD_EINVAL  # null edge arrays
    with pytest.raises(scd_amd.ScdError) as ei:
        engine.convert_device(1 << 31, 4096, 4096)  # rejected before any access
    assert ei.value.code == _lib.SCD_ELIMIT
