"""R-MAT scale 24 (n 3,966,900, m 212,844,450) on the MI355X against the
reference's own code.

The fixture (tests/golden/manifest.json, entry c3) holds what the reference
produced for this recipe: the SHA-256 of its convert tool's output file, and,
from src/{graph,scd,wcc}.cpp driven like src/main.cpp with one thread, the
SHA-256 of the labels and community statistics of the initial partition and
of every iteration, and the clustering coefficient and WCC bit patterns.  The
edge list is generated here (host generator), converted on the device,
downloaded, and then every stage runs on the device: triangle counts,
initial partition, initial WCC, the refinement iterations."""
import hashlib
import struct

import numpy as np
import pytest

import scd_amd
from golden_graphs import GRAPHS, make_edges

pytestmark = pytest.mark.gpu

WCC_RTOL = 1e-9


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unhex(h):
    return struct.unpack("<d", bytes.fromhex(h))[0]


@pytest.fixture(scope="module")
def c3(manifest):
    if "c3" not in manifest["graphs"]:
        pytest.fail("tests/golden/manifest.json has no c3 entry (make_golden.py c3)")
    ent = manifest["graphs"]["c3"]
    print("[gpu] c3: generating R-MAT scale 24 edges", flush=True)
    edges = make_edges(GRAPHS["c3"])
    eng = scd_amd.Engine(device=0)
    print(f"[gpu] c3: device convert of {len(edges)} edges", flush=True)
    g = eng.convert(edges)
    del edges
    yield eng, g, ent
    eng.close()


def test_c3_device_convert_bytes_equal_reference_convert(c3):
    eng, g, ent = c3
    assert (g.n, g.m) == (ent["n"], ent["m"])
    assert g.sha256() == ent["convert_sha256"]
    assert struct.pack("<d", g.clustering_coef).hex() == ent["cc_hex"]


def test_c3_triangles_equal_reference_convert(c3):
    eng, g, ent = c3
    print("[gpu] c3: triangle counts", flush=True)
    eng.transfer_graph(g)
    tri, td = eng.count_triangles()
    assert sha(tri) == ent["tri_sha256"] and sha(td) == ent["tri_deg_sha256"]


def test_c3_initial_partition_and_refinement_equal_reference(c3):
    eng, g, ent = c3
    print("[gpu] c3: initial partition", flush=True)
    eng.transfer_graph(g)
    lab, k = eng.initial_partition()
    assert k == ent["num_comms"]
    assert sha(lab) == ent["labels_sha256"][0]
    assert sha(np.stack(eng.stats())) == ent["stats_sha256"][0]
    best = eng.compute_wcc()
    assert best == pytest.approx(unhex(ent["initial_wcc_hex"]), rel=WCC_RTOL, abs=0)
    it = 0
    while True:
        it += 1
        print(f"[gpu] c3: iteration {it}", flush=True)
        wcc, lab = eng.improve_labels(return_labels=True)
        assert sha(lab) == ent["labels_sha256"][it], f"labels differ at iteration {it}"
        assert sha(np.stack(eng.stats())) == ent["stats_sha256"][it], f"statistics differ at iteration {it}"
        assert wcc == pytest.approx(unhex(ent["wcc_hex"][it - 1]), rel=WCC_RTOL, abs=0)
        if wcc > best:
            eng.store_labels()
        if wcc < best * 1.01:
            break
        best = wcc
    assert it == ent["iterations"]
    assert best == pytest.approx(unhex(ent["final_wcc_hex"]), rel=WCC_RTOL, abs=0)
    assert scd_amd.compress_count(eng.labels_from_device()) == ent["printed_num_communities"]
