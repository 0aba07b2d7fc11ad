"""Community labels on the host: the count the reference prints (against the
oracle's compress_labels restatement, itself pinned by the reference
fixtures), the corrected compaction (checked by its defining properties) and
the labels file."""
import ctypes
import os

import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from conftest import GOLDEN
from scd_amd import _lib

INVALID = 0xFFFFFFFF


def random_labels(rng, n, k, p_invalid):
    lab = rng.integers(0, k, n).astype(np.uint32) * np.uint32(7) + np.uint32(3)  # sparse label values
    lab[rng.random(n) < p_invalid] = INVALID
    return lab


@pytest.mark.parametrize("seed,n,k,p", [(1, 0, 1, 0.0), (2, 1, 1, 0.0), (3, 1, 1, 1.0), (4, 500, 20, 0.1),
                                        (5, 3000, 3000, 0.3), (6, 2000, 1, 0.0)])
def test_printed_count_matches_oracle(seed, n, k, p):
    lab = random_labels(np.random.default_rng(seed), n, k, p)
    assert scd_amd.compress_count(lab) == orc.compress_count(lab)


@pytest.mark.parametrize("seed,n,k,p", [(7, 1000, 30, 0.2), (8, 50, 50, 0.0), (9, 40, 5, 1.0), (10, 0, 1, 0.0)])
def test_compaction_properties(seed, n, k, p):
    lab = random_labels(np.random.default_rng(seed), n, k, p)
    out, nc = scd_amd.compact_labels(lab)
    valid = lab != INVALID
    # the same partition: equal compact labels <=> equal valid input labels
    for a in range(0, n, max(1, n // 50)):
        same_in = valid & valid[a] & (lab == lab[a])
        same_out = out == out[a]
        if valid[a]:
            assert np.array_equal(same_in, same_out)
        else:  # an INVALID vertex is a community of its own
            assert same_out.sum() == 1
    # numbered 0.. in order of first appearance
    _, first = np.unique(out, return_index=True)
    assert np.array_equal(out[np.sort(first)], np.arange(len(first)))
    assert nc == len(np.unique(lab[valid])) + int((~valid).sum())
    assert nc == (out.max() + 1 if n else 0)


def test_labels_file_round_trip(tmp_path):
    lab = random_labels(np.random.default_rng(11), 777, 40, 0.1)
    path = tmp_path / "l.bin"
    scd_amd.write_labels(path, lab)
    got, k = scd_amd.read_labels(path)
    exp, ek = scd_amd.compact_labels(lab)
    assert k == ek and np.array_equal(got, exp)
    raw = path.read_bytes()
    assert len(raw) == 8 + 4 * 777
    for bad in (raw[:-1], raw + b"\0", raw[:8] + (np.uint32(ek).tobytes()) + raw[12:]):
        (tmp_path / "bad.bin").write_bytes(bad)
        with pytest.raises(scd_amd.ScdError):
            scd_amd.read_labels(tmp_path / "bad.bin")


class GraphHead(ctypes.Structure):
    """The leading members of the reference's Graph (src/graph.hpp:25-27); the
    call surface's compress_labels reads only n."""
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("clustering_coef", ctypes.c_double),
                ("p", ctypes.c_void_p * 4)]


def test_refapi_compress_labels_equals_reference_table():
    """libscd_refapi.so's compress_labels (the reference name, C++ linkage)
    relabels every array of tests/golden/compress_table.npz -- written by the
    reference's own compress_labels, src/scd.cpp:32-55 -- to the same entries
    and count (n minus the number of distinct labels, INVALID counted apart)."""
    L = ctypes.CDLL(os.path.join(_lib.LIB_DIR, "libscd_refapi.so"))
    f = L._Z15compress_labelsPK5GraphPjS2_
    f.restype = None
    f.argtypes = [ctypes.POINTER(GraphHead), ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32)]
    t = np.load(os.path.join(GOLDEN, "compress_table.npz"))
    off = t["offsets"]
    for k in range(len(t["counts"])):
        a = np.ascontiguousarray(t["labels_in"][off[k]:off[k + 1]], np.uint32)
        head = GraphHead(len(a), 0, 0.0)
        count = ctypes.c_uint32()
        f(ctypes.byref(head), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ctypes.byref(count))
        assert count.value == t["counts"][k], k
        assert np.array_equal(a, t["labels_out"][off[k]:off[k + 1]]), k
        assert count.value == orc.compress_count(t["labels_in"][off[k]:off[k + 1]])
