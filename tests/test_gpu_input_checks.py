"""Uploads reject graphs whose adjacency entries, CSR offsets or labels would
make a kernel index out of range (SCD_EINVAL), and the context stays usable:
scd_upload, scd_batch_upload (lists leaving their graph), the device-resident
parts of scd_batch_upload_parts, and scd_server_submit."""
import ctypes

import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from scd_amd import _lib
from golden_graphs import graph

pytestmark = pytest.mark.gpu

EINVAL = -1
ESTATE = -6


def damaged(g, where, value):
    adj = g.adj.copy()
    adj[where] = value
    return scd_amd.Graph(g.n, g.m, g.indices.copy(), adj, g.tri.copy(), g.tri_deg.copy())


def still_works(eng, g):
    eng.transfer_graph(g)
    lab, k = orc.initial_partition(g)
    eng.labels_to_device(lab)
    s, e, i = eng.stats()
    ws, we, wi = orc.collect_stats(g, lab, k)
    assert np.array_equal(s, ws) and np.array_equal(e, we) and np.array_equal(i, wi)


@pytest.mark.parametrize("value", ["n", "max"])
@pytest.mark.parametrize("where", [0, "mid", "last"])
def test_upload_rejects_out_of_range_entry(engine, where, value):
    g = graph("pp8x40_s5")
    pos = {0: 0, "mid": len(g.adj) // 2, "last": len(g.adj) - 1}[where]
    bad = damaged(g, pos, g.n if value == "n" else 0xFFFFFFFF)
    print(f"[gpu] upload with adj[{pos}] = {bad.adj[pos]} (n = {g.n})", flush=True)
    with pytest.raises(scd_amd.ScdError) as ei:
        engine.transfer_graph(bad)
    assert ei.value.code == EINVAL and "adjacency" in ei.value.message
    with pytest.raises(scd_amd.ScdError):
        engine.improve_labels()  # no graph is left behind
    still_works(engine, g)


def test_batch_upload_rejects_list_leaving_its_graph():
    a = scd_amd.Arena.concat([graph("pp4x16_s4"), graph("k9"), graph("pp8x40_s5")])
    g = a.graph
    lo, hi = int(a.vtx_off[1]), int(a.vtx_off[2])
    adj = g.adj.copy()
    adj[g.indices[lo]] = hi  # graph 1's first vertex points into graph 2 (a valid vertex id)
    bad = scd_amd.Arena(scd_amd.Graph(g.n, g.m, g.indices, adj, g.tri, g.tri_deg), a.vtx_off, a.cc)
    with scd_amd.BatchEngine(device=0) as be:
        print("[gpu] batch upload with a cross-graph edge", flush=True)
        with pytest.raises(scd_amd.ScdError) as ei:
            be.upload(bad)
        assert ei.value.code == EINVAL
        be.upload(a)
        be.initial_partition(want_labels=False)
        best, iters, _ = be.refine(be.compute_wcc())
        assert (iters >= 1).all()


class _Hip:
    def __init__(self):
        self.lib = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.lib.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(a.nbytes, 4))) == 0
        assert self.lib.hipMemcpy(p, ctypes.c_void_p(a.ctypes.data), ctypes.c_size_t(a.nbytes), 1) == 0
        self.ptrs.append(p)
        return p.value

    def free(self):
        self.lib.hipDeviceSynchronize()
        for p in self.ptrs:
            self.lib.hipFree(p)


class ScdPartC(ctypes.Structure):
    _fields_ = [("n", ctypes.c_uint32), ("m", ctypes.c_uint32), ("indices", ctypes.c_void_p),
                ("adj", ctypes.c_void_p), ("tri", ctypes.c_void_p), ("tri_deg", ctypes.c_void_p),
                ("clustering_coef", ctypes.c_double)]


@pytest.mark.parametrize("fault", ["none", "adj", "indices_first", "indices_last", "indices_order", "label"])
def test_upload_parts_checks_each_part(fault):
    gs = [graph("pp4x16_s4"), graph("pp8x40_s5")]
    hip = _Hip()
    eng = scd_amd.Engine(device=0)
    try:
        parts = (ScdPartC * 2)()
        labels = []
        for k, g in enumerate(gs):
            idx, adj = g.indices.copy(), g.adj.copy()
            lab, _ = orc.initial_partition(g)
            if k == 1:
                if fault == "adj":
                    adj[7] = g.n
                elif fault == "indices_first":
                    idx[0] = 1
                elif fault == "indices_last":
                    idx[g.n] = idx[g.n] - 1
                elif fault == "indices_order":
                    idx[3], idx[4] = idx[4], idx[3]
                elif fault == "label":
                    lab[5] = g.n
            parts[k] = ScdPartC(g.n, g.m, hip.put(idx), hip.put(adj), hip.put(g.tri), hip.put(g.tri_deg),
                                g.clustering_coef)
            labels.append(hip.put(lab))
        cur = (ctypes.c_void_p * 2)(*labels)
        print(f"[gpu] upload_parts fault={fault}", flush=True)
        rc = _lib.hip().scd_batch_upload_parts(eng._ctx, 2, ctypes.cast(parts, ctypes.c_void_p),
                                               ctypes.cast(cur, ctypes.c_void_p), None)
        if fault == "none":
            assert rc == 0, _lib.hip().scd_last_error().decode()
        else:
            assert rc == EINVAL, (rc, _lib.hip().scd_last_error().decode())
            assert b"out of range" in _lib.hip().scd_last_error()
        still_works(eng, gs[0])
    finally:
        eng.close()
        hip.free()


def test_server_submit_rejects_out_of_range_entry():
    g = graph("pp8x40_s5")
    with scd_amd.Server(device=0) as srv:
        with pytest.raises(scd_amd.ScdError) as ei:
            srv.submit(damaged(g, len(g.adj) // 3, g.n + 5))
        assert ei.value.code == EINVAL
        rid = srv.submit(g)
        srv.run()
        assert srv.result(rid)["status"] == 0


def test_snapshot_does_not_outlive_its_graph(engine):
    """A snapshot belongs to the graph it was taken on: after another graph
    is uploaded, scd_restore reports that no snapshot exists instead of
    copying the old graph's labels (which may exceed the new n) into it."""
    big, small = graph("c1"), graph("pp8x40_s5")
    assert small.n < big.n
    engine.transfer_graph(big)
    lab, _ = orc.initial_partition(big)
    engine.labels_to_device(lab)
    engine.snapshot()
    engine.restore()  # same graph: fine
    engine.transfer_graph(small)
    engine.labels_to_device(orc.initial_partition(small)[0])
    with pytest.raises(scd_amd.ScdError) as ei:
        engine.restore()
    assert ei.value.code == ESTATE
    still_works(engine, small)
