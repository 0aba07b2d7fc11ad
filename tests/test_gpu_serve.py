"""Serving loop (MI355X): graphs arriving over time, admitted within an HBM
budget and refined with continuous batching, each get exactly the
single-graph path's result — best labels, best WCC, initial WCC, iteration
count — whether served alone or in a mix."""
import numpy as np
import pytest

import scd_amd
from golden_graphs import graph

pytestmark = pytest.mark.gpu

# one R-MAT scale 18 graph (n 104,290, m 3.4 M) among small planted and
# R-MAT graphs, as in the served workload's 90 / 10 mix
MIX_NAMES = ["pp4x16_s4", "c1", "rmat10", "rmat12", "pprand_s6", "er24_s1", "k9", "rmat18"]


@pytest.fixture(scope="module")
def mix():
    gs = [(name, graph(name)) for name in MIX_NAMES]
    arena = scd_amd.planted_batch(12, seed_base=101, n_min=64, n_max=3000, k_min=4, k_max=32,
                                  p_in=0.15, p_out=0.002)
    gs += [(f"planted{i}", arena.member(i)) for i in range(arena.count)]
    return gs


@pytest.fixture(scope="module")
def plain(engine, mix):
    out = {}
    for name, g in mix:
        engine.transfer_graph(g)
        engine.initial_partition(want_labels=False)
        w0 = engine.compute_wcc()
        best, iters, _ = engine.refine(w0)
        out[name] = (w0, best, iters, engine.labels_from_device())
    return out


def check(name, res, ref):
    w0, best, iters, labels = ref
    assert res["status"] == 0, name
    assert res["initial_wcc"] == w0, f"{name}: initial WCC"
    assert res["iterations"] == iters, f"{name}: iterations {res['iterations']} vs {iters}"
    assert res["best_wcc"] == best, f"{name}: best WCC"
    assert np.array_equal(res["labels"], labels), f"{name}: best labels"
    assert res["arrival_ms"] <= res["admit_ms"] <= res["start_ms"] <= res["done_ms"]


@pytest.mark.parametrize("graph_mode", [False, True])
def test_mix_equals_plain_path(mix, plain, graph_mode):
    rng = np.random.default_rng(5)
    arrivals = np.sort(rng.uniform(0, 40, len(mix)))
    order = rng.permutation(len(mix))
    print(f"[gpu] serving {len(mix)} graphs, graph mode {graph_mode}", flush=True)
    with scd_amd.Server(device=0, join_wait_ms=2.0, graph_mode=graph_mode) as srv:
        ids = {}
        for t, i in zip(arrivals, order):
            name, g = mix[i]
            ids[name] = srv.submit(g, arrival_ms=float(t))
        srv.run()
        for name, rid in ids.items():
            check(name, srv.result(rid), plain[name])
        st = srv.stats()
    assert st["joins"] >= 2 and st["peak_graphs"] >= 2, st


def test_alone_equals_in_mix(mix, plain):
    for name, g in mix[:3]:
        with scd_amd.Server(device=0) as srv:
            rid = srv.submit(g)
            srv.run()
            check(name, srv.result(rid), plain[name])


def test_budget_and_graph_cap(mix, plain):
    """A budget that holds only the largest graph serves the graphs one or
    two at a time; a graph over the budget fails alone, the others finish."""
    fp = {name: scd_amd.Server.footprint(g.n, g.m) for name, g in mix}
    big = max(fp, key=fp.get)
    budget = fp[big] + 1
    with scd_amd.Server(device=0, hbm_budget=budget, max_graphs=3, join_wait_ms=0.0) as srv:
        ids = {name: srv.submit(g, arrival_ms=0.0) for name, g in mix}
        srv.run()
        for name, rid in ids.items():
            check(name, srv.result(rid), plain[name])
        st = srv.stats()
    assert st["peak_graphs"] <= 3
    assert st["peak_bytes"] <= budget
    with scd_amd.Server(device=0, hbm_budget=fp[big] - 1) as srv:
        ids = {name: srv.submit(g) for name, g in mix}
        srv.run()
        for name, rid in ids.items():
            r = srv.result(rid)
            if name == big:
                assert r["status"] == scd_amd._lib.SCD_ELIMIT and r["labels"] is None
            else:
                check(name, r, plain[name])


def test_submit_rejects_bad_graphs():
    with scd_amd.Server(device=0) as srv:
        g = scd_amd.from_edge_pairs([(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 2)]).convert()
        bad = scd_amd.Graph(g.n, g.m, g.indices.copy(), g.adj, g.tri, g.tri_deg)
        bad.indices[-1] += 2
        with pytest.raises(scd_amd.ScdError):
            srv.submit(bad)


@pytest.mark.parametrize("intake_graphs,intake_mb,graph_mode", [(1, 0, True), (3, 0, False), (0, 1, True)])
def test_intake_batches_equal_plain_path(mix, plain, intake_graphs, intake_mb, graph_mode):
    """Small intake batches (one graph, three graphs, 1 MiB of footprint):
    many more preparations and arena assemblies, overlapping with captured
    or eager iterations — still each graph's plain result."""
    print(f"[gpu] serving with intake {intake_graphs} graphs / {intake_mb} MiB, graph mode {graph_mode}", flush=True)
    with scd_amd.Server(device=0, join_wait_ms=0.0, graph_mode=graph_mode, intake_graphs=intake_graphs,
                        intake_bytes=intake_mb << 20) as srv:
        ids = {name: srv.submit(g, arrival_ms=0.0) for name, g in mix}
        srv.run()
        for name, rid in ids.items():
            check(name, srv.result(rid), plain[name])
        assert srv.stats()["joins"] >= 2


def test_runs_again_with_pooled_blocks(mix, plain):
    """One server, two runs: the second run's graphs take the device blocks
    that the first run's graphs returned to the pool.  Every result equals
    the plain path's."""
    with scd_amd.Server(device=0, join_wait_ms=0.0) as srv:
        first = {name: srv.submit(g, arrival_ms=0.0) for name, g in mix}
        srv.run()
        second = {name: srv.submit(g, arrival_ms=0.0) for name, g in mix}
        srv.run()
        for ids in (first, second):
            for name, rid in ids.items():
                check(name, srv.result(rid), plain[name])


def test_one_graph_intake_batches_over_time(mix, plain):
    """Graphs arriving one by one as one-graph intake batches: the ones that
    find nothing refining are prepared in the main arena, the others in the
    preparation context while the arena refines; every result equals the
    plain path's."""
    with scd_amd.Server(device=0, join_wait_ms=0.0, intake_graphs=1) as srv:
        ids = {name: srv.submit(g, arrival_ms=0.5 * i) for i, (name, g) in enumerate(mix)}
        srv.run()
        st = srv.stats()
        for name, rid in ids.items():
            check(name, srv.result(rid), plain[name])
    assert st["joins"] >= 2
