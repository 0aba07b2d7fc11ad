"""Batched arenas (MI355X): every graph of a batch refined together gets
exactly what the single-graph path gives it alone — initial labels, label
counts, per-iteration WCC bits, iteration count, best WCC and best labels —
and the single-graph path itself matches the oracle."""
import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from golden_graphs import graph

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    # 48 graphs from the batched workload's distribution, capped at 3,000
    # vertices so the per-graph reference runs stay short
    return scd_amd.planted_batch(48, seed_base=7, n_min=64, n_max=3000, k_min=4, k_max=64, p_in=0.15, p_out=0.002)


@pytest.fixture(scope="module")
def wide_batch():
    # the batched workload's whole size range: n log-uniform in [64, 20000]
    return scd_amd.planted_batch(32, seed_base=4001, n_min=64, n_max=20000, k_min=4, k_max=64, p_in=0.15,
                                 p_out=0.002)


@pytest.mark.parametrize("which", ["small", "wide"])
def test_batch_equals_single_graph_runs(engine, small_batch, wide_batch, which):
    a = small_batch if which == "small" else wide_batch
    if which == "wide":
        assert a.count == 32 and int(np.diff(a.vtx_off).max()) > 5000
    print(f"[gpu] batch of {a.count}: n={a.graph.n} m={a.graph.m}", flush=True)
    with scd_amd.BatchEngine(device=0) as be:
        be.upload(a)
        lab0, k = be.initial_partition()
        w0 = be.compute_wcc()
        best, iters, done, hist = be.refine(w0, max_iters=100, history=True)
        lab_best = be.labels()
    for i in range(a.count):
        g = a.member(i)
        lo, hi = int(a.vtx_off[i]), int(a.vtx_off[i + 1])
        engine.transfer_graph(g)
        l0, ki = engine.initial_partition()
        assert ki == k[i] and np.array_equal(l0, lab0[lo:hi]), f"graph {i}: initial partition"
        wi = engine.compute_wcc()
        assert wi == w0[i], f"graph {i}: initial WCC"
        bi, it, h = engine.refine(wi)
        assert it == iters[i], f"graph {i}: iterations {it} vs {iters[i]}"
        assert np.array_equal(h, hist[i, :it]), f"graph {i}: WCC per iteration"
        assert bi == best[i]
        assert np.array_equal(engine.labels_from_device(), lab_best[lo:hi]), f"graph {i}: best labels"
    assert np.all(done > 0)


def test_batch_members_match_oracle(small_batch):
    """A few members through the oracle's driver (the CPU restatement of
    src/main.cpp): same iteration count, best labels and WCC."""
    a = small_batch
    with scd_amd.BatchEngine(device=0) as be:
        be.upload(a)
        be.initial_partition()
        best, iters, done = be.refine(be.compute_wcc())
        lab = be.labels()
    for i in (0, 5, 17, 31, 47):
        g = a.member(i)
        r = orc.run(g)
        lo, hi = int(a.vtx_off[i]), int(a.vtx_off[i + 1])
        assert r["iterations"] == iters[i]
        assert np.array_equal(r["best_labels"], lab[lo:hi])
        assert best[i] == pytest.approx(r["final_wcc"], rel=1e-9, abs=0)


def test_batch_rejects_empty_member(engine):
    g = scd_amd.from_edge_pairs([(0, 1), (1, 2), (2, 0)]).convert()
    arena = scd_amd.Arena.concat([g, g])
    arena.vtx_off[1] = 0  # graph 0 without vertices
    with scd_amd.BatchEngine(device=0) as be:
        with pytest.raises(scd_amd.ScdError):
            be.upload(arena)


def test_graphs_labels_one_download(small_batch):
    """scd_batch_graphs_labels hands back the labels of any list of graphs
    (best or current, unordered and repeated ids, the empty list) equal to
    the slices of the whole-arena download; an id past the arena is an
    error."""
    a = small_batch
    off = a.vtx_off
    with scd_amd.BatchEngine(device=0) as be:
        be.upload(a)
        be.initial_partition(want_labels=False)
        be.refine(be.compute_wcc(), max_iters=3)
        full = {True: be.labels(), False: be.current_labels()}
        for sel in ([], [0], [a.count - 1, 3, 3, 17, 0], list(range(a.count))):
            for best in (True, False):
                got = be.graph_labels(sel, best=best)
                assert len(got) == len(sel)
                for g, lab in zip(sel, got):
                    assert np.array_equal(lab, full[best][off[g]:off[g + 1]]), (sel, g, best)
        with pytest.raises(scd_amd.ScdError):
            be.graph_labels([a.count])


def test_refine_again_after_bins_shrank(small_batch):
    """scd_batch_refine bins only the graphs still refining once half of the
    binned vertices belong to stopped graphs, and bins every vertex again at
    its end: a second pass from a new initial partition repeats the first,
    and the current labels of every graph equal its labels at its last
    iteration in both."""
    a = small_batch
    with scd_amd.BatchEngine(device=0) as be:
        be.upload(a)
        runs = []
        for _ in range(2):
            be.initial_partition(want_labels=False)
            w0 = be.compute_wcc()
            best, iters, _, hist = be.refine(w0, max_iters=100, history=True)
            runs.append((w0, best, iters, hist, be.labels(), be.current_labels()))
    for x, y in zip(runs[0], runs[1]):
        assert np.array_equal(x, y)
    assert len(set(runs[0][2].tolist())) > 1  # graphs stopped at different iterations


def test_batch_hub_buckets_equal_single_graph_runs(engine):
    """An arena whose largest graph has hubs, refined with the hub buckets
    (lds_tile 2048, hub_table 2) while smaller graphs stop earlier: every
    graph gets what the single-graph path gives it with the default tuning."""
    names = ["rmat16", "pp8x40_s5", "rmat14", "c1"]
    gs = [graph(n) for n in names]
    assert gs[0].degree().max() > 2048
    a = scd_amd.Arena.concat(gs)
    print(f"[gpu] hub-bucket batch {names}: n={a.graph.n} m={a.graph.m}", flush=True)
    with scd_amd.BatchEngine(device=0) as be:
        be.upload(a)
        be.set_tuning(lds_tile=2048, hub_table=2, lds_hash=0)
        lab0, k = be.initial_partition()
        w0 = be.compute_wcc()
        best, iters, done, hist = be.refine(w0, max_iters=100, history=True)
        lab_best = be.labels()
    assert len(set(iters.tolist())) > 1  # graphs stop at different iterations
    for i, g in enumerate(gs):
        lo, hi = int(a.vtx_off[i]), int(a.vtx_off[i + 1])
        engine.transfer_graph(g)
        l0, ki = engine.initial_partition()
        assert ki == k[i] and np.array_equal(l0, lab0[lo:hi]), f"{names[i]}: initial partition"
        wi = engine.compute_wcc()
        assert wi == w0[i]
        bi, it, h = engine.refine(wi)
        assert it == iters[i] and np.array_equal(h, hist[i, :it]), f"{names[i]}: iterations / WCC"
        assert bi == best[i]
        assert np.array_equal(engine.labels_from_device(), lab_best[lo:hi]), f"{names[i]}: best labels"
