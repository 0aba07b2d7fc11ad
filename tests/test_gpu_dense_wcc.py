"""Internal triangles of dense communities by bitmap rows (tuning dense_wcc,
MI355X; 2 = on whatever the graph's size, so small graphs take the path):
per-vertex internal triangle counts, triangle degrees and WCC terms
bit-exact against the oracle (src/wcc.cpp:31-141) and against the list
kernels (dense_wcc 0), for the reference fixtures' label sequences of the
planted graphs, one community of 4,500 members (rows longer than one wave
load) and batched arenas."""
import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from conftest import full_fixture
from golden_graphs import graph

pytestmark = pytest.mark.gpu

DN_MIN_S, DN_MAX_S = 32, 16384


@pytest.fixture(autouse=True)
def _dense_default(engine):
    """Every test leaves the shared engine with the default (dense off)."""
    yield
    engine.set_tuning(dense_wcc=0)


def dense_labels(g, labels):
    """Communities the dense kernels take under these labels (the same
    criterion as the device classification)."""
    valid = labels != 0xFFFFFFFF
    k = int(labels[valid].max()) + 1 if valid.any() else 0
    s, _, i = orc.collect_stats(g, labels, k)
    s = s.astype(np.int64)
    z = i.astype(np.int64)
    return np.flatnonzero((s >= DN_MIN_S) & (s <= DN_MAX_S) & (16 * z >= s * (s - 1))), s


def details_both(engine, g, labels):
    out = []
    for dense in (2, 0):
        engine.set_tuning(dense_wcc=dense)
        engine.set_partition(labels)  # INVALID labels allowed
        out.append((engine.compute_wcc(details=True), engine.dense_wcc_info()))
    return out


def check_against_oracle(engine, g, labels, expect_dense):
    (d1, info1), (d0, info0) = details_both(engine, g, labels)
    _, sizes = dense_labels(g, labels)
    w, terms, it, itd = orc.compute_wcc(g, labels, sizes.astype(np.int32))
    for got in (d1, d0):
        wg, tg, itg, itdg = got
        assert np.array_equal(itg, it), np.flatnonzero(itg != it)[:10]
        assert np.array_equal(itdg, itd), np.flatnonzero(itdg != itd)[:10]
        assert np.array_equal(tg.view(np.uint64), terms.view(np.uint64))
        assert wg == pytest.approx(w, rel=1e-12, abs=0)
    assert d1[0] == d0[0]
    assert info0 == (0, 0)
    dl, s = dense_labels(g, labels)
    members = int(s[dl].sum())
    if expect_dense:
        assert members > 0, "no dense community in this case"
    # row units requested: s x ceil(s / 64) words per community in 32-word
    # units; all of them fit the buffer of 2 bytes per adjacency entry
    units = int(sum((int(x) * ((int(x) + 63) // 64) + 31) // 32 for x in s[dl]))
    capacity = min(max(2 * g.m // 128, 4096), 1 << 24)
    assert info1[0] == units, (info1, units)
    if units <= capacity:
        assert info1[1] == members, (info1, members)
    return members


@pytest.mark.parametrize("name", ["pp4x16_s4", "pp8x40_s5", "pprand_s6", "pprand_s7", "c1", "rmat12", "rmat14"])
def test_fixture_label_sequences(engine, name):
    """The reference's label sequences (their communities are small or
    sparse: the list kernels count them, and the classification agrees)."""
    g = graph(name)
    engine.transfer_graph(g)
    full = full_fixture(name)
    for it, lab in enumerate(full["labels"]):
        print(f"[gpu] dense wcc {name} labels {it}", flush=True)
        check_against_oracle(engine, g, np.ascontiguousarray(lab), expect_dense=False)


_synthetic = {}


def synthetic(size, p):
    """One planted block of `size` vertices at density p (converted)."""
    key = (size, p)
    if key not in _synthetic:
        _synthetic[key] = scd_amd.planted_partition(np.array([size], np.uint32), p, 0.0, 11).convert()
    return _synthetic[key]


def labels_of(kind, n):
    ids = np.arange(n, dtype=np.uint32)
    if kind == "one":
        return np.zeros(n, np.uint32)
    if kind == "halves":
        return ids % 2
    if kind == "one_with_invalid":  # every fifth vertex without a community
        lab = np.zeros(n, np.uint32)
        lab[::5] = 0xFFFFFFFF
        return lab
    if kind == "thirds_sizes":  # communities of unequal size: 1/2, 1/3, 1/6
        return np.where(ids % 6 < 3, 0, np.where(ids % 6 < 5, 1, 2)).astype(np.uint32)
    raise ValueError(kind)


# (graph, labelling, words per row of the largest dense community)
DENSE_CASES = [
    ("er40_s2", "one", 1), ("er64_s3", "one", 1), ("pp4x16_s4", "one", 1), ("pp4x16_s4", "halves", 1),
    ("s700", "one", 11), ("s700", "one_with_invalid", 9), ("s700", "thirds_sizes", 6), ("s2000", "one", 32),
    ("s4000", "one", 63), ("s4500", "one", 71),
]


@pytest.mark.parametrize("gname,kind,words", DENSE_CASES)
def test_dense_labelings(engine, gname, kind, words):
    """Labelings with dense communities: rows of 1 to 71 words (64 / Wp
    groups per wave up to 64 words, several words per lane beyond),
    INVALID vertices inside, communities of unequal size."""
    if gname.startswith("s"):
        size = int(gname[1:])
        g = synthetic(size, {700: 0.2, 2000: 0.08, 4000: 0.07, 4500: 0.07}[size])
    else:
        g = graph(gname)
    engine.transfer_graph(g)
    lab = labels_of(kind, g.n)
    print(f"[gpu] dense labelling {gname} {kind} n={g.n} m={g.m}", flush=True)
    members = check_against_oracle(engine, g, lab, expect_dense=True)
    dl, s = dense_labels(g, lab)
    assert (int(s[dl].max()) + 63) // 64 == words
    assert members > 0


def test_iterations_match_list_kernels(engine, manifest):
    """Refinement iterations (incremental WCC: only changed communities are
    recounted) give the same labels and WCC with the dense kernels as
    without."""
    name = "pprand_s7"
    g = graph(name)
    ent = manifest["graphs"][name]
    lab0 = np.ascontiguousarray(full_fixture(name)["labels"][0])
    runs = []
    for dense in (2, 0):
        engine.transfer_graph(g)
        engine.set_tuning(dense_wcc=dense)
        engine.labels_to_device(lab0)
        w = [engine.compute_wcc()]
        labs = []
        for _ in range(ent["iterations"]):
            wcc, lab = engine.improve_labels(return_labels=True)
            w.append(wcc)
            labs.append(lab)
        runs.append((w, labs))
    assert runs[0][0] == runs[1][0]
    for a, b in zip(runs[0][1], runs[1][1]):
        assert np.array_equal(a, b)


def test_batched_arena(engine):
    """A batched arena of planted graphs: best WCC, iterations and labels
    with the dense kernels equal those without."""
    arena = scd_amd.planted_batch(24, seed_base=71, n_min=500, n_max=6000, k_min=2, k_max=8, p_in=0.2,
                                  p_out=0.002)
    outs = []
    for dense in (2, 0):
        with scd_amd.BatchEngine(device=0) as be:
            be.upload(arena)
            be.set_tuning(dense_wcc=dense)
            be.initial_partition(want_labels=False)
            w0 = be.compute_wcc()
            best, iters, _ = be.refine(w0)
            info = be.eng.dense_wcc_info()
            outs.append((w0, best, iters, be.labels(), info))
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])
    assert np.array_equal(outs[0][3], outs[1][3])
    assert outs[1][4] == (0, 0)


def test_size_threshold(engine):
    """dense_wcc 1 takes the bitmap path from 2^20 adjacency entries on: not on 700 vertices (98,620 entries), on 4,500 (1.4 M)."""
    for size, p, used in ((700, 0.2, False), (4500, 0.07, True)):
        g = synthetic(size, p)
        engine.transfer_graph(g)
        engine.set_tuning(dense_wcc=1)
        engine.set_partition(np.zeros(g.n, np.uint32))
        engine.compute_wcc()
        units, members = engine.dense_wcc_info()
        assert (members == g.n) if used else (units, members) == (0, 0), (size, units, members)
