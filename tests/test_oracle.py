"""Pins the CPU oracle (oracle/scd_oracle.cpp) before it is trusted as the
checker: against the fixtures written by the reference's own code
(tests/golden, from oracle/_ref), closed forms and brute force."""
import hashlib
import itertools
import struct

import numpy as np
import pytest

import oracle_bind as orc
from conftest import GOLDEN, full_fixture
from golden_graphs import MEDIUM, SMALL, graph

INVALID = 0xFFFFFFFF


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def unhex(h):
    return struct.unpack("<d", bytes.fromhex(h))[0]


# ---------------------------------------------------------------------------
# golden vectors from the reference build
# ---------------------------------------------------------------------------
def test_improvement_table_bit_exact():
    """orc_improvement == the reference's calculate_improvement
    (src/scd.cpp:173-219) bit for bit, including int32-wrapping rows."""
    t = np.load(f"{GOLDEN}/improve_table.npz")
    got = np.array([orc.improvement(int(r), int(a), int(b), int(c), float(pi), float(pe))
                    for r, a, b, c, pi, pe in zip(t["r"], t["d_in"], t["d_out"], t["c_out"], t["p_in"], t["p_ext"])])
    assert np.array_equal(got.view(np.uint64), t["bits"])
    # the table must contain rows whose products wrap
    wraps = (t["d_out"].astype(np.int64) * (t["d_out"].astype(np.int64) - 1) > 2**31 - 1)
    assert wraps.sum() > 100


@pytest.mark.parametrize("name", SMALL + MEDIUM)
def test_oracle_run_matches_reference(name, manifest):
    """Labels and statistics of every iteration, iteration count, printed
    community count, clustering coefficient bits; WCC within 1e-12."""
    ent = manifest["graphs"][name]
    g = graph(name)
    r = orc.run(g)
    assert r["iterations"] == ent["iterations"]
    assert struct.pack("<d", r["cc"]).hex() == ent["cc_hex"]
    for it in range(r["iterations"] + 1):
        assert sha(r["labels"][it]) == ent["labels_sha256"][it], f"labels differ at iteration {it}"
    assert r["num_communities"] == ent["printed_num_communities"]
    assert r["initial_wcc"] == pytest.approx(unhex(ent["initial_wcc_hex"]), rel=1e-12, abs=0)
    assert r["final_wcc"] == pytest.approx(unhex(ent["final_wcc_hex"]), rel=1e-12, abs=0)
    for w, h in zip(r["wcc"], ent["wcc_hex"]):
        assert w == pytest.approx(unhex(h), rel=1e-12, abs=0)
    full = full_fixture(name)
    labels = r["labels"]
    for it in range(r["iterations"] + 1):
        s, e, i = orc.collect_stats(g, labels[it], ent["num_comms"])
        assert sha(np.stack([s, e, i])) == ent["stats_sha256"][it], f"stats differ at iteration {it}"
        if full is not None:
            assert np.array_equal(full["labels"][it], labels[it])


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name,k", [("k5", 5), ("k9", 9)])
def test_complete_graph(name, k):
    g = graph(name)
    tri, td = orc.triangles(g)
    assert np.all(tri == (k - 1) * (k - 2)) and np.all(td == k - 1)
    assert orc.clustering_coef(g) == 1.0
    labels, nl = orc.initial_partition(g)
    assert nl == 1 and np.all(labels == 0)
    s, e, i = orc.collect_stats(g, labels, 1)
    assert (s[0], e[0], i[0]) == (k, 0, k * (k - 1))
    # one community holding every triangle: each term is 1
    w, terms, it, itd = orc.compute_wcc(g, labels, s)
    assert w == 1.0 and np.all(terms == 1.0)


def test_bowtie_closed_form():
    """Hub 0 with triangles {0,1,2} and {0,3,4}: first in the (cc, deg,
    id)-descending order is vertex 4 (cc 1, deg 2, highest id), so {4, 0, 3}
    get label 0, then vertex 2 seeds {2, 1} as label 1.  Terms: vertex 0
    (2*4)/(4*(4+0)) = 1/2, vertices 3 and 4 (2*2)/(2*(2+0)) = 1, community
    {1, 2} has size 2 and scores 0; mean 2.5 / 5 = 0.5 exactly."""
    g = graph("bowtie")
    labels, nl = orc.initial_partition(g)
    assert list(labels) == [0, 1, 1, 0, 0] and nl == 2
    s, e, i = orc.collect_stats(g, labels, nl)
    w, terms, it, itd = orc.compute_wcc(g, labels, s)
    assert w == 0.5


def test_triangle_sum_is_six_times_triangles():
    for name in ("rmat10", "pp8x40_s5", "er64_s3"):
        g = graph(name)
        tri, _ = orc.triangles(g)
        assert np.array_equal(tri, g.tri)


# ---------------------------------------------------------------------------
# brute force on tiny graphs
# ---------------------------------------------------------------------------
def adjacency_sets(g):
    return [set(int(x) for x in g.adj[g.indices[v]:g.indices[v + 1]]) for v in range(g.n)]


@pytest.mark.parametrize("name", ["er24_s1", "er40_s2", "er64_s3", "pp4x16_s4"])
def test_triangles_brute_force(name):
    g = graph(name)
    N = adjacency_sets(g)
    tri = np.zeros(g.n, np.int64)
    tdeg = np.zeros(g.n, np.int64)
    for u in range(g.n):
        for v in N[u]:
            t = len(N[u] & N[v])
            tri[u] += t
            tdeg[u] += t > 0
    t2, d2 = orc.triangles(g)
    assert np.array_equal(t2, tri) and np.array_equal(d2, tdeg)
    assert np.array_equal(g.tri, tri)


@pytest.mark.parametrize("name", ["er24_s1", "er64_s3", "pp4x16_s4", "pp8x40_s5"])
def test_initial_partition_brute_force(name):
    """Sequential greedy over the (cc, deg, id)-descending order, restated
    with Python tuples; and its lexicographically-first-MIS form: seeds are
    exactly the vertices with no earlier seed neighbour."""
    g = graph(name)
    N = adjacency_sets(g)
    deg = g.degree()
    order = sorted(range(g.n), key=lambda p: (g.tri[p] / (deg[p] * (deg[p] - 1.0)), int(deg[p]), p), reverse=True)
    lab = [INVALID] * g.n
    k = 0
    seeds = []
    for p in order:
        if lab[p] != INVALID:
            continue
        for q in N[p]:
            if lab[q] == INVALID:
                lab[q] = k
        lab[p] = k
        seeds.append(p)
        k += 1
    labels, nl = orc.initial_partition(g)
    assert nl == k and list(labels) == lab
    rank = {p: i for i, p in enumerate(order)}
    seedset = set(seeds)
    for p in range(g.n):
        earlier_seed = any(q in seedset and rank[q] < rank[p] for q in N[p])
        assert (p in seedset) == (not earlier_seed)


def random_labels(g, rng, k, invalid_frac):
    lab = rng.integers(0, k, g.n).astype(np.uint32)
    lab[rng.random(g.n) < invalid_frac] = INVALID
    return lab


@pytest.mark.parametrize("name", ["er24_s1", "er64_s3", "pp4x16_s4", "pp8x40_s5"])
def test_move_rule_brute_force(name):
    """orc_improve against an exhaustive per-vertex evaluation built from the
    pinned improvement function: every distinct valid neighbour label is a
    candidate with its multiplicity; best = max improvement, ties to the
    smaller label; then the move rule of src/scd.cpp:321-330."""
    g = graph(name)
    rng = np.random.default_rng(5)
    cc = orc.clustering_coef(g)
    for trial in range(3):
        k = [3, 8, g.n][trial]
        lab = random_labels(g, rng, k, 0.1 * trial)
        s, e, i = orc.collect_stats(g, lab, k)
        got = orc.improve(g, lab, s, e, i, cc)
        for p in range(g.n):
            nb = g.adj[g.indices[p]:g.indices[p + 1]]
            deg = len(nb)
            my = int(lab[p])
            cnt = {}
            for q in nb:
                L = int(lab[q])
                if L != INVALID:
                    cnt[L] = cnt.get(L, 0) + 1
            best, best_l, rem = 0.0, 0, 0.0
            for L in sorted(cnt):
                f = cnt[L]
                size, ext, ie = int(s[L]), int(e[L]), int(i[L]) // 2
                if L == my:
                    size -= 1
                    ext += f - (deg - f)
                    ie -= f
                p_in = (2.0 * ie) / (size * (size - 1.0)) if size > 1 else 0.0
                imp = orc.improvement(size, f, deg - f, ext, p_in, cc)
                if L == my:
                    rem = -imp
                if imp > best:
                    best, best_l = imp, L
            ins = best + rem
            want = best_l if (ins > rem and ins > 0.0) else (INVALID if rem > 0 else my)
            assert got[p] == want, (trial, p)


def test_stats_invariants():
    g = graph("pp8x40_s5")
    rng = np.random.default_rng(1)
    lab = random_labels(g, rng, 30, 0.2)
    s, e, i = orc.collect_stats(g, lab, 30)
    valid = lab != INVALID
    deg = g.degree()
    assert s.sum() == valid.sum()
    assert (e.astype(np.int64) + i).sum() == deg[valid].sum()
    assert np.all(i % 2 == 0)


@pytest.mark.parametrize("name", ["er24_s1", "er64_s3", "pp8x40_s5"])
def test_wcc_brute_force(name):
    """Per-vertex internal triangles by set intersection, then the term of
    src/wcc.cpp:161-166; mean over all vertices."""
    g = graph(name)
    N = adjacency_sets(g)
    rng = np.random.default_rng(2)
    lab = random_labels(g, rng, 6, 0.1)
    s, _, _ = orc.collect_stats(g, lab, 6)
    w, terms, it, itd = orc.compute_wcc(g, lab, s)
    total = 0.0
    for p in range(g.n):
        L = int(lab[p])
        if L == INVALID or s[L] <= 2 or g.tri_deg[p] == 0:
            assert terms[p] == 0.0
            continue
        Nc = {q for q in N[p] if lab[q] == L}
        itri = sum(len(Nc & N[q]) for q in Nc)
        itdeg = sum(1 for q in Nc if Nc & N[q])
        assert (it[p], itd[p]) == (itri, itdeg)
        term = (itri * int(g.tri_deg[p])) / (int(g.tri[p]) * (int(g.tri_deg[p]) + (int(s[L]) - 1 - itdeg)))
        assert terms[p] == term
        total += term
    assert w == pytest.approx(total / g.n, rel=1e-14)


def test_neighbor_pairs_brute_force():
    g = graph("pp8x40_s5")
    rng = np.random.default_rng(3)
    lab = random_labels(g, rng, 12, 0.15)
    srt, off, rl, rc = orc.neighbor_label_pairs(g, lab)
    k = 0
    for p in range(g.n):
        nb = lab[g.adj[g.indices[p]:g.indices[p + 1]]]
        assert np.array_equal(srt[g.indices[p]:g.indices[p + 1]], np.sort(nb))
        vals, cnts = np.unique(nb[nb != INVALID], return_counts=True)
        assert off[p] == k
        assert np.array_equal(rl[k:k + len(vals)], vals) and np.array_equal(rc[k:k + len(vals)], cnts)
        k += len(vals)
    assert off[g.n] == k == len(rl)


def test_compress_count():
    lab = np.array([3, 3, INVALID, 7, 3, 7, 9], np.uint32)
    # n - #distinct valid labels = 7 - 3
    assert orc.compress_count(lab) == 4
    import scd_amd
    assert scd_amd.compress_count(lab) == 4


def test_lfmis_equivalence_exhaustive_tiny():
    """Greedy seeding == lexicographically-first MIS, checked over every
    order of a 6-vertex graph's vertices (brute force of the claim the GPU
    initial partition rests on)."""
    edges = [(0, 1), (1, 2), (2, 0), (2, 3), (3, 4), (4, 5), (5, 3), (1, 4)]
    N = [set() for _ in range(6)]
    for a, b in edges:
        N[a].add(b)
        N[b].add(a)
    for order in itertools.permutations(range(6)):
        lab = [None] * 6
        seeds = []
        for p in order:
            if lab[p] is not None:
                continue
            for q in N[p]:
                if lab[q] is None:
                    lab[q] = len(seeds)
            lab[p] = len(seeds)
            seeds.append(p)
        rank = {p: i for i, p in enumerate(order)}
        for p in range(6):
            if p in seeds:
                assert lab[p] == seeds.index(p)
            else:
                first = min((q for q in N[p] if q in seeds), key=lambda q: rank[q])
                assert lab[p] == seeds.index(first)
