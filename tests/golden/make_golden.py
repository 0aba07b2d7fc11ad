"""Writes the golden fixtures of tests/golden from the reference's own CPU
code, built from its sources by oracle/Makefile into oracle/_ref:

  _ref/convert            the reference convert tool (src/convert.cpp)
  _ref/ref_dump           src/{graph,scd,wcc,common}.cpp driven like
                          src/main.cpp, dumping labels, statistics and WCC of
                          every iteration
  _ref/ref_improve_table  src/scd.cpp's calculate_improvement on a table of
                          arguments

Outputs:
  manifest.json      per graph: generator recipe, SHA-256 of the reference
                     convert output, n, m, cc / WCC bit patterns, iteration
                     count, printed community count, SHA-256 of the labels and
                     statistics of every iteration
  full_<name>.npz    the complete label and statistics arrays of small graphs
  improve_table.npz  calculate_improvement arguments and result bits,
                     including rows whose int32 products wrap
  compress_table.npz compress_labels on label arrays: the relabelled
                     arrays and counts (_ref/ref_compress_table)

Run in the build container (it needs the reference build):
  make -C oracle ref && python tests/golden/make_golden.py [names...]
Graphs of golden_graphs.HUGE (R-MAT scale 24) are written only when named.
SCD_GOLDEN_CACHE=<dir>: a file <dir>/<name>.ref.bin (an earlier output of
_ref/convert on the recipe's edge list) replaces the convert run, and
<dir>/<name>.dump.bin (an earlier output of _ref/ref_dump on that file, one
thread) the dump run; both are still the reference programs' outputs, and
the first is still compared byte for byte with scd_convert's.
SCD_GOLDEN_NO_REF_CONVERT=1: for a HUGE graph without a cached reference
convert output, the graph file is scd_convert's (byte-equal to the reference
convert on every other recipe) and the entry says so ("convert_source");
everything else still comes from the reference's own code.  Temporary files
go to TMPDIR.
"""
import hashlib
import json
import os
import struct
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "het-scd_amd"))
sys.path.insert(0, os.path.dirname(HERE))

import scd_amd  # noqa: E402
from golden_graphs import GRAPHS, HUGE, make_edges  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
FULL_MAX_N = 20000


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def hexf(x):
    return struct.pack("<d", x).hex()


CACHE = os.environ.get("SCD_GOLDEN_CACHE")


def cached(name, suffix):
    if not CACHE:
        return None
    path = os.path.join(CACHE, f"{name}.{suffix}")
    return path if os.path.exists(path) else None


def ref_convert(edges, tmp, name=None):
    pre = cached(name, "ref.bin")
    if pre:
        with open(pre, "rb") as f:
            return f.read(), pre
    txt = os.path.join(tmp, "g.txt")
    out = os.path.join(tmp, "g.bin")
    edges.write_text(txt)
    subprocess.run([os.path.join(REF, "convert"), txt, out], check=True, stdout=subprocess.DEVNULL)
    with open(out, "rb") as f:
        return f.read(), out


def ref_dump(path, tmp, name=None):
    out = cached(name, "dump.bin")
    if not out:
        out = os.path.join(tmp, "dump.bin")
        # one thread: the reference's OpenMP sums (clustering coefficient,
        # src/graph.cpp:59; WCC, src/wcc.cpp:49) then run in vertex order, the
        # canonical order whose bits the oracle and the engine reproduce
        subprocess.run([os.path.join(REF, "ref_dump"), path, out], check=True, stdout=subprocess.DEVNULL,
                       env=dict(os.environ, OMP_NUM_THREADS="1"))
    b = open(out, "rb").read()
    n, nc, iters, printed = struct.unpack_from("<4I", b, 0)
    cc, init, final = struct.unpack_from("<3d", b, 16)
    pos = 40
    labels, stats, wcc = [], [], []
    for it in range(iters + 1):
        labels.append(np.frombuffer(b, np.uint32, n, pos))
        pos += 4 * n
        stats.append(np.frombuffer(b, np.int32, 3 * nc, pos).reshape(3, nc))
        pos += 12 * nc
        if it > 0:
            wcc.append(struct.unpack_from("<d", b, pos)[0])
            pos += 8
    assert pos == len(b)
    return dict(n=n, num_comms=nc, iterations=iters, printed=printed, cc=cc, initial_wcc=init, final_wcc=final,
                labels=labels, stats=stats, wcc=wcc)


def improve_table():
    rng = np.random.default_rng(20261203)
    rows = []
    for _ in range(4000):
        r = int(rng.integers(0, 400))
        d_in = int(rng.integers(0, 64))
        d_out = int(rng.integers(0, 2000))
        c_out = int(rng.integers(0, 5000))
        p_in = float(rng.random()) if rng.random() < 0.9 else 0.0
        p_ext = float(rng.random() * 0.6)
        rows.append((r, d_in, d_out, c_out, p_in, p_ext))
    # hub-sized arguments whose int32 products wrap (d_out*(d_out-1) etc.)
    for _ in range(1000):
        r = int(rng.integers(0, 70000))
        d_in = int(rng.integers(0, 70000))
        d_out = int(rng.integers(40000, 400000))
        c_out = int(rng.integers(0, 2**31 - 1))
        rows.append((r, d_in, d_out, c_out, float(rng.random()), float(rng.random())))
    # edge cases: empty and singleton communities, zero probabilities
    for r in (0, 1, 2, 3):
        for d_in in (0, 1, 2):
            for d_out in (0, 1, 5):
                rows.append((r, d_in, d_out, 3, 0.0, 0.25))
                rows.append((r, d_in, d_out, 0, 1.0, 0.0))
    # the table reads the doubles as u64 bit patterns in hex
    inp = "".join("%d %d %d %d %x %x\n" % (r, a, b, c, struct.unpack("<Q", struct.pack("<d", pi))[0],
                                          struct.unpack("<Q", struct.pack("<d", pe))[0])
                  for r, a, b, c, pi, pe in rows)
    out = subprocess.run([os.path.join(REF, "ref_improve_table")], input=inp, check=True, capture_output=True,
                         text=True).stdout.split()
    bits = np.array([int(x, 16) for x in out], dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "improve_table.npz"),
                        r=np.array([x[0] for x in rows], np.int32),
                        d_in=np.array([x[1] for x in rows], np.int32),
                        d_out=np.array([x[2] for x in rows], np.int32),
                        c_out=np.array([x[3] for x in rows], np.int32),
                        p_in=np.array([x[4] for x in rows], np.float64),
                        p_ext=np.array([x[5] for x in rows], np.float64),
                        bits=bits)
    return len(rows)


def compress_table():
    """The reference's compress_labels (src/scd.cpp:32-55) on label arrays
    with repeats, INVALID labels and sparse label values."""
    rng = np.random.default_rng(20261204)
    arrays = [np.array([7, 7, 0xFFFFFFFF, 3, 7], np.uint32), np.array([5], np.uint32),
              np.full(4, 0xFFFFFFFF, np.uint32), np.arange(6, dtype=np.uint32)[::-1].copy()]
    for _ in range(60):
        n = int(rng.integers(1, 400))
        a = rng.integers(0, int(rng.integers(1, 2 * n + 2)), n).astype(np.uint32)
        a[rng.random(n) < rng.random() * 0.3] = 0xFFFFFFFF
        arrays.append(a)
    inp = "".join(f"{len(a)}\n" + " ".join(str(int(x)) for x in a) + "\n" for a in arrays)
    out = subprocess.run([os.path.join(REF, "ref_compress_table")], input=inp, check=True, capture_output=True,
                         text=True).stdout
    rows = [l for l in out.splitlines() if l and not l.startswith("compress labels:")]
    assert len(rows) == len(arrays)
    counts, outs = [], []
    for a, row in zip(arrays, rows):
        v = [int(x) for x in row.split()]
        assert len(v) == len(a) + 1
        counts.append(v[0])
        outs.append(np.array(v[1:], np.uint32))
    off = np.concatenate([[0], np.cumsum([len(a) for a in arrays])]).astype(np.int64)
    np.savez_compressed(os.path.join(HERE, "compress_table.npz"), offsets=off, labels_in=np.concatenate(arrays),
                        labels_out=np.concatenate(outs), counts=np.array(counts, np.uint32))
    return len(arrays)


def main(names=None):
    man_path = os.path.join(HERE, "manifest.json")
    manifest = json.load(open(man_path)) if os.path.exists(man_path) else {"graphs": {}}
    with tempfile.TemporaryDirectory() as tmp:
        for name, recipe in GRAPHS.items():
            if (names and name not in names) or (not names and name in HUGE):
                continue
            print(f"[golden] {name}", flush=True)
            edges = make_edges(recipe)
            source = "reference convert (oracle/_ref/convert)"
            if name in HUGE and not cached(name, "ref.bin") and os.environ.get("SCD_GOLDEN_NO_REF_CONVERT"):
                ours = edges.convert()
                del edges
                path = os.path.join(tmp, "g.bin")
                ours.write(path)
                ref_bytes = ours.to_bytes()
                source = ("scd_convert (host restatement, byte-equal to the reference convert on every smaller "
                          "recipe; the reference convert was not run at this size)")
            else:
                ref_bytes, path = ref_convert(edges, tmp, name)
                ours = edges.convert()
                del edges
                assert ours.to_bytes() == ref_bytes, f"{name}: scd_convert differs from the reference convert"
            d = ref_dump(path, tmp, name)
            assert d["n"] == ours.n, f"{name}: the dump is not of this graph"
            ent = {
                "recipe": recipe,
                "convert_sha256": hashlib.sha256(ref_bytes).hexdigest(),
                "convert_source": source,
                "n": d["n"], "m": ours.m, "num_comms": d["num_comms"], "iterations": d["iterations"],
                "printed_num_communities": d["printed"],
                "cc_hex": hexf(d["cc"]), "initial_wcc_hex": hexf(d["initial_wcc"]),
                "final_wcc_hex": hexf(d["final_wcc"]), "wcc_hex": [hexf(w) for w in d["wcc"]],
                "labels_sha256": [sha(l) for l in d["labels"]],
                "stats_sha256": [sha(s) for s in d["stats"]],
                "tri_sha256": sha(ours.tri), "tri_deg_sha256": sha(ours.tri_deg),
                "full": d["n"] <= FULL_MAX_N,
            }
            if ent["full"]:
                np.savez_compressed(os.path.join(HERE, f"full_{name}.npz"), labels=np.stack(d["labels"]),
                                    stats=np.stack(d["stats"]))
            manifest["graphs"][name] = ent
        if not names or "improve_table" in names:
            manifest["improve_table_rows"] = improve_table()
        if not names or "compress_table" in names:
            manifest["compress_table_arrays"] = compress_table()
    with open(man_path, "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
