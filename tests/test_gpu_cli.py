"""The command line driver (MI355X): main-gpu on a graph file prints the
reference's lines and numbers, and writes the labels file and the
machine-readable run log."""
import json
import os
import re
import struct
import subprocess

import numpy as np
import pytest

import oracle_bind as orc
import scd_amd
from golden_graphs import graph

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "het-scd_amd", "bin", "main-gpu")


@pytest.mark.parametrize("name", ["pp4x16_s4", "c1", "rmat12"])
def test_main_gpu_outputs(manifest, name, tmp_path):
    ent = manifest["graphs"][name]
    g = graph(name)
    path = tmp_path / f"{name}.bin"
    g.write(path)
    env = dict(os.environ, SCD_LABELS_OUT=str(tmp_path / "labels.bin"), SCD_RUN_LOG=str(tmp_path / "run.jsonl"))
    print(f"[gpu] main-gpu {name}", flush=True)
    r = subprocess.run([BIN, str(path), "0"], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert int(re.search(r"num\. communities: (\d+)", out).group(1)) == ent["printed_num_communities"]
    assert int(re.search(r"iterations: (\d+)", out).group(1)) == ent["iterations"]
    final = struct.unpack("<d", bytes.fromhex(ent["final_wcc_hex"]))[0]
    log = [json.loads(l) for l in (tmp_path / "run.jsonl").read_text().splitlines()]
    assert log[0]["event"] == "setup" and log[0]["nodes"] == g.n and log[0]["edges"] == g.m
    its = [e for e in log if e["event"] == "iteration"]
    assert len(its) == ent["iterations"] and [e["iteration"] for e in its] == list(range(1, len(its) + 1))
    done = log[-1]
    assert done["event"] == "done" and done["iterations"] == ent["iterations"]
    assert done["best_wcc"] == pytest.approx(final, rel=1e-9, abs=0)
    assert done["printed_communities"] == ent["printed_num_communities"]
    # the labels file holds the compacted best labels of the oracle's run
    labels, k = scd_amd.read_labels(tmp_path / "labels.bin")
    exp, ek = scd_amd.compact_labels(orc.run(g)["best_labels"])
    assert k == ek == done["communities"] and np.array_equal(labels, exp)


CONVERT_BIN = os.path.join(os.path.dirname(BIN), "convert-gpu")
REF_CONVERT = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "convert")
TIMER = re.compile(r"^(.*): [-0-9.e+]+ ms$")


def _shape(stdout):
    """stdout lines with the timer values taken out (only they may differ)."""
    return [TIMER.sub(r"\1: <t> ms", l) for l in stdout.splitlines()]


@pytest.mark.parametrize("name", ["bowtie", "pp8x40_s5", "c1", "rmat14"])
def test_convert_gpu_cli(manifest, name, tmp_path):
    """convert-gpu on the text edge list: the reference convert's output file
    (its SHA-256 in the fixtures) and, where the reference's ./convert is
    built, its stdout line for line (timer values aside)."""
    import hashlib
    from golden_graphs import GRAPHS, make_edges
    ent = manifest["graphs"][name]
    txt = tmp_path / f"{name}.txt"
    make_edges(GRAPHS[name]).write_text(txt)
    out = tmp_path / "out.bin"
    print(f"[gpu] convert-gpu {name}", flush=True)
    r = subprocess.run([CONVERT_BIN, str(txt), str(out), "0"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert hashlib.sha256(out.read_bytes()).hexdigest() == ent["convert_sha256"]
    lines = _shape(r.stdout)
    assert f"final graph has {ent['n']} nodes and {ent['m']} edges" in lines
    if os.path.exists(REF_CONVERT):
        ref_out = tmp_path / "ref.bin"
        rr = subprocess.run([REF_CONVERT, str(txt), str(ref_out)], capture_output=True, text=True, timeout=120,
                            env=dict(os.environ, OMP_NUM_THREADS="4"))
        assert rr.returncode == 0
        assert lines == [l.replace(str(ref_out), str(out)) for l in _shape(rr.stdout)]


def test_convert_gpu_cli_errors(tmp_path):
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2\n2 3\n3 x\n")
    r = subprocess.run([CONVERT_BIN, str(bad), str(tmp_path / "o.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "line 3 is corrupt" in r.stderr
    r = subprocess.run([CONVERT_BIN, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "usage:" in r.stderr
