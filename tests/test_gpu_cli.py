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
