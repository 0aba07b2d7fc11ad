"""The reference's own drivers, unchanged, on this engine (MI355X).

oracle/_ref/main-gpu-amd is src/main-gpu.cpp compiled as it stands with the
reference's src/{graph,scd,wcc,common}.cpp and linked against
libscd_dropin.so, which provides the six cuda.hpp functions over the HIP
engine (cuda_init ... cuda_labels_from_device, cuda_labels_to_device
included).  Its results must be the reference CPU run's (oracle/_ref/main,
the same sources' ./main): nodes, edges, clustering coefficient, initial
WCC, every iteration's WCC, the iteration count, the final WCC and the
printed community count.  Both programs are built in the build container
from the reference tree; without them the test is skipped.

oracle/_ref/main-amd is the reference's CPU driver src/main.cpp, unchanged,
with only src/common.cpp and libscd_refapi.so: Graph::read and every
scd.hpp / wcc.hpp function it calls (create_initial_partition,
improve_partition, collect_stats, compute_wcc, compress_labels) run on the
device.  oracle/_ref/main-gpu-amd-dev is src/main-gpu.cpp with the same
library and libscd_dropin.so: its initial partition and initial WCC run on
the device as well."""
import os
import re
import subprocess

import pytest

from golden_graphs import graph

pytestmark = pytest.mark.gpu

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref")
DROPIN_MAIN = os.path.join(REF, "main-gpu-amd")
REFAPI_MAIN = os.path.join(REF, "main-amd")
REFAPI_GPU_MAIN = os.path.join(REF, "main-gpu-amd-dev")
CPU_MAIN = os.path.join(REF, "main")
NUM = r"([-0-9.e+]+|nan|inf)"


def parse(out):
    """The numbers both drivers print (src/main.cpp:49-93, src/main-gpu.cpp:63-110)."""
    def one(key):
        m = re.search(rf"^{re.escape(key)}: {NUM}( ms)?$", out, flags=re.M)
        assert m, f"no '{key}' line in\n{out}"
        return float(m.group(1))
    wccs = [float(x) for x in re.findall(rf"^WCC: {NUM}$", out, flags=re.M)]
    return {"nodes": one("nodes"), "edges": one("edges"), "avg_cc": one("avg. cc"),
            "initial_wcc": one("initial WCC"), "iteration_wcc": wccs[:-1], "final_wcc": wccs[-1],
            "iterations": one("iterations"), "communities": one("num. communities"),
            "ms_per_iteration": one("time per iteration")}


@pytest.mark.skipif(not (os.path.exists(DROPIN_MAIN) and os.path.exists(CPU_MAIN)),
                    reason="oracle/_ref not built (needs the reference tree at build time)")
@pytest.mark.parametrize("name,tuned", [("c1", False), ("rmat12", False), ("c2", False), ("c2", True)])
def test_unchanged_reference_main_gpu_matches_reference_cpu(name, tuned, tmp_path):
    """tuned: SCD_TUNE_CACHE names an (empty) autotuner file, so the drop-in
    tunes the graph's size class, appends it and runs with it."""
    g = graph(name)
    path = str(tmp_path / f"{name}.bin")
    g.write(path)
    env = dict(os.environ)
    env.pop("SCD_TUNE_CACHE", None)
    cache = tmp_path / "tune.txt"
    if tuned:
        env["SCD_TUNE_CACHE"] = str(cache)
    print(f"[gpu] reference main-gpu.cpp on libscd_dropin.so: {name} n={g.n} m={g.m} tuned={tuned}", flush=True)
    gpu = subprocess.run([DROPIN_MAIN, path, "0"], capture_output=True, text=True, timeout=300, env=env)
    assert gpu.returncode == 0, gpu.stderr
    cpu = subprocess.run([CPU_MAIN, path], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert cpu.returncode == 0, cpu.stderr
    a, b = parse(gpu.stdout), parse(cpu.stdout)
    for k in ("nodes", "edges", "iterations", "communities"):
        assert a[k] == b[k], (k, a[k], b[k])
    # six significant digits are printed: equal up to the last one
    for k in ("avg_cc", "initial_wcc", "final_wcc"):
        assert a[k] == pytest.approx(b[k], rel=2e-6), k
    assert len(a["iteration_wcc"]) == len(b["iteration_wcc"]) == b["iterations"]
    assert a["iteration_wcc"] == pytest.approx(b["iteration_wcc"], rel=2e-6)
    # the four phase lines of src/cuda.cu:357-360 once per iteration, non-negative
    phases = re.findall(r"^(improve partitioning|collect statistics|remove external edges|compute wcc): "
                        rf"{NUM} ms$", gpu.stdout, flags=re.M)
    assert len(phases) == 4 * b["iterations"]
    assert all(float(v) >= 0 for _, v in phases)
    assert "selected device:" in gpu.stdout and "memory:" in gpu.stdout
    if tuned:
        assert len(cache.read_text().splitlines()) == 1  # the size class, measured and appended
    print(f"[gpu] {name}: {a['ms_per_iteration']:.3f} ms per iteration (reference CPU, one thread: "
          f"{b['ms_per_iteration']:.3f})", flush=True)


@pytest.mark.skipif(not (os.path.exists(REFAPI_MAIN) and os.path.exists(REFAPI_GPU_MAIN) and os.path.exists(CPU_MAIN)),
                    reason="oracle/_ref not built (needs the reference tree at build time)")
@pytest.mark.parametrize("name", ["pp8x40_s5", "c1", "rmat12", "c2"])
def test_unchanged_reference_drivers_on_device_call_surface(name, tmp_path):
    g = graph(name)
    path = str(tmp_path / f"{name}.bin")
    g.write(path)
    cpu = subprocess.run([CPU_MAIN, path], capture_output=True, text=True, timeout=300,
                         env=dict(os.environ, OMP_NUM_THREADS="1"))
    assert cpu.returncode == 0, cpu.stderr
    b = parse(cpu.stdout)
    for exe, args in ((REFAPI_MAIN, [path]), (REFAPI_GPU_MAIN, [path, "0"])):
        print(f"[gpu] {os.path.basename(exe)} (scd.hpp / wcc.hpp on the device): {name} n={g.n} m={g.m}", flush=True)
        r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr
        a = parse(r.stdout)
        for k in ("nodes", "edges", "iterations", "communities"):
            assert a[k] == b[k], (exe, k, a[k], b[k])
        for k in ("avg_cc", "initial_wcc", "final_wcc"):
            assert a[k] == pytest.approx(b[k], rel=2e-6), (exe, k)
        assert a["iteration_wcc"] == pytest.approx(b["iteration_wcc"], rel=2e-6)
        # the reference functions' timer lines are printed as the reference's are
        for line in ("create initial partition:", "compute WCC:", "compress labels:"):
            assert line in r.stdout, (exe, line)
        if exe == REFAPI_MAIN:
            assert "improving partition:" in r.stdout and "collect community statistics:" in r.stdout
