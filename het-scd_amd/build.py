"""Builds the engine in-tree.

Outputs (git-ignored, shipped with the working tree):
  het-scd_amd/lib/libscd_host.so   binary format, generators, convert (g++)
  het-scd_amd/lib/libscd_hip.so    the gfx950 HIP engine behind include/scd_hip.h
  het-scd_amd/lib/libscd_dropin.so  the reference's cuda_* names over the C ABI
  het-scd_amd/lib/libscd_refapi.so  the reference's scd.hpp / wcc.hpp / Graph::read names
  het-scd_amd/lib/libscd_serve.so  the serving loop behind include/scd_serve.h
  het-scd_amd/bin/main-gpu         `main-gpu <graph> <device> [group size]`
  het-scd_amd/bin/convert-gpu      `convert-gpu <input file> <output file> [device]`
and, through oracle/Makefile, the CPU oracle (and the reference's own CPU
programs when the reference tree is present).

Usage: python het-scd_amd/build.py [--no-oracle] [-j N]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
CSRC = os.path.join(PKG, "csrc")
REF_SRC = "/root/reference/src"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

# -ffp-contract=off: no fused multiply-add, every fp64 operation rounds on its
# own as in the CPU path.  Packed fp32 instructions are disabled.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-I", INC,
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-I", INC]


def run(cmd, cwd=None):
    print("+ " + " ".join(cmd), flush=True)
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stdout.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    # keep compiler warnings visible, drop the host pass's note about the
    # GPU-only target feature
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if out:
        print(out)


def src(*parts):
    return os.path.join(CSRC, *parts)


def build_host():
    sources = [src("host", f) for f in ("graph_io.cpp", "gen.cpp", "labels_io.cpp")]
    run([CXX, *HOST_FLAGS, "-shared", "-o", os.path.join(LIB, "libscd_host.so"), *sources])


def build_hip():
    run([HIPCC, *HIP_FLAGS, "-shared", "-o", os.path.join(LIB, "libscd_hip.so"),
         src("hip", "engine.hip")])


def build_dropin_and_cli():
    dropin = os.path.join(LIB, "libscd_dropin.so")
    hip_rt = ["-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", "-L", "/opt/rocm/lib", "-lamdhip64",
              "-Wl,-rpath,/opt/rocm/lib"]
    run([CXX, *HOST_FLAGS, "-shared", "-o", dropin, src("dropin", "dropin.cpp"),
         "-L", LIB, "-lscd_hip", "-lscd_host", "-Wl,-rpath,$ORIGIN", *hip_rt])
    run([CXX, *HOST_FLAGS, "-shared", "-o", os.path.join(LIB, "libscd_refapi.so"), src("dropin", "refapi.cpp"),
         "-L", LIB, "-lscd_hip", "-lscd_host", "-Wl,-rpath,$ORIGIN"])
    run([CXX, *HOST_FLAGS, "-shared", "-o", os.path.join(LIB, "libscd_serve.so"), src("serve", "server.cpp"),
         "-L", LIB, "-lscd_hip", "-Wl,-rpath,$ORIGIN", *hip_rt])
    run([CXX, *HOST_FLAGS, "-o", os.path.join(BIN, "main-gpu"), src("cli", "main_gpu.cpp"),
         "-L", LIB, "-lscd_dropin", "-lscd_hip", "-lscd_host",
         "-Wl,-rpath,$ORIGIN/../lib"])
    run([CXX, *HOST_FLAGS, "-o", os.path.join(BIN, "convert-gpu"), src("cli", "convert_gpu.cpp"),
         "-L", LIB, "-lscd_hip", "-lscd_host", "-Wl,-rpath,$ORIGIN/../lib"])


def build_oracle():
    oracle = os.path.join(ROOT, "oracle")
    run(["make", "-s", "-C", oracle, "oracle"])
    if os.path.isdir(REF_SRC):
        run(["make", "-s", "-C", oracle, "ref"])


def build(oracle=True):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    with ThreadPoolExecutor(max_workers=3) as ex:
        futs = [ex.submit(build_host), ex.submit(build_hip)]
        if oracle:
            futs.append(ex.submit(build_oracle))
        for f in futs:
            f.result()
    if os.path.exists(src("dropin", "dropin.cpp")):
        build_dropin_and_cli()
    if oracle and os.path.isdir(REF_SRC):
        # the reference's own main-gpu.cpp, unchanged, against libscd_dropin.so
        run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-dropin"])


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    args = ap.parse_args()
    build(oracle=not args.no_oracle)
