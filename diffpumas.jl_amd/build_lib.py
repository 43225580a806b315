"""Build libdiffpumas_hip.so (gfx950) in-tree with hipcc.

Usage: python diffpumas.jl_amd/build_lib.py [-j N]

Each HIP translation unit is compiled separately (in parallel) into
diffpumas.jl_amd/_build/ and linked into diffpumas.jl_amd/libdiffpumas_hip.so.
Flags: no FP contraction (the traversal is bit-exact with the reference and
the transport follows its operation order), no fast-math, packed FP32
instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdiffpumas_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DP_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def compile_one(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{p.stderr[-6000:]}")
    return obj


def build(jobs=8, verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    return LIB


def build_oracle(verbose=True):
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    od = os.path.join(HERE, "..", "oracle")
    p = subprocess.run(["make", "-C", od, "-s"], capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{p.stderr[-4000:]}")
    if verbose:
        print("built oracle/liboracle_dp.so")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args()
    try:
        build(a.j)
        if not a.no_oracle:
            build_oracle()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
