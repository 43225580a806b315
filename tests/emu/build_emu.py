"""Host build of the library for the emulation tests: every HIP source of
diffpumas.jl_amd/csrc compiled as C++ with g++ against the HIP stand-in
(tests/emu/hip), linked into one shared library that the Python package loads
through DIFFPUMAS_HIP_LIB.  Kernels then run on the CPU with the lanes of a
block as threads (see hip/hip_runtime.h)."""
import concurrent.futures as cf
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, "diffpumas.jl_amd", "csrc")


def build(outdir, defines=(), extra=(), jobs=8):
    os.makedirs(outdir, exist_ok=True)
    srcs = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    flags = ["-x", "c++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread",
             "-I", HERE, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
             *[f"-D{d}" for d in defines], *extra]

    def one(f):
        obj = os.path.join(outdir, f + ".o")
        p = subprocess.run(["g++", *flags, "-c", os.path.join(CSRC, f), "-o", obj],
                           capture_output=True, text=True)
        if p.returncode:
            raise RuntimeError(f"g++ failed for {f}:\n{p.stderr[-4000:]}")
        return obj

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    lib = os.path.join(outdir, "libdiffpumas_emu.so")
    subprocess.run(["g++", "-shared", "-pthread", *extra, "-o", lib, *objs], check=True)
    return lib
