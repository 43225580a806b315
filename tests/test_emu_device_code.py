"""Device code of the library run on the host (tests/emu): the MC kernels
with the lanes of each block as threads, event by event against the oracle
(with the default hard-event pool and with a 4-record pool that forces the
inline fallback), the wave-cooperative hard-event evaluation in lockstep
against the inline evaluation, plain and under AddressSanitizer/UBSan, and the
GPU parity modules of the tetrahedral walk, the DEM ray exit, the direct CSDA
flux and the legacy samplers run unchanged against the host build (their
kernels reach every ballot and barrier with the whole block, which is what
the thread emulation models)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
sys.path.insert(0, EMU)

needs_gpp = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _run_check(lib, label):
    env = dict(os.environ)
    env["DIFFPUMAS_HIP_LIB"] = lib
    env["PYTHONPATH"] = os.pathsep.join([os.path.join(ROOT, "tests"),
                                         os.path.join(ROOT, "diffpumas.jl_amd"),
                                         env.get("PYTHONPATH", "")])
    r = subprocess.run([sys.executable, os.path.join(EMU, "mc_emu_check.py"), label],
                       capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout


@needs_gpp
@pytest.mark.parametrize("pool", [None, 4])
def test_mc_kernels_emulated_match_oracle(tmp_path, pool):
    import build_emu
    lib = build_emu.build(str(tmp_path / "lib"), defines=[f"DP_MC_POOL={pool}"] if pool else [])
    _run_check(lib, f"pool={pool or 'default'}")


@needs_gpp
@pytest.mark.parametrize("sanitize", [False, True])
def test_coop_hard_event_lockstep(tmp_path, sanitize):
    exe = tmp_path / "coop_emu"
    flags = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"] \
        if sanitize else ["-O2"]
    subprocess.run(["g++", *flags, "-std=c++17", "-pthread", "-ffp-contract=off", "-I", EMU,
                    "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "diffpumas.jl_amd", "csrc"), "-o", str(exe),
                    os.path.join(EMU, "coop_emu.cpp")], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert " 0 failures" in r.stdout


@needs_gpp
def test_gpu_parity_modules_on_host_build(tmp_path):
    import build_emu
    lib = build_emu.build(str(tmp_path / "lib"))
    env = dict(os.environ)
    env["DIFFPUMAS_HIP_LIB"] = lib
    mods = ["test_gpu_tetra.py", "test_gpu_dem.py", "test_gpu_direct.py",
            "test_gpu_legacy_samplers.py"]
    # the lane-quad tracers, on their small cases (the 3000-ray cases take minutes here)
    trace = [os.path.join(ROOT, "tests", "test_gpu_parity.py") + "::" + t
             for t in ("test_trace_cube_edge_cases", "test_trace_cube_empty",
                       "test_trace_edges_edge_cases")]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "gpu", "-p", "no:cacheprovider",
                        *[os.path.join(ROOT, "tests", m) for m in mods], *trace],
                       capture_output=True, text=True, timeout=900, env=env, cwd=ROOT)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-2000:]
