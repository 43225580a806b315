"""The device radix sort of dp_sort.hpp (used by the normal equations and the
density VJP) run on the host: lanes as threads with real barriers and
wave-wide ballots, checked against std::stable_sort (stability, equal keys,
partial tiles, 1-3 digit passes)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_radix_sort_emulated(tmp_path):
    exe = tmp_path / "sort_emu"
    src = os.path.join(ROOT, "tests", "emu", "sort_emu.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", os.path.join(ROOT, "tests", "emu"),
                    "-I", os.path.join(ROOT, "diffpumas.jl_amd", "csrc"), "-o", str(exe), src],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
