"""The C-ABI library loads and exports every symbol include/*.h declares
(CPU only: no compute call is made)."""
import ctypes
import glob
import os
import re

import diffpumas_hip._lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(dp_\w+)\s*\(", txt, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_symbols()
    assert {"dp_create", "dp_trace_cube", "dp_csda_cube", "dp_mc_flat_backward"} <= names


def test_library_exports_all_declared_symbols():
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(declared_symbols()) if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert declared_symbols() <= set(L.exported_symbols())
