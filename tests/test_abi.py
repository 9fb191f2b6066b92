"""The C-ABI library loads and exports every entry point include/blocksp_c.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "blocksp_c.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s*(bsp_[a-z0-9_]+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    assert "bsp_bsmv" in names and "bsp_pgmres" in names and len(names) >= 40


def test_library_loads_and_exports_every_symbol():
    import blocksp_amd as bsp
    lib = bsp.lib()
    for name in declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", bsp.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (bsp_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_python_bindings_cover_the_header():
    import blocksp_amd as bsp
    assert set(declared()) == set(bsp.exported_symbols())


def test_library_has_gfx950_code_object():
    import blocksp_amd as bsp
    blob = open(bsp.LIB_PATH, "rb").read()
    assert b"gfx950" in blob
