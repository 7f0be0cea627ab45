"""The C-ABI library loads and exports every symbol include/*.h declares (no compute calls)."""
import ctypes
import glob
import os
import re

import moris_amd as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s+\**\s*(mgpu_[A-Za-z0-9_]+)\s*\(", src, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    lib = M.load_library()
    names = declared_symbols()
    assert "mgpu_set_assemble" in names and "mgpu_set_create" in names
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(M.EXPORTED) <= names


def test_error_string_after_bad_call():
    lib = M.load_library()
    out = ctypes.c_void_p()
    rc = lib.mgpu_problem_build(None, ctypes.byref(out))
    assert rc == 1
    assert b"null" in lib.mgpu_last_error_string()
