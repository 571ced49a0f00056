"""The C-ABI library loads and exports every symbol include/conductorrigid.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from conductorrigid_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(cr_[a-z0-9_]+)\s*\(", text))
    return names


def test_library_exports_all_declared_symbols():
    lib = _lib.load()
    decl = declared_symbols()
    assert len(decl) >= 18
    missing = [n for n in sorted(decl) if not hasattr(lib, n)]
    assert not missing, missing
    assert decl == set(_lib.symbols())


def test_error_path_without_gpu_compute():
    lib = _lib.load()
    # a null context is rejected with a status code and a message, not a crash
    assert lib.cr_update_field(None, (ctypes.c_double * 3)(0, 0, 1)) != 0
    assert b"null" in lib.cr_last_error()
