"""The C-ABI library loads without a GPU and exports every entry point include/sparrow_amd.h declares."""
import ctypes
import os
import re

from sparrow_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "sparrow_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(sa_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "sa_scf" in syms and "sa_scf_batch" in syms and "sa_fock" in syms and len(syms) >= 10


def test_library_exports_every_declared_symbol():
    lib = _native.load_library()
    for s in declared_symbols():
        assert hasattr(lib, s), s
    assert set(declared_symbols()) == set(_native.SIGNATURES), "ctypes signatures out of sync with the header"


def test_error_path_without_context():
    lib = _native.load_library()
    # a null context is rejected with a message, not a crash (no GPU needed)
    n = ctypes.c_int(0)
    Z = (ctypes.c_int * 1)(1)
    rc = lib.sa_nao(None, b"PM6", 1, Z, ctypes.byref(n))
    assert rc != 0
    assert b"null" in lib.sa_last_error()
