"""The C ABI library loads and exports every entry point include/graeng.h
declares; the host-side loader binds exactly that set.  No compute calls
(no GPU needed)."""
import ctypes
import os
import re

import pytest

from graada3d_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graeng.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gra_[a-zA-Z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    assert "gra_assemble" in names and "gra_solve_cg" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libgraeng.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_loader_binds_exactly_the_header():
    assert sorted(_lib.EXPORTS) == declared()
    L = _lib.load()
    for n in _lib.EXPORTS:
        assert getattr(L, n).restype is not None or n == "gra_ctx_destroy"


def test_status_strings():
    L = _lib.load()
    for s in range(7):
        assert L.gra_status_string(s)
    assert L.gra_status_string(99) == b"unknown status"


def test_context_reports_missing_device_or_works():
    L = _lib.load()
    h = ctypes.c_void_p()
    st = L.gra_ctx_create(0, ctypes.byref(h))
    if st == 0:
        L.gra_ctx_destroy(h)
    else:
        assert st == 6  # GRA_NO_DEVICE
    assert L.gra_ctx_create(0, None) == 1
