"""The C-ABI library loads and exports every entry point include/ima_gpu.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os

import pytest

from ima2p_amd import _lib as L


def test_header_declares_entry_points():
    names = L.declared_symbols()
    assert "ima_score" in names and "ima_lmode_jointp" in names and "ima_swap" in names
    assert set(names) == set(L.PROTOS), set(names) ^ set(L.PROTOS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(L.LIB_PATH):
        pytest.fail("libima_gpu.so is not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(L.LIB_PATH)
    missing = [n for n in L.declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_abi_version_without_device():
    lib = L.load()
    assert lib.ima_abi_version() == 1


def test_no_device_is_an_error_not_a_fallback():
    lib = L.load()
    n = ctypes.c_int(-1)
    assert lib.ima_device_count(ctypes.byref(n)) == 0
    if n.value == 0:
        h = ctypes.c_void_p()
        assert lib.ima_ctx_create(0, ctypes.byref(h)) == 5  # IMA_ERR_NODEV
        assert not h.value
