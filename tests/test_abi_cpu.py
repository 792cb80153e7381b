"""The C-ABI library loads and exports every symbol include/pyci_engine.h declares
(no compute calls: this runs on hosts without a GPU)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT

LIB = os.path.join(ROOT, "pyci_amd", "libpyci_engine.so")
HDR = os.path.join(ROOT, "include", "pyci_engine.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"\b(pyci_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


def test_no_gpu_is_reported_not_faked():
    lib = ctypes.CDLL(LIB)
    cnt = ctypes.c_int(-1)
    assert lib.pyci_device_count(ctypes.byref(cnt)) == 0
    if cnt.value > 0:
        pytest.skip("a GPU is visible")
    ctx = ctypes.c_void_p()
    rc = lib.pyci_ctx_create(0, ctypes.byref(ctx))
    assert rc == 4  # PYCI_EHIP
    lib.pyci_last_error.restype = ctypes.c_char_p
    assert b"no GPU" in lib.pyci_last_error()


def test_python_module_raises_without_gpu():
    import pyci_amd as pyci

    if pyci.device_count() > 0:
        pytest.skip("a GPU is visible")
    w = pyci.fullci_wfn(4, 2, 2)
    with pytest.raises(RuntimeError):
        w.add_all_dets()
    assert len(w) == 0
