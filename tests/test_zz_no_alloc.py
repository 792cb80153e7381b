"""The allocation-free matrix-free product (pyci_op_matvec_device_noalloc) through
the C ABI, against pyci_op_matvec_device on the same device inputs: two calls in a
row (the second with a smaller operator and fewer vectors), and the size limit."""
import ctypes
import os
from math import comb

import numpy as np
import pytest

from pyci_amd.synthetic import seeded_integrals

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib():
    L = ctypes.CDLL(os.path.join(ROOT, "pyci_amd", "libpyci_engine.so"))
    vp, i64 = ctypes.c_void_p, ctypes.c_int64
    L.pyci_last_error.restype = ctypes.c_char_p
    L.pyci_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
    L.pyci_ctx_synchronize.argtypes = [vp]
    L.pyci_ctx_destroy.argtypes = [vp]
    L.pyci_ham_create.argtypes = [vp, i64, ctypes.c_double, vp, vp, ctypes.POINTER(vp)]
    L.pyci_wfn_create.argtypes = [vp, ctypes.c_int, i64, i64, i64, i64, vp, ctypes.POINTER(vp)]
    L.pyci_fill_all_dets.argtypes = [vp, ctypes.c_int, i64, i64, i64, vp]
    L.pyci_op_create_ex.argtypes = [vp, vp, i64, i64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]
    L.pyci_op_storage.argtypes = [vp, ctypes.POINTER(ctypes.c_int)]
    L.pyci_op_matvec_device.argtypes = [vp, i64, vp, vp]
    L.pyci_op_matvec_device_noalloc.argtypes = [vp, i64, vp, vp]
    for f in ("pyci_op_destroy", "pyci_wfn_destroy", "pyci_ham_destroy"):
        getattr(L, f).argtypes = [vp]
    return L


def _hip():
    # by its SONAME: the HIP runtime the engine is bound to (when torch was imported
    # first, its bundled copy; "libamdhip64.so" would load a second runtime)
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _ok(L, rc, what):
    assert rc == 0, f"{what}: {L.pyci_last_error().decode()}"


class _Problem:
    """Full FullCI space of seeded integrals with a matrix-free operator."""

    def __init__(self, L, ctx, nb, nu, nd):
        _, self.one, self.two = seeded_integrals(nb, 8, 40 + nb)
        self.nd = comb(nb, nu) * comb(nb, nd)
        self.dets = np.zeros((self.nd, 2), dtype=np.uint64)
        _ok(L, L.pyci_fill_all_dets(ctx, 1, nb, nu, nd, self.dets.ctypes.data), "fill")
        self.ham, self.wfn, self.op = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        _ok(L, L.pyci_ham_create(ctx, nb, 0.0, self.one.ctypes.data, self.two.ctypes.data, ctypes.byref(self.ham)),
            "ham")
        _ok(L, L.pyci_wfn_create(ctx, 1, nb, nu, nd, self.nd, self.dets.ctypes.data, ctypes.byref(self.wfn)), "wfn")
        _ok(L, L.pyci_op_create_ex(self.ham, self.wfn, -1, -1, 1, 1, ctypes.byref(self.op)), "op")
        st = ctypes.c_int()
        _ok(L, L.pyci_op_storage(self.op, ctypes.byref(st)), "storage")
        assert st.value == 1  # matrix-free

    def close(self, L):
        L.pyci_op_destroy(self.op)
        L.pyci_wfn_destroy(self.wfn)
        L.pyci_ham_destroy(self.ham)


def _products(L, h, ctx, prob, nvec, seed):
    """(existing entry point, allocation-free entry point) outputs on one device x."""
    n = prob.nd * nvec
    x = np.random.default_rng(seed).standard_normal(n)
    bufs = [ctypes.c_void_p() for _ in range(3)]
    for b in bufs:
        assert h.hipMalloc(ctypes.byref(b), 8 * n) == 0
    try:
        assert h.hipMemcpy(bufs[0], x.ctypes.data, 8 * n, 1) == 0
        _ok(L, L.pyci_op_matvec_device(prob.op, nvec, bufs[0], bufs[1]), "matvec_device")
        _ok(L, L.pyci_op_matvec_device_noalloc(prob.op, nvec, bufs[0], bufs[2]), "matvec_device_noalloc")
        _ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        ya, yb = np.empty(n), np.empty(n)
        assert h.hipMemcpy(ya.ctypes.data, bufs[1], 8 * n, 2) == 0
        assert h.hipMemcpy(yb.ctypes.data, bufs[2], 8 * n, 2) == 0
    finally:
        for b in bufs:
            h.hipFree(b)
    return ya, yb


def test_noalloc_matvec_matches_matvec_device():
    L, h = _lib(), _hip()
    ctx = ctypes.c_void_p()
    _ok(L, L.pyci_ctx_create(0, ctypes.byref(ctx)), "ctx")
    try:
        # first call: nbasis 10, 5+5 (63,504 determinants), 4 vectors
        big = _Problem(L, ctx, 10, 5, 5)
        ya, yb = _products(L, h, ctx, big, 4, 1)
        assert np.abs(ya).max() > 0
        np.testing.assert_array_equal(yb, ya)
        # second call, smaller: nbasis 8, 4+3 (3,920 determinants), 1 vector
        small = _Problem(L, ctx, 8, 4, 3)
        ya, yb = _products(L, h, ctx, small, 1, 2)
        np.testing.assert_array_equal(yb, ya)
        small.close(L)
        big.close(L)
    finally:
        L.pyci_ctx_destroy(ctx)


def test_noalloc_rejects_oversized_products():
    """nbasis 12, 6+6: one vector needs 546 MB of workspace, above the fixed 256 MiB."""
    L, h = _lib(), _hip()
    ctx = ctypes.c_void_p()
    _ok(L, L.pyci_ctx_create(0, ctypes.byref(ctx)), "ctx")
    try:
        prob = _Problem(L, ctx, 12, 6, 6)
        x, y = ctypes.c_void_p(), ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(x), 8 * prob.nd) == 0
        assert h.hipMalloc(ctypes.byref(y), 8 * prob.nd) == 0
        rc = L.pyci_op_matvec_device_noalloc(prob.op, 1, x, y)
        assert rc == 6 and b"workspace" in L.pyci_last_error()
        _ok(L, L.pyci_ctx_synchronize(ctx), "sync")
        h.hipFree(x)
        h.hipFree(y)
        prob.close(L)
    finally:
        L.pyci_ctx_destroy(ctx)
