"""Config-sized CPU-oracle computations for the golden fixtures (test infrastructure).

The oracle's CSR builder (oracle/pyci_oracle.c, a restatement of SparseOp::update /
add_row, sparseop.cpp:186-502) is called on row ranges from a thread pool (ctypes
releases the GIL and the builder keeps no global state), the lower+diagonal rows
are packed as int64 row pointers / int32 columns / float64 values, and products
use the oracle's threaded perform_op_symm restatement (sparseop.cpp:148-159).
Eigenpairs come from scipy's ARPACK (eigsh) on that product.
"""
import ctypes
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from oracle import oracle as O


def _lib():
    L = O.lib()
    if not hasattr(L, "_symv_ready"):
        L.oracle_symv_lower_mt.restype = ctypes.c_int
        L.oracle_symv_lower_mt.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L._symv_ready = True
    return L


class LowerCSR:
    """Lower+diagonal rows of a symmetric CI matrix (the reference's storage)."""

    def __init__(self, indptr, indices, data):
        self.indptr, self.indices, self.data = indptr, indices, data
        self.n = indptr.size - 1

    def diagonal(self):
        return self.data[self.indptr[1:] - 1]  # rows are sorted; the diagonal is last

    def symv(self, x, nthread=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty(self.n)
        rc = _lib().oracle_symv_lower_mt(self.n, self.indptr.ctypes.data, self.indices.ctypes.data,
                                         self.data.ctypes.data, x.ctypes.data, y.ctypes.data,
                                         nthread or os.cpu_count() or 1)
        if rc != 0:
            raise MemoryError("oracle_symv_lower_mt")
        return y


def build_lower(kind, nbasis, nup, ndn, dets, one_mo, two_mo, nthread=None, chunk_rows=None, log=print):
    """Oracle CSR (symmetric, lower+diagonal) of all rows, built in row chunks on threads."""
    nthread = nthread or os.cpu_count() or 1
    dets = np.ascontiguousarray(dets, dtype=np.uint64)
    ndet = dets.shape[0]
    chunk_rows = chunk_rows or max(256, ndet // (8 * nthread))
    ranges = [(r, min(ndet, r + chunk_rows)) for r in range(0, ndet, chunk_rows)]
    t0 = time.perf_counter()

    def one(rr):
        c = O.csr_build(kind, nbasis, nup, ndn, dets, one_mo, two_mo, nrow=rr[1], row0=rr[0])
        return c.indptr, c.indices.astype(np.int32), c.data

    with ThreadPoolExecutor(max_workers=nthread) as ex:
        parts = list(ex.map(one, ranges))
    nnz = sum(p[2].size for p in parts)
    indptr = np.zeros(ndet + 1, dtype=np.int64)
    indices = np.empty(nnz, dtype=np.int32)
    data = np.empty(nnz)
    off = 0
    for (r0, r1), (ip, ix, dv) in zip(ranges, parts):
        indptr[r0 + 1:r1 + 1] = ip[1:] + off
        indices[off:off + dv.size] = ix
        data[off:off + dv.size] = dv
        off += dv.size
    del parts
    if log:
        log(f"oracle CSR: {ndet} rows, {nnz} lower nonzeros, {time.perf_counter() - t0:.1f} s on {nthread} threads")
    return LowerCSR(indptr, indices, data)


def lowest(csr, k, tol=1e-12, ncv=None, log=print):
    """k lowest eigenpairs (ascending) of the symmetric operator by ARPACK."""
    import scipy.sparse.linalg as spl

    count = [0]

    def mv(v):
        count[0] += 1
        return csr.symv(np.asarray(v).reshape(-1))

    A = spl.LinearOperator((csr.n, csr.n), matvec=mv, dtype=np.float64)
    t0 = time.perf_counter()
    # a deterministic start vector: ARPACK's default draws a random one
    v0 = np.random.default_rng(12345).standard_normal(csr.n)
    w, v = spl.eigsh(A, k=k, which="SA", tol=tol, ncv=ncv or max(2 * k + 1, 20), v0=v0)
    o = np.argsort(w)
    if log:
        log(f"eigsh k={k}: {count[0]} products, {time.perf_counter() - t0:.1f} s")
    return w[o], v[:, o].T


def checksums(y, seed, nsample=2048):
    """Size-independent fingerprint of a long vector: values at seeded rows, its
    dot products with two seeded normal vectors, its 2-norm and sum."""
    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(y.size, size=min(nsample, y.size), replace=False))
    p = rng.standard_normal((2, y.size))
    return dict(rows=rows, vals=y[rows], proj=p @ y, norm=np.linalg.norm(y), sum=y.sum())


def projections(n, seed):
    """The two projection vectors of checksums(·, seed)."""
    rng = np.random.default_rng(seed)
    rng.choice(n, size=min(2048, n), replace=False)
    return rng.standard_normal((2, n))
