"""ctypes front-end to the CPU oracle (oracle/pyci_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product package never imports this module.

Every wrapper returns NumPy arrays in the reference's layouts:
  * determinants: uint64, one word per det (DOCI/GenCI) or (alpha, beta) pairs (FullCI)
  * CSR: data float64, indices int64, indptr int64 (sparseop.cpp:504-514)
"""
import ctypes
import os
import subprocess

import numpy as np

KIND = {"doci": 0, "fullci": 1, "genci": 2}

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force=False):
    """Compile the oracle with gcc (the same recipe as oracle/Makefile)."""
    src = os.path.join(_HERE, "pyci_oracle.c")
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(src):
        subprocess.check_call(["make", "-s", "-C", _HERE, "liboracle.so"])
    return _LIB


def lib():
    global _lib
    if _lib is None:
        build()
        L = ctypes.CDLL(_LIB)
        i64, u64p, dp, i64p = ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)
        L.oracle_binomial.restype = i64
        L.oracle_binomial.argtypes = [i64, i64]
        L.oracle_phase_single.restype = i64
        L.oracle_phase_single.argtypes = [ctypes.c_uint64, i64, i64]
        L.oracle_phase_double.restype = i64
        L.oracle_phase_double.argtypes = [ctypes.c_uint64, i64, i64, i64, i64]
        L.oracle_rank_colex.restype = i64
        L.oracle_rank_colex.argtypes = [i64, i64, ctypes.c_uint64]
        L.oracle_unrank_colex.restype = ctypes.c_uint64
        L.oracle_unrank_colex.argtypes = [i64, i64, i64]
        L.oracle_all_dets.restype = i64
        L.oracle_all_dets.argtypes = [ctypes.c_int, i64, i64, i64, u64p]
        L.oracle_add_excited_dets.restype = i64
        L.oracle_add_excited_dets.argtypes = [ctypes.c_int, i64, i64, i64, i64, u64p, u64p, i64, u64p, i64]
        L.oracle_csr_build.restype = ctypes.c_void_p
        L.oracle_csr_build.argtypes = [ctypes.c_int, i64, i64, i64, i64, u64p, dp, dp, i64, i64, i64, ctypes.c_int]
        L.oracle_csr_nnz.restype = i64
        L.oracle_csr_nnz.argtypes = [ctypes.c_void_p]
        L.oracle_csr_get.restype = None
        L.oracle_csr_get.argtypes = [ctypes.c_void_p, dp, i64p, i64p]
        L.oracle_csr_free.restype = None
        L.oracle_csr_free.argtypes = [ctypes.c_void_p]
        L.oracle_csr_matvec.restype = None
        L.oracle_csr_matvec.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.oracle_add_hci.restype = i64
        L.oracle_add_hci.argtypes = [ctypes.c_int, i64, i64, i64, i64, u64p, dp, dp, dp, ctypes.c_double, u64p, i64]
        L.oracle_rdms_doci.restype = None
        L.oracle_rdms_doci.argtypes = [i64, i64, u64p, dp, dp, dp]
        L.oracle_rdms_fullci.restype = None
        L.oracle_rdms_fullci.argtypes = [i64, i64, u64p, dp, dp, dp]
        L.oracle_rdms_genci.restype = None
        L.oracle_rdms_genci.argtypes = [i64, i64, u64p, dp, dp, dp]
        L.oracle_enpt2.restype = ctypes.c_double
        L.oracle_enpt2.argtypes = [ctypes.c_int, i64, i64, i64, i64, u64p, dp, dp, dp, ctypes.c_double,
                                   ctypes.c_double, i64p]
        for nm in ("oracle_trdms_doci", "oracle_trdms_fullci", "oracle_trdms_genci"):
            f = getattr(L, nm)
            f.restype = None
            f.argtypes = [i64, i64, u64p, i64, u64p, dp, dp, dp, dp]
        L.oracle_diagonal.restype = None
        L.oracle_diagonal.argtypes = [ctypes.c_int, i64, i64, i64, i64, u64p, dp, dp, dp]
        _lib = L
    return _lib


def _u64(a):
    a = np.ascontiguousarray(a, dtype=np.uint64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))


def _f64(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _i64(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def nkey(kind):
    return 2 if kind == "fullci" else 1


def binomial(n, k):
    return lib().oracle_binomial(n, k)


def phase_single(det, i, a):
    return lib().oracle_phase_single(int(det), i, a)


def phase_double(det, i1, i2, a1, a2):
    return lib().oracle_phase_double(int(det), i1, i2, a1, a2)


def rank_colex(nbasis, nocc, det):
    return lib().oracle_rank_colex(nbasis, nocc, int(det))


def unrank_colex(nbasis, nocc, rank):
    return lib().oracle_unrank_colex(nbasis, nocc, rank)


def all_dets(kind, nbasis, nup, ndn):
    """add_all_dets: the full space in colex order."""
    n = binomial(nbasis, nup) * (binomial(nbasis, ndn) if kind == "fullci" else 1)
    out = np.zeros(n * nkey(kind), dtype=np.uint64)
    lib().oracle_all_dets(KIND[kind], nbasis, nup, ndn, _u64(out)[1])
    return out.reshape(n, nkey(kind)) if kind == "fullci" else out


def add_excited_dets(kind, nbasis, nup, ndn, dets, ref, exc):
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    ref, rp = _u64(np.asarray(ref).reshape(-1))
    k = nkey(kind)
    n0 = dets.size // k
    # at most the number of rank-exc excitations of ref is added
    if kind == "fullci":
        nexc = sum(binomial(nup, a) * binomial(nbasis - nup, a) * binomial(ndn, exc - a) * binomial(nbasis - ndn, exc - a)
                   for a in range(exc + 1) if a <= nup and exc - a <= ndn)
    else:
        nexc = binomial(nup, exc) * binomial(nbasis - nup, exc)
    cap = n0 + nexc + 1
    out = np.zeros(cap * k, dtype=np.uint64)
    n = lib().oracle_add_excited_dets(KIND[kind], nbasis, nup, ndn, n0, dp, rp, exc, _u64(out)[1], cap)
    assert n <= cap
    out = out[: n * k]
    return out.reshape(n, 2) if kind == "fullci" else out


class CSR:
    """CSR matrix as the reference exports it (lower + diagonal when symmetric)."""

    def __init__(self, data, indices, indptr, shape, symmetric):
        self.data, self.indices, self.indptr = data, indices, indptr
        self.shape, self.symmetric = shape, symmetric

    def dense(self):
        a = np.zeros(self.shape)
        for i in range(self.shape[0]):
            for p in range(self.indptr[i], self.indptr[i + 1]):
                a[i, self.indices[p]] = self.data[p]
        if self.symmetric:
            a = a + np.tril(a, -1).T
        return a

    def scipy(self):
        import scipy.sparse as sp

        m = sp.csr_matrix((self.data, self.indices, self.indptr), shape=self.shape)
        if self.symmetric:
            m = m + sp.tril(m, -1).T
        return m.tocsr()


def csr_build(kind, nbasis, nup, ndn, dets, one_mo, two_mo, nrow=-1, ncol=-1, symmetric=True, row0=0):
    """SparseOp(ham, wfn, nrow, ncol, symmetric) restated (sparseop.cpp:57-63, 186-502)."""
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    ndet = dets.size // nkey(kind)
    nrow = ndet if nrow < 0 else nrow
    ncol = ndet if ncol < 0 else ncol
    one_mo, op1 = _f64(one_mo)
    two_mo, op2 = _f64(two_mo)
    L = lib()
    h = L.oracle_csr_build(KIND[kind], nbasis, nup, ndn, ndet, dp, op1, op2, row0, nrow, ncol, int(bool(symmetric)))
    try:
        nnz = L.oracle_csr_nnz(h)
        data = np.zeros(nnz)
        indices = np.zeros(nnz, dtype=np.int64)
        indptr = np.zeros(nrow - row0 + 1, dtype=np.int64)
        L.oracle_csr_get(h, _f64(data)[1], _i64(indices), _i64(indptr))
    finally:
        L.oracle_csr_free(h)
    return CSR(data, indices, indptr, (nrow - row0, ncol), bool(symmetric))


def matvec(csr, x):
    y = csr.scipy() @ np.asarray(x, dtype=np.float64)
    return y


def add_hci(kind, nbasis, nup, ndn, dets, one_mo, two_mo, coeffs, eps, cap=None):
    """add_hci restated serially; new dets in first-occurrence order (hci.cpp:49-279).
    `cap` bounds the number of new determinants (default: the full space, at most
    2^25); more raises."""
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    k = nkey(kind)
    ndet = dets.size // k
    if cap is None:
        cap = min(binomial(nbasis, nup) * (binomial(nbasis, ndn) if kind == "fullci" else 1), 1 << 25) + 1
    out = np.zeros(cap * k, dtype=np.uint64)
    one_mo, op1 = _f64(one_mo)
    two_mo, op2 = _f64(two_mo)
    coeffs, cp = _f64(coeffs)
    n = lib().oracle_add_hci(KIND[kind], nbasis, nup, ndn, ndet, dp, op1, op2, cp, float(eps), _u64(out)[1], cap)
    if n > cap:
        raise ValueError(f"add_hci appends {n} determinants, more than cap = {cap}")
    out = out[: n * k]
    return out.reshape(n, 2) if kind == "fullci" else out


def rdms(kind, nbasis, dets, coeffs):
    """compute_rdms: DOCI (d0, d2), FullCI (rdm1 (2,n,n), rdm2 (3,n,n,n,n)), GenCI (correct math)."""
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    coeffs, cp = _f64(coeffs)
    n = nbasis
    L = lib()
    if kind == "doci":
        d0, d2 = np.zeros((n, n)), np.zeros((n, n))
        L.oracle_rdms_doci(n, dets.size, dp, cp, _f64(d0)[1], _f64(d2)[1])
        return d0, d2
    if kind == "fullci":
        r1, r2 = np.zeros((2, n, n)), np.zeros((3, n, n, n, n))
        L.oracle_rdms_fullci(n, dets.size // 2, dp, cp, _f64(r1)[1], _f64(r2)[1])
        return r1, r2
    r1, r2 = np.zeros((n, n)), np.zeros((n, n, n, n))
    L.oracle_rdms_genci(n, dets.size, dp, cp, _f64(r1)[1], _f64(r2)[1])
    return r1, r2


def transition_rdms(kind, nbasis, dets1, dets2, c1, c2):
    """compute_transition_rdms: DOCI (d0, d2) and FullCI ((2,n,n), (3,n,n,n,n)) as
    rdm.cpp:634-899; GenCI ((n,n), (n,n,n,n)) by the operator algebra."""
    d1, p1 = _u64(np.asarray(dets1).reshape(-1))
    d2, p2 = _u64(np.asarray(dets2).reshape(-1))
    c1, q1 = _f64(c1)
    c2, q2 = _f64(c2)
    n, k = nbasis, nkey(kind)
    if kind == "doci":
        a, b = np.zeros((n, n)), np.zeros((n, n))
    elif kind == "fullci":
        a, b = np.zeros((2, n, n)), np.zeros((3, n, n, n, n))
    else:
        a, b = np.zeros((n, n)), np.zeros((n, n, n, n))
    getattr(lib(), "oracle_trdms_" + kind)(n, d1.size // k, p1, d2.size // k, p2, q1, q2, _f64(a)[1], _f64(b)[1])
    return a, b


def diagonal(kind, nbasis, nup, ndn, dets, one_mo, two_mo):
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    ndet = dets.size // nkey(kind)
    one_mo, op1 = _f64(one_mo)
    two_mo, op2 = _f64(two_mo)
    d = np.zeros(ndet)
    lib().oracle_diagonal(KIND[kind], nbasis, nup, ndn, ndet, dp, op1, op2, _f64(d)[1])
    return d


def enpt2(kind, nbasis, nup, ndn, dets, one_mo, two_mo, coeffs, energy, eps, ecore=0.0):
    """compute_enpt2 restated (enpt2.cpp:342-370): energy + sum_a (sum_i H_ai c_i)^2 /
    (energy - ecore - H_aa); a DOCI set enters as its FullCI (d, d) set
    (enpt2.cpp:382-386). Returns (energy, number of external determinants)."""
    if kind == "doci":
        d = np.asarray(dets, dtype=np.uint64).reshape(-1)
        dets = np.stack([d, d], axis=1)
        kind, nup, ndn = "fullci", nup, nup
    dets, dp = _u64(np.asarray(dets).reshape(-1))
    ndet = dets.size // nkey(kind)
    one_mo, op1 = _f64(one_mo)
    two_mo, op2 = _f64(two_mo)
    coeffs, cp = _f64(coeffs)
    nt = np.zeros(1, dtype=np.int64)
    corr = lib().oracle_enpt2(KIND[kind], nbasis, nup, ndn, ndet, dp, op1, op2, cp, float(energy - ecore),
                              float(eps), _i64(nt))
    return energy + corr, int(nt[0])


def lowest_eigs(csr, n=1, tol=1e-12):
    """Lowest eigenpairs of the oracle CSR (dense for small, ARPACK otherwise); ascending."""
    if csr.shape[0] <= 3000:
        w, v = np.linalg.eigh(csr.dense())
        return w[:n], v[:, :n].T
    import scipy.sparse.linalg as spl

    w, v = spl.eigsh(csr.scipy(), k=n, which="SA", tol=tol)
    o = np.argsort(w)
    return w[o], v[:, o].T
