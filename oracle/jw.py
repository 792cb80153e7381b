"""Independent brute-force second quantisation (Jordan-Wigner on bit strings).

TEST INFRASTRUCTURE ONLY. It shares no code or formula with the oracle's
Slater-Condon restatement: matrix elements come from literally applying
    H = sum_pq h_pq a+_p a_q + 1/2 sum_pqrs <pq|rs> a+_p a+_q a_s a_r
to occupation-number vectors, with a+_p carrying (-1)^(#occupied below p).
Spin-orbital order is alpha block (0..n-1) then beta block (n..2n-1), the
convention of the reference's spinize_rdms (utility.py:117-124).

Used to pin oracle CSR values, phases and ordering, and RDM index conventions,
on systems small enough for dense algebra (<= 12 spin-orbitals).
"""
import itertools

import numpy as np


def _apply(op_seq, states):
    """Apply a product of ladder operators (rightmost first) to an int array of
    occupation bit strings. op_seq: list of (orbital, create?) in left-to-right
    order. Returns (sign array, new states, valid mask)."""
    states = states.copy()
    sign = np.ones(states.shape, dtype=np.int64)
    valid = np.ones(states.shape, dtype=bool)
    for p, create in reversed(op_seq):
        bit = np.int64(1) << p
        occ = (states & bit) != 0
        valid &= (~occ) if create else occ
        below = states & (bit - 1)
        par = np.array([bin(int(x)).count("1") & 1 for x in below], dtype=np.int64)
        sign *= 1 - 2 * par
        states = states ^ bit
    return sign, states, valid


def spin_orbital_integrals(one_mo, two_mo):
    """Spatial (n) -> spin-orbital (2n) integrals; <pq|rs> nonzero only for
    spin(p) = spin(r) and spin(q) = spin(s)."""
    n = one_mo.shape[0]
    h = np.zeros((2 * n, 2 * n))
    h[:n, :n] = one_mo
    h[n:, n:] = one_mo
    g = np.zeros((2 * n,) * 4)
    for s1 in (0, n):
        for s2 in (0, n):
            g[s1:s1 + n, s2:s2 + n, s1:s1 + n, s2:s2 + n] = two_mo
    return h, g


def hamiltonian(h, g, basis):
    """Dense <I|H|J> over `basis` (array of spin-orbital bit strings)."""
    basis = np.asarray(basis, dtype=np.int64)
    index = {int(b): i for i, b in enumerate(basis)}
    m = len(basis)
    ns = h.shape[0]
    H = np.zeros((m, m))
    cols = np.arange(m)
    for p, q in itertools.product(range(ns), repeat=2):
        if h[p, q] == 0.0:
            continue
        s, st, ok = _apply([(p, True), (q, False)], basis)
        for j in cols[ok]:
            i = index.get(int(st[j]))
            if i is not None:
                H[i, j] += h[p, q] * s[j]
    for p, q, r, t in itertools.product(range(ns), repeat=4):
        if g[p, q, r, t] == 0.0 or p == q or r == t:
            continue
        s, st, ok = _apply([(p, True), (q, True), (t, False), (r, False)], basis)
        for j in cols[ok]:
            i = index.get(int(st[j]))
            if i is not None:
                H[i, j] += 0.5 * g[p, q, r, t] * s[j]
    return H


def operator_matrix(op_seq, basis):
    basis = np.asarray(basis, dtype=np.int64)
    index = {int(b): i for i, b in enumerate(basis)}
    m = len(basis)
    A = np.zeros((m, m))
    s, st, ok = _apply(op_seq, basis)
    for j in np.arange(m)[ok]:
        i = index.get(int(st[j]))
        if i is not None:
            A[i, j] += s[j]
    return A


def rdms(c, basis, ns):
    """Spin-orbital rdm1[p,q] = <a+_p a_q>, rdm2[p,q,r,s] = <a+_p a+_q a_s a_r>."""
    r1 = np.zeros((ns, ns))
    r2 = np.zeros((ns,) * 4)
    for p, q in itertools.product(range(ns), repeat=2):
        r1[p, q] = c @ operator_matrix([(p, True), (q, False)], basis) @ c
    for p, q, r, s in itertools.product(range(ns), repeat=4):
        if p == q or r == s:
            continue
        r2[p, q, r, s] = c @ operator_matrix([(p, True), (q, True), (s, False), (r, False)], basis) @ c
    return r1, r2


def fullci_basis(dets, n):
    """FullCI (alpha word, beta word) pairs -> spin-orbital bit strings."""
    dets = np.asarray(dets, dtype=np.uint64).reshape(-1, 2)
    return (dets[:, 0].astype(np.int64)) | (dets[:, 1].astype(np.int64) << n)


def doci_basis(dets, n):
    dets = np.asarray(dets, dtype=np.uint64).reshape(-1).astype(np.int64)
    return dets | (dets << n)
