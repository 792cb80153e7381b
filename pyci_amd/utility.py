"""Host-side integral and RDM post-processing (NumPy).

These mirror the reference's Python helpers (pyci/utility.py:30-150) by name,
arguments and results; they run on the host because they are O(n^4) reshapes
of small arrays, not part of the GPU hot path.
"""
import numpy as np

__all__ = [
    "make_senzero_integrals",
    "reduce_senzero_integrals",
    "spinize_rdms",
]


def make_senzero_integrals(one_mo, two_mo):
    """Seniority-zero pieces h, v, w of full integrals (utility.py:30-57):
    h[p] = t_pp, v[p,q] = <pp|qq>, w[p,q] = 2<pq|pq> - <pq|qp>."""
    one_mo = np.asarray(one_mo, dtype=np.float64)
    two_mo = np.asarray(two_mo, dtype=np.float64)
    n = one_mo.shape[0]
    p = np.arange(n)
    h = one_mo[p, p].copy()
    v = two_mo[p[:, None], p[:, None], p[None, :], p[None, :]].copy()
    w = 2.0 * two_mo[p[:, None], p[None, :], p[:, None], p[None, :]] - two_mo[p[:, None], p[None, :], p[None, :], p[:, None]]
    return h, v, w


def reduce_senzero_integrals(h, v, w, nocc):
    """Fold the one-body part into the two-body seniority-zero integrals
    (utility.py:60-91): rv = v + f diag(h), rw = w + f (h_p + h_q), f = 2/(2 nocc - 1)."""
    h = np.asarray(h, dtype=np.float64)
    f = 2.0 / (nocc * 2 - 1)
    rv = np.asarray(v, dtype=np.float64) + f * np.diag(h)
    rw = np.asarray(w, dtype=np.float64) + f * (h[:, None] + h[None, :])
    return rv, rw


def spinize_rdms(d1, d2):
    """DOCI (d0, d2) matrices or FullCI spin blocks -> generalized spin-orbital
    RDMs, alpha block 0..n-1 then beta block n..2n-1 (utility.py:94-150).

    rdm1[p,q] = <a+_p a_q>; rdm2[p,q,r,s] = <a+_p a+_q a_s a_r>."""
    d1 = np.asarray(d1, dtype=np.float64)
    d2 = np.asarray(d2, dtype=np.float64)
    n = d1.shape[1]
    m = 2 * n
    rdm1 = np.zeros((m, m))
    rdm2 = np.zeros((m, m, m, m))
    a, b = slice(0, n), slice(n, m)
    if d1.ndim == 2:
        p = np.arange(n)
        rdm1[p, p] = np.diag(d1)
        rdm1[p + n, p + n] = np.diag(d1)
        pp, qq = np.meshgrid(p, p, indexing="ij")
        # pair-hop and density-density pieces placed on the unsymmetrised tensor
        rdm2[pp, pp + n, qq, qq + n] += d1
        rdm2[pp + n, pp, qq + n, qq] += d1
        rdm2[pp, qq, pp, qq] += d2
        rdm2[pp + n, qq + n, pp + n, qq + n] += d2
        rdm2[pp, qq + n, pp, qq + n] += d2
        rdm2[pp + n, qq, pp + n, qq] += d2
        rdm2 = rdm2 - rdm2.transpose(1, 0, 2, 3)
        rdm2 = rdm2 - rdm2.transpose(0, 1, 3, 2)
        rdm2 *= 0.5
    else:
        rdm1[a, a] += d1[0]
        rdm1[b, b] += d1[1]
        rdm2[a, a, a, a] += d2[0]
        rdm2[b, b, b, b] += d2[1]
        rdm2[a, b, a, b] += d2[2]
        rdm2[b, a, b, a] += d2[2].transpose(1, 0, 3, 2)
        rdm2[a, b, b, a] -= d2[2].transpose(0, 1, 3, 2)
        rdm2[b, a, a, b] -= d2[2].transpose(1, 0, 2, 3)
    return rdm1, rdm2
