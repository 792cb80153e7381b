"""Seeded synthetic integrals for the benchmark and parity workloads.

Recipe (fixed so that every workload is reproducible from its seed):
  * one_mo: strict upper triangle drawn row-major from N(0, offdiag_std), mirrored;
    diagonal h_ii = hdiag0 + hstep * i.
  * P = 1..nchol Cholesky-like vectors L_P: upper triangle INCLUDING the diagonal,
    row-major, N(0, l_std) (optionally scaled by exp(-|i-j| / decay)), mirrored.
  * eri[i,j,k,l] = sum_P L_P[i,j] L_P[k,l]   (chemist order, 8-fold symmetric, PSD)
  * two_mo = eri.transpose(0, 2, 1, 3)       (physicist <pq|rs>, the layout the
    reference's SQuantOp expects; squantop.cpp:134 stores (ij|kl) at [i,k,j,l]).
"""
import numpy as np


def seeded_integrals(nbasis, nchol, seed, hdiag0=-2.0, hstep=0.25, offdiag_std=0.05, l_std=0.1, decay=None):
    n = nbasis
    rng = np.random.default_rng(seed)
    one = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    one[iu] = rng.normal(0.0, offdiag_std, size=iu[0].size)
    one = one + one.T
    one[np.diag_indices(n)] = hdiag0 + hstep * np.arange(n)
    iu0 = np.triu_indices(n, 0)
    scale = None
    if decay is not None:
        idx = np.arange(n)
        scale = np.exp(-np.abs(idx[:, None] - idx[None, :]) / decay)
    Ls = np.zeros((nchol, n, n))
    for p in range(nchol):
        L = np.zeros((n, n))
        L[iu0] = rng.normal(0.0, l_std, size=iu0[0].size)
        L = L + np.triu(L, 1).T
        if scale is not None:
            L = L * scale
        Ls[p] = L
    eri = np.einsum("pij,pkl->ijkl", Ls, Ls)
    two = np.ascontiguousarray(eri.transpose(0, 2, 1, 3))
    return 0.0, one, two


def config_integrals(name):
    """Integrals of the named workloads: 'fci8' (nbasis 8, 32 vectors, seed 1000),
    'fci12' (nbasis 12, 48 vectors, seed 1001), 'fci14' (nbasis 14, 48 vectors, seed 1002),
    'genci24' (24 spatial orbitals, 64 decaying vectors, seed 1003)."""
    if name == "fci8":
        return seeded_integrals(8, 32, 1000)
    if name == "fci12":
        return seeded_integrals(12, 48, 1001)
    if name == "fci14":
        return seeded_integrals(14, 48, 1002)
    if name == "genci24":
        return seeded_integrals(24, 64, 1003, hdiag0=-3.0, hstep=0.2, decay=6.0)
    raise ValueError(name)


def spin_orbital(one_mo, two_mo):
    """Spatial -> spin-orbital integrals (alpha 0..n-1, beta n..2n-1); <pq|rs> is
    nonzero only when spin(p) = spin(r) and spin(q) = spin(s)."""
    n = one_mo.shape[0]
    h = np.zeros((2 * n, 2 * n))
    h[:n, :n] = one_mo
    h[n:, n:] = one_mo
    g = np.zeros((2 * n,) * 4)
    for s1 in (0, n):
        for s2 in (0, n):
            g[s1:s1 + n, s2:s2 + n, s1:s1 + n, s2:s2 + n] = two_mo
    return h, g


def batch_problems(count=64, seed=4000, nchol=32):
    """The uneven-batch workload: `count` FullCI problems with nbasis ~ U{8..12} and
    m ~ U{3..6} electrons of each spin, drawn (with lambda ~ U[0, 1]) from
    default_rng(seed) in the order nbasis, m, lambda per problem; integrals
    (1 - lambda) A + lambda B, with A and B the recipe above (h_ii = -2 + 0.25 i,
    `nchol` vectors) at that nbasis with seeds 2000 + k and 3000 + k."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        nb = int(rng.integers(8, 13))
        m = int(rng.integers(3, 7))
        lam = float(rng.random())
        _, oa, ta = seeded_integrals(nb, nchol, 2000 + k)
        _, ob, tb = seeded_integrals(nb, nchol, 3000 + k)
        out.append(dict(nbasis=nb, nocc=m, lam=lam, one=(1 - lam) * oa + lam * ob, two=(1 - lam) * ta + lam * tb))
    return out
