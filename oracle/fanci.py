"""CPU restatement of the FanCI geminal objectives (test infrastructure only).

AP1roG and APIG overlaps (ap1rog.cpp:55-220, apig.cpp:55-195) as permanents
evaluated by brute force over permutations (independent of the reference's Ryser
formula), their derivatives as permanents of minors, and the objective /
Jacobian of fanci.cpp:129-217 over the oracle's rectangular CSR.
"""
import itertools

import numpy as np

from oracle import oracle as O


def permanent(a):
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[0]
    if n == 0:
        return 1.0
    return float(sum(np.prod([a[i, s[i]] for i in range(n)]) for s in itertools.permutations(range(n))))


def _bits(d):
    return [p for p in range(64) if (int(d) >> p) & 1]


def overlaps(kind, nbasis, nocc, dets, x):
    """<d|Psi(x)> for pair determinants `dets` (one uint64 each)."""
    out = np.zeros(len(dets))
    if kind == "ap1rog":
        c = np.asarray(x, dtype=np.float64)[: nocc * (nbasis - nocc)].reshape(nocc, nbasis - nocc)
        hf = (1 << nocc) - 1
        for k, d in enumerate(dets):
            holes = _bits(hf & ~int(d))
            parts = [p - nocc for p in _bits(int(d) & ~hf)]
            out[k] = permanent(c[np.ix_(holes, parts)])
    else:
        c = np.asarray(x, dtype=np.float64)[: nbasis * nocc].reshape(nbasis, nocc)
        for k, d in enumerate(dets):
            out[k] = permanent(c[_bits(d), :])
    return out


def d_overlaps(kind, nbasis, nocc, dets, x):
    """d<d|Psi>/dx_k, shape (ndet, nparam): permanents of minors."""
    if kind == "ap1rog":
        nv = nbasis - nocc
        c = np.asarray(x, dtype=np.float64)[: nocc * nv].reshape(nocc, nv)
        hf = (1 << nocc) - 1
        out = np.zeros((len(dets), nocc * nv))
        for k, d in enumerate(dets):
            holes = _bits(hf & ~int(d))
            parts = [p - nocc for p in _bits(int(d) & ~hf)]
            for ih, h in enumerate(holes):
                for ip, p in enumerate(parts):
                    rows = holes[:ih] + holes[ih + 1:]
                    cols = parts[:ip] + parts[ip + 1:]
                    out[k, h * nv + p] = permanent(c[np.ix_(rows, cols)])
        return out
    c = np.asarray(x, dtype=np.float64)[: nbasis * nocc].reshape(nbasis, nocc)
    out = np.zeros((len(dets), nbasis * nocc))
    for k, d in enumerate(dets):
        occ = _bits(d)
        for io, orb in enumerate(occ):
            for g in range(nocc):
                rows = occ[:io] + occ[io + 1:]
                cols = [j for j in range(nocc) if j != g]
                out[k, orb * nocc + g] = permanent(c[np.ix_(rows, cols)])
    return out


class Objective:
    """AP1roGObjective / APIGObjective restated over the oracle's CSR."""

    def __init__(self, kind, nbasis, nocc, dets, one_mo, two_mo, nproj, idx_det=(), val_det=(), idx_par=(),
                 val_par=()):
        self.kind, self.nbasis, self.nocc = kind, nbasis, nocc
        self.dets = np.asarray(dets, dtype=np.uint64).reshape(-1)
        self.nproj, self.nconn = nproj, self.dets.size
        self.nparam = nocc * (nbasis - nocc) if kind == "ap1rog" else nocc * nbasis
        csr = O.csr_build("doci", nbasis, nocc, nocc, self.dets, one_mo, two_mo, nrow=nproj, ncol=self.nconn,
                          symmetric=False)
        self.A = csr.scipy()
        self.idx_det, self.val_det = list(idx_det), list(val_det)
        self.idx_par, self.val_par = list(idx_par), list(val_par)
        self.nequation = nproj + len(self.idx_det) + len(self.idx_par)

    def overlap(self, x):
        return overlaps(self.kind, self.nbasis, self.nocc, self.dets, x)

    def d_overlap(self, x):
        return d_overlaps(self.kind, self.nbasis, self.nocc, self.dets, x)

    def objective(self, x):
        x = np.asarray(x, dtype=np.float64)
        e, c = x[self.nparam], self.overlap(x)
        f = np.zeros(self.nequation)
        f[: self.nproj] = self.A @ c - e * c[: self.nproj]
        for i, (j, v) in enumerate(zip(self.idx_det, self.val_det)):
            f[self.nproj + i] = c[j] - v
        for i, (j, v) in enumerate(zip(self.idx_par, self.val_par)):
            f[self.nproj + len(self.idx_det) + i] = x[j] - v
        return f

    def jacobian(self, x):
        x = np.asarray(x, dtype=np.float64)
        e, c, D = x[self.nparam], self.overlap(x), self.d_overlap(x)
        J = np.zeros((self.nequation, self.nparam + 1))
        J[: self.nproj, : self.nparam] = self.A @ D - e * D[: self.nproj]
        J[: self.nproj, self.nparam] = -c[: self.nproj]
        for i, j in enumerate(self.idx_det):
            J[self.nproj + i, : self.nparam] = D[j]
        for i, j in enumerate(self.idx_par):
            J[self.nproj + len(self.idx_det) + i, j] = 1.0
        return J
