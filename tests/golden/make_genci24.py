"""Writes tests/golden/data/genci24.npz: two heat-bath selection steps of the
GenCI workload of 48 spin-orbitals and 12 electrons, computed by the CPU oracle.

Integrals: pyci_amd.synthetic.config_integrals("genci24") (24 spatial orbitals,
64 decaying vectors, seed 1003) in spin-orbital form (alpha 0..23, beta 24..47).
Start: genci_wfn(48, 12, 0) with the Hartree-Fock determinant (bits 0-5 and
24-29), then its singles and doubles (add_excitations(wfn, 0, 1, 2, ref=hf),
excitation_ci.py:23-39; 42,013 determinants).

  S0 = HF + SD; (E0, c0) = lowest root (oracle CSR + ARPACK, tol 1e-12)
  S1 = S0 + add_hci(S0, c0, eps1)   (hci.cpp:184-279, first-occurrence order)
  (E1, c1) = lowest root on S1
  S2 = S1 + add_hci(S1, c1, eps2);  E2 = lowest root on S2
  pool = add_hci(S0, c0, eps1 / 5)  (the learned selector's candidate pool)

The coefficient vectors handed to add_hci are stored in float32 and used as
float64 by both sides (the oracle here, the GPU in the test), so the selection
threshold sees the same numbers on both. Determinant sets are stored as counts
and SHA-256 digests of the uint64 arrays. Run from the repository root:
python tests/golden/make_genci24.py (about a minute on 8 threads).
"""
import hashlib
import os
import sys
import time

import numpy as np
import scipy.sparse.linalg as sla

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pyci_amd.synthetic import config_integrals, spin_orbital  # noqa: E402
from tests.golden import oracle_big as OB  # noqa: E402

NB, NE = 48, 12
EPS1, EPS2, POOL = 1.0e-3, 5.0e-4, 5.0


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def hartreefock():
    return np.array([((1 << 6) - 1) | (((1 << 6) - 1) << 24)], dtype=np.uint64)


def lowest(dets, one, two):
    A = OB.build_lower("genci", NB, NE, 0, dets, one, two)
    n = A.n
    op = sla.LinearOperator((n, n), matvec=A.symv, dtype=np.float64)
    v0 = np.full(n, 1.0e-3)
    v0[0] = 1.0
    w, v = sla.eigsh(op, k=1, which="SA", tol=1e-12, v0=v0)
    c = v[:, 0] * np.sign(v[0, 0])
    r = np.linalg.norm(A.symv(c) - w[0] * c)
    print(f"  {n} dets: E = {w[0]:.12f}, residual {r:.1e}", flush=True)
    return float(w[0]), c


def main():
    t0 = time.perf_counter()
    _, h, g = config_integrals("genci24")
    one, two = spin_orbital(h, g)
    hf = hartreefock()
    s0 = O.add_excited_dets("genci", NB, NE, 0, hf, hf, 1)
    s0 = O.add_excited_dets("genci", NB, NE, 0, s0, hf, 2)
    e0, c0 = lowest(s0, one, two)
    c0 = c0.astype(np.float32)
    new1 = O.add_hci("genci", NB, NE, 0, s0, one, two, c0.astype(np.float64), EPS1)
    pool = O.add_hci("genci", NB, NE, 0, s0, one, two, c0.astype(np.float64), EPS1 / POOL)
    s1 = np.concatenate([s0, new1])
    e1, c1 = lowest(s1, one, two)
    c1 = c1.astype(np.float32)
    new2 = O.add_hci("genci", NB, NE, 0, s1, one, two, c1.astype(np.float64), EPS2)
    s2 = np.concatenate([s1, new2])
    e2, _ = lowest(s2, one, two)
    out = os.path.join(ROOT, "tests", "golden", "data", "genci24.npz")
    np.savez_compressed(
        out, eps=np.array([EPS1, EPS2, POOL]), energies=np.array([e0, e1, e2]), c0=c0, c1=c1,
        counts=np.array([s0.size, new1.size, new2.size, pool.size]),
        sha=np.array([sha(s0), sha(new1), sha(new2), sha(pool)]),
        first=np.stack([new1[:8], new2[:8], pool[:8]]), last=np.stack([new1[-8:], new2[-8:], pool[-8:]]))
    print(f"wrote {out}: sets {s0.size} + {new1.size} + {new2.size}, pool {pool.size}, "
          f"{time.perf_counter() - t0:.0f} s", flush=True)


if __name__ == "__main__":
    main()
