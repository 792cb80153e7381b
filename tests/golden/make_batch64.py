"""Writes tests/golden/data/batch64.npz: CPU-oracle golden roots for the 64-problem
uneven batch (pyci_amd.synthetic.batch_problems(): FullCI, nbasis 8..12, 3..6
electrons of each spin, 784..853,776 determinants, integrals (1-l) A + l B).

For every problem the oracle's full space (add_all_dets restated) and lower CSR
(SparseOp restated) are built on the CPU and the 2 lowest eigenvalues taken by
ARPACK at tol 1e-12 (dense eigh below 3,000 determinants). Also stored: the
sha256 of every problem's integrals, so a change of the recipe is caught.

Run in the build container (about 20 minutes on 8 cores):
    python tests/golden/make_batch64.py
"""
import hashlib
import os
import sys
import time
from math import comb

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pyci_amd.synthetic import batch_problems  # noqa: E402
from tests.golden import oracle_big as B  # noqa: E402

NROOT = 2


def main():
    t0 = time.perf_counter()
    probs = batch_problems()
    evals = np.zeros((len(probs), NROOT))
    resid = np.zeros((len(probs), NROOT))
    shas = []
    nds = np.zeros(len(probs), dtype=np.int64)
    for p, q in enumerate(probs):
        nb, m = q["nbasis"], q["nocc"]
        nd = comb(nb, m) ** 2
        nds[p] = nd
        one, two = np.ascontiguousarray(q["one"]), np.ascontiguousarray(q["two"])
        shas.append(hashlib.sha256(one.tobytes() + two.tobytes()).hexdigest())
        dets = O.all_dets("fullci", nb, m, m)
        csr = B.build_lower("fullci", nb, m, m, dets, one, two, log=None)
        if nd < 3000:
            a = np.zeros((nd, nd))
            for i in range(nd):
                s, e = csr.indptr[i], csr.indptr[i + 1]
                a[i, csr.indices[s:e]] = csr.data[s:e]
            a = a + np.tril(a, -1).T
            w, v = np.linalg.eigh(a)
            w, v = w[:NROOT], v[:, :NROOT].T
        else:
            w, v = B.lowest(csr, NROOT, tol=1e-12, log=None)
        evals[p] = w
        resid[p] = [np.linalg.norm(csr.symv(v[k]) - w[k] * v[k]) for k in range(NROOT)]
        print(f"problem {p}: nbasis {nb}, {m}+{m}, {nd} dets, roots {w}, max residual {resid[p].max():.1e}, "
              f"{time.perf_counter() - t0:.0f} s", flush=True)
        del csr
    path = os.path.join(ROOT, "tests", "golden", "data", "batch64.npz")
    np.savez_compressed(path, evals=evals, resid=resid, ndet=nds, sha=np.array(shas))
    print(f"wrote {path} in {time.perf_counter() - t0:.0f} s")


if __name__ == "__main__":
    main()
