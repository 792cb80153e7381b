"""Writes tests/golden/data/fci12.npz: CPU-oracle golden values for FullCI nbasis 12,
6 alpha + 6 beta electrons (853,776 determinants), seeded integrals (48 vectors,
seed 1001) -- the full-size workload of the GPU parity tests and the benchmark.

Contents (all from the CPU oracle; nothing from the GPU):
  * sha256 of one_mo / two_mo (pins the integral recipe)
  * e8: the 8 lowest eigenvalues (ARPACK on the oracle's lower CSR, tol 1e-12)
  * resid: ||H c - e c|| of each returned vector
  * y = H x for x = default_rng(7).standard_normal(nd): values at 2048 seeded rows,
    two seeded projections, norm, sum (see oracle_big.checksums, seed 11)
  * diag: the same fingerprint of the diagonal (seed 13)
  * rdm1 (2,12,12) of the ground state, and of rdm2 (3,12,12,12,12): the energy
    rebuilt from the RDMs, its traces, and two seeded projections (seed 17)

Run in the build container (about 5 minutes on 8 cores, ~20 GB of memory):
    python tests/golden/make_fci12.py
"""
import hashlib
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from pyci_amd.synthetic import config_integrals  # noqa: E402
from tests.golden import oracle_big as B  # noqa: E402


def rdm_energy(one, two, r1, r2):
    """E = sum h.gamma + 1/2 sum <pq|rs> Gamma for FullCI spin-resolved RDMs in the
    reference's layout (rdm1[s][p,q] = <a+_ps a_qs>; rdm2[0] aaaa, [1] bbbb,
    [2] abab = <a+_pa a+_qb a_sb a_ra>)."""
    e1 = np.einsum("pq,pq", one, r1[0] + r1[1])
    e2 = 0.5 * (np.einsum("pqrs,pqrs", two, r2[0]) + np.einsum("pqrs,pqrs", two, r2[1])) + np.einsum(
        "pqrs,pqrs", two, r2[2])
    return e1 + e2


def main():
    t0 = time.perf_counter()
    ecore, one, two = config_integrals("fci12")
    nb, nup, ndn = 12, 6, 6
    dets = O.all_dets("fullci", nb, nup, ndn)
    csr = B.build_lower("fullci", nb, nup, ndn, dets, one, two)
    nd = csr.n
    x = np.random.default_rng(7).standard_normal(nd)
    y = csr.symv(x)
    diag = csr.diagonal()
    w, v = B.lowest(csr, 8, tol=1e-12, ncv=40)
    resid = np.array([np.linalg.norm(csr.symv(v[k]) - w[k] * v[k]) for k in range(8)])
    print("roots", w, "residuals", resid)
    t1 = time.perf_counter()
    r1, r2 = O.rdms("fullci", nb, dets, v[0])
    print(f"oracle RDMs: {time.perf_counter() - t1:.1f} s")
    erdm = rdm_energy(one, two, r1, r2)
    print("E from RDMs", erdm, "E0", w[0], "diff", erdm - w[0])
    ys = B.checksums(y, 11)
    ds = B.checksums(diag, 13)
    p2 = np.random.default_rng(17).standard_normal((2, r2.size))
    out = dict(
        one_sha=hashlib.sha256(np.ascontiguousarray(one).tobytes()).hexdigest(),
        two_sha=hashlib.sha256(np.ascontiguousarray(two).tobytes()).hexdigest(),
        nd=nd, nnz_lower=csr.data.size, e8=w, resid=resid,
        y_rows=ys["rows"], y_vals=ys["vals"], y_proj=ys["proj"], y_norm=ys["norm"], y_sum=ys["sum"],
        d_rows=ds["rows"], d_vals=ds["vals"], d_proj=ds["proj"], d_norm=ds["norm"], d_sum=ds["sum"],
        rdm1=r1, rdm2_proj=p2 @ r2.reshape(-1), rdm2_tr=np.array([np.einsum("pqpq", r2[k]) for k in range(3)]),
        rdm_energy=erdm,
    )
    path = os.path.join(ROOT, "tests", "golden", "data", "fci12.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes) in {time.perf_counter() - t0:.0f} s")


if __name__ == "__main__":
    main()
