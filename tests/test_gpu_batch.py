"""Batched uneven FullCI solves (pyci_batch_solve) against per-problem solves and
the oracle."""
import numpy as np
import pytest

import pyci_amd as pyci
from pyci_amd.synthetic import batch_problems, seeded_integrals

pytestmark = pytest.mark.gpu

SHAPES = [(6, 3, 3), (7, 3, 2), (8, 3, 3), (6, 2, 2), (8, 4, 4), (7, 4, 3), (5, 2, 2)]


def problems(shapes, seed0=10):
    out = []
    for i, (nb, nu, nd) in enumerate(shapes):
        e, one, two = seeded_integrals(nb, 8, seed0 + i)
        out.append((pyci.hamiltonian(0.3 * i, one, two), nu, nd, one, two))
    return out


def oracle_roots(oracle, nb, nu, nd, one, two, n, ecore):
    dets = oracle.all_dets("fullci", nb, nu, nd)
    w = oracle.lowest_eigs(oracle.csr_build("fullci", nb, nu, nd, dets, one, two), n)[0]
    return np.sort(w) + ecore


@pytest.mark.parametrize("precision", ["f64", "mixed"])
@pytest.mark.parametrize("n", [1, 2, 4])
def test_batch_matches_oracle(oracle, precision, n):
    probs = problems(SHAPES)
    ev, info = pyci.batch_solve([(h, nu, nd) for h, nu, nd, _, _ in probs], n=n, tol=1e-10, precision=precision)
    assert ev.shape == (len(probs), n)
    assert all(info["converged"])
    for p, (h, nu, nd, one, two) in enumerate(probs):
        ref = oracle_roots(oracle, h.nbasis, nu, nd, one, two, n, h.ecore)
        got = ev[p]
        if n > 1:
            assert np.all(np.diff(got) <= 0)  # descending, as solve()
        np.testing.assert_allclose(np.sort(got), ref, rtol=0, atol=1e-9)
    lat = np.asarray(info["latency_ms"])
    assert np.all(lat > 0) and np.all(lat <= info["total_ms"])
    assert np.all(np.asarray(info["sigma_vectors"]) >= n)
    assert info["rows_per_block"] in (256, 1024, 4096)


def test_batch_matches_single_solves_on_uneven_workload():
    """A slice of the 64-problem integral scan (nbasis 8..12, 3..6 electrons of each spin)."""
    ps = batch_problems(count=64)
    pick = [p for p in ps if p["nbasis"] <= 10][:8]
    hams = [pyci.hamiltonian(0.0, q["one"], q["two"]) for q in pick]
    ev, info = pyci.batch_solve([(h, q["nocc"], q["nocc"]) for h, q in zip(hams, pick)], n=2, tol=1e-9,
                                precision="mixed")
    assert all(info["converged"])
    for h, q, e in zip(hams, pick, ev):
        w = pyci.fullci_wfn(q["nbasis"], q["nocc"], q["nocc"])
        w.add_all_dets()
        ref, _ = pyci.sparse_op(h, w).solve(n=2, tol=1e-10)
        np.testing.assert_allclose(e, ref, rtol=0, atol=1e-8)


def test_batch_errors():
    probs = problems(SHAPES[:2])
    with pytest.raises(ValueError):
        pyci.batch_solve([(h, nu, nd) for h, nu, nd, _, _ in probs], n=9)
    with pytest.raises(ValueError):
        pyci.batch_solve([(h, nu, nd) for h, nu, nd, _, _ in probs], precision="f16")
    with pytest.raises(ValueError):
        pyci.batch_solve([])
    e, one, two = seeded_integrals(3, 4, 1)
    with pytest.raises(ValueError):  # 9 determinants: smaller than the subspace
        pyci.batch_solve([(pyci.hamiltonian(0.0, one, two), 1, 1)], n=2)


def test_no_root_skipped_with_singlet_start():
    """Problems 35 and 58 of the 64-problem scan have a second root that the
    closed-shell lowest-diagonal start vectors do not reach (a k-vector start
    converged to roots 1 and 3): both solvers must return the two lowest roots,
    checked against the lowest two of a six-root FP64 solve."""
    ps = batch_problems(count=64)
    pick = [ps[35], ps[58]]
    hams = [pyci.hamiltonian(0.0, q["one"], q["two"]) for q in pick]
    ev, info = pyci.batch_solve([(h, q["nocc"], q["nocc"]) for h, q in zip(hams, pick)], n=2, tol=1e-8,
                                precision="mixed")
    for h, q, e in zip(hams, pick, ev):
        w = pyci.fullci_wfn(q["nbasis"], q["nocc"], q["nocc"])
        w.add_all_dets()
        op = pyci.sparse_op(h, w)
        six, _ = op.solve(n=6, tol=1e-10)
        ref = np.sort(six)[:2]
        np.testing.assert_allclose(np.sort(e), ref, rtol=0, atol=1e-7)
        two, _ = op.solve(n=2, tol=1e-10)
        np.testing.assert_allclose(np.sort(two), ref, rtol=0, atol=1e-9)


@pytest.mark.parametrize("precision", ["mixed", "f64"])
def test_batch_routes_sizes_to_both_paths(precision):
    """Problems of 150,000+ determinants take the single-problem solver beside the
    replayed graph; mid-size ones (65,536..150,000) stay in the graph on its fused
    pass; small ones on its batched kernels. Every root equals a per-problem FP64
    solve of the same operator."""
    shapes = [(11, 4, 4), (12, 4, 4), (9, 4, 4), (8, 3, 3), (11, 4, 4)]
    probs = []
    for i, (nb, nu, nd) in enumerate(shapes):
        _, one, two = seeded_integrals(nb, 16, 900 + i)
        probs.append((pyci.hamiltonian(0.1 * i, one, two), nu, nd))
    ev, info = pyci.batch_solve(probs, n=2, tol=1e-8, precision=precision)
    assert info["single_path"] == 1  # only the 245,025-determinant problem
    assert all(info["converged"])
    for (h, nu, nd), e in zip(probs, ev):
        w = pyci.fullci_wfn(h.nbasis, nu, nd)
        w.add_all_dets()
        ref, _ = pyci.sparse_op(h, w).solve(n=2, tol=1e-10)
        np.testing.assert_allclose(e, ref, rtol=0, atol=1e-8)
