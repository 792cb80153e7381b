"""The benchmark configurations at full size on the GPU, against CPU-oracle goldens.

  * FullCI nbasis 12, 6+6 (853,776 determinants), seeded integrals (48 vectors,
    seed 1001): H x on the CSR and the matrix-free operator, FP64 solves on both,
    compute_rdms, and the mixed-precision 8-root block solve -- all against
    tests/golden/data/fci12.npz (tests/golden/make_fci12.py: the oracle's lower
    CSR, its threaded symmetric product and ARPACK at tol 1e-12).
  * FullCI nbasis 14, 7+7 (11,778,624 determinants; the CSR would not fit in HBM):
    the 8-root mixed solve checked through size-independent properties -- FP64
    matrix-free residuals, orthonormality, agreement with an FP64 solve.
  * The 64-problem uneven batch (nbasis 8..12, 3..6 electrons of each spin):
    every root of one batch_solve call, FP64 and mixed, against
    tests/golden/data/batch64.npz (the oracle's CSR + ARPACK for every problem).
  * GenCI over 48 spin-orbitals with 12 electrons (24 spatial orbitals, 64
    decaying vectors, seed 1003): HF + singles + doubles, two heat-bath steps and
    the learned selector's step, against tests/golden/data/genci24.npz.
Reference path: sparseop.cpp:96-146 (perform_op, solve_ci), rdm.cpp:269-530,
hci.cpp:184-279.
"""
import hashlib

import numpy as np
import numpy.testing as npt
import pytest

import pyci_amd as pyci
from pyci_amd.synthetic import batch_problems, config_integrals
from tests.conftest import datafile

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fci12():
    g = dict(np.load(datafile("fci12.npz")))
    ecore, one, two = config_integrals("fci12")
    assert hashlib.sha256(one.tobytes()).hexdigest() == str(g["one_sha"])
    assert hashlib.sha256(two.tobytes()).hexdigest() == str(g["two_sha"])
    ham = pyci.hamiltonian(ecore, one, two)
    wfn = pyci.fullci_wfn(12, 6, 6)
    wfn.add_all_dets()
    assert len(wfn) == int(g["nd"])
    return g, ham, wfn, one, two


def _check_vector(y, g, key, rtol):
    """y against the golden fingerprint (rows, two seeded projections, norm, sum)."""
    scale = np.linalg.norm(y)
    npt.assert_allclose(y[g[key + "_rows"]], g[key + "_vals"], rtol=0, atol=rtol * np.abs(g[key + "_vals"]).max())
    rng = np.random.default_rng({"y": 11, "d": 13}[key])
    rng.choice(y.size, size=min(2048, y.size), replace=False)
    p = rng.standard_normal((2, y.size))
    npt.assert_allclose(p @ y, g[key + "_proj"], rtol=0, atol=rtol * scale * np.sqrt(y.size))
    npt.assert_allclose(scale, g[key + "_norm"], rtol=rtol)
    npt.assert_allclose(y.sum(), g[key + "_sum"], rtol=0, atol=rtol * scale * np.sqrt(y.size))


@pytest.mark.timeout(300)
@pytest.mark.parametrize("storage", ["csr", "direct"])
def test_fci12_sigma_and_diagonal(fci12, storage):
    g, ham, wfn, _, _ = fci12
    op = pyci.sparse_op(ham, wfn, storage=storage)
    assert op.storage == storage
    if storage == "csr":
        assert op.size == int(g["nnz_lower"])
    x = np.random.default_rng(7).standard_normal(len(wfn))
    _check_vector(op(x), g, "y", 1e-12)
    _check_vector(op.diagonal(), g, "d", 1e-13)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("storage", ["csr", "direct"])
def test_fci12_solve_and_rdms(fci12, storage):
    g, ham, wfn, one, two = fci12
    op = pyci.sparse_op(ham, wfn, storage=storage)
    es, cs = op.solve(n=1, tol=1e-10)
    assert abs(es[0] - g["e8"][0]) < 1e-10, es[0] - g["e8"][0]
    if storage != "direct":
        return
    r1, r2 = pyci.compute_rdms(wfn, cs[0])
    npt.assert_allclose(r1, g["rdm1"], rtol=0, atol=1e-9)
    assert abs(np.trace(r1[0]) + np.trace(r1[1]) - 12.0) < 1e-10
    tr = np.array([np.einsum("pqpq", r2[k]) for k in range(3)])
    npt.assert_allclose(tr, g["rdm2_tr"], rtol=0, atol=1e-9)
    p2 = np.random.default_rng(17).standard_normal((2, r2.size))
    npt.assert_allclose(p2 @ r2.reshape(-1), g["rdm2_proj"], rtol=0, atol=1e-9 * np.sqrt(r2.size))
    e1 = np.einsum("pq,pq", one, r1[0] + r1[1])
    e2 = 0.5 * (np.einsum("pqrs,pqrs", two, r2[0]) + np.einsum("pqrs,pqrs", two, r2[1])) + np.einsum(
        "pqrs,pqrs", two, r2[2])
    assert abs(e1 + e2 - g["e8"][0]) < 1e-9


@pytest.mark.timeout(300)
def test_fci12_mixed_block_solve_8_roots(fci12):
    """Mixed precision (FP32 sigma, FP64 correction) block solve, nroot 8: every
    root within 1e-8 Eh of the oracle's (returned in the reference's descending
    order), FP64 residuals within the tolerance."""
    g, ham, wfn, _, _ = fci12
    op = pyci.sparse_op(ham, wfn)
    tol = 1e-7
    es, cs = op.solve(n=8, tol=tol, precision="mixed")
    assert np.all(np.diff(es) <= 0)
    npt.assert_allclose(es[::-1], g["e8"], rtol=0, atol=1e-8)
    for k in range(8):
        r = op(cs[k]) - (es[k] - ham.ecore) * cs[k]
        assert np.linalg.norm(r) <= 1.5 * tol * abs(es[k])


@pytest.mark.timeout(600)
def test_fci14_mixed_block_solve_properties():
    """nbasis 14, 7+7, nroot 8, mixed: FP64 matrix-free residuals within the
    tolerance, orthonormal vectors, roots within 1e-8 Eh of an FP64 solve at
    tol 1e-10 of the same operator."""
    ecore, one, two = config_integrals("fci14")
    ham = pyci.hamiltonian(ecore, one, two)
    wfn = pyci.fullci_wfn(14, 7, 7)
    wfn.add_all_dets()
    assert len(wfn) == 11778624
    op = pyci.sparse_op(ham, wfn)
    assert op.storage == "direct"
    tol = 1e-7
    es, cs = op.solve(n=8, tol=tol, precision="mixed")
    assert np.all(np.diff(es) <= 0)
    gram = cs @ cs.T
    assert np.abs(gram - np.eye(8)).max() <= 1e-10
    for k in range(8):
        r = op(cs[k]) - (es[k] - ecore) * cs[k]
        assert np.linalg.norm(r) <= 1.5 * tol * abs(es[k])
    del cs
    ef, _ = op.solve(n=8, tol=1e-10, precision="f64")
    npt.assert_allclose(es, ef, rtol=0, atol=1e-8)


@pytest.fixture(scope="module")
def batch64():
    g = dict(np.load(datafile("batch64.npz")))
    probs = batch_problems()
    for q, sha in zip(probs, g["sha"]):
        h = hashlib.sha256(np.ascontiguousarray(q["one"]).tobytes() + np.ascontiguousarray(q["two"]).tobytes())
        assert h.hexdigest() == str(sha)
    return g, probs


@pytest.mark.timeout(300)
@pytest.mark.parametrize("precision", ["f64", "mixed"])
def test_batch64_all_problems_match_oracle(batch64, precision):
    """One batch_solve call over all 64 problems (784..853,776 determinants, 18 of
    them of 65,536+ on the fused matrix-free path): both roots of every problem
    within 1e-8 Eh of the oracle's."""
    g, probs = batch64
    items = [(pyci.hamiltonian(0.0, q["one"], q["two"]), q["nocc"], q["nocc"]) for q in probs]
    ev, info = pyci.batch_solve(items, n=2, tol=1e-7, precision=precision)
    assert ev.shape == (64, 2) and all(info["converged"])
    assert np.all(np.diff(ev, axis=1) <= 0)  # descending, as solve()
    npt.assert_allclose(ev[:, ::-1], g["evals"], rtol=0, atol=1e-8)
    lat = np.asarray(info["latency_ms"])
    assert np.all(lat > 0) and np.all(lat <= info["total_ms"])


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def genci24():
    """48 spin-orbitals, 12 electrons: genci_wfn(48, 12, 0) with an explicit
    Hartree-Fock reference and its singles and doubles (42,013 determinants)."""
    from pyci_amd.selection import hartreefock_genci
    from pyci_amd.synthetic import spin_orbital

    g = dict(np.load(datafile("genci24.npz")))
    _, h, gg = config_integrals("genci24")
    one, two = spin_orbital(h, gg)
    ham = pyci.hamiltonian(0.0, one, two)
    hf = hartreefock_genci(24, 6, 6)

    def start():
        w = pyci.genci_wfn(48, 12, 0)
        pyci.add_excitations(w, 0, 1, 2, ref=hf)
        return w

    return g, ham, start


@pytest.mark.timeout(300)
def test_genci24_two_heat_bath_steps_match_oracle(genci24):
    """Two add_hci -> update -> solve steps (eps 1e-3, then 5e-4) on the GenCI
    workload: each appended set (count, order, SHA-256) equals the oracle's, and
    each variational energy is within 1e-8 Eh of the oracle's ARPACK root
    (tests/golden/make_genci24.py; hci.cpp:184-279, sparseop.cpp:429-502)."""
    g, ham, start = genci24
    eps1, eps2, _ = g["eps"]
    wfn = start()
    d0 = wfn.to_det_array().reshape(-1)
    assert d0.size == int(g["counts"][0]) and _sha(d0) == str(g["sha"][0])
    op = pyci.sparse_op(ham, wfn)
    es, _ = op.solve(n=1, tol=1e-10)
    assert abs(es[0] - g["energies"][0]) < 1e-8
    n0 = len(wfn)
    added = pyci.add_hci(ham, wfn, g["c0"].astype(np.float64), eps=eps1)
    assert added == int(g["counts"][1])
    d1 = wfn.to_det_array().reshape(-1)
    assert _sha(d1[n0:]) == str(g["sha"][1])
    op.update(ham, wfn)
    es, _ = op.solve(n=1, tol=1e-10)
    assert abs(es[0] - g["energies"][1]) < 1e-8
    n1 = len(wfn)
    added = pyci.add_hci(ham, wfn, g["c1"].astype(np.float64), eps=eps2)
    assert added == int(g["counts"][2])
    assert _sha(wfn.to_det_array().reshape(-1)[n1:]) == str(g["sha"][2])
    op.update(ham, wfn)
    es, _ = op.solve(n=1, tol=1e-10)
    assert abs(es[0] - g["energies"][2]) < 1e-8


@pytest.mark.timeout(300)
def test_genci24_learned_step_selects_from_oracle_pool(genci24):
    """The learned selector's step from the same start: its candidate pool (plain
    add_hci at eps / pool) equals the oracle's, it appends as many determinants as
    the plain step, and they are the network's top scores over the pool, in
    pool order; its energy is below the start space's and is printed beside the plain one."""
    from oracle import mlp as M

    g, ham, start = genci24
    eps1, _, pool = g["eps"]
    c0 = g["c0"].astype(np.float64)
    w_pool = start()
    n0 = len(w_pool)
    assert pyci.add_hci(ham, w_pool, c0, eps=eps1 / pool) == int(g["counts"][3])
    cand = w_pool.to_det_array().reshape(-1)[n0:]
    assert _sha(cand) == str(g["sha"][3])
    wfn = start()
    sel = pyci.mlp_selector(48, seed=0)
    loss = sel.train(wfn, c0, steps=200, batch=8192, lr=1e-3, cmin=1e-10, seed=0)
    assert np.isfinite(loss)
    y = np.asarray(sel.predict(w_pool), dtype=np.float32)[n0:]
    expect = cand[M.topk_select(y, int(g["counts"][1]))]
    added = pyci.add_hci(ham, wfn, c0, eps=eps1, selector=sel, pool=pool)
    assert added == int(g["counts"][1])
    npt.assert_array_equal(wfn.to_det_array().reshape(-1)[n0:], expect)
    op = pyci.sparse_op(ham, wfn)
    es, _ = op.solve(n=1, tol=1e-10)
    print(f"genci24 step 1 at {len(wfn)} dets: learned {es[0]:.10f}, plain {g['energies'][1]:.10f}")
    assert es[0] <= g["energies"][0] + 1e-9  # the start space is a subset


@pytest.mark.timeout(300)
def test_genci24_long_rows_sorted(genci24):
    """Rows of a selected set are sorted by column in three tiers (up to 1,024
    entries, up to 4,096, longer: a segmented radix sort over a gathered copy). The
    HF + singles + doubles start of the GenCI workload has a row of 42,013 entries
    (HF connects to every other determinant) and rows in the two upper tiers (the
    first is covered by the small sets of test_gpu_csr.py): in the full-row
    operator every row is strictly increasing, HF's row has all 42,013 columns, and
    binary-search reads of it agree with the other rows' entries."""
    g, ham, start = genci24
    wfn = start()
    nd = len(wfn)
    op = pyci.sparse_op(ham, wfn, symmetric=False)
    indptr, indices = np.asarray(op.indptr()), np.asarray(op.indices())
    lens = np.diff(indptr)
    assert lens[0] == nd
    assert np.any(lens > 4096) and np.any((lens > 1024) & (lens <= 4096))
    step = np.diff(indices)
    inside = np.ones(step.size, dtype=bool)
    inside[indptr[1:-1] - 1] = False  # the step across a row boundary
    assert np.all(step[inside] > 0)
    # spin-changing doubles couple to HF with exactly zero (spin-orbital integrals),
    # so only the agreement is required, with enough nonzero samples to mean something
    rng = np.random.default_rng(11)
    nonzero = 0
    for j in rng.choice(np.arange(1, nd), size=400, replace=False):
        a, b = op.get_element(0, int(j)), op.get_element(int(j), 0)
        assert a == b
        nonzero += b != 0.0
    assert nonzero >= 40
