"""GPU parity: CSR build, export, get_element, matvec and diagonal versus the oracle.

Column indices and row pointers must match exactly; matrix elements use the
reference's formulas and summation order, so they are compared bit for bit
(rtol 1e-12 is the stated bar, exact equality is what is asserted)."""
import numpy as np
import numpy.testing as npt
import pytest

import pyci_amd as pyci
from oracle import fcidump
from pyci_amd.synthetic import seeded_integrals, config_integrals, spin_orbital
from tests.conftest import datafile

pytestmark = pytest.mark.gpu


def make_wfn(kind, nb, nu, nd, dets=None):
    cls = {"doci": pyci.doci_wfn, "fullci": pyci.fullci_wfn, "genci": pyci.genci_wfn}[kind]
    w = cls(nb, nu, nd)
    if dets is None:
        w.add_all_dets()
    else:
        for d in np.asarray(dets, dtype=np.uint64).reshape(len(dets), -1):
            w.add_det(d)
    return w


def check_csr(op, ref):
    assert op.shape == ref.shape
    assert op.size == ref.data.size
    npt.assert_array_equal(op.indptr(), ref.indptr)
    npt.assert_array_equal(op.indices(), ref.indices)
    npt.assert_array_equal(op.data(), ref.data)


CASES = [
    ("fullci", 6, 3, 2, 21),
    ("fullci", 6, 2, 2, 22),
    ("doci", 8, 3, 3, 23),
    ("genci", 8, 3, 0, 24),
]


@pytest.mark.parametrize("kind,nb,nu,nd,seed", CASES)
@pytest.mark.parametrize("symmetric", [True, False])
def test_full_space_csr(oracle, kind, nb, nu, nd, seed, symmetric):
    if kind == "genci":
        _, one, two = seeded_integrals(nb // 2, 6, seed)
        one, two = spin_orbital(one, two)
    else:
        _, one, two = seeded_integrals(nb, 6, seed)
    w = make_wfn(kind, nb, nu, nd)
    dets = oracle.all_dets(kind, nb, nu, nd)
    npt.assert_array_equal(w.to_det_array().reshape(dets.shape), dets)
    ham = pyci.hamiltonian(0.0, one, two)
    op = pyci.sparse_op(ham, w, symmetric=symmetric)
    check_csr(op, oracle.csr_build(kind, nb, nu, nd, dets, one, two, symmetric=symmetric))
    npt.assert_array_equal(op.diagonal(), oracle.diagonal(kind, nb, nu, nd, dets, one, two))


@pytest.mark.parametrize("symmetric", [True, False])
def test_subset_wfn_hash_index(oracle, symmetric):
    """A selected (non-colex) determinant set goes through the GPU hash index."""
    _, one, two = seeded_integrals(7, 6, 31)
    w = pyci.fullci_wfn(7, 3, 2)
    pyci.add_excitations(w, 0, 1, 2)
    dets = w.to_det_array().reshape(-1, 2)
    nrow = len(w) - 10
    ham = pyci.hamiltonian(0.0, one, two)
    op = pyci.sparse_op(ham, w, nrow, symmetric=symmetric)
    check_csr(op, oracle.csr_build("fullci", 7, 3, 2, dets, one, two, nrow=nrow, symmetric=symmetric))


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("symmetric", [True, False])
def test_genci_fixed_lower_half_count(oracle, mixed, symmetric):
    """GenCI sets whose determinants all hold the same number of occupied spin orbitals
    in the lower half (one Sz in the alpha-then-beta ordering) do not enumerate the
    excitations that change that number, which cannot find a column; one determinant of
    another count (mixed) turns that off. General spin-orbital integrals (no spin blocks),
    so the skipped excitations would have nonzero elements: both CSRs against the oracle,
    bit for bit, with row ranges (nrow) on the symmetric one."""
    nb, ne = 12, 4
    _, one, two = seeded_integrals(nb, 6, 33)
    alld = oracle.all_dets("genci", nb, ne, 0).reshape(-1)
    low = np.array([bin(int(d) & 63).count("1") for d in alld])
    sel = alld[low == 2]
    if mixed:
        sel = np.concatenate([sel[:100], alld[low == 3][:1], sel[100:]])
    dets = sel.reshape(-1, 1)
    w = make_wfn("genci", nb, ne, 0, dets)
    ham = pyci.hamiltonian(0.0, one, two)
    if symmetric:
        nrow = len(sel) - 7
        op = pyci.sparse_op(ham, w, nrow, symmetric=True)
        check_csr(op, oracle.csr_build("genci", nb, ne, 0, dets, one, two, nrow=nrow, symmetric=True))
    else:
        op = pyci.sparse_op(ham, w, symmetric=False)
        check_csr(op, oracle.csr_build("genci", nb, ne, 0, dets, one, two, symmetric=False))


def test_rectangular_matvec_equals_get_element_sums(oracle):
    """test_routines.py:57-80 on be_ccpvdz FullCI HF+singles+doubles."""
    ham = pyci.secondquant_op(datafile("be_ccpvdz.fcidump"))
    w = pyci.fullci_wfn(ham.nbasis, 2, 2)
    pyci.add_excitations(w, 0, 1, 2)
    nrow = len(w) - 10
    op = pyci.sparse_op(ham, w, nrow, symmetric=False)
    assert op.shape == (nrow, len(w))
    y = op(np.ones(op.shape[1]))
    assert y.ndim == 1 and y.shape[0] == nrow
    rows = np.arange(0, nrow, 37)
    z = np.array([sum(op.get_element(int(i), j) for j in range(op.shape[1])) for i in rows[:6]])
    npt.assert_allclose(y[rows[:6]], z, rtol=1e-12, atol=1e-12)
    ref = oracle.csr_build("fullci", ham.nbasis, 2, 2, w.to_det_array().reshape(-1, 2), ham.one_mo, ham.two_mo,
                           nrow=nrow, symmetric=False)
    npt.assert_allclose(y, ref.scipy() @ np.ones(op.shape[1]), rtol=1e-12, atol=1e-12)


def test_workload_fci8_bit_exact(oracle):
    """Seeded nbasis=8, 4+4 FullCI (4,900 dets, 886,900 stored lower nonzeros)."""
    _, one, two = config_integrals("fci8")
    w = make_wfn("fullci", 8, 4, 4)
    ham = pyci.hamiltonian(0.0, one, two)
    op = pyci.sparse_op(ham, w)
    assert op.size == 886900
    ref = oracle.csr_build("fullci", 8, 4, 4, oracle.all_dets("fullci", 8, 4, 4), one, two)
    check_csr(op, ref)
    x = np.random.default_rng(0).normal(size=op.shape[1])
    npt.assert_allclose(op(x), ref.scipy() @ x, rtol=1e-11, atol=1e-11)
    out = np.zeros(op.shape[0])
    op.matvec(x, out)
    npt.assert_allclose(out, ref.scipy() @ x, rtol=1e-11, atol=1e-11)


@pytest.mark.parametrize("fn,kind,occ", [("be_ccpvdz", "fullci", (2, 2)), ("li2_ccpvdz", "doci", (3, 3))])
def test_molecule_csr_bit_exact(oracle, fn, kind, occ):
    ec, one, two = fcidump.read(datafile(fn + ".fcidump"))
    ham = pyci.secondquant_op(datafile(fn + ".fcidump"))
    w = make_wfn(kind, ham.nbasis, *occ)
    op = pyci.sparse_op(ham, w)
    assert op.ecore == ec
    check_csr(op, oracle.csr_build(kind, ham.nbasis, *occ, oracle.all_dets(kind, ham.nbasis, *occ), one, two))


def test_get_element_and_symmetry(oracle):
    _, one, two = seeded_integrals(6, 6, 40)
    w = make_wfn("fullci", 6, 2, 2)
    ham = pyci.hamiltonian(0.0, one, two)
    op = pyci.sparse_op(ham, w)
    dense = oracle.csr_build("fullci", 6, 2, 2, oracle.all_dets("fullci", 6, 2, 2), one, two).dense()
    for i, j in [(0, 0), (17, 3), (100, 99), (224, 0), (5, 200), (224, 224)]:
        expect = dense[i, j] if j <= i else 0.0  # symmetric storage: upper reads as 0
        assert op.get_element(i, j) == expect
    full = pyci.sparse_op(ham, w, symmetric=False)
    for i, j in [(5, 200), (3, 17)]:
        assert full.get_element(i, j) == dense[i, j]


def test_update_appends_rows(oracle):
    """sparse_op.update after new determinants (sparseop.cpp:175-201), both modes."""
    _, one, two = seeded_integrals(6, 6, 41)
    ham = pyci.hamiltonian(0.0, one, two)
    for symmetric in (True, False):
        w = pyci.fullci_wfn(6, 2, 2)
        pyci.add_excitations(w, 0, 1)
        op = pyci.sparse_op(ham, w, symmetric=symmetric)
        n0 = len(w)
        pyci.add_excitations(w, 2)
        op.update(ham, w)
        dets = w.to_det_array().reshape(-1, 2)
        ref = oracle.csr_build("fullci", 6, 2, 2, dets, one, two, symmetric=symmetric)
        if symmetric:
            check_csr(op, ref)
        else:
            # old rows keep their old column bound, new rows see every column
            old = oracle.csr_build("fullci", 6, 2, 2, dets, one, two, nrow=n0, ncol=n0, symmetric=False)
            new = oracle.csr_build("fullci", 6, 2, 2, dets, one, two, row0=n0, symmetric=False)
            npt.assert_array_equal(op.indices(), np.concatenate([old.indices, new.indices]))
            npt.assert_array_equal(op.data(), np.concatenate([old.data, new.data]))
            npt.assert_array_equal(op.indptr(), np.concatenate([old.indptr, new.indptr[1:] + old.indptr[-1]]))
