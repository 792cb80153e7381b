"""GPU heat-bath selection and RDMs versus the oracle and the reference's own
known answers and closed forms (hci.cpp, rdm.cpp, test_routines.py)."""
import numpy as np
import numpy.testing as npt
import pytest
from scipy.special import comb

import pyci_amd as pyci
from pyci_amd.synthetic import seeded_integrals, spin_orbital
from tests.conftest import datafile

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind,nb,nu,nd", [("fullci", 7, 3, 2), ("doci", 9, 3, 3), ("genci", 10, 4, 0)])
def test_add_hci_matches_oracle_order(oracle, kind, nb, nu, nd):
    if kind == "genci":
        _, one, two = seeded_integrals(nb // 2, 6, 60)
        one, two = spin_orbital(one, two)
        ref = np.array([0b0001100011], dtype=np.uint64)
    else:
        _, one, two = seeded_integrals(nb, 6, 61)
        ref = None
    cls = {"doci": pyci.doci_wfn, "fullci": pyci.fullci_wfn, "genci": pyci.genci_wfn}[kind]
    w = cls(nb, nu, nd)
    pyci.add_excitations(w, 0, 1, ref=ref)
    ham = pyci.hamiltonian(0.0, one, two)
    rng = np.random.default_rng(7)
    for eps in (3e-3, 1e-3, 0.0):
        dets = w.to_det_array().reshape(len(w), -1)
        c = rng.normal(size=len(w))
        expect = oracle.add_hci(kind, nb, nu, nd, dets, one, two, c, eps).reshape(-1, dets.shape[1])
        n = pyci.add_hci(ham, w, c, eps=eps)
        assert n == len(expect)
        npt.assert_array_equal(w.to_det_array().reshape(len(w), -1)[len(dets):], expect)
        # the host index sees the appended determinants
        for k in range(0, n, max(1, n // 7)):
            assert w.index_det(w[len(dets) + k]) == len(dets) + k


def test_run_hci_reaches_fci():
    """test_routines.py:442-470 for be_ccpvdz FullCI."""
    ham = pyci.secondquant_op(datafile("be_ccpvdz.fcidump"))
    wfn = pyci.fullci_wfn(ham.nbasis, 2, 2)
    wfn.add_hartreefock_det()
    op = pyci.sparse_op(ham, wfn)
    es, cs = op.solve(n=1, tol=1.0e-6)
    added, niter = 1, 0
    while added:
        added = pyci.add_hci(ham, wfn, cs[0], eps=1.0e-4)
        op.update(ham, wfn)
        es, cs = op.solve(n=1, tol=1.0e-6)
        niter += 1
    assert niter > 1
    assert len(wfn) <= comb(wfn.nbasis, 2, exact=True) ** 2
    added = 1
    while added:
        added = pyci.add_hci(ham, wfn, cs[0], eps=0.0)
        op.update(ham, wfn)
        es, cs = op.solve(n=1, tol=1.0e-6)
    assert len(wfn) == comb(wfn.nbasis, 2, exact=True) ** 2
    npt.assert_allclose(es[0], -14.617409507, rtol=0.0, atol=2.0e-9)


@pytest.mark.parametrize("kind,nb,nu,nd", [("fullci", 6, 3, 2), ("doci", 8, 3, 3), ("genci", 10, 4, 0)])
def test_rdms_match_oracle(oracle, kind, nb, nu, nd):
    cls = {"doci": pyci.doci_wfn, "fullci": pyci.fullci_wfn, "genci": pyci.genci_wfn}[kind]
    w = cls(nb, nu, nd)
    w.add_all_dets()
    c = np.random.default_rng(9).normal(size=len(w))
    c /= np.linalg.norm(c)
    r1, r2 = pyci.compute_rdms(w, c)
    o1, o2 = oracle.rdms(kind, nb, w.to_det_array().reshape(len(w), -1), c)
    npt.assert_allclose(r1, o1, rtol=0, atol=1e-12)
    npt.assert_allclose(r2, o2, rtol=0, atol=1e-12)
    # subset wave function (hash index path)
    s = cls(nb, nu, nd)
    pyci.add_excitations(s, 0, 1, 2)
    cs = np.random.default_rng(10).normal(size=len(s))
    r1, r2 = pyci.compute_rdms(s, cs)
    o1, o2 = oracle.rdms(kind, nb, s.to_det_array().reshape(len(s), -1), cs)
    npt.assert_allclose(r1, o1, rtol=0, atol=1e-12)
    npt.assert_allclose(r2, o2, rtol=0, atol=1e-12)


def test_rdm_closed_forms():
    """test_routines.py:494-532 and 695-800."""
    w = pyci.fullci_wfn(2, 1, 1)
    w.add_all_dets()
    c = np.sqrt(np.arange(1.0, 5.0))
    c /= np.linalg.norm(c)
    _, d = pyci.compute_rdms(w, c)
    npt.assert_allclose([d[2, 0, 0, 0, 0], d[2, 0, 1, 0, 1], d[2, 1, 0, 1, 0], d[2, 1, 1, 1, 1]], c ** 2, atol=1e-12)
    npt.assert_allclose(d[2, 1, 1, 0, 1], c[3] * c[1], atol=1e-12)
    assert np.abs(d[0]).max() < 1e-12 and np.abs(d[1]).max() < 1e-12
    w = pyci.fullci_wfn(3, 2, 1)
    w.add_all_dets()
    c = np.sqrt(np.arange(1.0, 10.0))
    c /= np.linalg.norm(c)
    _, d = pyci.compute_rdms(w, c)
    npt.assert_allclose(d[0, 0, 1, 1, 2], c[0] * c[6] + c[1] * c[7] + c[2] * c[8], atol=1e-12)
    npt.assert_allclose(d[2, 2, 2, 0, 0], -c[0] * c[8], atol=1e-12)
    npt.assert_allclose(d[2, 2, 2, 2, 2], c[8] ** 2 + c[5] ** 2, atol=1e-12)


@pytest.mark.parametrize("fn,wtype,occs", [("be_ccpvdz", pyci.fullci_wfn, (2, 2)), ("be_ccpvdz", pyci.doci_wfn, (2, 2))])
def test_compute_rdms_energy(fn, wtype, occs):
    """test_routines.py:83-133: traces and the energy rebuilt from RDMs."""
    ham = pyci.secondquant_op(datafile(fn + ".fcidump"))
    wfn = wtype(ham.nbasis, *occs)
    wfn.add_all_dets()
    es, cs = pyci.sparse_op(ham, wfn).solve(n=1, ncv=30, tol=1.0e-9)
    d1, d2 = pyci.compute_rdms(wfn, cs[0])
    if wtype is pyci.doci_wfn:
        npt.assert_allclose(np.trace(d1), wfn.nocc_up, atol=1e-9)
        npt.assert_allclose(np.sum(d2), wfn.nocc_up * (wfn.nocc_up - 1), atol=1e-9)
        k0, k2 = pyci.reduce_senzero_integrals(ham.h, ham.v, ham.w, wfn.nocc_up)
        e = ham.ecore + np.einsum("ij,ij", k0, d1) + np.einsum("ij,ij", k2, d2)
        npt.assert_allclose(e, es[0], atol=1e-9)
    rdm1, rdm2 = pyci.spinize_rdms(d1, d2)
    with np.load(datafile(fn + "_spinres.npz")) as f:
        one_mo, two_mo = f["one_mo"], f["two_mo"]
    assert np.all(np.abs(rdm1 - rdm1.T) < 1e-5)
    for i in range(wfn.nbasis * 2):
        assert np.all(np.abs(rdm2[i, i]) < 1e-5) and np.all(np.abs(rdm2[:, :, i, i]) < 1e-5)
    e = ham.ecore + np.einsum("ij,ij", one_mo, rdm1) + 0.25 * np.einsum("ijkl,ijkl", two_mo, rdm2)
    npt.assert_allclose(e, es[0], rtol=0.0, atol=1.0e-9)
