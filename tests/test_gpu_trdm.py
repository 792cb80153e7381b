"""GPU compute_transition_rdms (rdm.cpp:634-899) against the CPU oracle (its
restatement, pinned element by element by Jordan-Wigner operator matrices), and
<Psi|...|Psi> against compute_rdms."""
import numpy as np
import numpy.testing as npt
import pytest

import pyci_amd as pyci
from pyci_amd.synthetic import seeded_integrals

pytestmark = pytest.mark.gpu

CLS = {"doci": pyci.doci_wfn, "fullci": pyci.fullci_wfn, "genci": pyci.genci_wfn}


def _sets(kind, nb, nu, nd, seed):
    """Two different sets: a CISD-like set and a random half of the full space."""
    ref = np.array([0b000011000011], dtype=np.uint64) if kind == "genci" else None
    w1 = CLS[kind](nb, nu, nd)
    pyci.add_excitations(w1, 0, 1, 2, ref=ref)
    full = CLS[kind](nb, nu, nd)
    full.add_all_dets()
    rng = np.random.default_rng(seed)
    dets = full.to_det_array().reshape(len(full), -1)
    pick = rng.permutation(len(full))[: len(full) // 2]
    w2 = CLS[kind](nb, nu, nd)
    for d in dets[pick]:
        w2.add_det(d if kind == "fullci" else d[0:1])
    return w1, w2, rng


@pytest.mark.parametrize("kind,nb,nu,nd", [("fullci", 7, 3, 2), ("doci", 9, 4, 4), ("genci", 12, 4, 0)])
def test_transition_rdms_match_oracle(oracle, kind, nb, nu, nd):
    w1, w2, rng = _sets(kind, nb, nu, nd, 31)
    c1, c2 = rng.normal(size=len(w1)), rng.normal(size=len(w2))
    a1, a2 = pyci.compute_transition_rdms(w1, w2, c1, c2)
    d1 = w1.to_det_array().reshape(len(w1), -1)
    d2 = w2.to_det_array().reshape(len(w2), -1)
    o1, o2 = oracle.transition_rdms(kind, nb, d1, d2, c1, c2)
    assert a1.shape == o1.shape and a2.shape == o2.shape
    scale = max(np.abs(o1).max(), np.abs(o2).max())
    npt.assert_allclose(a1, o1, rtol=0, atol=1e-12 * scale)
    npt.assert_allclose(a2, o2, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("kind,nb,nu,nd", [("fullci", 6, 3, 3), ("doci", 8, 3, 3), ("genci", 10, 4, 0)])
def test_transition_rdms_of_one_state_are_its_rdms(kind, nb, nu, nd):
    w = CLS[kind](nb, nu, nd)
    w.add_all_dets()
    c = np.random.default_rng(3).normal(size=len(w))
    c /= np.linalg.norm(c)
    a1, a2 = pyci.compute_transition_rdms(w, w, c, c)
    r1, r2 = pyci.compute_rdms(w, c)
    npt.assert_allclose(a1, r1, rtol=0, atol=1e-12)
    npt.assert_allclose(a2, r2, rtol=0, atol=1e-12)


def test_transition_rdms_between_eigenstates():
    """Two roots of one Hamiltonian: the 1-TRDM is traceless (orthogonal states)
    and the Hamiltonian's transition element sum h.g1 + 1/2 <pq|rs> G2 vanishes."""
    _, one, two = seeded_integrals(7, 8, 33)
    ham = pyci.hamiltonian(0.0, one, two)
    w = pyci.fullci_wfn(7, 3, 3)
    w.add_all_dets()
    es, cs = pyci.sparse_op(ham, w, storage="csr").solve(n=2, tol=1e-12)
    g1, g2 = pyci.compute_transition_rdms(w, w, cs[0], cs[1])
    assert abs(np.trace(g1[0]) + np.trace(g1[1])) < 1e-10
    e = np.einsum("pq,pq", one, g1[0] + g1[1]) + 0.5 * (
        np.einsum("pqrs,pqrs", two, g2[0]) + np.einsum("pqrs,pqrs", two, g2[1])) + np.einsum("pqrs,pqrs", two, g2[2])
    assert abs(e) < 1e-9


def test_transition_rdms_shape_errors():
    a = pyci.fullci_wfn(5, 2, 2)
    a.add_all_dets()
    b = pyci.fullci_wfn(5, 3, 2)
    b.add_all_dets()
    with pytest.raises(ValueError):
        pyci.compute_transition_rdms(a, b, np.ones(len(a)), np.ones(len(b)))
    with pytest.raises(ValueError):
        pyci.compute_transition_rdms(a, a, np.ones(len(a)), np.ones(3))
