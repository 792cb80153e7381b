"""Host-side logic of pyci_amd (no GPU): containers, FCIDUMP I/O, rank_det,
add_excited_dets order, RDM/integral utilities. Expectations follow the
reference's test_wavefunction.py / test_hamiltonian.py and the oracle."""
import ctypes
import os
from filecmp import cmp as compare
from tempfile import NamedTemporaryFile

import numpy as np
import numpy.testing as npt
import pytest
from scipy.special import comb

import pyci_amd as pyci
from tests.conftest import ROOT, datafile


def test_raises():
    npt.assert_raises(ValueError, pyci.doci_wfn, 10, 11, 11)
    npt.assert_raises(ValueError, pyci.fullci_wfn, 10, 11, 1)
    npt.assert_raises(ValueError, pyci.fullci_wfn, 10, 0, 1)
    npt.assert_raises(ValueError, pyci.fullci_wfn, 10, 2, 3)
    npt.assert_raises(ValueError, pyci.doci_wfn, 10, 3, 2)
    npt.assert_raises(ValueError, pyci.genci_wfn, 10, 3, 2)
    w = pyci.fullci_wfn(5, 3, 3)
    w.add_hartreefock_det()
    assert w[0].flags["OWNDATA"]
    assert w[0].tolist() == [[7], [7]]


@pytest.mark.parametrize("nbasis, nocc", [(16, 8), (64, 1), (64, 4), (65, 1), (65, 4), (129, 3)])
def test_doci_add_excited_dets_counts(nbasis, nocc):
    wfn = pyci.doci_wfn(nbasis, nocc, nocc)
    if comb(nbasis, nocc, exact=True) > 20000:
        pytest.skip("large")
    length = 0
    for i in range(nocc + 1):
        length += comb(nocc, i, exact=True) * comb(wfn.nvir_up, i, exact=True)
        wfn.add_excited_dets(i)
        assert len(wfn) == length
    assert len(wfn) == comb(nbasis, nocc, exact=True)
    d = wfn.to_det_array()
    assert all(pyci.popcnt(x) == nocc for x in d[:50])
    occs = wfn.to_occ_array()
    w2 = pyci.doci_wfn(nbasis, nocc, nocc, occs)
    npt.assert_array_equal(w2.to_det_array(), d)
    w3 = pyci.doci_wfn(nbasis, nocc, nocc, d)
    npt.assert_array_equal(w3.to_occ_array(), occs)


@pytest.mark.parametrize("nbasis, nu, nd", [(8, 3, 3), (64, 1, 1), (64, 2, 1), (65, 2, 1), (129, 2, 1)])
def test_fullci_excited_and_roundtrips(nbasis, nu, nd):
    ndet = comb(nbasis, nu, exact=True) * comb(nbasis, nd, exact=True)
    wfn = pyci.fullci_wfn(nbasis, nu, nd)
    wfn.reserve(ndet)
    for i in range(nu + nd + 1):
        wfn.add_excited_dets(i)
    assert len(wfn) == ndet
    d1 = wfn.to_det_array()
    assert d1.shape == (ndet, 2, (nbasis + 63) // 64)
    w2 = pyci.fullci_wfn(nbasis, nu, nd, d1)
    npt.assert_array_equal(w2.to_det_array(), d1)
    o1 = wfn.to_occ_array()
    assert o1.shape[2] == nu
    w3 = pyci.fullci_wfn(nbasis, nu, nd, o1)
    npt.assert_array_equal(w3.to_occ_array()[:, 0, :nu], o1[:, 0, :nu])
    npt.assert_array_equal(w3.to_occ_array()[:, 1, :nd], o1[:, 1, :nd])
    f1, f2 = NamedTemporaryFile(), NamedTemporaryFile()
    wfn.to_file(f1.name)
    w4 = pyci.fullci_wfn(f1.name)
    w4.to_file(f2.name)
    assert compare(f1.name, f2.name, shallow=False)
    w5 = pyci.fullci_wfn(wfn)
    npt.assert_array_equal(w5.to_det_array(), d1)
    for k in (0, ndet // 2, ndet - 1):
        assert wfn.index_det(d1[k]) == k


def test_add_excited_dets_order_matches_oracle(oracle):
    hf = np.array([0b000111, 0b000011], dtype=np.uint64)
    w = pyci.fullci_wfn(6, 3, 2)
    dets = np.zeros((0, 2), dtype=np.uint64)
    for e in (0, 1, 2, 3):
        w.add_excited_dets(e, ref=hf)
        dets = oracle.add_excited_dets("fullci", 6, 3, 2, dets, hf, e)
        npt.assert_array_equal(w.to_det_array().reshape(-1, 2), dets)
    g = pyci.genci_wfn(10, 4, 0)
    ref = np.array([0b0000110011], dtype=np.uint64)
    gd = np.zeros(0, dtype=np.uint64)
    for e in (2, 0, 1):
        g.add_excited_dets(e, ref=ref)
        gd = oracle.add_excited_dets("genci", 10, 4, 0, gd, ref, e)
    npt.assert_array_equal(g.to_det_array().reshape(-1), gd)


def test_fullci_doci_genci_conversions():
    d = pyci.doci_wfn(4, 2, 2)
    d.add_excited_dets(0)
    d.add_excited_dets(1)
    f = pyci.fullci_wfn(d)
    assert len(f) == len(d)
    fa = f.to_det_array()
    npt.assert_array_equal(fa[:, 0, 0], fa[:, 1, 0])
    g = pyci.genci_wfn(f)
    assert g.nbasis == 8 and g.nocc == 4 and g.nocc_dn == 0
    ga = g.to_det_array()[:, 0]
    npt.assert_array_equal(ga, fa[:, 0, 0] | (fa[:, 1, 0] << np.uint64(4)))


def test_add_det_dedupe_and_dets_from_wfn():
    w = pyci.genci_wfn(6, 2, 0)
    assert w.add_det(np.array([3], dtype=np.uint64)) == 0
    assert w.add_det(np.array([3], dtype=np.uint64)) == -1
    assert w.add_occs(np.array([0, 5])) == 1
    assert w.index_det(np.array([33], dtype=np.uint64)) == 1
    assert w.index_det(np.array([5], dtype=np.uint64)) == -1
    v = pyci.genci_wfn(6, 2, 0)
    v.add_det(np.array([33], dtype=np.uint64))
    v.add_det(np.array([6], dtype=np.uint64))
    w.add_dets_from_wfn(v)
    assert w.to_det_array()[:, 0].tolist() == [3, 33, 6]


def test_rank_det_equals_reference_spookyhash():
    ref = os.path.join(ROOT, "oracle", "_ref", "libspooky_ref.so")
    if not os.path.exists(ref):
        pytest.skip("reference SpookyV2 build (oracle/_ref) not present")
    L = ctypes.CDLL(ref)
    L.spooky_ref_hash128.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_uint64)]
    rng = np.random.default_rng(0)
    for nw in (1, 2, 3, 4, 12, 23, 24, 25, 36, 60):
        g = pyci.genci_wfn(64 * nw, 0, 0)
        for _ in range(10):
            d = rng.integers(0, 2 ** 63, size=nw, dtype=np.uint64)
            h1, h2 = ctypes.c_uint64(0x23a23cf5033c3c81), ctypes.c_uint64(0xb3816f6a2c68e530)
            L.spooky_ref_hash128(d.ctypes.data, d.nbytes, ctypes.byref(h1), ctypes.byref(h2))
            assert tuple(g.rank_det(d)) == (h1.value, h2.value)
    w = pyci.fullci_wfn(5, 2, 2)
    w.add_hartreefock_det()
    w.add_excited_dets(1)
    r = w.rank_det(w[3])
    assert w.index_det_from_rank(r) == 3


@pytest.mark.parametrize("fn", ["be_ccpvdz", "h4_sto3g", "lih_sto6g"])
def test_fcidump_roundtrip(fn):
    f1 = NamedTemporaryFile()
    h1 = pyci.secondquant_op(datafile(fn + ".fcidump"))
    h1.to_file(f1.name)
    h2 = pyci.secondquant_op(f1.name)
    for a in ("ecore", "one_mo", "two_mo", "h", "v", "w"):
        npt.assert_allclose(getattr(h2, a), getattr(h1, a), rtol=0, atol=1e-12)
    from oracle import fcidump

    ec, one, two = fcidump.read(datafile(fn + ".fcidump"))
    npt.assert_array_equal(h1.one_mo, one)
    npt.assert_array_equal(h1.two_mo, two)
    assert h1.ecore == ec
    hh, vv, ww = pyci.make_senzero_integrals(one, two)
    npt.assert_array_equal(h1.h, hh)
    npt.assert_array_equal(h1.v, vv)
    npt.assert_allclose(h1.w, ww, rtol=0, atol=1e-15)


def test_secondquant_from_arrays():
    from pyci_amd.synthetic import seeded_integrals

    e, one, two = seeded_integrals(5, 4, 3)
    h = pyci.hamiltonian(e, one, two)
    assert h.nbasis == 5
    npt.assert_array_equal(h.two_mo, two)
    with pytest.raises(ValueError):
        pyci.hamiltonian(0.0, one, two[:2])


def test_spinize_and_reduce(oracle):
    """spinize_rdms of FullCI blocks equals the GenCI (spin-orbital) RDMs of the
    same state; reduce_senzero_integrals keeps the DOCI energy expression."""
    dets = oracle.all_dets("fullci", 3, 2, 1)
    c = np.random.default_rng(5).normal(size=len(dets))
    c /= np.linalg.norm(c)
    r1, r2 = oracle.rdms("fullci", 3, dets, c)
    R1, R2 = pyci.spinize_rdms(r1, r2)
    g = (dets[:, 0] | (dets[:, 1] << np.uint64(3))).astype(np.uint64)
    G1, G2 = oracle.rdms("genci", 6, g, c)
    npt.assert_allclose(R1, G1, atol=1e-14)
    npt.assert_allclose(R2, G2, atol=1e-14)
    h = np.array([1.0, 2.0, 3.0])
    rv, rw = pyci.reduce_senzero_integrals(h, np.eye(3), np.ones((3, 3)), 2)
    npt.assert_allclose(rv, np.eye(3) + 2.0 / 3.0 * np.diag(h))
    npt.assert_allclose(rw, np.ones((3, 3)) + 2.0 / 3.0 * (h[:, None] + h[None, :]))
