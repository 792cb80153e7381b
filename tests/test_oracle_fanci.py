"""Pins of the FanCI oracle (oracle/fanci.py) without a GPU: permanent closed
forms, derivatives by finite differences, and the reference's own ground-state
energies (pyci/fanci/test/test_ap1rog.py:47-55, test_apig.py:47-55) reached by
least squares on the oracle's objective and Jacobian."""
import math

import numpy as np
import numpy.testing as npt
import pytest
from scipy.optimize import least_squares

from oracle import fanci as OF
from oracle import fcidump
from tests.conftest import datafile


def test_permanent_closed_forms():
    for n in range(1, 6):
        assert OF.permanent(np.ones((n, n))) == math.factorial(n)
        d = np.arange(1.0, n + 1)
        assert OF.permanent(np.diag(d)) == np.prod(d)
    a = np.array([[1.0, 2.0], [3.0, 4.0]])
    assert OF.permanent(a) == 1 * 4 + 2 * 3
    assert OF.permanent(np.zeros((0, 0))) == 1.0


def _dets(nb, no, order):
    import itertools

    out = [sum(1 << p for p in c) for c in itertools.combinations(range(nb), no)]
    return np.array(out if order is None else [out[i] for i in order], dtype=np.uint64)


@pytest.mark.parametrize("kind", ["ap1rog", "apig"])
def test_overlap_derivatives_finite_difference(kind):
    nb, no = 6, 2
    dets = _dets(nb, no, None)
    npar = no * (nb - no) if kind == "ap1rog" else nb * no
    x = np.random.default_rng(3).normal(size=npar)
    D = OF.d_overlaps(kind, nb, no, dets, x)
    h = 1e-6
    for k in range(npar):
        e = np.zeros(npar)
        e[k] = h
        fd = (OF.overlaps(kind, nb, no, dets, x + e) - OF.overlaps(kind, nb, no, dets, x - e)) / (2 * h)
        npt.assert_allclose(D[:, k], fd, rtol=0, atol=1e-8)


SYSTEMS = [(1, "h2_hf_sto6g", 0.71317683129, -1.83843, 2, -1.1459130128588935),
           (2, "lih_hf_sto6g", 0.995317634356, -8.94728, 15, -7.968213475280394)]


@pytest.mark.parametrize("nocc,system,nucnuc,e_hf,nproj,expected", SYSTEMS)
def test_ap1rog_oracle_reaches_reference_energy(nocc, system, nucnuc, e_hf, nproj, expected):
    """The P space of the reference's default AP1roG problem: pair singles of HF,
    then HF (fanci.py:681-688), S space = P plus its pair singles."""
    from oracle import oracle as O

    ecore, one, two = fcidump.read(datafile(system + ".fcidump"))
    nb = one.shape[0]
    hf = np.array([(1 << nocc) - 1], dtype=np.uint64)
    p = O.add_excited_dets("doci", nb, nocc, nocc, hf[:0], hf, 1)
    p = O.add_excited_dets("doci", nb, nocc, nocc, p, hf, 0)[:nproj]
    s = p
    for d in p:
        s = O.add_excited_dets("doci", nb, nocc, nocc, s, np.array([d], dtype=np.uint64), 1)
    obj = OF.Objective("ap1rog", nb, nocc, s, one, two, nproj)
    x0 = np.zeros(obj.nparam + 1)
    x0[-1] = e_hf
    res = least_squares(obj.objective, x0, jac=obj.jacobian)
    npt.assert_allclose(res.x[-1] + nucnuc, expected, rtol=1e-5, atol=1e-8)
