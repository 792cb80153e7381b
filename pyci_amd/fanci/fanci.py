"""FanCI base class: the interface of pyci/fanci/fanci.py.

A FanCI problem solves the projected Schrodinger equations

    f_n(p, E) = <n|H|Psi(p)> - E <n|Psi(p)>,   n in the projection ("P") space,

plus optional constraints, for the wave-function parameters p and the energy E.
<n|H|Psi> = sum_m <n|H|m> c_m(p) runs over the auxiliary ("S") space: the P space
and every determinant connected to it by H. The CI operator is the rectangular
sparse_op(ham, wfn, nrow=nproj, ncol=len(wfn), symmetric=False).
"""
from abc import ABCMeta, abstractmethod
from collections import OrderedDict
from typing import Any, Callable, Dict, Sequence, Tuple

import numpy as np
from scipy.optimize import least_squares, root

import pyci_amd as pyci

__all__ = ["FanCI", "fill_wavefunction"]


class FanCI(metaclass=ABCMeta):
    """FanCI problem (pyci/fanci/fanci.py:24-645)."""

    def __init__(self, ham, wfn, nproj: int, nparam: int, norm_param: Sequence[Tuple[int, float]] = None,
                 norm_det: Sequence[Tuple[int, float]] = None,
                 constraints: Dict[str, Tuple[Callable, Callable]] = None, fill: str = "excitation") -> None:
        if not isinstance(ham, pyci.secondquant_op):
            raise TypeError(f"Invalid `ham` type `{type(ham)}`; must be `pyci.hamiltonian`")
        if not isinstance(wfn, pyci.wavefunction):
            raise TypeError(f"Invalid `wfn` type `{type(wfn)}`; must be `pyci.wavefunction`")
        if not isinstance(nproj, int) or isinstance(nproj, bool):
            raise TypeError(f"Invalid `nproj` type `{type(nproj)}`; must be int")
        if not isinstance(nparam, int) or isinstance(nparam, bool):
            raise TypeError(f"Invalid `nparam` type `{type(nparam)}`; must be int")
        if nproj <= 0 or nparam <= 0:
            raise ValueError("`nproj` and `nparam` must be positive")
        if constraints is None:
            constraints = OrderedDict()
        elif isinstance(constraints, dict):
            constraints = OrderedDict(constraints)
        else:
            raise TypeError(f"Invalid `constraints` type `{type(constraints)}`; must be dictionary")
        self._nparam = nparam
        for index, value in norm_param or ():
            constraints[f"p_{{{index}}} - v_{{{index}}}"] = self.make_param_constraint(index, value)
        for index, value in norm_det or ():
            constraints[f"<\\psi_{{{index}}}|\\Psi> - v_{{{index}}}"] = self.make_det_constraint(index, value)
        wfn = fill_wavefunction(wfn, nproj, fill)
        self._nequation = nproj + len(constraints)
        self._nproj = nproj
        self._constraints = constraints
        self._ham = ham
        self._wfn = wfn
        self._ci_op = pyci.sparse_op(ham, wfn, nrow=nproj, ncol=len(wfn), symmetric=False)
        self._sspace = wfn.to_occ_array()
        self._pspace = self._sspace[:nproj]
        self._sspace.setflags(write=False)
        self._pspace.setflags(write=False)

    # ---- properties --------------------------------------------------------
    nequation = property(lambda self: self._nequation, doc="Number of nonlinear equations.")
    nproj = property(lambda self: self._nproj, doc="Number of determinants in the projection space.")
    nparam = property(lambda self: self._nparam, doc="Number of parameters (the energy included).")
    constraints = property(lambda self: tuple(self._constraints.keys()), doc="Constraint names.")
    ham = property(lambda self: self._ham, doc="Hamiltonian.")
    wfn = property(lambda self: self._wfn, doc="Wave function (P space first, then the rest of S).")
    ci_op = property(lambda self: self._ci_op, doc="Rectangular CI operator (nproj x len(wfn)).")
    pspace = property(lambda self: self._pspace, doc="Occupations of the projection space.")
    sspace = property(lambda self: self._sspace, doc="Occupations of the auxiliary space.")
    nbasis = property(lambda self: self._wfn.nbasis)
    nocc_up = property(lambda self: self._wfn.nocc_up)
    nocc_dn = property(lambda self: self._wfn.nocc_dn)
    nvir_up = property(lambda self: self._wfn.nvir_up)
    nvir_dn = property(lambda self: self._wfn.nvir_dn)

    # ---- optimisation ------------------------------------------------------
    def optimize(self, x0: np.ndarray, mode: str = "lstsq", use_jac: bool = False, **kwargs: Any):
        """Solve f(x) = 0 by scipy's least_squares ("lstsq") or root ("root")."""
        x0 = np.asarray(x0, dtype=pyci.c_double)
        if x0.size != self._nparam:
            raise ValueError("length of `x0` does not match `param`")
        opts = dict(kwargs)
        if use_jac:
            opts["jac"] = self.compute_jacobian
        if mode == "lstsq":
            return least_squares(self.compute_objective, x0, **opts)
        if mode == "root":
            return root(self.compute_objective, x0, **opts)
        raise ValueError("invalid mode parameter")

    def add_constraint(self, name: str, f: Callable, dfdx: Callable) -> None:
        self._constraints[name] = f, dfdx
        self._nequation = self._nproj + len(self._constraints)

    def remove_constraint(self, name: str) -> None:
        del self._constraints[name]
        self._nequation = self._nproj + len(self._constraints)

    def make_param_constraint(self, i: int, val: float) -> Tuple[Callable, Callable]:
        """p_i - v and its gradient delta_ki."""

        def f(x):
            return x[i] - val

        def dfdx(x):
            y = np.zeros(self._nparam, dtype=pyci.c_double)
            y[i] = 1.0
            return y

        return f, dfdx

    def make_det_constraint(self, i: int, val: float) -> Tuple[Callable, Callable]:
        """<psi_i|Psi> - v and its gradient d<psi_i|Psi>/dp_k (energy column 0)."""

        def f(x):
            return self.compute_overlap(x[:-1], self._sspace[np.newaxis, i])[0] - val

        def dfdx(x):
            y = np.zeros(self._nparam, dtype=pyci.c_double)
            y[:-1] = self.compute_overlap_deriv(x[:-1], self._sspace[np.newaxis, i])[0]
            return y

        return f, dfdx

    # ---- objective and Jacobian (generic form, one operator product) --------
    def compute_objective(self, x: np.ndarray) -> np.ndarray:
        """f_n = <n|H|Psi> - E <n|Psi> over P, then the constraints."""
        x = np.asarray(x, dtype=pyci.c_double)
        e = x[-1]
        ovlp = self.compute_overlap(x[:-1], "S")
        f = np.empty(self._nequation, dtype=pyci.c_double)
        f[: self._nproj] = self._ci_op(ovlp) - e * ovlp[: self._nproj]
        for i, (fc, _) in enumerate(self._constraints.values()):
            f[self._nproj + i] = fc(x)
        return f

    def compute_jacobian(self, x: np.ndarray) -> np.ndarray:
        """J_nk = <n|H| d Psi/dp_k> - E d<n|Psi>/dp_k, last column -<n|Psi>."""
        x = np.asarray(x, dtype=pyci.c_double)
        e = x[-1]
        jac = np.empty((self._nequation, self._nparam), order="F", dtype=pyci.c_double)
        d_ovlp = self.compute_overlap_deriv(x[:-1], "S")
        jac[: self._nproj, -1] = -self.compute_overlap(x[:-1], "P")
        for k in range(self._nparam - 1):
            jac[: self._nproj, k] = self._ci_op(np.ascontiguousarray(d_ovlp[:, k])) - e * d_ovlp[: self._nproj, k]
        for i, (_, dfc) in enumerate(self._constraints.values()):
            jac[self._nproj + i] = dfc(x)
        return jac

    @abstractmethod
    def compute_overlap(self, x: np.ndarray, occs_array) -> np.ndarray:
        """Overlaps <n|Psi(x)> for the occupations in occs_array (or "P" / "S")."""

    @abstractmethod
    def compute_overlap_deriv(self, x: np.ndarray, occs_array) -> np.ndarray:
        """Derivatives d<n|Psi(x)>/dp_k, shape (ndet, nparam - 1)."""


def fill_wavefunction(wfn, nproj: int, fill: str):
    """A copy of wfn holding the P space (filled by excitation level, or by
    seniority for DOCI, up to nproj determinants) followed by the determinants
    connected to it (pyci/fanci/fanci.py:647-713)."""
    if isinstance(wfn, pyci.doci_wfn):
        e_max = min(wfn.nocc_up, wfn.nvir_up)
        connections = (1,)
    elif isinstance(wfn, (pyci.fullci_wfn, pyci.genci_wfn)):
        e_max = min(wfn.nocc, wfn.nvir)
        connections = (1, 2)
    else:
        raise TypeError(f"invalid `wfn` type `{type(wfn)}`; must be `pyci.wavefunction`")
    wfn = wfn.__class__(wfn)
    if fill == "excitation":
        for nexc in range(e_max + 1):
            if len(wfn) >= nproj:
                break
            pyci.add_excitations(wfn, nexc)
    elif fill == "seniority":
        if not isinstance(wfn, pyci.doci_wfn):
            raise ValueError("seniority filling is available for DOCI wave functions only")
        if len(wfn) < nproj:
            wfn.add_all_dets()
    elif fill is not None:
        raise ValueError(f"Invalid `fill` value: '{fill}'")
    if len(wfn) < nproj:
        raise ValueError(f"unable to generate `nproj={nproj}` determinants")
    if len(wfn) > nproj:
        wfn = wfn.__class__(wfn.nbasis, wfn.nocc_up, wfn.nocc_dn, wfn.to_det_array(nproj))
    for det in wfn.to_det_array(nproj):
        pyci.add_excitations(wfn, *connections, ref=det)
    return wfn
