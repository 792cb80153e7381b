"""AP1roG FanCI problem (pyci/fanci/ap1rog.py): the antisymmetrized product of
1-reference-orbital geminals,

    |Psi> = prod_{i <= nocc} (a+_{i alpha} a+_{i beta} + sum_{a > nocc} c_{ia} a+_{a alpha} a+_{a beta}) |0>,

whose overlap with a pair determinant is the permanent of c restricted to its
excited holes and particles. Parameters: c (nocc x nvir, row-major), then E."""
import numpy as np

import pyci_amd as pyci
from pyci_amd._pyci import AP1roGObjective
from pyci_amd.fanci._geminal import GeminalFanCI

__all__ = ["AP1roG"]


class AP1roG(GeminalFanCI):
    _objective_class = AP1roGObjective

    def __init__(self, ham, nocc: int, nproj: int = None, wfn=None, **kwargs) -> None:
        if not isinstance(ham, pyci.secondquant_op):
            raise TypeError(f"Invalid `ham` type `{type(ham)}`; must be `pyci.hamiltonian`")
        self._setup(ham, nocc, nocc * (ham.nbasis - nocc) + 1, nproj, wfn, kwargs)

    @staticmethod
    def _default_wfn(ham, nocc):
        # P space: the pair single excitations of the reference (then the
        # reference determinant itself is appended by the excitation fill)
        wfn = pyci.doci_wfn(ham.nbasis, nocc, nocc)
        wfn.add_excited_dets(1)
        return wfn
