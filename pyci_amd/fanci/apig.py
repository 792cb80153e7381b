"""APIG FanCI problem (pyci/fanci/apig.py): the antisymmetrized product of
interacting geminals,

    |Psi> = prod_{i <= nocc} (sum_p c_{pi} a+_{p alpha} a+_{p beta}) |0>,

whose overlap with a pair determinant is the permanent of c restricted to its
occupied orbitals. Parameters: c (nbasis x nocc, row-major), then E."""
import pyci_amd as pyci
from pyci_amd._pyci import APIGObjective
from pyci_amd.fanci._geminal import GeminalFanCI

__all__ = ["APIG"]


class APIG(GeminalFanCI):
    _objective_class = APIGObjective

    def __init__(self, ham, nocc: int, nproj: int = None, wfn=None, **kwargs) -> None:
        if not isinstance(ham, pyci.secondquant_op):
            raise TypeError(f"Invalid `ham` type `{type(ham)}`; must be `pyci.hamiltonian`")
        self._setup(ham, nocc, ham.nbasis * nocc + 1, nproj, wfn, kwargs)

    @staticmethod
    def _default_wfn(ham, nocc):
        return pyci.doci_wfn(ham.nbasis, nocc, nocc)
