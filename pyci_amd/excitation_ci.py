"""Excitation CI: add all determinants of given excitation levels.

Mirrors the reference's pyci.excitation_ci.add_excitations (excitation_ci.py:23-39).
"""

__all__ = ["add_excitations"]


def add_excitations(wfn, *excitations, ref=None):
    """Add every determinant at each excitation level in `excitations` relative to
    `ref` (default: the Hartree-Fock determinant) to `wfn`, in the given order."""
    for e in excitations:
        wfn.add_excited_dets(e, ref=ref)
