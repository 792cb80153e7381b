"""pyci_amd: an MI355X-native configuration-interaction engine with PyCI's API.

`import pyci_amd as pyci` gives the reference package's names for the hot path
(pyci/__init__.py:18-100): integrals (`secondquant_op`, alias `hamiltonian`),
wave functions (`doci_wfn`, `fullci_wfn`, `genci_wfn`), the sparse CI operator
(`sparse_op` with `.solve`), `add_excitations`, `add_hci` and `compute_rdms`.
The native pieces are `_pyci` (pybind11 host module) over `libpyci_engine.so`
(HIP kernels for gfx950 behind a C ABI, include/pyci_engine.h). Operations that
run on the GPU raise RuntimeError when no GPU is present; there is no CPU path.
"""
from pyci_amd._pyci import __version__, c_long, c_ulong, c_double
from pyci_amd._pyci import secondquant_op, wavefunction, one_spin_wfn, two_spin_wfn
from pyci_amd._pyci import doci_wfn, fullci_wfn, genci_wfn, sparse_op
from pyci_amd._pyci import get_num_threads, set_num_threads, popcnt, ctz
from pyci_amd._pyci import compute_rdms, compute_transition_rdms, compute_enpt2, add_hci, device_count, mlp_selector, MLP_NPARAM, batch_solve

from pyci_amd.utility import make_senzero_integrals, reduce_senzero_integrals, spinize_rdms
from pyci_amd.excitation_ci import add_excitations

hamiltonian = secondquant_op

__all__ = [
    "__version__",
    "c_long",
    "c_ulong",
    "c_double",
    "secondquant_op",
    "hamiltonian",
    "wavefunction",
    "one_spin_wfn",
    "two_spin_wfn",
    "doci_wfn",
    "fullci_wfn",
    "genci_wfn",
    "sparse_op",
    "get_num_threads",
    "set_num_threads",
    "popcnt",
    "ctz",
    "add_hci",
    "compute_rdms",
    "compute_enpt2",
    "compute_transition_rdms",
    "make_senzero_integrals",
    "reduce_senzero_integrals",
    "spinize_rdms",
    "add_excitations",
    "device_count",
    "mlp_selector",
    "MLP_NPARAM",
    "batch_solve",
]
