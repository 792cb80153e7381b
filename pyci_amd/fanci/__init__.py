"""FanCI: projected-Schrodinger-equation wave functions over a CI determinant space
(the reference's pyci.fanci package). The geminal wave functions AP1roG and APIG
evaluate their overlaps, objective and Jacobian on the GPU (pyci_amd._pyci
AP1roGObjective / APIGObjective)."""
from pyci_amd.fanci.fanci import FanCI, fill_wavefunction
from pyci_amd.fanci.ap1rog import AP1roG
from pyci_amd.fanci.apig import APIG

__all__ = ["FanCI", "AP1roG", "APIG", "fill_wavefunction"]
