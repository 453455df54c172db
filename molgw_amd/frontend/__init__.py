"""Native integral and RI-HF front end: basis sets in the reference's file
format, Gaussian integrals, and the closed-shell RI-HF that produces the
inputs of the G0W0 path."""
from .basis import BasisSet, read_basis_file
from .scf import (DeviceRIHF, Molecule, RIHFResult, alkane, frozen_core_states, run_rihf, run_rihf_device,
                  water_cluster)

__all__ = ['BasisSet', 'read_basis_file', 'DeviceRIHF', 'Molecule', 'RIHFResult', 'alkane', 'frozen_core_states',
           'run_rihf', 'run_rihf_device', 'water_cluster']
