"""CC4S file interchange: Coulomb vertex and eigenenergies.

Reads the two files MOLGW accepts with ``init_hamiltonian='CC4S_FILES'``:

* ``<root>_CoulombVertex.yaml`` / ``.elements`` -- complex128 vertex
  Gamma(G, i, j) written with ``halfGrid: 1``; MOLGW splits it into
  2*naux real auxiliary functions (real parts first, then imaginary parts)
  and reads the binary records with the state index ``istate`` outer and
  ``jstate`` inner (src/m_eri_ao_mo.f90:1192-1331, non-MPI branch at
  :1250-1259).
* ``<root>_EigenEnergies.yaml`` / ``.elements`` -- text energies; the
  occupation is 2 below ``fermiEnergy + 1e-8`` and 0 above, C = identity
  and the Fock matrix is diag(energy) (src/m_io.f90:4048-4120).
"""
from __future__ import annotations

import os
import re

import numpy as np


def _yaml_value(path: str, key: str) -> list[str]:
    """Values that follow ``key:`` in a CC4S yaml (the reference's
    yaml_search_keyword reads the first occurrence(s) of the key)."""
    vals = []
    with open(path, 'r') as fh:
        for line in fh:
            m = re.match(r'^\s*-?\s*' + re.escape(key) + r'\s*:\s*(.*)$', line)
            if m:
                vals.append(m.group(1).strip())
    if not vals:
        raise ValueError(f'{path}: key {key!r} not found')
    return vals


def read_eigenenergies(root: str):
    """Return (energy[nstate], occupation[nstate], efermi)."""
    ypath = root + '_EigenEnergies.yaml'
    efermi = float(_yaml_value(ypath, 'fermiEnergy')[0])
    nstate = int(_yaml_value(ypath, 'length')[0])
    energy = np.loadtxt(root + '_EigenEnergies.elements', dtype=np.float64).reshape(-1)[:nstate]
    occupation = np.where(energy < efermi + 1.0e-8, 2.0, 0.0)
    if occupation.sum() < 0.001:
        raise ValueError('read_eigenenergies: Fermi energy is such that there is no occupied state')
    return energy, occupation, efermi


def read_coulomb_vertex(root: str) -> np.ndarray:
    """Return eri_mo[2*naux, nstate, nstate] (real parts, then imaginary parts)."""
    ypath = root + '_CoulombVertex.yaml'
    if int(_yaml_value(ypath, 'halfGrid')[0]) != 1:
        raise ValueError('read_coulomb_vertex: only works for half grid (real wavefunctions)')
    lengths = [int(v) for v in _yaml_value(ypath, 'length')]
    naux, nstate = lengths[0], lengths[1]
    path = root + '_CoulombVertex.elements'
    expected = naux * nstate * nstate * 16
    if os.path.getsize(path) != expected:
        raise ValueError(f'{path}: size {os.path.getsize(path)} != {expected}')
    raw = np.fromfile(path, dtype=np.complex128).reshape(nstate, nstate, naux)
    # record (istate, jstate) -> eri_3center_mo(:, istate, jstate)
    out = np.empty((2 * naux, nstate, nstate), dtype=np.float64)
    out[:naux] = np.transpose(raw.real, (2, 0, 1))
    out[naux:] = np.transpose(raw.imag, (2, 0, 1))
    return out
