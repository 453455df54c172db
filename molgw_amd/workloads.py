"""Synthetic RI G0W0 inputs.

The recipe is the synthetic imaginary-frequency kernel of the benchmark
definition, usable at any size:

* C: random orthonormal FP64 matrix, QR of an N(0,1) matrix from the seed;
* energies: occupied uniform in [-1.2, -0.3] Ha, virtual uniform in
  [0.05, 3.0] Ha, each sorted;
* B(P, mu nu) = sum_k L(P, k) G(k, mu nu) with G ~ N(0,1) symmetric in
  (mu, nu) and L = U diag(1/(1+k)^1.5) V^T (seeded orthogonal U, V);
* B is handed over in MOLGW's packed layout eri_3center(npair, nauxil) with
  diagonal pairs stored x1/2 (src/m_eri_calculate.f90:1840-1841) and the
  upper-triangle pair list (k <= l, column by column).

Molecular shapes (nbf, nauxil, nocc) of the cc-pVDZ / cc-pVTZ systems used
by the benchmark are given by ``alkane_shape`` and ``water_cluster_shape``
(basis-function counts of basis/{H,C,O}_cc-pV{D,T}Z{,-RI} of the reference).
"""
from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np


@dataclass
class System:
    name: str
    nbf: int
    naux: int
    nocc: int
    eri_3center: np.ndarray     # (npair, naux), Fortran order
    index_basis: np.ndarray     # (2, npair), 1-based
    c_matrix: np.ndarray        # (nbf, nstate)
    energy: np.ndarray          # (nstate,)
    occupation: np.ndarray      # (nstate,)

    @property
    def nstate(self) -> int:
        return self.c_matrix.shape[1]


# basis-function counts per element (pure functions): (orbital, RI) for cc-pVDZ, cc-pVTZ
_BASIS = {
    'cc-pVDZ': {'H': (5, 14), 'C': (14, 56), 'O': (14, 56)},
    'cc-pVTZ': {'H': (14, 30), 'C': (30, 81), 'O': (30, 81)},
}
_ELECTRONS = {'H': 1, 'C': 6, 'O': 8}


def shape_of(counts: dict, basis: str):
    nbf = sum(n * _BASIS[basis][el][0] for el, n in counts.items())
    naux = sum(n * _BASIS[basis][el][1] for el, n in counts.items())
    nocc = sum(n * _ELECTRONS[el] for el, n in counts.items()) // 2
    return nbf, naux, nocc


def alkane_shape(n: int, basis: str = 'cc-pVDZ'):
    """(nbf, naux, nocc) of CnH(2n+2)."""
    return shape_of({'C': n, 'H': 2 * n + 2}, basis)


def water_cluster_shape(m: int, basis: str = 'cc-pVDZ'):
    """(nbf, naux, nocc) of (H2O)m."""
    return shape_of({'O': m, 'H': 2 * m}, basis)


def upper_pairs(nbf: int) -> np.ndarray:
    """index_basis (2, npair), 1-based, pairs k <= l ordered column by column."""
    l, k = np.triu_indices(nbf)[1], np.triu_indices(nbf)[0]
    order = np.lexsort((k, l))
    return np.vstack([k[order] + 1, l[order] + 1]).astype(np.int32)


def orthogonal(rng: np.random.Generator, n: int) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))[None, :]


def synthetic_system(nbf: int, naux: int, nocc: int, seed: int = 0, name: str = 'synthetic',
                     scale: float = 1.0) -> System:
    """Seeded synthetic system in MOLGW's packed RI layout (host numpy).

    ``scale`` multiplies B: 1.0 is the benchmark recipe (which gives
    (ia|ia) ~ 2 Ha, an unphysically strong coupling); ~0.3 gives molecular
    magnitudes ((ia|ia) ~ 0.2 Ha, QP roots inside the +-0.25 Ha window)."""
    rng = np.random.default_rng(seed)
    c = orthogonal(rng, nbf)
    e_occ = np.sort(rng.uniform(-1.2, -0.3, nocc))
    e_virt = np.sort(rng.uniform(0.05, 3.0, nbf - nocc))
    energy = np.concatenate([e_occ, e_virt])
    occupation = np.concatenate([np.full(nocc, 2.0), np.zeros(nbf - nocc)])
    u = orthogonal(rng, naux)
    v = orthogonal(rng, naux)
    sv = 1.0 / (1.0 + np.arange(naux)) ** 1.5
    lmat = (u * sv[None, :]) @ v.T
    ib = upper_pairs(nbf)
    npair = ib.shape[1]
    g = rng.standard_normal((naux, npair))
    # G symmetric in (mu, nu): one value per unordered pair, the full matrix
    # element; packed storage halves the diagonal pairs
    b = (lmat @ g).T
    if scale != 1.0:
        b *= scale
    diag = ib[0] == ib[1]
    b[diag, :] *= 0.5
    return System(name, nbf, naux, nocc, np.asfortranarray(b), ib, np.asfortranarray(c), energy, occupation)


def dense_ri(sys: System) -> np.ndarray:
    """Dense symmetric B(P, mu, nu) of a System (for checks at small sizes)."""
    k = sys.index_basis[0] - 1
    l = sys.index_basis[1] - 1
    out = np.zeros((sys.naux, sys.nbf, sys.nbf))
    out[:, k, l] = sys.eri_3center.T
    out[:, l, k] = sys.eri_3center.T
    d = k == l
    out[:, k[d], l[d]] = 2.0 * sys.eri_3center[d].T
    return out


BASIS_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), 'tests', 'golden', 'basis')


def molecular_system_device(mol, name: str, basis: str = 'cc-pVDZ', auxil_basis: str = 'cc-pVDZ-RI',
                            basis_dir: str | None = None, engine=None):
    """A real molecule's G0W0 inputs from the GPU front end (device
    integrals, RI-HF with J and K through the resident RI tensor,
    Sigma_x - Vxc from its Fock matrix): a DeviceSystem whose packed RI tensor
    eri_3center(npair, naux_kept) stays in HBM (a libmgw device buffer), with
    the orbitals, energies, occupations and the diagonal of Sigma_x - Vxc on
    the host."""
    from .frontend import run_rihf_device
    from .frontend.basis import BasisSet
    from .gw import DeviceBuffer, Engine
    bdir = basis_dir or BASIS_DIR
    own = engine is None
    eng = Engine(0) if own else engine
    try:
        nbf = BasisSet(mol.atoms, mol.coords, basis, bdir).nbf
        npair = nbf * (nbf + 1) // 2
        # the kept auxiliary count is known after the metric; allocate for all
        # (column-major (npair, naux): the kept columns come first)
        naux_bas = BasisSet(mol.atoms, mol.coords, auxil_basis, bdir).nbf
        buf = DeviceBuffer(8 * naux_bas * npair, eng.device)
        hf = run_rihf_device(mol, basis, auxil_basis, bdir, eng, eri_3center_out=buf.data_ptr())
        eng.set_mo(hf.c_matrix, hf.energy_mo, hf.occupation)
        _, exc = eng.exchange_m_vxc(hf.hamiltonian_fock)
        eng.synchronize()
    finally:
        if own:
            eng.close()
    sysd = DeviceSystem(name, hf.nbf, hf.naux, hf.nocc, buf, hf.index_basis, np.asfortranarray(hf.c_matrix),
                        hf.energy_mo, hf.occupation)
    sysd.exchange_m_vxc = exc
    sysd.scf_energy = hf.energy
    return sysd


@dataclass
class DeviceSystem:
    """A system whose packed RI tensor lives in device memory: a libmgw
    DeviceBuffer whose first 8 npair naux bytes hold eri_3center(npair, naux)
    column-major."""
    name: str
    nbf: int
    naux: int
    nocc: int
    eri_3center_dev: object      # molgw_amd.gw.DeviceBuffer
    index_basis: np.ndarray
    c_matrix: np.ndarray
    energy: np.ndarray
    occupation: np.ndarray
    exchange_m_vxc: np.ndarray | None = None   # diagonal of Sigma_x - Vxc (None: converged HF, zero)
    scf_energy: float | None = None

    @property
    def nstate(self) -> int:
        return self.c_matrix.shape[1]

    @property
    def npair(self) -> int:
        return self.index_basis.shape[1]

    def eri_3center_host(self) -> np.ndarray:
        """eri_3center(npair, naux) copied to the host (Fortran order)."""
        out = np.empty((self.npair, self.naux), order='F')
        return self.eri_3center_dev.download(out)


def synthetic_system_device(nbf: int, naux: int, nocc: int, seed: int = 0, name: str = 'synthetic',
                            scale: float = 1.0, chunk: int = 32768) -> DeviceSystem:
    """Same recipe as ``synthetic_system`` for sizes whose RI tensor does not
    fit the host comfortably: B = L G is formed on the host a block of pairs
    at a time (G drawn per block from the same seeded stream, so values
    differ from the host recipe) and copied into a device buffer column by
    column of the (npair, naux) layout."""
    from .gw import DeviceBuffer
    rng = np.random.default_rng(seed)
    c = orthogonal(rng, nbf)
    e_occ = np.sort(rng.uniform(-1.2, -0.3, nocc))
    e_virt = np.sort(rng.uniform(0.05, 3.0, nbf - nocc))
    energy = np.concatenate([e_occ, e_virt])
    occupation = np.concatenate([np.full(nocc, 2.0), np.zeros(nbf - nocc)])
    u = orthogonal(rng, naux)
    v = orthogonal(rng, naux)
    sv = 1.0 / (1.0 + np.arange(naux)) ** 1.5
    lmat = (u * sv[None, :]) @ v.T
    ib = upper_pairs(nbf)
    npair = ib.shape[1]
    diag = ib[0] == ib[1]
    buf = DeviceBuffer(8 * naux * npair)
    for p0 in range(0, npair, chunk):
        p1 = min(npair, p0 + chunk)
        blk = lmat @ rng.standard_normal((naux, p1 - p0))      # (naux, pairs) C order
        if scale != 1.0:
            blk *= scale
        blk[:, diag[p0:p1]] *= 0.5
        for q in range(naux):                                  # column q of (npair, naux)
            buf.upload(blk[q], 8 * (p0 + q * npair))
    return DeviceSystem(name, nbf, naux, nocc, buf, ib, np.asfortranarray(c), energy, occupation)
