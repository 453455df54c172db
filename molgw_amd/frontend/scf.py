"""RI Hartree-Fock front end: the molecule, its integrals, the SCF, and the
inputs of the G0W0 path (packed eri_3center_tmp, the 2-center metric, C,
energies, occupations).

Follows the reference's RI-HF flow (src/m_scf_loop.f90; Hartree and exchange
through the RI tensor, src/m_hamiltonian_twobodies.f90): the orbital basis
is filtered by the overlap eigenvalues (setup_x_matrix,
src/m_hamiltonian_tools.f90:1171-1219, min_overlap 1e-5), the metric by
eigenvalues above 1e-6 (src/m_eri_calculate.f90:1049-1262), and the pair
tensor carries the 1/2 on diagonal pairs (src/m_eri_calculate.f90:1840-1841).
The SCF itself is a textbook closed-shell Roothaan-Hall iteration with DIIS
from the core-Hamiltonian guess; at convergence its energy and orbitals are
those of any converged RI-HF run.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import integrals as I
from .basis import ANGSTROM_BOHR, BasisSet, atomic_number
from ..workloads import upper_pairs

MIN_OVERLAP = 1.0e-5        # src/input_variables.yaml:1804-1811
TOO_LOW_EIGENVAL = 1.0e-6   # src/m_eri_calculate.f90 (metric filter)


@dataclass
class Molecule:
    atoms: list          # atomic numbers
    coords: np.ndarray   # bohr, (natom, 3)

    @classmethod
    def from_lines(cls, lines, unit: str = 'angstrom'):
        """Atoms as in a MOLGW input: ``symbol-or-Z x y z`` per line."""
        atoms, xyz = [], []
        for ln in lines:
            f = ln.split()
            if not f:
                continue
            atoms.append(atomic_number(f[0]))
            xyz.append([float(v) for v in f[1:4]])
        xyz = np.array(xyz, dtype=np.float64)
        if unit.lower().startswith('ang'):
            xyz = xyz * ANGSTROM_BOHR
        return cls(atoms, xyz)

    @property
    def nelectron(self) -> int:
        return int(sum(self.atoms))


def frozen_core_states(atoms, frozencore: str = 'MOLGW') -> int:
    """ncore_G = ncore_W of the frozencore keyword (src/m_selfenergy_tools.f90:82-87,
    src/m_spectral_function.f90:186-191): the sum over the atoms of
    element_core ('MOLGW'/'YES': none up to Be, 1 up to Mg, 5 up to Zn, 9 up
    to Cd, else 18) or element_core_gaussian ('GAUSSIAN': none up to He, 1 up
    to Ne, then as MOLGW) of src/m_elements.f90:40-100 (all-electron atoms);
    0 for 'NO'."""
    rule = frozencore.strip().upper()
    if rule in ('NO', ''):
        return 0
    if rule not in ('MOLGW', 'YES', 'GAUSSIAN'):
        raise ValueError(f'unknown frozencore {frozencore!r}')
    first = 2 if rule == 'GAUSSIAN' else 4
    n = 0
    for z in atoms:
        n += 0 if z <= first else 1 if z <= (10 if rule == 'GAUSSIAN' else 12) else 5 if z <= 30 else \
            9 if z <= 48 else 18
    return n


@dataclass
class RIHFResult:
    energy: float
    components: dict
    c_matrix: np.ndarray        # (nbf, nstate)
    energy_mo: np.ndarray       # (nstate,)
    occupation: np.ndarray      # (nstate,)
    eri_3center_tmp: np.ndarray  # packed (npair, naux_bas), diagonal pairs x1/2, before V^-1/2
    eri_2center: np.ndarray     # (naux_bas, naux_bas)
    index_basis: np.ndarray     # (2, npair), 1-based
    nbf: int
    iterations: int


def _pack(eri3: np.ndarray, index_basis: np.ndarray) -> np.ndarray:
    k, l = index_basis[0] - 1, index_basis[1] - 1
    out = eri3[k, l, :].copy()
    out[k == l] *= 0.5
    return out


def run_rihf(mol: Molecule, basis: str, auxil_basis: str, basis_dir: str, gaussian_type: str = 'PURE',
             tol: float = 1e-10, max_iter: int = 100, diis: int = 8, engine: str = 'native') -> RIHFResult:
    """Closed-shell RI-HF. Returns the converged energy, its parts (kinetic,
    nucleus, Hartree, exchange, nucleus-nucleus) and the G0W0 inputs.
    ``engine``: 'native' (ints.cpp, OpenMP; the default) or 'numpy' (the
    integrals module, the same algorithm, which pins the native one)."""
    if mol.nelectron % 2:
        raise ValueError('closed shell only')
    bs = BasisSet(mol.atoms, mol.coords, basis, basis_dir, gaussian_type)
    aux = BasisSet(mol.atoms, mol.coords, auxil_basis, basis_dir, gaussian_type)
    charges = [float(z) for z in mol.atoms]
    if engine == 'native':
        from . import native as ints
    elif engine == 'numpy':
        ints = I
    else:
        raise ValueError(f'unknown integral engine {engine!r}')
    s, t, v = ints.one_electron(bs, charges, mol.coords)
    m2 = ints.eri_2center(aux)
    e3 = ints.eri_3center(bs, aux)
    enuc = I.nuclear_repulsion(charges, mol.coords)
    # metric V^-1/2 over the kept eigenvalues, and B(mu nu, P)
    w, u = np.linalg.eigh(m2)
    keep = w > TOO_LOW_EIGENVAL
    sqrtinv = u[:, keep] / np.sqrt(w[keep])
    b = e3 @ sqrtinv                                            # (n, n, naux_kept)
    # orthogonalisation X = U s^-1/2 over the kept overlap eigenvalues
    sw, su = np.linalg.eigh(s)
    xk = sw > MIN_OVERLAP
    x = su[:, xk] / np.sqrt(sw[xk])
    nstate = x.shape[1]
    nocc = mol.nelectron // 2
    h = t + v

    def fock(d):
        j = np.einsum('mnP,P->mn', b, np.einsum('lsP,ls->P', b, d), optimize=True)
        bd = np.einsum('mlP,ls->msP', b, d, optimize=True)
        k = np.einsum('msP,snP->mn', bd, b, optimize=True)
        return h + j - 0.5 * k, j, k

    def solve(f):
        e, cp = np.linalg.eigh(x.T @ f @ x)
        return e, x @ cp

    eps, c = solve(h)
    d = 2.0 * c[:, :nocc] @ c[:, :nocc].T
    e_old = 0.0
    fs, errs = [], []
    it = 0
    for it in range(1, max_iter + 1):
        f, j, k = fock(d)
        energy = 0.5 * np.sum(d * (h + f)) + enuc
        err = x.T @ (f @ d @ s - s @ d @ f) @ x
        fs.append(f)
        errs.append(err)
        if len(fs) > diis:
            fs.pop(0)
            errs.pop(0)
        if len(fs) > 1:
            nd = len(fs)
            bm = -np.ones((nd + 1, nd + 1))
            bm[nd, nd] = 0.0
            for a_ in range(nd):
                for b_ in range(nd):
                    bm[a_, b_] = np.sum(errs[a_] * errs[b_])
            rhs = np.zeros(nd + 1)
            rhs[nd] = -1.0
            coef = np.linalg.solve(bm, rhs)[:nd]
            f_use = sum(cf * ff for cf, ff in zip(coef, fs))
        else:
            f_use = f
        eps, c = solve(f_use)
        d = 2.0 * c[:, :nocc] @ c[:, :nocc].T
        if abs(energy - e_old) < tol and np.abs(err).max() < 1e-7:
            break
        e_old = energy
    # final orbitals from the converged Fock matrix itself
    f, j, k = fock(d)
    eps, c = solve(f)
    d = 2.0 * c[:, :nocc] @ c[:, :nocc].T
    f, j, k = fock(d)
    energy = 0.5 * np.sum(d * (h + f)) + enuc
    comps = {'kinetic': float(np.sum(d * t)), 'nucleus': float(np.sum(d * v)),
             'hartree': float(0.5 * np.sum(d * j)), 'exchange': float(-0.25 * np.sum(d * k)),
             'nucleus_nucleus': enuc}
    occ = np.zeros(nstate)
    occ[:nocc] = 2.0
    ib = upper_pairs(bs.nbf)
    return RIHFResult(float(energy), comps, c, eps, occ, _pack(e3, ib), m2, ib, bs.nbf, it)


@dataclass
class DeviceRIHF:
    """Closed-shell RI-HF whose RI tensor was built and stays resident on the
    GPU (Engine): the G0W0 inputs without any host copy of the 3-centre
    integrals."""
    energy: float
    components: dict
    c_matrix: np.ndarray        # (nbf, nstate)
    energy_mo: np.ndarray       # (nstate,)
    occupation: np.ndarray      # (nstate,)
    hamiltonian_fock: np.ndarray  # (nbf, nbf): h + J + exchange at the final density
    s_matrix: np.ndarray        # (nbf, nbf)
    index_basis: np.ndarray     # (2, npair), 1-based
    nbf: int
    naux: int                   # kept auxiliary functions
    naux_bas: int
    iterations: int

    @property
    def nstate(self) -> int:
        return self.c_matrix.shape[1]

    @property
    def nocc(self) -> int:
        return int(np.count_nonzero(self.occupation > 1e-5))


def run_rihf_device(mol: Molecule, basis: str, auxil_basis: str, basis_dir: str, engine, gaussian_type: str = 'PURE',
                    tol: float = 1e-10, max_iter: int = 100, diis: int = 8, eri_3center_out: int | None = None):
    """Closed-shell RI-HF with the integrals and the Hartree/exchange builds
    on the GPU: (P|Q) and (mu nu|P) by the device integral kernels, V^-1/2
    and B = (mu nu|Q) V^-1/2 on the device (B then stays resident in
    ``engine``, a molgw_amd.gw.Engine), and J and K through the resident B
    at every iteration (setup_hartree_ri / setup_exchange_ri). The
    one-electron integrals, the orthogonalisation, the diagonalisations and
    DIIS run on the host (nbf x nbf). The same SCF as run_rihf, so at
    convergence the energy and orbitals are those of the host route.
    ``eri_3center_out`` (optional device address) receives the packed RI
    tensor eri_3center(npair, naux_kept)."""
    if mol.nelectron % 2:
        raise ValueError('closed shell only')
    from . import native
    bs = BasisSet(mol.atoms, mol.coords, basis, basis_dir, gaussian_type)
    aux = BasisSet(mol.atoms, mol.coords, auxil_basis, basis_dir, gaussian_type)
    charges = [float(z) for z in mol.atoms]
    s, t, v = native.one_electron(bs, charges, mol.coords)
    enuc = I.nuclear_repulsion(charges, mol.coords)
    m2 = engine.eri_2center_ao(aux)
    kept, _, _ = engine.inverse_sqrt_eri_2center(m2, copy=False)
    ib = upper_pairs(bs.nbf)
    engine.eri_3center_ao(bs, aux, ib, copy=False)
    engine.build_eri_3center_resident(ib, bs.nbf, out_ptr=eri_3center_out)
    sw, su = np.linalg.eigh(s)
    xk = sw > MIN_OVERLAP
    x = su[:, xk] / np.sqrt(sw[xk])
    nstate = x.shape[1]
    nocc = mol.nelectron // 2
    h = t + v

    def fock(c):
        j, kx = engine.hartree_exchange_ri(c[:, :nocc], spin_fact=2.0)   # sqrt(occupation / spin_fact) = 1
        return h + j + kx, j, kx

    def solve(f):
        e, cp = np.linalg.eigh(x.T @ f @ x)
        return e, x @ cp

    eps, c = solve(h)
    e_old = 0.0
    fs, errs = [], []
    it = 0
    for it in range(1, max_iter + 1):
        d = 2.0 * c[:, :nocc] @ c[:, :nocc].T
        f, j, kx = fock(c)
        energy = 0.5 * np.sum(d * (h + f)) + enuc
        err = x.T @ (f @ d @ s - s @ d @ f) @ x
        fs.append(f)
        errs.append(err)
        if len(fs) > diis:
            fs.pop(0)
            errs.pop(0)
        if len(fs) > 1:
            nd = len(fs)
            bm = -np.ones((nd + 1, nd + 1))
            bm[nd, nd] = 0.0
            for a_ in range(nd):
                for b_ in range(nd):
                    bm[a_, b_] = np.sum(errs[a_] * errs[b_])
            rhs = np.zeros(nd + 1)
            rhs[nd] = -1.0
            coef = np.linalg.solve(bm, rhs)[:nd]
            f_use = sum(cf * ff for cf, ff in zip(coef, fs))
        else:
            f_use = f
        eps, c = solve(f_use)
        if abs(energy - e_old) < tol and np.abs(err).max() < 1e-7:
            break
        e_old = energy
    # final orbitals from the Fock matrix of the converged density itself
    f, j, kx = fock(c)
    eps, c = solve(f)
    d = 2.0 * c[:, :nocc] @ c[:, :nocc].T
    f, j, kx = fock(c)
    energy = 0.5 * np.sum(d * (h + f)) + enuc
    comps = {'kinetic': float(np.sum(d * t)), 'nucleus': float(np.sum(d * v)),
             'hartree': float(0.5 * np.sum(d * j)), 'exchange': float(0.5 * np.sum(d * kx)),
             'nucleus_nucleus': enuc}
    occ = np.zeros(nstate)
    occ[:nocc] = 2.0
    return DeviceRIHF(float(energy), comps, c, eps, occ, f, s, ib, bs.nbf, kept, aux.nbf, it)


def water_cluster(m: int, seed: int = 0, oo_min: float = 2.7, volume_per_molecule: float = 45.0,
                  inter_min: float = 1.5) -> Molecule:
    """(H2O)m from rigid monomers of the water geometry O (0, 0, 0),
    H (+-0.757, 0.587, 0) Angstrom, each with a seeded random orientation
    (a uniformly distributed rotation), with O centres placed by seeded
    rejection sampling in a sphere of volume m * volume_per_molecule A^3
    with every O-O distance >= oo_min. A trial placement is also rejected
    when an atom comes within inter_min Angstrom of an atom of another
    monomer (no hydrogen closer than a hydrogen bond allows)."""
    rng = np.random.default_rng(seed)
    mono = np.array([[0.0, 0.0, 0.0], [0.757, 0.587, 0.0], [-0.757, 0.587, 0.0]])
    radius = (3.0 * m * volume_per_molecule / (4.0 * math.pi)) ** (1.0 / 3.0)
    placed = []
    tries = 0
    while len(placed) < m:
        tries += 1
        if tries > 200000:
            raise RuntimeError(f'water_cluster: could not place {m} molecules')
        o = rng.uniform(-radius, radius, 3)
        if np.dot(o, o) > radius * radius:
            continue
        q, r = np.linalg.qr(rng.standard_normal((3, 3)))
        rot = q * np.sign(np.diag(r))
        if np.linalg.det(rot) < 0:
            rot[:, 0] = -rot[:, 0]
        xyz = o + mono @ rot.T
        ok = True
        for other in placed:
            if np.linalg.norm(other[0] - o) < oo_min:
                ok = False
                break
            if np.min(np.linalg.norm(other[:, None, :] - xyz[None, :, :], axis=2)) < inter_min:
                ok = False
                break
        if ok:
            placed.append(xyz)
    atoms, coords = [], []
    for xyz in placed:
        atoms += [8, 1, 1]
        coords += list(xyz)
    return Molecule(atoms, np.array(coords) * ANGSTROM_BOHR)


def alkane(n: int, cc: float = 1.54, ch: float = 1.09) -> Molecule:
    """All-trans n-alkane CnH2n+2 from bond rules: C-C cc, C-H ch (Angstrom),
    tetrahedral angles, the carbon zigzag in the xy plane."""
    th = math.acos(-1.0 / 3.0)                    # tetrahedral angle
    dx, dy = cc * math.sin(th / 2), cc * math.cos(th / 2)
    cs = [np.array([k * dx, dy * (k % 2), 0.0]) for k in range(n)]
    atoms, xyz = [6] * n, list(cs)
    ez = np.array([0.0, 0.0, 1.0])
    for k, c in enumerate(cs):
        nb = [cs[j] for j in (k - 1, k + 1) if 0 <= j < n]
        us = [(b - c) / np.linalg.norm(b - c) for b in nb]
        if len(us) == 2:
            bis = -(us[0] + us[1])
            bis /= np.linalg.norm(bis)
            dirs = [bis * math.cos(th / 2) + s * ez * math.sin(th / 2) for s in (1, -1)]
        else:
            u = us[0] if us else np.array([1.0, 0.0, 0.0])
            e1 = np.cross(ez, u)
            e1 /= np.linalg.norm(e1)
            if len(cs) > 1 and np.dot(e1, np.array([0.0, 1.0, 0.0])) * (1 if k % 2 else -1) > 0:
                e1 = -e1
            dirs = [-u / 3.0 + math.sqrt(8.0) / 3.0 * (math.cos(p) * e1 + math.sin(p) * ez)
                    for p in (0.0, 2 * math.pi / 3, 4 * math.pi / 3)]
            if not us:                               # methane: a fourth hydrogen
                dirs.append(u)
        for d in dirs:
            atoms.append(1)
            xyz.append(c + ch * d)
    return Molecule(atoms, np.array(xyz) * ANGSTROM_BOHR)
