"""Gaussian basis sets in the reference's file format and layout.

File format (read by init_basis_set, src/m_basis_set.f90:90-364): the number
of contracted shells, then per shell a line ``ngaussian am`` followed by
ngaussian lines ``exponent coefficient`` (coefficients of normalised
primitives). Files are named ``<Element>_<basis name>`` in a basis directory.

Each contracted Cartesian function is normalised to one on its own
(init_basis_function, src/m_basis_set.f90:1191-1238). Pure functions span
the harmonic polynomials of degree l (real solid harmonics): here the
2l+1 functions of a shell are the orthonormal (on the sphere) basis of that
space that a Lowdin orthonormalisation of the kernel of the Laplacian gives,
each normalised to one; they differ from the reference's solid harmonics by
an orthogonal rotation within the shell, which leaves energies, overlap
eigenvalues and Coulomb-metric eigenvalues unchanged.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

import numpy as np

ANGSTROM_BOHR = 1.0 / 0.52917721092          # src/m_definitions.f90 (bohr_A)

ELEMENTS = ['H', 'He', 'Li', 'Be', 'B', 'C', 'N', 'O', 'F', 'Ne', 'Na', 'Mg', 'Al', 'Si', 'P', 'S', 'Cl', 'Ar']


def element_symbol(z: int) -> str:
    return ELEMENTS[z - 1]


def atomic_number(sym: str) -> int:
    s = sym.strip()
    if s.isdigit():
        return int(s)
    return ELEMENTS.index(s[0].upper() + s[1:].lower()) + 1


def cart_components(l: int):
    """(nx, ny, nz) of a Cartesian shell, x power decreasing first."""
    return [(l - i, i - k, k) for i in range(l + 1) for k in range(i + 1)]


def _double_factorial(n: int) -> int:
    return 1 if n <= 0 else n * _double_factorial(n - 2)


def _sphere_moment(i: int, j: int, k: int) -> float:
    """Integral of x^i y^j z^k over the unit sphere."""
    if i % 2 or j % 2 or k % 2:
        return 0.0
    return 2.0 * math.gamma((i + 1) / 2) * math.gamma((j + 1) / 2) * math.gamma((k + 1) / 2) \
        / math.gamma((i + j + k + 3) / 2)


def harmonic_transform(l: int) -> np.ndarray:
    """(2l+1, ncart) coefficients of an orthonormal basis (sphere inner product)
    of the degree-l harmonic polynomials in the Cartesian monomials."""
    comps = cart_components(l)
    if l < 2:
        t = np.eye(len(comps))
    else:
        lower = {c: n for n, c in enumerate(cart_components(l - 2))}
        lap = np.zeros((len(lower), len(comps)))
        for n, (i, j, k) in enumerate(comps):
            for d, (a, b, c) in enumerate(((i, j, k), (j, i, k), (k, i, j))):
                if a >= 2:
                    e = [i, j, k]
                    e[d] -= 2
                    lap[lower[tuple(e)], n] += a * (a - 1)
        # the Laplacian maps degree l onto degree l-2 (full rank): its
        # null space is the 2l+1 harmonic polynomials
        _, _, vt = np.linalg.svd(lap)
        t = vt[lap.shape[0]:]
    m = np.array([[_sphere_moment(a[0] + b[0], a[1] + b[1], a[2] + b[2]) for b in comps] for a in comps])
    g = t @ m @ t.T
    w, u = np.linalg.eigh(g)
    return (u @ np.diag(w ** -0.5) @ u.T) @ t


@dataclass
class Shell:
    """A contracted shell: l, centre (bohr), exponents, primitive weights d_g
    (file coefficient times the l-dependent part of the primitive norm) and
    the (nfunc, ncart) map from the Cartesian components to the final
    normalised functions (filled by BasisSet)."""
    l: int
    center: np.ndarray
    exps: np.ndarray
    coefs: np.ndarray
    atom: int
    transform: np.ndarray = field(default=None)

    @property
    def ncart(self) -> int:
        return (self.l + 1) * (self.l + 2) // 2

    @property
    def nfunc(self) -> int:
        return self.transform.shape[0]


def read_basis_file(path: str):
    """[(l, exponents, coefficients)] of one element's basis file."""
    with open(path) as fh:
        tokens = fh.read().split()
    pos = 0
    nshell = int(tokens[pos]); pos += 1
    out = []
    for _ in range(nshell):
        ng, am = int(tokens[pos]), int(tokens[pos + 1]); pos += 2
        if am == 10:
            raise ValueError(f'{path}: shared-exponent SP shells are not supported (split them)')
        e = np.array([float(tokens[pos + 2 * g].replace('D', 'E')) for g in range(ng)])
        c = np.array([float(tokens[pos + 2 * g + 1].replace('D', 'E')) for g in range(ng)])
        pos += 2 * ng
        out.append((am, e, c))
    return out


class BasisSet:
    """The shells of a molecule in one basis (orbital or auxiliary), with the
    reference's function order: atoms in input order, shells in file order,
    functions of a shell contiguous."""

    def __init__(self, atoms, coords_bohr, basis_name: str, basis_dir: str, gaussian_type: str = 'PURE'):
        self.gaussian_type = gaussian_type.upper()
        self.shells: list[Shell] = []
        for ia, (z, xyz) in enumerate(zip(atoms, coords_bohr)):
            path = os.path.join(basis_dir, f'{element_symbol(z)}_{basis_name}')
            for l, e, c in read_basis_file(path):
                d = c * (2.0 * e / math.pi) ** 0.75 * (4.0 * e) ** (l / 2.0)
                self.shells.append(Shell(l, np.asarray(xyz, dtype=np.float64), e, d, ia))
        from .integrals import shell_overlap_cart
        for sh in self.shells:
            s = shell_overlap_cart(sh, sh)
            if self.gaussian_type == 'CART':
                sh.transform = np.diag(1.0 / np.sqrt(np.diag(s)))
            else:
                h = harmonic_transform(sh.l)
                n = np.einsum('ia,ab,ib->i', h, s, h)
                sh.transform = h / np.sqrt(n)[:, None]
        self.offsets = np.cumsum([0] + [sh.nfunc for sh in self.shells])
        self.nbf = int(self.offsets[-1])

    def __len__(self):
        return self.nbf
