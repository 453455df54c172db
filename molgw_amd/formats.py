"""Files around the G0W0 path, in the reference's layouts.

* ENERGY_QP: the QP energies a run leaves behind and a later run (evGW,
  restart, scissor) reads back (src/m_io.f90:3645-3733).
* molgw.yaml sections that ``molgw.py`` parses for GW results:
  ``<approx> energies`` (src/m_hamiltonian_tools.f90:532-560) and
  ``<approx> selfenergy`` (src/m_selfenergy_tools.f90:501-553).
* SCREENED_COULOMB: the poles and residues of W (write_spectral_function /
  read_spectral_function, src/m_spectral_function.f90:414-620), as a
  sequential unformatted Fortran file (4-byte record markers, gfortran's
  default).
* dump_gw_ingredients (src/m_gw_selfenergy_analytic.f90:801-881):
  energies.dat, omegas.dat, w.bin, v.dat, v.bin.

Energies are in Hartree in the arrays and in eV in the yaml, as in the
reference (Ha_eV of src/m_definitions.f90).
"""
from __future__ import annotations

import os

import numpy as np

from .gw import HA_EV


# ---------------------------------------------------------------- ENERGY_QP
def write_energy_qp(path: str, energy_qp) -> None:
    """write_energy_qp (src/m_io.f90:3645-3681): nspin, nstate, then one
    line per state ``istate e(spin 1) [e(spin 2)]`` (1-based, Hartree),
    whitespace separated (Fortran list-directed output)."""
    e = np.asarray(energy_qp, dtype=np.float64)
    if e.ndim == 1:
        e = e[:, None]
    nstate, nspin = e.shape
    if nspin not in (1, 2):
        raise ValueError('energy_qp must have one or two spin columns')
    with open(path, 'w') as fh:
        fh.write(f'{nspin:12d}\n{nstate:12d}\n')
        for i in range(nstate):
            fh.write(f'{i + 1:12d}' + ''.join(f' {v:25.17e}' for v in e[i]) + '\n')


def read_energy_qp(nstate: int, nspin: int = 1, directory: str = '.'):
    """read_energy_qp (src/m_io.f90:3685-3733). Looks for ENERGY_QP, then
    energy_qp. Returns (energy_qp[nstate, nspin] or None, status) with the
    reference's status codes: 0 read, 1 no file, 2 wrong dimensions, -1 a
    state index of -1 (the scissor-operator form)."""
    for name in ('ENERGY_QP', 'energy_qp'):
        path = os.path.join(directory, name)
        if os.path.exists(path):
            break
    else:
        return None, 1
    with open(path) as fh:
        tokens = fh.read().split()
    nspin_read, nstate_read = int(tokens[0]), int(tokens[1])
    if nspin_read != nspin or nstate_read != nstate:
        return None, 2
    e = np.zeros((nstate, nspin))
    pos = 2
    for i in range(nstate):
        j = int(tokens[pos])
        if j == -1:
            return None, -1
        e[i] = [float(t.replace('D', 'E').replace('d', 'e')) for t in tokens[pos + 1:pos + 1 + nspin]]
        pos += 1 + nspin
    return e, 0


# ---------------------------------------------------------------- yaml
def _es18_8(x: float) -> str:
    """Fortran ES18.8: one digit before the point, 8 after, a two-digit (or
    wider) signed exponent, right-justified in 18 columns."""
    if x == 0.0:
        return f'{"0.00000000E+00":>18}'
    m, ex = f'{x:.8E}'.split('E')
    return f'{m}E{int(ex):+03d}'.rjust(18)


def _state_key(i: int, indent: int) -> str:
    # '(Nx,a6,a,...)' with ADJUSTL(char6): the index left-adjusted in 6 columns
    return ' ' * indent + f'{i:<6d}' + ':'


def yaml_energies(title: str, energy, lb: int | None = None, ub: int | None = None) -> str:
    """dump_out_energy_yaml (src/m_hamiltonian_tools.f90:532-560) for
    energy[nstate, nspin] (or [nstate]) in Hartree, states lb..ub (1-based)."""
    e = np.asarray(energy, dtype=np.float64)
    if e.ndim == 1:
        e = e[:, None]
    nstate, nspin = e.shape
    lo = max(1, lb or 1)
    hi = min(nstate, ub or nstate)
    out = [f'\n{title.strip().lower()}:', '    unit: eV']
    for s in range(nspin):
        out.append(f'    spin channel{s + 1:2d}:')
        for i in range(lo, hi + 1):
            out.append(_state_key(i, 8) + ' ' + _es18_8(e[i - 1, s] * HA_EV))
    return '\n'.join(out) + '\n'


def yaml_qp_selfenergy(calcname: str, energy0, exchange_m_vxc, energy_qp, zz, nsemin: int, nsemax: int) -> str:
    """output_qp_energy_yaml (src/m_selfenergy_tools.f90:501-553) with the QP
    energies and Z given: correlation Sigma_c(E_qp) = E_qp - e0 - (Sigma_x - Vxc),
    exchange minus vxc, and the renormalization factor, states nsemin..nsemax."""
    def cols(a):
        a = np.asarray(a, dtype=np.float64)
        return a[:, None] if a.ndim == 1 else a
    e0, exc, eqp, z = cols(energy0), cols(exchange_m_vxc), cols(energy_qp), cols(zz)
    nspin = e0.shape[1]
    out = [f'\n{calcname.strip().lower()} selfenergy:', '    correlation:', '        unit: eV']
    for s in range(nspin):
        out.append(f'        spin channel{s + 1:2d}:')
        for p in range(nsemin, nsemax + 1):
            sig = eqp[p - 1, s] - e0[p - 1, s] - exc[p - 1, s]
            out.append(_state_key(p, 12) + ' ' + _es18_8(sig * HA_EV))
    out += ['    exchange minus vxc:', '        unit: eV']
    for s in range(nspin):
        out.append(f'        spin channel{s + 1:2d}:')
        for p in range(nsemin, nsemax + 1):
            out.append(_state_key(p, 12) + ' ' + _es18_8(exc[p - 1, s] * HA_EV))
    out.append('    renormalization factor:')
    for s in range(nspin):
        out.append(f'        spin channel{s + 1:2d}:')
        for p in range(nsemin, nsemax + 1):
            out.append(_state_key(p, 12) + ' ' + _es18_8(z[p - 1, s]))
    return '\n'.join(out) + '\n'


def append_gw_yaml(path: str, energy0, exchange_m_vxc, energy_qp, zz, nsemin: int, nsemax: int,
                   approx: str = 'GW') -> None:
    """Append the two sections the reference writes after a one-shot G0W0
    (src/m_selfenergy_evaluation.f90:373-375) to a molgw.yaml file."""
    with open(path, 'a') as fh:
        fh.write(yaml_qp_selfenergy(approx, energy0, exchange_m_vxc, energy_qp, zz, nsemin, nsemax))
        fh.write(yaml_energies(f'{approx} energies', energy_qp, nsemin, nsemax))


# ---------------------------------------------------------------- SCREENED_COULOMB
def _record(fh, payload: bytes) -> None:
    n = np.array([len(payload)], dtype='<i4').tobytes()
    fh.write(n + payload + n)


def _records(path: str):
    with open(path, 'rb') as fh:
        data = fh.read()
    pos = 0
    while pos < len(data):
        n = int(np.frombuffer(data, dtype='<i4', count=1, offset=pos)[0])
        payload = data[pos + 4:pos + 4 + n]
        tail = int(np.frombuffer(data, dtype='<i4', count=1, offset=pos + 4 + n)[0])
        if tail != n:
            raise ValueError(f'{path}: record markers differ ({n} vs {tail})')
        yield payload
        pos += n + 8


def write_screened_coulomb(path: str, pole, w_s, postscf_name: str = 'GW') -> None:
    """write_spectral_function (src/m_spectral_function.f90:414-445, the
    serial branch): postscf_name (character(len=100)), nprodbasis_total,
    npole_reso, pole(:), then w_s(:, s) for every pole, one record each."""
    pole = np.asarray(pole, dtype='<f8').reshape(-1)
    w_s = np.asarray(w_s, dtype='<f8')
    if w_s.shape[1] != len(pole):
        raise ValueError('w_s must be (nprodbasis, npole)')
    with open(path, 'wb') as fh:
        _record(fh, postscf_name.ljust(100)[:100].encode('ascii'))
        _record(fh, np.array([w_s.shape[0]], dtype='<i4').tobytes())
        _record(fh, np.array([len(pole)], dtype='<i4').tobytes())
        _record(fh, pole.tobytes())
        for s in range(len(pole)):
            _record(fh, np.ascontiguousarray(w_s[:, s]).tobytes())


def read_screened_coulomb(path: str):
    """read_spectral_function (src/m_spectral_function.f90:508-560, the
    serial branch). Returns (postscf_name, pole[npole], w_s[nprodbasis, npole])."""
    recs = _records(path)
    name = next(recs).decode('ascii').rstrip()
    nprod = int(np.frombuffer(next(recs), dtype='<i4')[0])
    npole = int(np.frombuffer(next(recs), dtype='<i4')[0])
    pole = np.frombuffer(next(recs), dtype='<f8').copy()
    if len(pole) != npole:
        raise ValueError(f'{path}: {len(pole)} poles in a file that announces {npole}')
    w_s = np.zeros((nprod, npole), order='F')
    for s in range(npole):
        col = np.frombuffer(next(recs), dtype='<f8')
        if len(col) != nprod:
            raise ValueError(f'{path}: residue {s + 1} has {len(col)} values, expected {nprod}')
        w_s[:, s] = col
    return name, pole, w_s


# ---------------------------------------------------------------- GW ingredients
def _ld(x) -> str:
    """One list-directed value per line (the layout of list-directed output
    is the compiler's; these values read back exactly)."""
    if isinstance(x, (int, np.integer)):
        return f'{int(x):12d}\n'
    return f' {float(x): .17E}\n'


def dump_gw_ingredients(directory: str, energy_g, pole, w_spin, v_spin) -> None:
    """dump_gw_ingredients (src/m_gw_selfenergy_analytic.f90:801-881):
    energies.dat (ng, then energy(p, 1) for the G states), omegas.dat (npole,
    then the poles), w.bin and v.bin (stream files: for every spin channel and
    every q, w(:, :, q) = (npole, ng) and v(:, :, q) = (nauxil, ng) in
    column-major order) and v.dat (nauxil, ng, ng). ``w_spin``/``v_spin`` are
    lists over the spin channels of the arrays of Engine.gw_ingredients;
    ``energy_g`` the spin-1 energies of the G states."""
    os.makedirs(directory, exist_ok=True)
    e = np.asarray(energy_g, dtype=np.float64).reshape(-1)
    pole = np.asarray(pole, dtype=np.float64).reshape(-1)
    naux, ng = v_spin[0].shape[0], v_spin[0].shape[1]
    with open(os.path.join(directory, 'energies.dat'), 'w') as fh:
        fh.write(_ld(len(e)) + ''.join(_ld(x) for x in e))
    with open(os.path.join(directory, 'omegas.dat'), 'w') as fh:
        fh.write(_ld(len(pole)) + ''.join(_ld(x) for x in pole))
    with open(os.path.join(directory, 'w.bin'), 'wb') as fh:
        for w in w_spin:
            fh.write(np.asarray(w, dtype='<f8').tobytes(order='F'))
    with open(os.path.join(directory, 'v.dat'), 'w') as fh:
        fh.write(_ld(naux) + _ld(ng) + _ld(ng))
    with open(os.path.join(directory, 'v.bin'), 'wb') as fh:
        for v in v_spin:
            fh.write(np.asarray(v, dtype='<f8').tobytes(order='F'))
