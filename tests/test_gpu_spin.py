"""Spin channels (nspin = 2) through the C-ABI against the oracle.

W sums the transitions of both channels; the self-energy and QP equation
run per channel (mgw_set_mo_spin, mgw_set_spin, mgw_g0w0's channel loop).
Tolerances: QP energies 1e-8 Ha, E_RPA / E_GW 1e-10 Ha, sigma 1e-10 relative.
"""
import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd.gw import selfenergy_set_state_range
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def polarized(nbf=36, naux=100, nocc=6, seed=51):
    """Two channels with different orbitals and energies and one electron
    fewer in channel 2 (an open-shell doublet)."""
    s = synthetic_system(nbf, naux, nocc, seed=seed, scale=0.1)
    rng = np.random.default_rng(seed + 1)
    rot = np.linalg.qr(np.eye(nbf) + 0.05 * rng.standard_normal((nbf, nbf)))[0]
    c2 = np.stack([s.c_matrix, s.c_matrix @ rot], axis=2)
    e2 = np.stack([s.energy, np.sort(s.energy + 0.03 * rng.standard_normal(nbf))], axis=1)
    o2 = np.stack([s.occupation / 2.0, s.occupation / 2.0], axis=1)
    o2[nocc - 1, 1] = 0.0
    return s, c2, e2, o2


@pytest.mark.parametrize('technique', ['one_shot', 'imaginary_pade'])
def test_spin_polarized_matches_oracle(engine, technique):
    s, c2, e2, o2 = polarized()
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(c2, e2, o2)
    rng = selfenergy_set_state_range(o2)
    exc = np.stack([np.linspace(-0.01, 0.01, s.nbf), np.linspace(0.01, -0.01, s.nbf)], axis=1)
    kw = dict(nomega_sigma=7, exchange_m_vxc=exc)
    if technique == 'imaginary_pade':
        kw.update(nomega_chi_imag=16, nomega_sigma_calc=8)
    out = engine.g0w0(rng, technique=technique, spin_fact=1.0, **kw)
    ref = O.g0w0_spin(s.eri_3center, s.index_basis, c2, e2, o2, technique=technique, spin_fact=1.0,
                      ranges=O.selfenergy_set_state_range(o2), **kw)
    print(f'{technique}: max |dE_QP| {np.max(np.abs(out.energy_qp - ref["energy_qp"])):.2e} Ha, '
          f'E_RPA {out.en_rpa:.10f} vs {ref["en_rpa"]:.10f}')
    assert out.energy_qp.shape == (s.nbf, 2)
    assert np.max(np.abs(out.energy_qp - ref['energy_qp'])) < 1e-8
    assert np.max(np.abs(out.zz - ref['zz'])) < 1e-7
    assert rel_err(out.sigma, ref['sigma']) < 1e-9
    assert abs(out.en_rpa - ref['en_rpa']) < 1e-10
    assert abs(out.en_gw - ref['en_gw']) < 1e-9
    # the two channels really differ
    assert np.max(np.abs(out.energy_qp[:, 0] - out.energy_qp[:, 1])) > 1e-3


@pytest.mark.parametrize('technique', ['one_shot', 'imaginary_pade'])
def test_closed_shell_as_two_channels_equals_one_channel(engine, technique):
    s = synthetic_system(30, 90, 5, seed=53, scale=0.1)
    kw = dict(nomega_sigma=7)
    if technique == 'imaginary_pade':
        kw.update(nomega_chi_imag=12, nomega_sigma_calc=6)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    one = engine.g0w0(technique=technique, **kw)
    c2 = np.stack([s.c_matrix, s.c_matrix], axis=2)
    engine.set_mo(c2, np.stack([s.energy] * 2, axis=1), np.stack([s.occupation / 2.0] * 2, axis=1))
    two = engine.g0w0(technique=technique, spin_fact=1.0, **kw)
    for sp in range(2):
        assert np.max(np.abs(two.energy_qp[:, sp] - one.energy_qp)) < 1e-10
    assert abs(two.en_rpa - one.en_rpa) < 1e-10
    assert abs(two.en_gw - one.en_gw) < 1e-10


def test_set_spin_selects_the_channel_of_the_transform(engine):
    s, c2, e2, o2 = polarized(nbf=24, naux=60, nocc=4, seed=57)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(c2, e2, o2)
    for sp in (1, 2):
        engine.set_spin(sp)
        engine.destroy_eri_3center_mo()
        got = engine.calculate_eri_3center_mo(1, 6, 2, 9)
        ref = O.calculate_eri_3center_mo(s.eri_3center, s.index_basis, c2, 1, 6, 2, 9, spin=sp - 1)
        assert rel_err(got, ref) < 1e-12
    engine.set_spin(1)
