"""Contour deformation self-energy (gw_selfenergy_contour) through the C-ABI
against the oracle.

W in the low-rank form of sf_vsqrt_chi_vsqrt_rpa(low_rank) on the imaginary
and the real grid (batched MFMA chi0, rocSOLVER dsyevd, eigenpairs with
|lambda/(1-lambda)| <= 1e-3 dropped), then the imaginary-axis integral and
the residues at the nearest real-grid point. Tolerances: sigma 1e-9
relative, QP energies 1e-8 Ha, Z 1e-7, de_max 1e-14 Ha.
"""
import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd._lib import MgwError
from molgw_amd.gw import selfenergy_set_state_range
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize('shape', [(36, 100, 6, 6), (90, 300, 10, 20)])
def test_contour_selfenergy_matches_oracle(engine, shape):
    nbf, naux, nocc, nreal = shape
    s = synthetic_system(nbf, naux, nocc, seed=61, scale=0.1)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    rng = selfenergy_set_state_range(s.occupation, selfenergy_state_range=4)
    omega = 0.01 * np.arange(-3, 4)
    e0 = s.energy[rng.nsemin - 1:rng.nsemax]
    print(f'contour nbf {nbf} naux {naux} nocc {nocc} nomega_chi_real {nreal}', flush=True)
    sig, de_max = engine.gw_selfenergy_contour(rng, omega, e0, 16, nreal)
    ref = O.g0w0_contour(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation, 16, nreal,
                         nomega_sigma=7, ranges=O.selfenergy_set_state_range(s.occupation[:, None],
                                                                             selfenergy_state_range=4))
    print(f'  de_max {de_max:.12f} vs {ref["de_max"]:.12f}; sigma rel err {rel_err(sig, ref["sigma"][:, :, 0]):.2e}')
    assert abs(de_max - ref['de_max']) < 1e-14
    assert rel_err(sig, ref['sigma'][:, :, 0]) < 1e-9
    assert np.all(sig.imag == 0.0)


def test_contour_g0w0_matches_oracle(engine):
    s = synthetic_system(40, 110, 6, seed=63, scale=0.1)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    exc = np.linspace(-0.02, 0.02, s.nbf)
    kw = dict(nomega_sigma=9, step_sigma=0.02, exchange_m_vxc=exc)
    out = engine.g0w0(technique='contour', nomega_chi_imag=12, nomega_chi_real=8, **kw)
    ref = O.g0w0_contour(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation, 12, 8, **kw)
    print(f'contour G0W0: max |dE_QP| {np.max(np.abs(out.energy_qp - ref["energy_qp"][:, 0])):.2e} Ha')
    assert np.max(np.abs(out.energy_qp - ref['energy_qp'][:, 0])) < 1e-8
    assert np.max(np.abs(out.zz - ref['zz'][:, 0])) < 1e-7
    assert np.max(np.abs(out.energy_qp_lin - ref['energy_qp_lin'][:, 0])) < 1e-8
    assert rel_err(out.sigma, ref['sigma'][:, :, 0]) < 1e-9
    assert out.en_rpa == 0.0 and out.en_gw == 0.0
    # the correction is real
    assert np.max(np.abs(out.energy_qp - s.energy)) > 1e-3


def _two_channels(seed, open_shell):
    s = synthetic_system(30, 90, 5, seed=seed, scale=0.1)
    r = np.random.default_rng(seed + 1)
    rot = np.linalg.qr(np.eye(s.nbf) + 0.05 * r.standard_normal((s.nbf, s.nbf)))[0]
    c2 = np.stack([s.c_matrix, s.c_matrix @ rot], axis=2)
    e2 = np.stack([s.energy, np.sort(s.energy + 0.03 * r.standard_normal(s.nbf))], axis=1)
    o2 = np.stack([s.occupation / 2.0, s.occupation / 2.0], axis=1)
    if open_shell:
        o2[4, 1] = 0.0
    return s, c2, e2, o2


def test_contour_spin_polarized_matches_oracle(engine):
    """Two channels with different orbitals and energies."""
    s, c2, e2, o2 = _two_channels(65, False)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(c2, e2, o2)
    rng = selfenergy_set_state_range(o2, selfenergy_state_range=3)
    kw = dict(nomega_sigma=7)
    out = engine.g0w0(rng, technique='contour', nomega_chi_imag=10, nomega_chi_real=6, spin_fact=1.0, **kw)
    ref = O.g0w0_contour(s.eri_3center, s.index_basis, c2, e2, o2, 10, 6, spin_fact=1.0,
                         ranges=O.selfenergy_set_state_range(o2, selfenergy_state_range=3), **kw)
    print(f'contour spin: max |dE_QP| {np.max(np.abs(out.energy_qp - ref["energy_qp"])):.2e} Ha')
    assert out.energy_qp.shape == (s.nbf, 2)
    assert np.max(np.abs(out.energy_qp - ref['energy_qp'])) < 1e-8
    assert rel_err(out.sigma, ref['sigma']) < 1e-9
    assert np.max(np.abs(out.energy_qp[:, 0] - out.energy_qp[:, 1])) > 1e-3
    # the standalone entry point gives every channel's sigma
    e0 = e2[rng.nsemin - 1:rng.nsemax, :]
    sig, de_max = engine.gw_selfenergy_contour(rng, 0.01 * np.arange(-3, 4), e0, 10, 6, spin_fact=1.0)
    assert abs(de_max - ref['de_max']) < 1e-14
    assert rel_err(sig, ref['sigma']) < 1e-9


def test_contour_open_shell_follows_the_reference(engine):
    """A state occupied in one channel and empty in the other lies between
    nhomo_G and nlumo_G: the reference's de_max scan skips it while both
    residue branches take it, so a residue can fall outside the real grid,
    where the reference stops. The GPU path stops where the oracle does and
    agrees with it otherwise."""
    s, c2, e2, o2 = _two_channels(69, True)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(c2, e2, o2)
    rng = selfenergy_set_state_range(o2, selfenergy_state_range=3)
    assert rng.nlumo_G == rng.nhomo_G + 2
    kw = dict(nomega_sigma=7)
    try:
        ref = O.g0w0_contour(s.eri_3center, s.index_basis, c2, e2, o2, 10, 6, spin_fact=1.0,
                             ranges=O.selfenergy_set_state_range(o2, selfenergy_state_range=3), **kw)
    except ValueError as exc:
        assert 'out of range' in str(exc)
        with pytest.raises(MgwError, match='outside the real W grid'):
            engine.g0w0(rng, technique='contour', nomega_chi_imag=10, nomega_chi_real=6, spin_fact=1.0, **kw)
        print('contour open shell: both stop (residue beyond the real grid)')
        return
    out = engine.g0w0(rng, technique='contour', nomega_chi_imag=10, nomega_chi_real=6, spin_fact=1.0, **kw)
    assert np.max(np.abs(out.energy_qp - ref['energy_qp'])) < 1e-8


def test_contour_rejects_bad_grids_and_reduced_precision(engine):
    s = synthetic_system(20, 60, 4, seed=67, scale=0.1)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    with pytest.raises(MgwError, match='nomega_chi_real >= 2'):
        engine.g0w0(technique='contour', nomega_chi_imag=8, nomega_chi_real=1, nomega_sigma=5)
    with pytest.raises(MgwError, match='FP64'):
        engine.g0w0(technique='contour', nomega_chi_imag=8, nomega_chi_real=4, nomega_sigma=5,
                    precision='fp16x3')
    # the context still works after the errors
    out = engine.g0w0(technique='contour', nomega_chi_imag=8, nomega_chi_real=4, nomega_sigma=5)
    assert np.all(np.isfinite(out.energy_qp))
