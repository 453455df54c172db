"""Native integral + RI-HF front end against the reference's own test goldens
(tests/inputs/testsuite of MOLGW; basis files from its basis/ directory,
kept under tests/golden/basis) and against closed forms.

Goldens (Ha unless noted, tolerance as in the testsuite):
  ne_rings.in     Ne, HF cc-pVDZ / cc-pVDZ-RI: SCF -128.4900916481 (1e-7),
                  HOMO -22.639065 eV, LUMO 46.110304 eV (1e-4)
  co_gw_dm.in     CO, HF cc-pVDZ_gaussian (Cartesian) / cc-pVDZ-RI: SCF -112.7480846685
  lih_bse-i.in    LiH (3 bohr), HF cc-pVQZ (Cartesian) / cc-pVQZ-RI: SCF -7.9872035011
  ch4_gwtdhf.in   CH4, HF aug-cc-pVTZ (pure, diffuse) / aug-cc-pVTZ-RI: SCF -40.2135693443
  lih_gw.in       the same LiH, one-shot G0W0@HF: HOMO-LUMO gap 8.053158 eV (1e-4),
                  checked through the oracle's G0W0 on the front end's integrals
  lih_gw_frozencore_molgw.in / _gaussian.in  the same with frozencore 'molgw'
                  (no Li core frozen) / 'gaussian' (Li 1s frozen in G and W),
                  selfenergy_state_range 1: GW HOMO -8.270983 / -8.265769 eV
"""
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from molgw_amd.frontend import Molecule, frozen_core_states, run_rihf
from molgw_amd.frontend import integrals as I
from molgw_amd.frontend.basis import BasisSet, Shell

BASIS = os.path.join(os.path.dirname(__file__), 'golden', 'basis')
HA_EV = 27.21138505


@pytest.fixture(scope='module')
def lih_qz():
    m = Molecule.from_lines(['Li 0.000000 0.0 1.50', 'H 0.000000 0.0 -1.50'], unit='bohr')
    return run_rihf(m, 'cc-pVQZ', 'cc-pVQZ-RI', BASIS, gaussian_type='CART')


def test_boys_function_matches_quadrature():
    t = np.array([0.0, 1e-6, 0.3, 0.999, 1.0, 2.5, 11.0, 40.0])
    f = I.boys(12, t)
    for n in (0, 3, 12):
        for k, tv in enumerate(t):
            ref = quad(lambda u: u ** (2 * n) * math.exp(-tv * u * u), 0.0, 1.0, epsabs=1e-15, epsrel=1e-13)[0]
            assert abs(f[n, k] - ref) < 1e-13 * max(1.0, ref)


def test_s_gaussian_closed_forms():
    """One normalised s Gaussian (exponent a) on a nucleus Z at its centre:
    <T> = 3a/2, <V> = -2Z sqrt(2a/pi); two such s functions at distance R:
    (s s| s) and (s|s) from the error-function closed forms."""
    a, z = 0.8, 3.0
    sh = Shell(0, np.zeros(3), np.array([a]), np.array([(2 * a / math.pi) ** 0.75]), 0)
    sh.transform = np.eye(1)
    assert abs(I.shell_overlap_cart(sh, sh)[0, 0] - 1.0) < 1e-14
    assert abs(I.shell_kinetic_cart(sh, sh)[0, 0] - 1.5 * a) < 1e-13
    assert abs(I.shell_nuclear_cart(sh, sh, [z], [np.zeros(3)])[0, 0] + 2 * z * math.sqrt(2 * a / math.pi)) < 1e-13
    b, r = 1.3, np.array([0.0, 0.4, 1.1])
    sb = Shell(0, r, np.array([b]), np.array([(2 * b / math.pi) ** 0.75]), 1)
    # (a|b) for normalised s Gaussians: (2/sqrt(pi))^... = erf(sqrt(mu) R) / R times the charge norms
    qa, qb = (math.pi / a) ** 1.5 * (2 * a / math.pi) ** 0.75, (math.pi / b) ** 1.5 * (2 * b / math.pi) ** 0.75
    mu = a * b / (a + b)
    rr = np.linalg.norm(r)
    ref = qa * qb * math.erf(math.sqrt(mu) * rr) / rr
    assert abs(I.shell_eri2_cart(sh, sb)[0, 0] - ref) < 1e-13
    # (ss|b) with both bra functions on the origin: a charge of exponent 2a at 0
    pa = 2 * a
    qaa = (math.pi / pa) ** 1.5 * (2 * a / math.pi) ** 1.5
    mu2 = pa * b / (pa + b)
    ref3 = qaa * qb * math.erf(math.sqrt(mu2) * rr) / rr
    assert abs(I.shell_eri3_group(sh, sh, [sb])[0][0, 0, 0] - ref3) < 1e-13


def test_normalised_functions_and_symmetric_integrals():
    m = Molecule.from_lines(['O 0.0 0.0 0.1', 'H 0.0 0.76 -0.5', 'H 0.0 -0.76 -0.5'])
    for gt in ('PURE', 'CART'):
        bs = BasisSet(m.atoms, m.coords, 'cc-pVDZ', BASIS, gt)
        s, t, v = I.one_electron(bs, [8.0, 1.0, 1.0], m.coords)
        np.testing.assert_allclose(np.diag(s), 1.0, atol=1e-13)
        assert np.linalg.eigvalsh(t).min() > 0
        assert bs.nbf == (24 if gt == 'PURE' else 25)
    aux = BasisSet(m.atoms, m.coords, 'cc-pVDZ-RI', BASIS)
    m2 = I.eri_2center(aux)
    np.testing.assert_allclose(m2, m2.T, atol=0)
    assert np.linalg.eigvalsh(m2).min() > 0


def test_rihf_neon_golden():
    r = run_rihf(Molecule.from_lines(['Ne 0 0 0']), 'cc-pVDZ', 'cc-pVDZ-RI', BASIS)
    assert abs(r.energy - (-128.4900916481)) < 1e-7
    assert abs(r.energy_mo[4] * HA_EV - (-22.639065)) < 1e-4
    assert abs(r.energy_mo[5] * HA_EV - 46.110304) < 1e-4


def test_rihf_co_cartesian_golden():
    m = Molecule.from_lines(['8 .000000 .000000 .493003', '6 .000000 .000000 -.657337'])
    r = run_rihf(m, 'cc-pVDZ_gaussian', 'cc-pVDZ-RI', BASIS, gaussian_type='CART')
    assert r.nbf == 30
    assert abs(r.energy - (-112.7480846685)) < 1e-7


def test_rihf_methane_augmented_triple_zeta_golden():
    m = Molecule.from_lines(['6 .000000 .000000 .000000', '1 .629118 .629118 .629118', '1 -.629118 -.629118 .629118',
                             '1 .629118 -.629118 -.629118', '1 -.629118 .629118 -.629118'])
    r = run_rihf(m, 'aug-cc-pVTZ', 'aug-cc-pVTZ-RI', BASIS)
    assert r.nbf == 138
    assert abs(r.energy - (-40.2135693443)) < 1e-7


def test_rihf_lih_quadruple_zeta_golden(lih_qz):
    assert lih_qz.nbf == 105
    assert abs(lih_qz.energy - (-7.9872035011)) < 1e-7


def test_g0w0_gap_of_lih_through_the_oracle(lih_qz):
    """The oracle's one-shot G0W0 (postscf='GW' defaults: 51 points, 0.01 Ha,
    eta 1e-3, all states) on the front end's LiH integrals gives the
    reference's HOMO-LUMO gap: both the front end and the oracle's G0W0
    restatement are pinned by it."""
    from oracle import gw_oracle as O
    sq, w = O.inverse_sqrt_eri_2center(lih_qz.eri_2center)
    assert sq.shape[1] < sq.shape[0]                   # Cartesian RI: some metric eigenvalues dropped
    b = lih_qz.eri_3center_tmp @ sq
    res = O.g0w0_casida(b, lih_qz.index_basis, lih_qz.c_matrix, lih_qz.energy_mo, lih_qz.occupation)
    gap = (res.energy_qp[2] - res.energy_qp[1]) * HA_EV
    assert abs(gap - 8.053158) < 1e-4


@pytest.mark.parametrize('rule,homo', [('molgw', -8.270983), ('gaussian', -8.265769)])
def test_g0w0_frozen_core_goldens_of_lih_through_the_oracle(lih_qz, rule, homo):
    from oracle import gw_oracle as O
    nc = frozen_core_states([3, 1], rule)
    assert nc == (0 if rule == 'molgw' else 1)
    sq, _ = O.inverse_sqrt_eri_2center(lih_qz.eri_2center)
    b = lih_qz.eri_3center_tmp @ sq
    rng = O.selfenergy_set_state_range(lih_qz.occupation[:, None], ncoreg=nc, ncorew=nc, selfenergy_state_range=1)
    res = O.g0w0_casida(b, lih_qz.index_basis, lih_qz.c_matrix, lih_qz.energy_mo, lih_qz.occupation, ranges=rng)
    assert abs(res.energy_qp[1] * HA_EV - homo) < 1e-4


def test_frozen_core_rules():
    assert frozen_core_states([8, 1, 1], 'molgw') == 1
    assert frozen_core_states([8, 1, 1], 'gaussian') == 1
    assert frozen_core_states([3, 1], 'molgw') == 0 and frozen_core_states([3, 1], 'gaussian') == 1
    assert frozen_core_states([14, 6], 'molgw') == 6 and frozen_core_states([14, 6], 'no') == 0


def test_native_engine_equals_the_numpy_engine():
    """The native integral engine (ints.cpp) against the numpy one that the
    goldens above pin: Cartesian and pure, up to h functions (LiH cc-pVQZ-RI)."""
    from molgw_amd.frontend import native as N
    cases = [(['8 .0 .0 .493003', '6 .0 .0 -.657337'], 'angstrom', 'cc-pVDZ_gaussian', 'cc-pVDZ-RI', 'CART'),
             (['O 0.0 0.0 0.1', 'H 0.0 0.76 -0.5', 'H 0.0 -0.76 -0.5'], 'angstrom', 'cc-pVDZ', 'cc-pVDZ-RI', 'PURE'),
             (['Li 0 0 1.5', 'H 0 0 -1.5'], 'bohr', 'cc-pVQZ', 'cc-pVQZ-RI', 'CART')]
    for lines, unit, bas, auxb, gt in cases:
        m = Molecule.from_lines(lines, unit=unit)
        bs = BasisSet(m.atoms, m.coords, bas, BASIS, gt)
        ax = BasisSet(m.atoms, m.coords, auxb, BASIS, gt)
        q = [float(z) for z in m.atoms]
        for x, y in zip(N.one_electron(bs, q, m.coords), I.one_electron(bs, q, m.coords)):
            assert np.abs(x - y).max() < 1e-13 * np.abs(y).max()
        e3 = I.eri_3center(bs, ax)
        assert np.abs(N.eri_3center(bs, ax) - e3).max() < 1e-13 * np.abs(e3).max()
        m2 = I.eri_2center(ax)
        assert np.abs(N.eri_2center(ax) - m2).max() < 1e-13 * np.abs(m2).max()


def test_rihf_engines_agree():
    r1 = run_rihf(Molecule.from_lines(['Ne 0 0 0']), 'cc-pVDZ', 'cc-pVDZ-RI', BASIS, engine='numpy')
    r2 = run_rihf(Molecule.from_lines(['Ne 0 0 0']), 'cc-pVDZ', 'cc-pVDZ-RI', BASIS, engine='native')
    assert abs(r1.energy - r2.energy) < 1e-11
