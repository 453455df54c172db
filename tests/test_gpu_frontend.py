"""A real molecule end to end on the GPU: the native front end's integrals
and RI-HF orbitals of LiH (cc-pVQZ / cc-pVQZ-RI, Cartesian, 3 bohr), the RI
tensor built on the device from the 2-center metric and the pair tensor
(mgw_inverse_sqrt_eri_2center, mgw_build_eri_3center), then mgw_g0w0.

The one-shot G0W0 HOMO-LUMO gap is the reference's lih_gw.in golden,
8.053158 eV (tolerance 1e-4 eV, as in its testsuite); the one-shot, the
imaginary-axis + Pade and the contour-deformation QP energies match the
oracle on the same inputs to 1e-8 Ha.
"""
import os

import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd.frontend import Molecule, frozen_core_states, run_rihf
from molgw_amd.gw import selfenergy_set_state_range

pytestmark = pytest.mark.gpu

BASIS = os.path.join(os.path.dirname(__file__), 'golden', 'basis')
HA_EV = 27.21138505


@pytest.fixture(scope='module')
def lih():
    m = Molecule.from_lines(['Li 0.000000 0.0 1.50', 'H 0.000000 0.0 -1.50'], unit='bohr')
    r = run_rihf(m, 'cc-pVQZ', 'cc-pVQZ-RI', BASIS, gaussian_type='CART')
    sq, _ = O.inverse_sqrt_eri_2center(r.eri_2center)
    return r, r.eri_3center_tmp @ sq


def _load(engine, r):
    kept, lowest, _ = engine.inverse_sqrt_eri_2center(r.eri_2center)
    engine.build_eri_3center(r.eri_3center_tmp, r.index_basis, r.nbf)
    engine.set_mo(r.c_matrix, r.energy_mo, r.occupation)
    return kept


def test_lih_g0w0_gap_golden(engine, lih):
    r, b = lih
    print(f'LiH RI-HF {r.energy:.10f} Ha, nbf {r.nbf}', flush=True)
    kept = _load(engine, r)
    assert kept == b.shape[1]
    out = engine.g0w0(technique='one_shot')
    gap = (out.energy_qp[2] - out.energy_qp[1]) * HA_EV
    ref = O.g0w0_casida(b, r.index_basis, r.c_matrix, r.energy_mo, r.occupation)
    print(f'  G0W0 gap {gap:.6f} eV (golden 8.053158); max |dE_QP| vs oracle '
          f'{np.max(np.abs(out.energy_qp - ref.energy_qp)):.2e} Ha')
    assert abs(gap - 8.053158) < 1e-4
    assert np.max(np.abs(out.energy_qp - ref.energy_qp)) < 1e-8


def test_lih_imaginary_and_contour_match_oracle(engine, lih):
    r, b = lih
    _load(engine, r)
    rng = selfenergy_set_state_range(r.occupation, selfenergy_state_range=2)
    orng = O.selfenergy_set_state_range(r.occupation[:, None], selfenergy_state_range=2)
    pade = engine.g0w0(rng, technique='imaginary_pade', nomega_chi_imag=32, nomega_sigma_calc=16)
    ref_p = O.g0w0_pade(b, r.index_basis, r.c_matrix, r.energy_mo, r.occupation, 32, nomega_sigma_calc=16,
                        ranges=orng)
    cont = engine.g0w0(rng, technique='contour', nomega_chi_imag=16, nomega_chi_real=6)
    ref_c = O.g0w0_contour(b, r.index_basis, r.c_matrix, r.energy_mo, r.occupation, 16, 6, ranges=orng)
    dp = np.max(np.abs(pade.energy_qp - ref_p.energy_qp))
    dc = np.max(np.abs(cont.energy_qp - ref_c['energy_qp'][:, 0]))
    print(f'  LiH Pade gap {(pade.energy_qp[2] - pade.energy_qp[1]) * HA_EV:.6f} eV, contour gap '
          f'{(cont.energy_qp[2] - cont.energy_qp[1]) * HA_EV:.6f} eV; vs oracle {dp:.2e} / {dc:.2e} Ha')
    assert dp < 1e-8
    assert dc < 1e-8


@pytest.mark.parametrize('rule,homo', [('molgw', -8.270983), ('gaussian', -8.265769)])
def test_lih_frozen_core_goldens(engine, lih, rule, homo):
    """lih_gw_frozencore_molgw.in / lih_gw_frozencore_gaussian.in: the GW HOMO
    with the frozen-core ranges of each rule, selfenergy_state_range 1."""
    r, b = lih
    _load(engine, r)
    nc = frozen_core_states([3, 1], rule)
    rng = selfenergy_set_state_range(r.occupation, ncoreg=nc, ncorew=nc, selfenergy_state_range=1)
    out = engine.g0w0(rng, technique='one_shot')
    print(f'  frozencore {rule}: GW HOMO {out.energy_qp[1] * HA_EV:.6f} eV (golden {homo})')
    assert abs(out.energy_qp[1] * HA_EV - homo) < 1e-4


def test_water_and_ethane_real_integrals_match_oracle(engine):
    """Pure (spherical) cc-pVDZ / cc-pVDZ-RI on real molecules: water, and
    ethane from the alkane builder. The one-shot and the imaginary-axis QP
    energies of the GPU path equal the oracle's on the same front-end inputs."""
    from molgw_amd.frontend import alkane
    water = Molecule.from_lines(['O 0.0 0.0 0.1173', 'H 0.0 0.7572 -0.4692', 'H 0.0 -0.7572 -0.4692'])
    for name, mol in (('water', water), ('ethane', alkane(2))):
        r = run_rihf(mol, 'cc-pVDZ', 'cc-pVDZ-RI', BASIS)
        sq, _ = O.inverse_sqrt_eri_2center(r.eri_2center)
        b = r.eri_3center_tmp @ sq
        _load(engine, r)
        one = engine.g0w0(technique='one_shot')
        ref1 = O.g0w0_casida(b, r.index_basis, r.c_matrix, r.energy_mo, r.occupation)
        rng = selfenergy_set_state_range(r.occupation, selfenergy_state_range=3)
        orng = O.selfenergy_set_state_range(r.occupation[:, None], selfenergy_state_range=3)
        imag = engine.g0w0(rng, technique='imaginary_pade', nomega_chi_imag=32, nomega_sigma_calc=16)
        ref2 = O.g0w0_pade(b, r.index_basis, r.c_matrix, r.energy_mo, r.occupation, 32, nomega_sigma_calc=16,
                           ranges=orng)
        d1 = np.max(np.abs(one.energy_qp - ref1.energy_qp))
        d2 = np.max(np.abs(imag.energy_qp - ref2.energy_qp))
        nh = int(np.sum(r.occupation > 0))
        print(f'  {name}: RI-HF {r.energy:.8f} Ha, nbf {r.nbf}; G0W0 HOMO {one.energy_qp[nh - 1] * HA_EV:.4f} eV; '
              f'vs oracle {d1:.1e} / {d2:.1e} Ha')
        assert d1 < 1e-8 and d2 < 1e-8
