"""The integral front end on the GPU: (mu nu|P) and (P|Q) from the device
kernels (mgw_eri_3center_ao, mgw_eri_2center_ao) against the host engine
(ints.cpp, which tests/test_frontend.py pins to the numpy engine and, through
the RI-HF, to the reference's SCF goldens), and the device RI-HF (Hartree and
exchange through the resident RI tensor, mgw_hartree_exchange_ri) against
the reference's SCF goldens (tests/inputs/testsuite of MOLGW):
  ne_rings.in    Ne  HF cc-pVDZ / cc-pVDZ-RI                   -128.4900916481
  co_gw_dm.in    CO  HF cc-pVDZ_gaussian (Cartesian) / cc-pVDZ-RI -112.7480846685
  ch4_gwtdhf.in  CH4 HF aug-cc-pVTZ / aug-cc-pVTZ-RI            -40.2135693443
(tolerance 1e-7 Ha, as in the testsuite).

Tolerances: integrals 1e-12 relative to the largest; device and host RI-HF
energies 1e-9 Ha.
"""
import os

import numpy as np
import pytest

from molgw_amd.frontend import Molecule, run_rihf, run_rihf_device, water_cluster, alkane
from molgw_amd.frontend import native as N
from molgw_amd.frontend.basis import BasisSet
from molgw_amd.gw import Engine
from molgw_amd.workloads import upper_pairs

pytestmark = pytest.mark.gpu

BASIS = os.path.join(os.path.dirname(__file__), 'golden', 'basis')

WATER = ['O 0.0 0.0 0.0', 'H 0.757 0.587 0.0', 'H -0.757 0.587 0.0']
CO = ['8 .000000 .000000 .493003', '6 .000000 .000000 -.657337']
CH4 = ['6 .000000 .000000 .000000', '1 .629118 .629118 .629118', '1 -.629118 -.629118 .629118',
       '1 .629118 -.629118 -.629118', '1 -.629118 .629118 -.629118']


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize('lines,bas,auxb,gt', [
    (WATER, 'cc-pVDZ', 'cc-pVDZ-RI', 'PURE'),
    (CO, 'cc-pVDZ_gaussian', 'cc-pVDZ-RI', 'CART'),
    (CH4, 'aug-cc-pVTZ', 'aug-cc-pVTZ-RI', 'PURE'),
])
def test_integrals_match_the_host_engine(engine, lines, bas, auxb, gt):
    m = Molecule.from_lines(lines)
    bs = BasisSet(m.atoms, m.coords, bas, BASIS, gt)
    ax = BasisSet(m.atoms, m.coords, auxb, BASIS, gt)
    print(f'{bas}/{auxb} {gt}: nbf {bs.nbf}, naux {ax.nbf}, l_max {max(s.l for s in bs.shells)}/'
          f'{max(s.l for s in ax.shells)}', flush=True)
    m2 = N.eri_2center(ax)
    d2 = engine.eri_2center_ao(ax)
    assert rel(d2, m2) < 1e-12
    np.testing.assert_array_equal(d2, d2.T)
    ib = upper_pairs(bs.nbf)
    e3 = N.eri_3center(bs, ax)
    k, l = ib[0] - 1, ib[1] - 1
    ref = e3[k, l, :].copy()
    ref[k == l] *= 0.5
    d3 = engine.eri_3center_ao(bs, ax, ib)
    assert rel(d3, ref) < 1e-12


def test_unsupported_angular_momentum_is_refused(engine):
    from molgw_amd._lib import MgwError
    m = Molecule.from_lines(['Li 0 0 1.5', 'H 0 0 -1.5'], unit='bohr')
    bs = BasisSet(m.atoms, m.coords, 'cc-pVQZ', BASIS, 'CART')      # g functions on Li
    ax = BasisSet(m.atoms, m.coords, 'cc-pVQZ-RI', BASIS, 'CART')
    with pytest.raises(MgwError) as exc:
        engine.eri_3center_ao(bs, ax, upper_pairs(bs.nbf))
    assert exc.value.code == 1


@pytest.mark.parametrize('lines,bas,auxb,gt,golden', [
    (['Ne 0 0 0'], 'cc-pVDZ', 'cc-pVDZ-RI', 'PURE', -128.4900916481),
    (CO, 'cc-pVDZ_gaussian', 'cc-pVDZ-RI', 'CART', -112.7480846685),
    (CH4, 'aug-cc-pVTZ', 'aug-cc-pVTZ-RI', 'PURE', -40.2135693443),
])
def test_device_rihf_goldens(lines, bas, auxb, gt, golden):
    with Engine(0) as eng:
        r = run_rihf_device(Molecule.from_lines(lines), bas, auxb, BASIS, eng, gaussian_type=gt)
    print(f'{bas}: device RI-HF {r.energy:.10f} Ha ({r.iterations} iterations), golden {golden}')
    assert abs(r.energy - golden) < 1e-7


def test_device_rihf_equals_the_host_rihf():
    m = water_cluster(3, seed=5)
    host = run_rihf(m, 'cc-pVDZ', 'cc-pVDZ-RI', BASIS)
    with Engine(0) as eng:
        dev = run_rihf_device(m, 'cc-pVDZ', 'cc-pVDZ-RI', BASIS, eng)
    print(f'(H2O)3: host {host.energy:.12f}, device {dev.energy:.12f} Ha')
    assert abs(dev.energy - host.energy) < 1e-9
    np.testing.assert_allclose(dev.energy_mo, host.energy_mo, atol=1e-8)


def test_hartree_exchange_match_dense_contractions(engine):
    """J and K through the resident RI tensor against dense einsums of the
    same tensor (host), for a random set of orbitals."""
    m = Molecule.from_lines(WATER)
    bs = BasisSet(m.atoms, m.coords, 'cc-pVDZ', BASIS)
    ax = BasisSet(m.atoms, m.coords, 'cc-pVDZ-RI', BASIS)
    m2 = N.eri_2center(ax)
    e3 = N.eri_3center(bs, ax)
    w, u = np.linalg.eigh(m2)
    b = e3 @ (u[:, w > 1e-6] / np.sqrt(w[w > 1e-6]))
    ib = upper_pairs(bs.nbf)
    engine.inverse_sqrt_eri_2center(m2, copy=False)
    engine.eri_3center_ao(bs, ax, ib, copy=False)
    engine.build_eri_3center_resident(ib, bs.nbf)
    x = np.random.default_rng(3).standard_normal((bs.nbf, 5))
    j, k = engine.hartree_exchange_ri(x, spin_fact=2.0)
    d = 2.0 * x @ x.T
    jr = np.einsum('mnP,P->mn', b, np.einsum('klP,kl->P', b, d))
    kr = -np.einsum('mkP,kl,lnP->mn', b, x @ x.T, b)
    assert rel(j, jr) < 1e-12
    assert rel(k, kr) < 1e-12


def test_exchange_m_vxc_and_g0w0_from_a_nonconverged_start():
    """setup_exchange_m_vxc on the GPU against the oracle, for a converged
    RI-HF (zero to the SCF convergence) and a deliberately short SCF (two
    iterations: the Fock matrix of the final density no longer commutes with
    the orbitals), then the one-shot G0W0 with that Sigma_x - Vxc against the
    oracle's (QP energies 1e-8 Ha)."""
    from oracle import gw_oracle as O
    m = Molecule.from_lines(WATER)
    with Engine(0) as eng:
        conv = run_rihf_device(m, 'cc-pVDZ', 'cc-pVDZ-RI', BASIS, eng)
        eng.set_mo(conv.c_matrix, conv.energy_mo, conv.occupation)
        mat, dg = eng.exchange_m_vxc(conv.hamiltonian_fock)
        assert np.abs(mat).max() < 1e-6
        short = run_rihf_device(m, 'cc-pVDZ', 'cc-pVDZ-RI', BASIS, eng, max_iter=2)
        eng.set_mo(short.c_matrix, short.energy_mo, short.occupation)
        mat, dg = eng.exchange_m_vxc(short.hamiltonian_fock)
        ref = O.setup_exchange_m_vxc(short.c_matrix, short.hamiltonian_fock, short.energy_mo)
        print(f'short SCF: max |Sigma_x - Vxc| {np.abs(ref).max():.3e} Ha, GPU error {np.abs(mat - ref).max():.2e}')
        assert np.abs(ref).max() > 1e-4
        assert np.abs(mat - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())
        np.testing.assert_array_equal(dg, np.diag(mat))
        out = eng.g0w0(exchange_m_vxc=dg)
        b = _packed_ri(eng, m)
    res = O.g0w0_casida(b, short.index_basis, short.c_matrix, short.energy_mo, short.occupation,
                        exchange_m_vxc=dg)
    print(f'G0W0 with Sigma_x - Vxc: max |dE_QP| vs oracle {np.max(np.abs(out.energy_qp - res.energy_qp)):.2e} Ha')
    assert np.max(np.abs(out.energy_qp - res.energy_qp)) < 1e-8


def _packed_ri(eng, m):
    """The packed RI tensor eri_3center(npair, naux_kept) of molecule m, built
    again on the device from the device integrals and copied to the host."""
    bs = BasisSet(m.atoms, m.coords, 'cc-pVDZ', BASIS)
    ax = BasisSet(m.atoms, m.coords, 'cc-pVDZ-RI', BASIS)
    eng2 = Engine(0)
    try:
        eng2.inverse_sqrt_eri_2center(eng2.eri_2center_ao(ax), copy=False)
        ib = upper_pairs(bs.nbf)
        eng2.eri_3center_ao(bs, ax, ib, copy=False)
        return eng2.build_eri_3center_resident(ib, bs.nbf, copy=True)
    finally:
        eng2.close()
