"""FP16x3 AO->MO transforms (the default of an MGW_FP16X3 solve;
MGW_AO2MO_FP16X3 = 0 keeps them FP64):
both products of the transform split their operands into two fp16 parts of
power-of-two-scaled values (max |B| from the RI unpack, max |C|, max |H|
from the first product) with FP32 accumulation over K = nbf.

* chi0 of the imaginary grid against the FP64 build, relative to max |chi0|
  (chi0 recovered from chi as 1 - (1 + chi)^-1): the transforms add their
  rounding to X on top of the chi0 split's (8e-6, tests/test_gpu_chi_fused.py):
  each transform's product is off by ~3 u^2 (u = 2^-11) plus 3 nbf / 32 FP32
  accumulations of 2^-23 (1e-5 at nbf 64, two transforms, and chi0 is
  quadratic in X): within 5e-5.
* QP energies against FP64 within 1 meV (the metric's condition), on the
  molecular-magnitude synthetic system, the benchmark recipe and water3 /
  alkane5 from the GPU front end.
"""
import os

import numpy as np
import pytest

from molgw_amd.gw import Engine, selfenergy_set_state_range, imaginary_quad_grid
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu

HA_MEV = 27211.38505


class _env:
    def __init__(self, **kw):
        self.kw = kw

    def __enter__(self):
        self.old = {k: os.environ.get(k) for k in self.kw}
        os.environ.update({k: str(v) for k, v in self.kw.items()})

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _chi0(chi):
    n = chi.shape[0]
    eye = np.eye(n)
    return np.stack([eye - np.linalg.inv(eye + chi[:, :, f]) for f in range(chi.shape[2])], axis=2)


def test_chi0_with_fp16x3_transforms_against_fp64():
    s = synthetic_system(64, 190, 9, seed=5, scale=0.3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        rng = selfenergy_set_state_range(s.occupation)
        om, wt = imaginary_quad_grid(12)
        e64, _, chi64 = eng.polarizability_grid(rng, om, wt, return_chi=True)
        with _env(MGW_AO2MO_FP16X3=0):
            e0, _, chi_0 = eng.polarizability_grid(rng, om, wt, return_chi=True, precision='fp16x3')
        with _env(MGW_AO2MO_FP16X3=1):
            e1, _, chi_1 = eng.polarizability_grid(rng, om, wt, return_chi=True, precision='fp16x3')
    c64 = _chi0(chi64)
    scale = np.max(np.abs(c64))
    err0 = np.max(np.abs(_chi0(chi_0) - c64)) / scale
    err1 = np.max(np.abs(_chi0(chi_1) - c64)) / scale
    print(f'chi0 vs FP64 (relative to max |chi0|): FP64 transforms {err0:.2e}, FP16x3 transforms {err1:.2e}; '
          f'E_RPA {e0 - e64:.2e} / {e1 - e64:.2e} Ha')
    assert err1 < 5e-5


@pytest.mark.parametrize('recipe', ['molecular', 'benchmark'])
def test_qp_with_fp16x3_transforms_within_1_mev(recipe):
    if recipe == 'molecular':
        s = synthetic_system(64, 190, 9, seed=5, scale=0.3)
    else:
        s = synthetic_system(48, 140, 8, seed=3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        kw = dict(technique='imaginary_pade', nomega_chi_imag=32, nomega_sigma_calc=16, return_sigma=False)
        ref = eng.g0w0(**kw)
        with _env(MGW_AO2MO_FP16X3=0):
            a = eng.g0w0(precision='fp16x3', **kw)
        with _env(MGW_AO2MO_FP16X3=1):
            b = eng.g0w0(precision='fp16x3', **kw)
    ea = np.max(np.abs(a.energy_qp - ref.energy_qp)) * HA_MEV
    eb = np.max(np.abs(b.energy_qp - ref.energy_qp)) * HA_MEV
    print(f'{recipe}: max |dE_QP| vs FP64: FP64 transforms {ea:.4f} meV, FP16x3 transforms {eb:.4f} meV')
    assert eb <= 1.0


def test_real_molecules_with_fp16x3_transforms():
    from molgw_amd.frontend import alkane, water_cluster
    from molgw_amd.workloads import molecular_system_device
    for s in (molecular_system_device(water_cluster(3, seed=3), 'water3'),
              molecular_system_device(alkane(5), 'alkane5')):
        with Engine(0) as eng:
            eng.set_eri_3center_device(s.eri_3center_dev.data_ptr(), s.npair, s.naux, s.index_basis, s.nbf)
            eng.set_mo(s.c_matrix, s.energy, s.occupation)
            rng = selfenergy_set_state_range(s.occupation)
            rng.nsemin = max(rng.nsemin, rng.nhomo_G - 7)
            rng.nsemax = min(rng.nsemax, rng.nhomo_G + 8)
            kw = dict(technique='imaginary_pade', nomega_chi_imag=32, nomega_sigma_calc=16, return_sigma=False,
                      exchange_m_vxc=s.exchange_m_vxc)
            ref = eng.g0w0(rng, **kw)
            with _env(MGW_AO2MO_FP16X3=1):
                b = eng.g0w0(rng, precision='fp16x3', **kw)
        sl = slice(rng.nsemin - 1, rng.nsemax)
        err = np.max(np.abs(b.energy_qp[sl] - ref.energy_qp[sl])) * HA_MEV
        print(f'{s.name}: FP16x3 transforms and chi0 vs FP64, states {rng.nsemin}..{rng.nsemax}: {err:.4f} meV')
        assert err <= 1.0


def test_plane_transform_equals_split_on_load():
    """The second product of the FP16x3 transform on pre-split planes (H
    split once per matrix, the transform's columns of C once) against
    splitting both operands on every tile load (MGW_AO2MO_PLANES=0): the same
    split of each element, the same products in the same order per
    accumulator, so chi and the QP energies are identical bit for bit
    (naux 500, nbf 160: several row, column and K tiles)."""
    s = synthetic_system(160, 500, 20, seed=9, scale=0.3)
    out = {}
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        rng = selfenergy_set_state_range(s.occupation)
        om, wt = imaginary_quad_grid(8)
        for planes in (1, 0):
            with _env(MGW_AO2MO_PLANES=planes):
                e, _, chi = eng.polarizability_grid(rng, om, wt, return_chi=True, precision='fp16x3')
                qp = eng.g0w0(technique='imaginary_pade', nomega_chi_imag=8, nomega_sigma_calc=8,
                              precision='fp16x3', return_sigma=False).energy_qp
            out[planes] = (e, chi, qp)
    np.testing.assert_array_equal(out[1][1], out[0][1])
    assert out[1][0] == out[0][0]
    np.testing.assert_array_equal(out[1][2], out[0][2])
