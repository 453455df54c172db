"""The fused-split chi0 build (the FP32 product and the hi/lo split done in
the SYRK's operand load) against the per-frequency planes it replaces, and
its two K schedules against each other and against FP64.

* With the K split (FP32 partial tiles of 64 slabs summed in FP64) the fused
  kernel does the same arithmetic in the same order as the planes path: the
  result is identical bit for bit.
* Without the K split (FP32 MFMA sums flushed into FP64 every 4 slabs) it
  agrees with the K-split result to 1e-6 relative and with the FP64 build
  to the FP16x3 accuracy (1e-5 relative to the largest element of chi0).
"""
import os

import numpy as np
import pytest

from molgw_amd.gw import Engine, selfenergy_set_state_range, imaginary_quad_grid
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def _chi(eng, s, precision, env):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        rng = selfenergy_set_state_range(s.occupation)
        om, wt = imaginary_quad_grid(12)
        erpa, egw, chi = eng.polarizability_grid(rng, om, wt, return_chi=True, precision=precision)
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return erpa, chi


@pytest.mark.parametrize('precision', ['fp16x3', 'bf16x3'])
def test_fused_split_equals_the_planes(precision):
    s = synthetic_system(60, 300, 12, seed=41, scale=0.3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        e_planes, chi_planes = _chi(eng, s, precision, {'MGW_CHI_FUSED': '0'})
        e_fused, chi_fused = _chi(eng, s, precision, {'MGW_CHI_FUSED': '1', 'MGW_CHI_FUSED_KSPLIT': '1'})
        e_flush, chi_flush = _chi(eng, s, precision, {'MGW_CHI_FUSED': '1', 'MGW_CHI_FUSED_KSPLIT': '0'})
        e64, chi64 = _chi(eng, s, 'fp64', {})
    np.testing.assert_array_equal(chi_fused, chi_planes)
    assert e_fused == e_planes
    scale = np.max(np.abs(chi64))
    print(f'{precision}: flush vs K split {np.max(np.abs(chi_flush - chi_fused)) / scale:.2e}, '
          f'vs FP64 {np.max(np.abs(chi_flush - chi64)) / scale:.2e} (relative to max |chi|)')
    assert np.max(np.abs(chi_flush - chi_fused)) < 1e-6 * scale
    tol = 1e-5 if precision == 'fp16x3' else 1e-3
    assert np.max(np.abs(chi_flush - chi64)) < tol * scale
    assert np.max(np.abs(chi_fused - chi64)) < tol * scale


def test_cholesky_leaf_kernels_agree():
    """The four-wave LDS leaf of the batched inverse Cholesky against the
    one-wave register leaf (same factorisation, other operation order):
    chi and the traces agree to round-off."""
    s = synthetic_system(60, 330, 12, seed=43, scale=0.3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        e_lds, chi_lds = _chi(eng, s, 'fp64', {'MGW_CHOL_LEAF': '1'})
        e_wave, chi_wave = _chi(eng, s, 'fp64', {'MGW_CHOL_LEAF': '0'})
    scale = np.max(np.abs(chi_wave))
    print(f'LDS leaf vs register leaf: chi {np.max(np.abs(chi_lds - chi_wave)) / scale:.2e}, '
          f'E_RPA {abs(e_lds - e_wave):.2e}')
    assert np.max(np.abs(chi_lds - chi_wave)) < 1e-12 * scale
    assert abs(e_lds - e_wave) < 1e-12 * abs(e_wave)
