"""The fused-split chi0 build (the FP32 product and the hi/lo split done in
the SYRK's operand load) against the per-frequency planes it replaces, and
its two K schedules against each other and against FP64.

* With the K split (FP32 partial tiles of 64 slabs summed in FP64) the fused
  kernel does the same arithmetic in the same order as the planes path: the
  result is identical bit for bit.
* Without the K split (FP32 MFMA sums flushed into FP64 every 4 slabs) the
  products are the same and only the FP32 sums are shorter, so it must be at
  least as accurate against the FP64 build as the planes path (within 1.5x,
  a margin for the statistics of the rounding). chi = (1 - chi0)^-1 chi0
  amplifies chi0's rounding by the condition number of 1 - chi0, so the
  errors are stated on chi0 = 1 - (1 + chi)^-1, recovered on the host in
  FP64. Bounds, relative to max |chi0| (a diagonal element, whose 576
  transition products all have one sign): a two-plane split represents each
  operand to u^2 (u = 2^-11 fp16, 2^-8 bf16), the three products drop lo*lo,
  so a product is off by at most ~3 u^2 (7e-7 fp16, 4.6e-5 bf16); y = x w is
  rounded to FP32 (2e-7); and each MFMA accumulation into an FP32 register
  may round by 2^-23 of the running sum, 3 products x 18 slabs = 54 of them
  before the FP64 sum (6.4e-6): FP16x3 within 8e-6, BF16x3 within 1e-4.
  (Measured on MI355X: 3.2e-6 and 3.9e-6.)
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
    c64, cpl, cfl = _chi0(chi64), _chi0(chi_planes), _chi0(chi_flush)
    scale = np.max(np.abs(c64))
    err_planes = np.max(np.abs(cpl - c64)) / scale
    err_flush = np.max(np.abs(cfl - c64)) / scale
    print(f'{precision}: chi0 vs FP64 (relative to max |chi0|): planes / fused K split {err_planes:.2e}, '
          f'fused FP64 flush {err_flush:.2e}; E_RPA {e_planes - e64:.2e} / {e_flush - e64:.2e} Ha')
    tol = 8e-6 if precision == 'fp16x3' else 1e-4
    assert err_planes < tol and err_flush < tol
    assert err_flush <= 1.5 * err_planes + 1e-15


@pytest.mark.parametrize('precision', ['fp16x3', 'bf16x3'])
@pytest.mark.parametrize('ksplit', ['1', '0'])
def test_one_sided_split_against_fp64(precision, ksplit):
    """The one-sided split (A = X split once, B = X diag(s_f) split in the
    SYRK): other roundings than the planes path (no square root of the
    weight, scales per operand), the same two-plane accuracy: chi0 within the
    bounds above and within 1.5x the planes path's error against FP64."""
    s = synthetic_system(60, 300, 12, seed=41, scale=0.3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        e_planes, chi_planes = _chi(eng, s, precision, {'MGW_CHI_FUSED': '0'})
        e_half, chi_half = _chi(eng, s, precision, {'MGW_CHI_FUSED': '2', 'MGW_CHI_FUSED_KSPLIT': ksplit})
        e64, chi64 = _chi(eng, s, 'fp64', {})
    c64 = _chi0(chi64)
    scale = np.max(np.abs(c64))
    err_planes = np.max(np.abs(_chi0(chi_planes) - c64)) / scale
    err_half = np.max(np.abs(_chi0(chi_half) - c64)) / scale
    print(f'{precision} K split {ksplit}: chi0 vs FP64 (relative to max |chi0|): planes {err_planes:.2e}, '
          f'one-sided {err_half:.2e}; E_RPA {e_planes - e64:.2e} / {e_half - e64:.2e} Ha')
    tol = 8e-6 if precision == 'fp16x3' else 1e-4
    assert err_half < tol
    assert err_half <= 1.5 * err_planes + 1e-15


def _chi0(chi):
    """chi0 = 1 - (1 + chi)^-1 for every frequency (chi is naux x naux x nomega)."""
    n = chi.shape[0]
    eye = np.eye(n)
    return np.stack([eye - np.linalg.inv(eye + chi[:, :, f]) for f in range(chi.shape[2])], axis=2)


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


@pytest.mark.parametrize('naux,f16_min', [(330, '64'), (600, '128')])
def test_cholesky_epilogue_planes_equal_split_planes(naux, f16_min):
    """The FP16x3 Cholesky products that write L21 and P straight into the
    hi/lo planes of their later uses (the same split applied to the FP64
    value in the epilogue) against the products that write FP64 and split
    it afterwards: chi identical bit for bit (64-wide products on naux 330;
    128-wide ones at the 600 level's split)."""
    s = synthetic_system(60, naux, 12, seed=47, scale=1.0)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        e0, chi0 = _chi(eng, s, 'fp16x3', {'MGW_CHOL_F16_MIN': f16_min, 'MGW_CHOL_EPI_PLANES': '0'})
        e1, chi1 = _chi(eng, s, 'fp16x3', {'MGW_CHOL_F16_MIN': f16_min, 'MGW_CHOL_EPI_PLANES': '1'})
    np.testing.assert_array_equal(chi1, chi0)
    assert e1 == e0


@pytest.mark.parametrize('variant', ['1', '2', '3', '4', '5', '6', '7', '8'])
def test_syrk_lds_copies_equal_the_register_prefetch(variant):
    """The SYRK variants that copy slabs straight into LDS (one or two
    buffers; 128-wide tiles on 4 waves of 64 x 64 or 2 waves of 64 x 128,
    or 256-wide tiles on 8 waves of 64 x 128; three or four buffers with
    two or three slabs in flight) do the same products in the same order
    per accumulator, over the same K chunks, as the register-prefetch
    kernel: chi is identical bit for bit, FP16x3 and BF16x3 (naux 300: one
    full and two partial 256-wide tiles; one K chunk of 18 slabs, so the
    pipelines' tails are exercised)."""
    s = synthetic_system(60, 300, 12, seed=41, scale=0.3)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        for precision in ('fp16x3', 'bf16x3'):
            e0, chi0 = _chi(eng, s, precision, {'MGW_CHI_SYRK': '0'})
            e1, chi1 = _chi(eng, s, precision, {'MGW_CHI_SYRK': variant})
            np.testing.assert_array_equal(chi1, chi0)
            assert e1 == e0


@pytest.mark.parametrize('scale', [0.3, 1.0])
def test_fp16x3_cholesky_products_against_fp64(scale):
    """The batched inverse Cholesky of an FP16x3 solve with its large products
    in FP16x3 (operand bounds from max diag(1 - chi0): |L^-1| <= 1,
    |L21| <= sqrt(d), |A21| <= d), forced down to 64-wide products, against
    the same factorisation in FP64 on the same FP16x3 chi0 (bitwise the same
    input): chi0 = 1 - (1 + chi)^-1 recovered on the host, relative to
    max |chi0|, within the FP16x3 bound of the chi0 build itself (8e-6), so
    the factorisation adds no more than the build's own rounding; E_RPA
    within 1e-6 relative. Molecular-magnitude couplings (scale 0.3) and the
    benchmark recipe's strong coupling (scale 1.0)."""
    s = synthetic_system(60, 330, 12, seed=43, scale=scale)
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        e_lo, chi_lo = _chi(eng, s, 'fp16x3', {'MGW_CHOL_F16_MIN': '64'})
        e_hi, chi_hi = _chi(eng, s, 'fp16x3', {'MGW_CHOL_F16_MIN': '0'})
        e64, chi64 = _chi(eng, s, 'fp64', {})
    c_lo, c_hi, c64 = _chi0(chi_lo), _chi0(chi_hi), _chi0(chi64)
    scale = np.max(np.abs(c_hi))
    d_chol = np.max(np.abs(c_lo - c_hi)) / scale
    d_lo64 = np.max(np.abs(c_lo - c64)) / scale
    d_hi64 = np.max(np.abs(c_hi - c64)) / scale
    print(f'FP16x3 Cholesky products: chi0 vs FP64 factorisation {d_chol:.2e} (relative to max |chi0|); '
          f'against the FP64 solve {d_lo64:.2e} (FP64 factorisation {d_hi64:.2e}); '
          f'E_RPA {e_lo - e_hi:.2e} Ha (E_RPA {e_hi:.6f})')
    assert d_chol < 8e-6
    assert abs(e_lo - e_hi) < 1e-6 * abs(e_hi)
