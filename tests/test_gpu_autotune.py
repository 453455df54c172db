"""GEMM tile autotuner: tuned products give the same results (the K order of
each element's sum does not depend on the tile shape), and the table is
persisted to MGW_TUNE_FILE."""
import os
import subprocess
import sys

import numpy as np
import pytest

from molgw_amd.gw import autotune
from molgw_amd.workloads import synthetic_system
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def test_tuned_results_are_identical(engine):
    s = synthetic_system(96, 260, 12, seed=41, scale=0.3)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    ref = engine.g0w0()
    try:
        autotune(True)
        tuned = engine.g0w0()
        again = engine.g0w0()
    finally:
        autotune(False)
    np.testing.assert_array_equal(tuned.energy_qp, ref.energy_qp)
    np.testing.assert_array_equal(again.sigma, ref.sigma)


def test_tuned_split_kernels_are_identical(engine):
    """The tuner's choices for the split-operand kernels (the K-split chi0
    SYRK variants, the W.B dot pipelines) do the same products in the same
    order: an FP16x3 imaginary-axis G0W0 tuned from an empty table, and
    again from the table it filled, equals the untuned one bit for bit."""
    s = synthetic_system(96, 260, 12, seed=41, scale=0.3)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    kw = dict(technique='imaginary_pade', nomega_chi_imag=12, nomega_sigma_calc=8, precision='fp16x3')
    ref = engine.g0w0(**kw)
    try:
        autotune(True)
        tuned = engine.g0w0(**kw)
        again = engine.g0w0(**kw)
    finally:
        autotune(False)
    np.testing.assert_array_equal(tuned.energy_qp, ref.energy_qp)
    np.testing.assert_array_equal(again.energy_qp, ref.energy_qp)
    np.testing.assert_array_equal(again.sigma, ref.sigma)


def test_table_is_persisted(tmp_path):
    table = tmp_path / 'tune.txt'
    code = ('import sys; sys.path.insert(0, %r)\n'
            'from molgw_amd.gw import Engine, autotune\n'
            'from molgw_amd.workloads import synthetic_system\n'
            's = synthetic_system(60, 170, 8, seed=3, scale=0.3)\n'
            'autotune(True)\n'
            'with Engine(0) as e:\n'
            '    e.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)\n'
            '    e.set_mo(s.c_matrix, s.energy, s.occupation)\n'
            '    e.g0w0()\n'
            '    e.g0w0(technique="imaginary_pade", nomega_chi_imag=8, nomega_sigma_calc=4, precision="fp16x3")\n'
            ) % ROOT
    env = dict(os.environ, MGW_TUNE_FILE=str(table))
    proc = subprocess.run([sys.executable, '-c', code], env=env, capture_output=True, text=True, timeout=240)
    assert proc.returncode == 0, proc.stderr[-2000:]
    rows = [list(map(int, line.split())) for line in table.read_text().splitlines() if line.strip()]
    assert rows and all(len(r) == 8 for r in rows)
    # GEMM tiles: 4 candidates; the chi0 SYRK (family 105): 7; W.B (107): 2
    ncand = {105: 7, 107: 2}
    assert all(0 <= r[-1] < ncand.get(r[0], 4) for r in rows)
    assert {105, 107} <= {r[0] for r in rows}
