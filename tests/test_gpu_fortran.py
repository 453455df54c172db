"""The Fortran drop-in: a Fortran program runs MOLGW's one-shot flow through
the iso_c_binding module m_mgw (molgw_amd/fortran) and libmgw on the GPU;
its QP energies must match the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd.build import FDRIVER
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def write_input(path, s):
    with open(path, 'wb') as fh:
        np.array([s.nbf, s.naux, s.nstate], dtype=np.int32).tofile(fh)
        np.array([s.eri_3center.shape[0]], dtype=np.int64).tofile(fh)
        np.asarray(s.eri_3center, dtype=np.float64).ravel(order='F').tofile(fh)
        np.asarray(s.index_basis, dtype=np.int32).ravel(order='F').tofile(fh)
        np.asarray(s.c_matrix, dtype=np.float64).ravel(order='F').tofile(fh)
        s.energy.astype(np.float64).tofile(fh)
        s.occupation.astype(np.float64).tofile(fh)


def test_fortran_driver_matches_oracle(tmp_path):
    if not os.path.exists(FDRIVER):
        pytest.skip('Fortran driver not built (amdflang absent)')
    s = synthetic_system(30, 90, 6, seed=21, scale=0.3)
    fin, fout = tmp_path / 'in.bin', tmp_path / 'out.bin'
    write_input(fin, s)
    proc = subprocess.run([FDRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    raw = np.fromfile(fout, dtype=np.float64)
    n = s.nstate
    eqp, zz, elin = raw[:n], raw[n:2 * n], raw[2 * n:3 * n]
    en_rpa, en_gw = raw[3 * n], raw[3 * n + 1]
    eqp_one = raw[3 * n + 2:4 * n + 2]
    eqp_contour = raw[4 * n + 2:5 * n + 2]
    ref = O.g0w0_casida(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation)
    assert abs(en_rpa - ref.en_rpa) < 1e-10
    assert abs(en_gw - ref.en_gw) < 1e-10
    assert np.max(np.abs(eqp - ref.energy_qp)) < 1e-8
    assert np.max(np.abs(elin - ref.energy_qp_lin)) < 1e-8
    np.testing.assert_array_equal(eqp, eqp_one)
    # contour deformation through its iso_c_binding wrapper
    refc = O.g0w0_contour(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation, 12, 6)
    assert np.max(np.abs(eqp_contour - refc['energy_qp'][:, 0])) < 1e-8
    # QSGW matrix and GW ingredients through their iso_c_binding wrappers
    off = 5 * n + 2
    qs_mo = raw[off:off + n * n].reshape((n, n), order='F')
    off += n * n
    qs_ao = raw[off:off + s.nbf ** 2].reshape((s.nbf, s.nbf), order='F')
    off += s.nbf ** 2
    npole = ref.pole.size
    w = raw[off:off + npole * n * n].reshape((npole, n, n), order='F')
    off += npole * n * n
    v = raw[off:off + s.naux * n * n].reshape((s.naux, n, n), order='F')
    off += s.naux * n * n
    ws = raw[off:off + s.naux * npole].reshape((s.naux, npole), order='F')
    off += s.naux * npole
    xmv = raw[off:off + n * n].reshape((n, n), order='F')
    assert off + n * n == raw.size
    # Sigma_x - Vxc through its wrapper: the planted shifts on the diagonal, nothing off it
    np.testing.assert_allclose(np.diag(xmv), 0.001 * np.arange(1, n + 1), atol=1e-12)
    assert np.abs(xmv - np.diag(np.diag(xmv))).max() < 1e-12
    rng = O.selfenergy_set_state_range(s.occupation[:, None])
    eri = O.calculate_eri_3center_mo(s.eri_3center, s.index_basis, s.c_matrix, 1, n, 1, n)
    ref_qs = O.gw_selfenergy_qs(eri, ws, ref.pole, s.energy, s.occupation, rng)
    scale = np.max(np.abs(ref_qs))
    assert np.max(np.abs(qs_mo - ref_qs)) < 1e-10 * scale
    assert np.max(np.abs(qs_ao - s.c_matrix @ ref_qs @ s.c_matrix.T)) < 1e-10 * scale
    assert np.max(np.abs(v - eri)) < 1e-12 * np.max(np.abs(eri))
    ref_w = np.einsum('Ps,Ppq->spq', ws, eri)      # w_s of the same run (eigenvector signs)
    assert np.max(np.abs(w - ref_w)) < 1e-11 * np.max(np.abs(ref_w))
