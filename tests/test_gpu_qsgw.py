"""QSGW self-energy matrix (gw_selfenergy_qs + apply_qs_approximation)
through the C-ABI against the oracle. Tolerances: Sigma_MO and Sigma_AO
1e-10 relative to their largest element."""
import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd.gw import selfenergy_set_state_range
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize('window', [None, (3, 14)])
def test_qs_matrix_matches_oracle(engine, window):
    s = synthetic_system(34, 96, 6, seed=73, scale=0.3)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    rng = selfenergy_set_state_range(s.occupation)
    if window:
        rng.nsemin, rng.nsemax = window
    _, _, pole, ws = engine.polarizability(rng, return_ws=True)
    a = np.random.default_rng(2).standard_normal((s.nbf, s.nbf))
    smat = np.eye(s.nbf) + 0.05 * (a + a.T)
    mo, ao = engine.gw_selfenergy_qs(rng, s_matrix=smat)
    orng = O.selfenergy_set_state_range(s.occupation[:, None])
    orng.nsemin, orng.nsemax = rng.nsemin, rng.nsemax
    eri_pi = O.calculate_eri_3center_mo(s.eri_3center, s.index_basis, s.c_matrix, rng.nsemin, rng.nsemax,
                                        rng.ncore_G + 1, rng.nvirtual_G - 1)
    ref = O.gw_selfenergy_qs(eri_pi, ws, pole, s.energy, s.occupation, orng)
    win = slice(rng.nsemin - 1, rng.nsemax)
    print(f'window {window}: Sigma_MO rel err {rel_err(mo[win, win], ref):.2e}')
    assert rel_err(mo[win, win], ref) < 1e-10
    outside = np.ones(mo.shape, bool)
    outside[win, win] = False
    assert np.all(mo[outside] == 0.0)
    cw = s.c_matrix[:, win]
    ref_ao = smat @ cw @ ref @ cw.T @ smat
    assert rel_err(ao, ref_ao) < 1e-10
