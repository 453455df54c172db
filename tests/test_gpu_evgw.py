"""evGW (GnW0, GnWn) through the C-ABI against the oracle's restatement of
selfenergy_evaluation with EVSC (src/m_selfenergy_evaluation.f90:78-182)."""
import numpy as np
import pytest

from oracle import gw_oracle as O
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize('approx', ['GnWn', 'GnW0'])
def test_evgw_matches_oracle(engine, approx):
    s = synthetic_system(40, 130, 6, seed=7, scale=0.1)
    exc = np.linspace(-0.01, 0.01, s.nstate)
    engine.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    engine.set_mo(s.c_matrix, s.energy, s.occupation)
    got = engine.evgw(4, approx, exchange_m_vxc=exc)
    ref, _ = O.evgw(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation, 4, approx,
                    exchange_m_vxc=exc)
    for k, (a, b) in enumerate(zip(got, ref)):
        err = np.abs(a - b).max()
        print(f'{approx} step {k + 1}: max |dE| {err:.2e} Ha, HOMO {a[5]:.8f}, LUMO {a[6]:.8f}')
        assert err < 1e-8
    # the engine is left with the set_mo energies: a one-shot G0W0 after evGW is unchanged
    out = engine.g0w0()
    g = O.g0w0_casida(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation)
    assert np.abs(out.energy_qp - g.energy_qp).max() < 1e-8
