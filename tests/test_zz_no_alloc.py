"""The no-allocation AO -> MO entry point (mgw_calculate_eri_3center_mo_noalloc)
against mgw_calculate_eri_3center_mo on the same resident tensors: the same
GEMMs in the same summation order, so the blocks agree bit for bit. Two
calls in a row share the GPU's scratch array, the second on a smaller
molecule; a block whose half transform exceeds the scratch is refused."""
import numpy as np
import pytest

from molgw_amd._lib import MgwError
from molgw_amd.gw import DeviceBuffer, Engine
from molgw_amd.workloads import synthetic_system

pytestmark = pytest.mark.gpu


def _both(eng, s, mlo, mhi, nlo, nhi):
    eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
    eng.set_mo(s.c_matrix, s.energy, s.occupation)
    ref = eng.calculate_eri_3center_mo(mlo, mhi, nlo, nhi)
    eng.destroy_eri_3center_mo()
    out = DeviceBuffer(ref.nbytes)
    eng.calculate_eri_3center_mo_noalloc(mlo, mhi, nlo, nhi, out)
    eng.synchronize()
    got = out.download(np.empty(ref.shape, order='F'))
    out.free()
    return ref, got


def test_noalloc_ao2mo_equals_the_allocating_entry_point():
    big = synthetic_system(96, 320, 12, seed=61)
    small = synthetic_system(40, 130, 5, seed=62)
    with Engine(0) as eng:
        # (P|a i): the Casida / chi0 block; then (P|p m) of a 16-state window
        for s, (mlo, mhi, nlo, nhi) in ((big, (13, 96, 1, 12)), (big, (1, 96, 5, 20)),
                                        (small, (6, 40, 1, 5)), (small, (1, 40, 1, 3))):
            ref, got = _both(eng, s, mlo, mhi, nlo, nhi)
            print(f'nbf {s.nbf} naux {s.naux} block ({mlo}..{mhi}, {nlo}..{nhi}): '
                  f'max |difference| {np.max(np.abs(got - ref)):.1e}', flush=True)
            np.testing.assert_array_equal(got, ref)


def test_noalloc_ao2mo_refuses_blocks_beyond_the_scratch():
    s = synthetic_system(160, 400, 140, seed=63)        # 400 x 160 x 140 = 8.96e6 doubles > 8 Mi
    with Engine(0) as eng:
        eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        out = DeviceBuffer(8 * 400 * 140 * 160)
        with pytest.raises(MgwError, match='exceeds the scratch'):
            eng.calculate_eri_3center_mo_noalloc(1, 160, 1, 140, out)
        out.free()
