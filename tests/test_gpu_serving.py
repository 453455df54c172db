"""Continuous batching: solves overlap on the GPU, each result is the same
alone and in the mix (tolerance 1e-10 Ha; the kernels are deterministic),
and admission respects the HBM budget."""
import numpy as np
import pytest

from molgw_amd.serving import Request, Server, summarize
from molgw_amd.workloads import synthetic_system, water_cluster_shape, alkane_shape

pytestmark = pytest.mark.gpu


def small_requests(precision='bf16x3', technique='imaginary_pade'):
    shapes = [alkane_shape(1), water_cluster_shape(1), alkane_shape(2), water_cluster_shape(2),
              alkane_shape(3), water_cluster_shape(3)]
    reqs = []
    for k, (nbf, naux, nocc) in enumerate(shapes):
        s = synthetic_system(nbf, naux, nocc, seed=100 + k, scale=0.3)
        reqs.append(Request(k, s, arrival=0.002 * k, technique=technique, precision=precision,
                            nomega_chi_imag=16, nomega_sigma_calc=8))
    return reqs


@pytest.mark.parametrize('technique', ['imaginary_pade', 'one_shot'])
def test_each_result_same_alone_and_in_the_mix(technique):
    reqs = small_requests(technique=technique)
    with Server(nslots=3, hbm_budget=8 << 30) as srv:
        mixed = srv.run(reqs, realtime=True)
        # overlap happened: some solve started before another ended
        spans = sorted((r.start, r.end) for r in mixed)
        assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))
    with Server(nslots=1, hbm_budget=8 << 30) as srv:
        alone = [srv.run([r], realtime=False)[0] for r in reqs]
    for a, b in zip(mixed, alone):
        assert a.rid == b.rid
        assert np.max(np.abs(a.energy_qp - b.energy_qp)) <= 1e-10
        assert abs(a.en_rpa - b.en_rpa) <= 1e-10 * max(1.0, abs(b.en_rpa))
    summary = summarize(mixed)
    assert summary['solves'] == len(reqs) and summary['p99_ms'] >= summary['p50_ms']


def test_admission_keeps_within_budget():
    reqs = small_requests()
    from molgw_amd.serving import footprint_bytes
    biggest = max(footprint_bytes(r) for r in reqs)
    # room for one request at a time only: solves must not overlap
    with Server(nslots=3, hbm_budget=int(biggest * 1.2)) as srv:
        res = srv.run(reqs, realtime=False)
    spans = sorted((r.start, r.end) for r in res)
    assert all(spans[i + 1][0] >= spans[i][1] for i in range(len(spans) - 1))
    assert len(res) == len(reqs)


def test_request_larger_than_budget_is_refused():
    reqs = small_requests()[:1]
    with Server(nslots=1, hbm_budget=1 << 20) as srv:
        with pytest.raises(ValueError):
            srv.run(reqs)


@pytest.mark.parametrize('technique,precision,shape', [
    ('imaginary_pade', 'bf16x6', alkane_shape(4)), ('imaginary_pade', 'fp64', alkane_shape(4)),
    ('imaginary_pade', 'fp16x3', alkane_shape(4)), ('one_shot', 'fp64', alkane_shape(4)),
    ('imaginary_pade', 'fp16x3', water_cluster_shape(12))])
def test_footprint_bounds_the_measured_peak(technique, precision, shape):
    """Admission relies on mgw_footprint_bytes_nse: it must bound the HBM
    the solve actually holds (the context's ledger peak), with the request's
    self-energy window (16 states: of 106 for the alkane, of 288 for the
    water cluster, where the per-molecule buffers outweigh the fixed
    1 GiB allowance)."""
    from molgw_amd.gw import Engine
    from molgw_amd.serving import footprint_bytes
    s = synthetic_system(*shape, seed=77, scale=0.3)
    req = Request(0, s, technique=technique, precision=precision, nomega_chi_imag=16, nomega_sigma_calc=8)
    with Server(nslots=1, hbm_budget=64 << 30) as srv:
        srv.run([req], realtime=False)
        peak = srv.slots[0].memory()[1]
    fp = footprint_bytes(req)
    print(f'{technique} {precision}: peak {peak / 2**20:.1f} MiB, footprint {fp / 2**20:.1f} MiB')
    assert peak <= fp


def test_device_pool_stays_within_the_budget():
    """Admission counts the slots' ledgers and promised footprints, and the
    device pool releases what it holds above the budget at each
    synchronisation. On a mixed-size stream whose budget forces trims, the
    pool holds no more than the budget once the run is over, and at no time
    more than the budget plus one request's footprint (buffers freed while a
    larger molecule's buffers grow are released at that slot's next
    synchronisation). Memory that other contexts of this process hold (the
    session's engine) is measured before the run and taken off."""
    from molgw_amd.serving import footprint_bytes
    shapes = [alkane_shape(6), water_cluster_shape(2), alkane_shape(8), water_cluster_shape(7),
              alkane_shape(3), water_cluster_shape(9), alkane_shape(7)]
    reqs = [Request(k, synthetic_system(*sh, seed=300 + k, scale=0.3), technique='imaginary_pade',
                    precision='fp16x3', nomega_chi_imag=16, nomega_sigma_calc=8)
            for k, sh in enumerate(shapes)]
    fps = sorted(footprint_bytes(r) for r in reqs)
    # room for one request at a time (these small footprints are mostly the
    # fixed 1 GiB allowance): the idle slots' cached buffers must be
    # released to admit the next one
    budget = int(1.02 * fps[-1])
    with Server(nslots=3, hbm_budget=budget) as srv:
        for e in srv.slots:
            e.trim()
        base, _ = srv.slots[0].pool_memory(reset_high=True)
        res = srv.run(reqs, realtime=False)
        for e in srv.slots:
            e.synchronize()
        cur, high = srv.slots[0].pool_memory()
        cur, high = cur - base, high - base
        trims = srv.trims
    assert len(res) == len(reqs)
    assert trims >= 1
    print(f'budget {budget / 2**20:.0f} MiB, pool now {cur / 2**20:.0f} MiB, high-water {high / 2**20:.0f} MiB '
          f'above the {base / 2**20:.0f} MiB held before the run, {trims} trims')
    assert cur <= budget
    assert high <= budget + fps[-1]
