#!/usr/bin/env python3
"""Benchmark: G0W0 QP solves per second on one MI355X.

Workload (default ``octane``): the plain FP64 G0W0@HF drop-in at the shape of
all-trans n-octane in cc-pVTZ / cc-pVTZ-RI (nbf=492, nauxil=1188, 33 occupied,
npole = 33 x 459 = 15147), analytic Casida W, 51-point real grid. The image has
no Gaussian-integral engine, so the RI tensor, C and the energies are synthetic
(seeded recipe of molgw_amd.workloads at that shape); every kernel does the
work of the real molecule's shapes.

A step is one full G0W0 solve (AO->MO, Casida build + eigensolve, w_s,
AO->MO for Sigma, pole self-energy on 51 frequencies for all 492 states, QP
equation) from the RI tensor, C and energies already resident in HBM.

Prints ONE JSON line (rank 0). The dominant hand-written kernel's roofline
comes from HIP events recorded inside libmgw on the stream the kernels run
on. The CPU baseline is the oracle (numpy/OpenBLAS FP64 port of the
reference algorithm) timed on a bounded, stage-sampled piece of one solve
and scaled to one solve (see ``cpu_baseline``).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'G0W0 QP solves/sec + p50/p99 latency at ≤1 meV QP error vs FP64, mixed stream'
MFMA_F64_PEAK_TFLOPS = 78.6
VALU_F64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (nbf, naux, nocc, technique, description)
    'octane': (492, 1188, 33, 'one_shot',
               'n-octane-shaped FP64 G0W0@HF drop-in, cc-pVTZ/cc-pVTZ-RI shape (nbf=492, naux=1188, '
               'nocc=33, npole=15147), analytic Casida W, 51-point real grid, all 492 states'),
    'water': (24, 84, 5, 'one_shot',
              'water-shaped FP64 G0W0@HF, cc-pVDZ/cc-pVDZ-RI shape (nbf=24, naux=84, nocc=5)'),
    'synthetic-imag': (1200, 3600, 120, 'imaginary_pade',
                       'synthetic imaginary-frequency kernel (nbf=1200, naux=3600, nocc=120), '
                       '32 GL frequencies, Pade continuation, 16 states HOMO-7..LUMO+8'),
    'lih-qz': (105, 244, 2, 'one_shot',
               'LiH (3 bohr) FP64 G0W0@HF with real integrals from the native front end: cc-pVQZ / cc-pVQZ-RI, '
               'Cartesian (nbf=105, naux=244 kept of 251), analytic Casida W, 51-point real grid, all states; '
               'the reference golden (lih_gw.in) gap is 8.053158 eV'),
    'octane-real': (492, 1188, 33, 'one_shot',
                    'all-trans n-octane FP64 G0W0@HF with real integrals and RI-HF orbitals from the native front end '
                    '(geometry from bond rules: C-C 1.54 A, C-H 1.09 A, tetrahedral), cc-pVTZ / cc-pVTZ-RI '
                    '(nbf=492, naux=1188, nocc=33, npole=15147), analytic Casida W, 51-point real grid, all 492 states'),
    'contour': (192, 672, 40, 'contour',
                '(H2O)8-shaped FP64 G0W0 by contour deformation, cc-pVDZ/cc-pVDZ-RI shape (nbf=192, naux=672, '
                'nocc=40), W on 16 imaginary and 6 real frequencies, 51-point real grid, 16 states '
                'HOMO-7..LUMO+8'),
}


def stream_bench(workload, args):
    """Mixed-size request streams (the Poisson serving stream and the
    64-alkane batch): an FP64 pass that serves one request at a time (the
    FP64 drop-in, also the QP reference), then the serving pass (mixed
    precision, concurrent slots, continuous batching). A step is one request;
    the warm-up requests are served before the timed stream."""
    from molgw_amd.frontend import alkane, water_cluster
    from molgw_amd.gw import Engine, device_memory
    from molgw_amd.serving import Request, Server, summarize, poisson_stream, alkane_batch
    from molgw_amd.workloads import molecular_system_device
    if workload == 'serving':
        spec = poisson_stream(args.requests, args.rate, seed=args.seed)
        realtime = True
    else:
        spec = [(0.0, 'alkane', n) for n in alkane_batch(args.requests, seed=args.seed)]
        realtime = False
    # the molecules, from the GPU front end (not timed): alkanes from bond
    # rules, water clusters seeded per request (rigid monomers, random
    # orientations, O-O >= 2.7 A in a sphere)
    t0 = time.perf_counter()
    systems, cache = [], {}
    with Engine(0) as fe:
        for k, (_, kind, size) in enumerate(spec):
            key = (kind, size) if kind == 'alkane' else (kind, size, args.seed * 100003 + k)
            if key not in cache:
                mol = alkane(size) if kind == 'alkane' else water_cluster(size, seed=key[2])
                cache[key] = molecular_system_device(mol, f'{kind}{size}', engine=fe)
            systems.append(cache[key])
    free, _ = device_memory(0)
    budget = int(0.85 * free)
    print(f'# {workload}: {len(spec)} requests, {len(cache)} molecules from the GPU front end in '
          f'{time.perf_counter() - t0:.1f}s, HBM budget {budget / 2**30:.0f} GiB', file=sys.stderr, flush=True)

    def requests(precision):
        return [Request(k, systems[k], arrival=t, technique='imaginary_pade', precision=precision,
                        nomega_chi_imag=32, nomega_sigma_calc=16) for k, (t, kind, size) in enumerate(spec)]

    with Server(nslots=1, hbm_budget=budget, graphs=not args.no_graphs) as srv:
        srv.run(requests('fp64')[:1], realtime=False)                 # warm-up
        ref = srv.run(requests('fp64'), realtime=False)
    ref_sum = summarize(ref)
    print(f'# fp64 one at a time: {ref_sum}', file=sys.stderr, flush=True)
    with Server(nslots=args.slots, hbm_budget=budget, graphs=not args.no_graphs) as srv:
        if args.warmup:
            srv.run(requests(args.precision)[:args.warmup], realtime=False)  # untimed warm-up requests
        # capacity: the whole stream available at t = 0, as in the FP64 pass
        srv.profile(True)
        w0 = (time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
        res = srv.run(requests(args.precision), realtime=False)
        w1 = (time.clock_gettime_ns(time.CLOCK_MONOTONIC), time.clock_gettime_ns(time.CLOCK_BOOTTIME))
        # the capacity pass's clock window, to select its kernels in a kernel trace
        print(f'# capacity window ns: monotonic {w0[0]} {w1[0]} boottime {w0[1]} {w1[1]}', file=sys.stderr, flush=True)
        prof = srv.profile_read()
        srv.profile(False)
        # latency: the same requests arriving as the Poisson stream
        trims_cap = srv.trims
        live = srv.run(requests(args.precision), realtime=True) if realtime else res
        esc = srv.escalations()
        print(f'# admission trims: capacity pass {trims_cap}, live pass {srv.trims - trims_cap}', file=sys.stderr,
              flush=True)
    # the same requests one at a time, after the timed passes: each kernel
    # class alone on the device, so its HIP-event time is its kernels' time
    # (as a kernel trace of the pass shows them)
    with Server(nslots=1, hbm_budget=budget, graphs=not args.no_graphs) as srv:
        srv.run(requests(args.precision), realtime=False)   # warm: buffers of every shape allocated
        srv.profile(True)
        i0 = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
        srv.run(requests(args.precision), realtime=False)
        i1 = time.clock_gettime_ns(time.CLOCK_MONOTONIC)
        print(f'# isolated window ns: monotonic {i0} {i1}', file=sys.stderr, flush=True)
        prof_iso = srv.profile_read()
        srv.profile(False)
    sm = summarize(res)
    sm['profile_isolated'] = prof_iso
    lv = summarize(live)
    sm['profile'] = prof
    sm['escalations'] = esc
    sm['live'] = lv
    print('# serving per-class device ms (summed over slots): '
          + ', '.join(f'{k}={v["ms"]:.1f}' for k, v in prof.items()), file=sys.stderr, flush=True)
    err = 0.0
    for run in (res, live):
        for a, b in zip(run, ref):
            sl = slice(a.nsemin - 1, a.nsemax)
            err = max(err, float(np.max(np.abs(a.energy_qp[sl] - b.energy_qp[sl]))))
    print(f'# serving capacity: {sm["solves_per_s"]:.2f} solves/s, {sm["busy_slots_mean"]:.1f} slots solving on '
          f'average, {100 * sm["ingest_share"]:.0f} % of their time loading RI tensors and MO data; '
          f'live stream at {args.rate}/s: {lv}; '
          f'max |dE_QP| = {err * 27211.38505:.4f} meV', file=sys.stderr, flush=True)
    cpu = None
    if not args.no_cpu_baseline:
        print('# cpu baseline (oracle on a sample of the stream) ...', file=sys.stderr, flush=True)
        cpu = cpu_baseline_stream(spec, systems)
        print(f'# cpu: {cpu["value"]:.4f} solves/s; {cpu["sample"]}', file=sys.stderr, flush=True)
    sm['cpu_baseline'] = cpu
    return sm, ref_sum, err * 27211.38505, spec


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def gpu_bench(system, technique, steps, warmup, kw, precision='fp64'):
    from molgw_amd.gw import Engine
    eng = Engine(0)
    if hasattr(system, 'eri_3center_dev'):
        eng.set_eri_3center_device(system.eri_3center_dev.data_ptr(), system.npair, system.naux,
                                   system.index_basis, system.nbf)
        system.eri_3center_dev.free()                    # the engine holds its own dense copy
        system.eri_3center_dev = None
    else:
        eng.set_eri_3center(system.eri_3center, system.index_basis, system.nbf)
    eng.set_mo(system.c_matrix, system.energy, system.occupation)
    ref = None
    if precision != 'fp64':
        # accuracy condition of the metric: QP energies against the FP64 solve
        ref = eng.g0w0(technique=technique, return_sigma=False, **kw)
    for _ in range(warmup):
        eng.g0w0(technique=technique, return_sigma=False, precision=precision, **kw)
    eng.synchronize()
    eng.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        out = eng.g0w0(technique=technique, return_sigma=False, precision=precision, **kw)
        lat.append(time.perf_counter() - ts)   # g0w0 returns after the stream is drained
    eng.synchronize()
    t1 = time.perf_counter()
    prof = eng.profile_read()
    eng.profile(False)
    mem = eng.memory()
    eng.close()
    err_mev = None
    if ref is not None:
        lo, hi = kw['rng'].nsemin - 1 if 'rng' in kw else 0, kw['rng'].nsemax if 'rng' in kw else len(out.energy_qp)
        err_mev = float(np.max(np.abs(out.energy_qp[lo:hi] - ref.energy_qp[lo:hi]))) * 27.21138505e3
    return t1 - t0, lat, prof, out, mem, err_mev


PEAK_TFLOPS = {'fp64': 78.6, 'fp32': 157.3, 'bf16x3': 2500.0 / 3.0, 'bf16x6': 2500.0 / 6.0,
               'fp16x3': 2500.0 / 3.0}
# arithmetic of the kernel classes that run below FP64 in each mode (the rest stay FP64)
LOWP_CLASSES = {'fp32': {'gemm_ao2mo': 'fp32', 'gemm_chi0': 'fp32', 'gemm_casida': 'fp32'},
                'bf16x3': {'gemm_ao2mo': 'bf16x3', 'gemm_chi0': 'bf16x3', 'gemm_casida': 'bf16x3',
                           'gemm_wb': 'bf16x3'},
                'bf16x6': {'gemm_chi0': 'bf16x6', 'gemm_wb': 'bf16x6'},
                'fp16x3': {'gemm_ao2mo': 'fp16x3', 'gemm_chi0': 'fp16x3', 'gemm_wb': 'fp16x3'}}


def roofline_of(prof, precision='fp64', traffic_key=None, by='ms'):
    """Dominant hand-written kernel class (the rocSOLVER eigensolver excluded),
    by device time (by='ms') or by algorithmic work (by='flops': used for the
    concurrent serving pass, where an event-timed class also counts the time
    other slots' kernels hold the GPU).

    achieved = algorithmic flops per launch / average launch time. For the
    split products the ceiling is the published dense BF16/FP16 MFMA peak
    (2.5 PFLOP/s) divided by the products each algorithmic product costs (3
    for BF16x3 and FP16x3, 6 for BF16x6)."""
    own = {k: v for k, v in prof.items() if k != 'eigensolver'}
    if not own:
        return None, None
    name = max(own, key=lambda k: own[k][by])
    v = own[name]
    t_avg = v['ms'] / v['launches'] * 1e-3
    achieved = v['flops'] / v['launches'] / t_avg / 1e12
    if name == 'sigma_poles':
        r = {'bound': 'valu', 'achieved': achieved, 'peak': VALU_F64_PEAK_TFLOPS, 'unit': 'TFLOP/s'}
    else:
        prec = LOWP_CLASSES.get(precision, {}).get(name, 'fp64')
        r = {'bound': 'mfma', 'achieved': achieved, 'peak': PEAK_TFLOPS[prec], 'unit': 'TFLOP/s',
             'arith': prec}
    r['frac'] = r['achieved'] / r['peak']
    r['traffic'] = _measured_traffic(name, traffic_key)
    r['kernel'] = name
    r['avg_launch_ms'] = v['ms'] / v['launches']
    r['launches'] = v['launches']
    return name, r


def _measured_traffic(kclass, key=None):
    """HBM bytes per launch of a kernel class from the committed rocprofv3
    --pmc summary (profiles/traffic.json, keyed by workload/precision), if any."""
    path = os.path.join(ROOT, 'profiles', 'traffic.json')
    try:
        with open(path) as fh:
            table = json.load(fh)
    except (OSError, ValueError):
        return None
    return table.get(key or '', {}).get(kclass, {}).get('bytes_per_launch')


def cpu_baseline(system, budget_note=True):
    """Stage-sampled CPU oracle time for one solve of ``system`` (one-shot).

    Stages run at full size: AO->MO for W and for Sigma, the Casida build
    (A+B by a BLAS-3 product) and the w_s product (timed on a same-shape
    random product). Sampled and scaled: the DSYEV eigensolve on the leading
    n_s x n_s block of M (scaled by (npole/n_s)^3) and the pole self-energy
    for 2 of the i states (scaled linearly by n_i / 2).
    """
    import scipy.linalg as sla
    from oracle import gw_oracle as O
    from oracle import build_c
    if not os.path.exists(build_c.OUT):
        raise RuntimeError('oracle C helpers not built: run __graft_entry__.build()')
    s = system
    t = {}
    occ = s.occupation
    rng = O.selfenergy_set_state_range(occ[:, None])
    ttable = O.transition_table(occ[:, None], 2.0, rng.ncore_W, rng.nvirtual_W)
    npole = len(ttable)
    t0 = time.perf_counter()
    eri_w = O.calculate_eri_3center_mo(s.eri_3center, s.index_basis, s.c_matrix, 1, s.nstate, 1, s.nstate)
    t['ao2mo_w'] = time.perf_counter() - t0
    x = O.casida_ia_matrix(eri_w, ttable, 1)
    del eri_w
    t0 = time.perf_counter()
    i = ttable[:, 0] - 1
    a = ttable[:, 1] - 1
    de = s.energy[a] - s.energy[i]
    apb = 4.0 * (x.T @ x)
    apb[np.diag_indices_from(apb)] += de
    sq = np.sqrt(de)
    m = sq[:, None] * apb * sq[None, :]
    t['casida_build'] = time.perf_counter() - t0
    del apb
    ns = min(npole, 1536)
    t0 = time.perf_counter()
    sla.eigh(m[:ns, :ns], driver='ev')
    t['eigensolve'] = (time.perf_counter() - t0) * (npole / ns) ** 3
    zr = np.random.default_rng(0).standard_normal((npole, npole))
    del m
    t0 = time.perf_counter()
    ws = x @ zr
    t['w_s'] = time.perf_counter() - t0
    del zr
    ws *= 1.0 / np.sqrt(np.abs(ws).max() + 1.0)
    pole = np.sort(np.random.default_rng(1).uniform(0.5, 3.0, npole))
    t0 = time.perf_counter()
    eri_pi = O.calculate_eri_3center_mo(s.eri_3center, s.index_basis, s.c_matrix, rng.nsemin, rng.nsemax,
                                        rng.ncore_G + 1, rng.nvirtual_G - 1)
    t['ao2mo_sigma'] = time.perf_counter() - t0
    ni = rng.nvirtual_G - 1 - rng.ncore_G
    nsample = 2
    sub = O.StateRanges(rng.nhomo_G - 1, rng.nhomo_G + 2, rng.nhomo_G, rng.nlumo_G, rng.nsemin, rng.nsemax,
                        rng.ncore_W, rng.nvirtual_W)
    omega = 0.01 * np.arange(-25, 26)
    t0 = time.perf_counter()
    O.gw_selfenergy_c(eri_pi[:, :, rng.nhomo_G - 1:rng.nhomo_G + 1], ws, pole, s.energy, occ, sub, omega,
                      s.energy[rng.nsemin - 1:rng.nsemax], 0.001)
    t['sigma'] = (time.perf_counter() - t0) * ni / nsample
    total = sum(t.values())
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = int(os.environ.get('OMP_NUM_THREADS', cores))
    return {'value': 1.0 / total, 'unit': 'solves/s', 'cores': min(cores, threads), 'kind': 'port',
            'sample': ('oracle (numpy/OpenBLAS, pole sum in C/OpenMP), one solve estimated from stages: AO->MO x2, '
                       'Casida build and '
                       f'w_s product at full size; DSYEV on the leading {ns}x{ns} block of M scaled by '
                       f'(npole/{ns})^3; pole self-energy for {nsample} of {ni} i-states scaled linearly'),
            'stage_seconds': {k: round(v, 3) for k, v in t.items()},
            'seconds_per_solve': total}


def _real_system(hf):
    """The RI tensor of a front-end RI-HF result built on the device (metric
    inverse square root + B), as a system for gpu_bench."""
    from types import SimpleNamespace
    from molgw_amd.gw import Engine
    eng = Engine(0)
    kept, _, _ = eng.inverse_sqrt_eri_2center(hf.eri_2center)
    b = eng.build_eri_3center(hf.eri_3center_tmp, hf.index_basis, hf.nbf)
    eng.close()
    return SimpleNamespace(eri_3center=b, index_basis=hf.index_basis, nbf=hf.nbf, naux=kept, nstate=len(hf.energy_mo),
                           c_matrix=hf.c_matrix, energy=hf.energy_mo, occupation=hf.occupation, scf_energy=hf.energy)


def real_system_alkane(n):
    """All-trans CnH2n+2 at cc-pVTZ / cc-pVTZ-RI from the native front end. Not timed."""
    from molgw_amd.frontend import alkane, run_rihf
    t0 = time.perf_counter()
    hf = run_rihf(alkane(n), 'cc-pVTZ', 'cc-pVTZ-RI', os.path.join(ROOT, 'tests', 'golden', 'basis'))
    print(f'# front end: RI-HF {hf.energy:.8f} Ha in {time.perf_counter() - t0:.0f}s ({hf.iterations} iterations)',
          file=sys.stderr, flush=True)
    return _real_system(hf)


def real_system_lih():
    """LiH at 3 bohr, cc-pVQZ / cc-pVQZ-RI Cartesian (the reference's lih_gw.in)
    from the native front end. Not timed."""
    from molgw_amd.frontend import Molecule, run_rihf
    m = Molecule.from_lines(['Li 0.000000 0.0 1.50', 'H 0.000000 0.0 -1.50'], unit='bohr')
    return _real_system(run_rihf(m, 'cc-pVQZ', 'cc-pVQZ-RI', os.path.join(ROOT, 'tests', 'golden', 'basis'),
                                 gaussian_type='CART'))


def cpu_baseline_contour(system, kw, gpu_out):
    """The oracle's whole contour-deformation solve (g0w0_contour: numpy/OpenBLAS
    FP64, the reference's loops) of the same system and ranges, timed once;
    its QP energies must match the GPU solve's."""
    from oracle import gw_oracle as O
    s, rng = system, kw['rng']
    ranges = O.StateRanges(rng.ncore_G, rng.nvirtual_G, rng.nhomo_G, rng.nlumo_G, rng.nsemin, rng.nsemax,
                           rng.ncore_W, rng.nvirtual_W)
    t0 = time.perf_counter()
    ref = O.g0w0_contour(s.eri_3center, s.index_basis, s.c_matrix, s.energy, s.occupation, kw['nomega_chi_imag'],
                         kw['nomega_chi_real'], ranges=ranges)
    dt = time.perf_counter() - t0
    diff = float(np.max(np.abs(ref['energy_qp'][:, 0] - gpu_out.energy_qp)))
    if diff > 1e-8:
        raise RuntimeError(f'contour: GPU and oracle QP energies differ by {diff:.2e} Ha')
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = int(os.environ.get('OMP_NUM_THREADS', cores))
    return {'value': 1.0 / dt, 'unit': 'solves/s', 'cores': min(cores, threads), 'kind': 'port',
            'sample': 'oracle g0w0_contour (numpy/OpenBLAS FP64), one whole solve of the same system; '
                      f'its QP energies agree with the GPU solve to {diff:.1e} Ha',
            'seconds_per_solve': dt}


def _stream_flops(nbf, naux, nocc, nomega=32, nse=16):
    """Operation count of one imaginary-axis solve as the CPU oracle runs it:
    AO->MO (X and the Sigma block), the chi0 SYRK, the per-frequency
    eigenvalues + inverse of 1 - chi0 (~4 naux^3) and the W.B product."""
    nstate = nbf
    npole = nocc * (nstate - nocc)
    return (2.0 * naux * nbf * nbf * (nocc + nse) + naux * naux * npole * nomega
            + 4.0 * naux ** 3 * nomega + 2.0 * naux * naux * nstate * nse * nomega)


def _blas_info():
    """BLAS library and version numpy calls (the oracle's dense algebra)."""
    try:
        cfg = np.show_config(mode='dicts')
        b = cfg['Build Dependencies']['blas']
        return f"{b.get('name', '?')} {b.get('version', '')}".strip()
    except Exception:                       # older numpy: no dict mode
        return 'numpy BLAS (unknown)'


def _timed(fn, repeats):
    """Median wall time of ``repeats`` runs after one untimed warm-up run."""
    fn()
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts


def cpu_baseline_stream(spec, systems, window=8, repeats=3):
    """CPU oracle (imaginary axis + Pade, FP64; numpy/OpenBLAS) on three of
    the stream's own molecules: the smallest, the median and the largest
    request by operation count. Each is timed after one warm-up run as the
    median of ``repeats`` runs. The smallest and the median run the whole
    32-frequency solve; the largest is timed at 1 and 2 frequencies and its
    32-frequency time follows from the measured cost per frequency (every
    frequency repeats the same chi0 build, inversion and self-energy work).
    The stream's time is the sum over its requests of the measured time of
    the nearest timed molecule in operation count, scaled by the operation
    count (AO->MO, chi0, 4 naux^3 per frequency, W.B). Value: requests per
    second served one at a time on the host."""
    from oracle import gw_oracle as O
    order = sorted(range(len(systems)), key=lambda k: _stream_flops(systems[k].nbf, systems[k].naux,
                                                                     systems[k].nocc))
    picks = [order[0], order[len(order) // 2], order[-1]]
    measured = []      # (flops, seconds for 32 frequencies, description)
    for rank, k in enumerate(picks):
        s = systems[k]
        f = _stream_flops(s.nbf, s.naux, s.nocc)
        if measured and f == measured[-1][0]:
            continue
        eri = s.eri_3center_host()
        rng = O.selfenergy_set_state_range(s.occupation[:, None])
        rng.nsemin = max(rng.nsemin, rng.nhomo_G - window + 1)
        rng.nsemax = min(rng.nsemax, rng.nhomo_G + window)

        def solve(nw):
            return lambda: O.g0w0_pade(eri, s.index_basis, s.c_matrix, s.energy, s.occupation, nw,
                                       nomega_sigma_calc=16, ranges=rng, exchange_m_vxc=s.exchange_m_vxc)
        if rank < 2:
            t, ts = _timed(solve(32), repeats)
            measured.append((f, t, f'{s.name} (nbf {s.nbf}, naux {s.naux}) 32 frequencies: median {t:.2f}s of '
                                   + ', '.join(f'{x:.2f}' for x in ts)))
        else:
            t1, _ = _timed(solve(1), repeats)
            t2, _ = _timed(solve(2), repeats)
            t = t1 + 31.0 * (t2 - t1)
            measured.append((f, t, f'{s.name} (nbf {s.nbf}, naux {s.naux}): median {t1:.2f}s at 1 frequency, '
                                   f'{t2:.2f}s at 2, so {t:.1f}s at 32'))
        print(f'# cpu: {measured[-1][2]}', file=sys.stderr, flush=True)
        del eri
    seconds = 0.0
    for sy in systems:
        f = _stream_flops(sy.nbf, sy.naux, sy.nocc)
        fm, tm, _ = min(measured, key=lambda m: abs(math.log(m[0] / f)))
        seconds += tm * f / fm
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    threads = int(os.environ.get('OPENBLAS_NUM_THREADS', os.environ.get('OMP_NUM_THREADS', cores)))
    return {'value': len(spec) / seconds, 'unit': 'solves/s', 'cores': min(cores, threads), 'kind': 'port',
            'sample': ('oracle (numpy FP64 with ' + _blas_info() + f', {min(cores, threads)} threads; imaginary '
                       'axis + Pade, window HOMO-7..LUMO+8) on three of the stream\'s molecules, one warm-up run and '
                       f'the median of {repeats}: ' + '; '.join(m[2] for m in measured)
                       + f'; the stream\'s {len(spec)} requests scaled from the nearest of them by operation count'),
            'seconds_for_stream': seconds}


RI_BUILD = (492, 1188, 'n-octane-shaped RI tensor build, cc-pVTZ/cc-pVTZ-RI shape (nbf=492, npair=121278, '
            'nauxil=1188): inverse square root of the 2-center metric (FP64 eigensolve, eigenvalues > 1e-6 kept) '
            'and B = (mu nu|Q) V^-1/2 (FP64)')


def ri_build_bench(args):
    """Metric V^-1/2 and B build (src/m_eri_calculate.f90:1049-1262, :1877-1892)
    at the octane shape, (mu nu|Q) resident in HBM. A step is one metric
    eigendecomposition plus one B build over all pairs."""
    from molgw_amd.gw import DeviceBuffer, Engine
    nbf, naux, desc = RI_BUILD
    k, l = np.triu_indices(nbf)
    ib = np.vstack([k + 1, l + 1]).astype(np.int32)
    npair = ib.shape[1]
    rng = np.random.default_rng(args.seed)
    # metric with a Gaussian-RI-like spectrum (1e-8 .. 1e2): a few eigenvalues dropped
    q, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    lam = np.geomspace(1e-8, 1e2, naux)
    v = np.asfortranarray((q * lam) @ q.T)
    v = 0.5 * (v + v.T)
    tmp_h = np.asfortranarray(0.1 * rng.standard_normal((npair, naux)))     # (mu nu|Q), column-major
    tmp = DeviceBuffer(tmp_h.nbytes)
    tmp.upload(tmp_h)
    eng = Engine(0)
    for _ in range(args.warmup):
        eng.inverse_sqrt_eri_2center(v, copy=False)
        eng.build_eri_3center_device(tmp.data_ptr(), 1, npair, npair, ib, nbf)
    eng.synchronize()
    eng.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        kept, low, _ = eng.inverse_sqrt_eri_2center(v, copy=False)
        eng.build_eri_3center_device(tmp.data_ptr(), 1, npair, npair, ib, nbf)
        lat.append(time.perf_counter() - ts)
    eng.synchronize()
    elapsed = time.perf_counter() - t0
    prof = eng.profile_read()
    eng.profile(False)
    eng.close()
    print(f'# ri-build: {args.steps} builds in {elapsed:.3f}s, kept {kept} of {naux}; per-class device ms: '
          + ', '.join(f'{c}={x["ms"]:.1f}' for c, x in prof.items()), file=sys.stderr, flush=True)
    roof = None
    if 'gemm_other' in prof:
        x = prof['gemm_other']
        t_avg = x['ms'] / x['launches'] * 1e-3
        ach = x['flops'] / x['launches'] / t_avg / 1e12
        roof = {'bound': 'mfma', 'achieved': ach, 'peak': MFMA_F64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                'frac': ach / MFMA_F64_PEAK_TFLOPS, 'traffic': None, 'kernel': 'gemm_other (B build)',
                'avg_launch_ms': x['ms'] / x['launches'], 'launches': x['launches']}
    cpu = None
    if not args.no_cpu_baseline:
        # oracle on the full metric and on a sample of the pairs, scaled by npair / sample
        from oracle import gw_oracle as O
        sample = 16384
        tmp_h = np.ascontiguousarray(tmp_h[:sample])
        c0 = time.perf_counter()
        sq, _ = O.inverse_sqrt_eri_2center(v)
        c1 = time.perf_counter()
        O.build_eri_3center(tmp_h, sq)
        c2 = time.perf_counter()
        per = (c1 - c0) + (c2 - c1) * npair / sample
        try:
            cores = len(os.sched_getaffinity(0))
        except AttributeError:
            cores = os.cpu_count()
        cpu = {'value': 1.0 / per, 'unit': 'builds/s', 'cores': min(cores, int(os.environ.get('OMP_NUM_THREADS', cores))),
               'kind': 'port', 'sample': f'oracle (numpy eigh + matmul, OpenBLAS): full metric eigensolve, B for '
                                         f'{sample} of {npair} pairs scaled linearly',
               'seconds_per_build': per}
    line = {
        'metric': 'RI tensor builds/s (metric inverse square root + B)', 'value': args.steps / elapsed,
        'unit': 'builds/s', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic (seeded metric with a 1e-8..1e2 spectrum, '
                                                       'seeded (mu nu|Q))',
        'config': {'workload': desc, 'nbf': nbf, 'npair': npair, 'nauxil': naux, 'nauxil_kept': kept},
        'latency_ms': {'p50': 1e3 * _percentile(lat, 50), 'p99': 1e3 * _percentile(lat, 99)},
        'roofline': roof, 'cpu_baseline': cpu,
        'stages_ms_per_build': {c: x['ms'] / args.steps for c, x in prof.items()},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed steps: solves (single-system workloads, default 5) or requests of the stream '
                         '(serving: default 20; batch64: default 64)')
    ap.add_argument('--warmup', type=int, default=None,
                    help='untimed warm-up steps (default 1; serving: 5 requests; batch64: one per slot)')
    ap.add_argument('--workload', default='serving', choices=sorted(WORKLOADS) + ['serving', 'batch64', 'ri-build'])
    ap.add_argument('--requests', type=int, default=None, help='stream / batch length (same as --steps)')
    ap.add_argument('--rate', type=float, default=20.0, help='Poisson arrival rate, requests/s (serving)')
    ap.add_argument('--slots', type=int, default=16, help='concurrent solve slots (serving, batch64)')
    ap.add_argument('--autotune', action='store_true',
                    help='time the FP64 GEMM tile candidates of each new shape class on first use (mgw_autotune)')
    ap.add_argument('--hw-queues', type=int, default=16,
                    help='hardware queues of this process (GPU_MAX_HW_QUEUES, set before the HIP runtime starts; '
                         'HIP defaults to 4, which the slots\' streams then share)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-graphs', action='store_true',
                    help='launch the Cholesky sequence eagerly instead of replaying its HIP graph (serving slots); '
                         'for profiler runs')
    ap.add_argument('--seed', type=int, default=0)
    ap.add_argument('--precision', default=None, choices=['fp64', 'fp32', 'bf16x3', 'bf16x6', 'fp16x3'],
                    help='arithmetic of the AO->MO transforms, the chi0/Casida build and the '
                         'imaginary-axis self-energy product (default: fp16x3 for the streams, fp64 otherwise)')
    args = ap.parse_args()
    # one hardware queue per slot stream (measured on MI355X over the serving
    # stream: 16 slots on 16 queues 47.8 solves/s, 8 slots on the default 4
    # queues 46.4, means of three runs); the HIP runtime reads it at start-up,
    # which is the first libmgw call below
    if args.hw_queues > 0:
        os.environ['GPU_MAX_HW_QUEUES'] = str(min(args.hw_queues, 32))
    if args.autotune:
        from molgw_amd.gw import autotune
        autotune(True)
    if args.gpus != 1:
        sys.exit('bench.py measures one GPU: every box here has one, and the path has no multi-GPU code; '
                 'run with --gpus 1')
    stream = args.workload in ('serving', 'batch64')
    if args.precision is None:
        args.precision = 'fp16x3' if stream else 'fp64'
    if args.requests is None:
        args.requests = args.steps if args.steps is not None else (64 if args.workload == 'batch64' else 20)
    if args.steps is None:
        args.steps = args.requests if stream else 5
    if args.warmup is None:
        args.warmup = (5 if args.workload == 'serving' else args.slots) if stream else 1
    if args.workload == 'ri-build':
        ri_build_bench(args)
        return
    if stream:
        prec = args.precision
        sm, ref, err, spec = stream_bench(args.workload, args)
        desc = ('Poisson stream of G0W0 requests: alkanes CnH2n+2 (n in 2..24) and water clusters (H2O)m '
                f'(m in 1..32), cc-pVDZ/cc-pVDZ-RI shapes, {args.requests} requests at {args.rate}/s'
                if args.workload == 'serving' else
                f'batch of {args.requests} alkanes CnH2n+2, n uniform in 1..16, cc-pVDZ/cc-pVDZ-RI shapes')
        line = {
            'metric': METRIC, 'value': sm['solves_per_s'], 'unit': 'solves/s', 'n_gpus': 1,
            'steps': sm['solves'], 'warmup': args.warmup, 'ms_per_step': 1e3 / sm['solves_per_s'],
            'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': (f'{prec} products, f64 accumulation; f64 Cholesky leaves and QP' if prec == 'fp16x3'
                      else f'{prec} products, f64 accumulation; f64 inversion and QP'),
            'qp_error_vs_fp64_mev': err,
            'data': 'real molecules: integrals, RI tensor and RI-HF orbitals from the GPU front end (no synthetic data)',
            'config': {'workload': desc, 'technique': 'imaginary_pade', 'nomega_chi_imag': 32,
                       'nomega_sigma_calc': 16, 'states': 'HOMO-7..LUMO+8', 'slots': args.slots,
                       'precision': prec,
                       'value': 'capacity: the whole stream available at t = 0 (as the FP64 pass)'},
            'latency_ms': {'p50': sm['live']['p50_ms'], 'p99': sm['live']['p99_ms'],
                           'stream': (f'Poisson arrivals at {args.rate}/s, {sm["live"]["solves_per_s"]:.2f} solves/s '
                                      'delivered' if args.workload == 'serving' else 'batch at t = 0')},
            'fp64_one_at_a_time': {'solves_per_s': ref['solves_per_s'], 'p50_ms': ref['p50_ms'],
                                   'p99_ms': ref['p99_ms']},
            'speedup_vs_fp64_one_at_a_time': sm['solves_per_s'] / ref['solves_per_s'],
            'roofline': roofline_of(sm['profile'], prec, f'{args.workload}/{prec}', by='flops')[1],
            'roofline_isolated': dict(roofline_of(sm['profile_isolated'], prec, f'{args.workload}/{prec}',
                                                  by='flops')[1],
                                      **{'pass': 'the same requests one at a time after the timed passes'}),
            'cpu_baseline': sm['cpu_baseline'],
            'stages_device_ms_summed_over_slots': {k: v['ms'] for k, v in sm['profile'].items()},
            'stages_ms_per_solve_isolated': {k: round(v['ms'] / len(spec), 3) for k, v in sm['profile_isolated'].items()},
            'fp64_rebuilt_frequencies': sm['escalations'],
        }
        print(json.dumps(line), flush=True)
        return
    from molgw_amd.workloads import synthetic_system, synthetic_system_device
    nbf, naux, nocc, technique, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    if args.workload == 'lih-qz':
        system = real_system_lih()
        nbf, naux = system.nbf, system.naux
    elif args.workload == 'octane-real':
        system = real_system_alkane(8)
        nbf, naux = system.nbf, system.naux
    elif naux * nbf * nbf > 1e9:
        system = synthetic_system_device(nbf, naux, nocc, seed=args.seed, name=args.workload)
    else:
        system = synthetic_system(nbf, naux, nocc, seed=args.seed, name=args.workload)
    print(f'# generated {args.workload}: nbf={nbf} naux={naux} nocc={nocc} in {time.perf_counter() - t0:.1f}s',
          file=sys.stderr, flush=True)
    kw = {}
    if technique in ('imaginary_pade', 'contour'):
        from molgw_amd.gw import selfenergy_set_state_range
        rng = selfenergy_set_state_range(system.occupation)
        rng.nsemin, rng.nsemax = rng.nhomo_G - 7, rng.nhomo_G + 8
        kw = (dict(rng=rng, nomega_chi_imag=32, nomega_sigma_calc=16) if technique == 'imaginary_pade' else
              dict(rng=rng, nomega_chi_imag=16, nomega_chi_real=6))
    elapsed, lat, prof, out, mem, err_mev = gpu_bench(system, technique, args.steps, args.warmup, kw,
                                                      args.precision)
    print(f'# gpu: {args.steps} solves in {elapsed:.3f}s; per-class device ms: '
          + ', '.join(f'{k}={v["ms"]:.1f}' for k, v in prof.items()), file=sys.stderr, flush=True)
    kname, roof = roofline_of(prof, args.precision, f'{args.workload}/{args.precision}')
    cpu = None
    if not args.no_cpu_baseline and technique == 'one_shot' and hasattr(system, 'eri_3center'):
        print('# cpu baseline (stage-sampled oracle) ...', file=sys.stderr, flush=True)
        cpu = cpu_baseline(system)
        print(f'# cpu: {cpu["seconds_per_solve"]:.1f} s per solve, stages {cpu["stage_seconds"]}',
              file=sys.stderr, flush=True)
    elif not args.no_cpu_baseline and technique == 'contour':
        print('# cpu baseline (oracle, one whole solve) ...', file=sys.stderr, flush=True)
        cpu = cpu_baseline_contour(system, kw, out)
        print(f'# cpu: {cpu["seconds_per_solve"]:.1f} s per solve', file=sys.stderr, flush=True)
    value = args.steps / elapsed
    line = {
        'metric': METRIC,
        'value': value,
        'unit': 'solves/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': 1e3 * elapsed / args.steps,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': {'fp64': 'f64', 'fp32': 'f32 products, f64 accumulation',
                  'bf16x3': 'bf16x3 products, f64 accumulation',
                  'bf16x6': 'bf16x6 products (chi0, W.B), f64 transforms, f64 accumulation',
                  'fp16x3': 'fp16x3 products (transforms, chi0, Cholesky products, W.B), f64 accumulation'}[
                      args.precision],
        'qp_error_vs_fp64_mev': err_mev,
        'data': 'synthetic (seeded RI tensor, orthonormal C and energies at the molecule\'s shape; '
                'no integral engine in the image)',
        'config': {'workload': desc, 'technique': technique, 'precision': args.precision,
                   'nbf': nbf, 'naux': naux, 'nocc': nocc, 'nstate': nbf},
        'latency_ms': {'p50': 1e3 * _percentile(lat, 50), 'p99': 1e3 * _percentile(lat, 99)},
        'roofline': roof,
        'stages_ms_per_solve': {k: v['ms'] / args.steps for k, v in prof.items()},
        'hbm_peak_bytes': mem[1],
        'cpu_baseline': cpu,
        'result_check': {'homo_qp_ha': float(out.energy_qp[nocc - 1]), 'lumo_qp_ha': float(out.energy_qp[nocc]),
                         'en_rpa': out.en_rpa, 'en_gw': out.en_gw},
    }
    if args.workload == 'octane-real':
        line['data'] = 'real integrals and RI-HF orbitals from the native front end (no synthetic data)'
        line['result_check']['scf_energy_ha'] = system.scf_energy
    if args.workload == 'lih-qz':
        line['data'] = 'real integrals and RI-HF orbitals from the native front end (no synthetic data)'
        gap = float(out.energy_qp[nocc] - out.energy_qp[nocc - 1]) * 27.21138505
        line['result_check'].update({'scf_energy_ha': system.scf_energy, 'gap_ev': gap, 'gap_golden_ev': 8.053158,
                                     'gap_within_1e-4_ev': abs(gap - 8.053158) < 1e-4})
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
