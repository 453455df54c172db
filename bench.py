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
}


def _percentile(v, q):
    return float(np.percentile(np.asarray(v), q)) if len(v) else None


def gpu_bench(system, technique, steps, warmup, kw):
    from molgw_amd.gw import Engine
    eng = Engine(0)
    if hasattr(system, 'eri_3center_dev'):
        eng.set_eri_3center_device(system.eri_3center_dev.data_ptr(), system.npair, system.naux,
                                   system.index_basis, system.nbf)
        system.eri_3center_dev = None                    # the engine holds its own dense copy
        import torch
        torch.cuda.empty_cache()
    else:
        eng.set_eri_3center(system.eri_3center, system.index_basis, system.nbf)
    eng.set_mo(system.c_matrix, system.energy, system.occupation)
    for _ in range(warmup):
        eng.g0w0(technique=technique, return_sigma=False, **kw)
    eng.synchronize()
    eng.profile(True)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        out = eng.g0w0(technique=technique, return_sigma=False, **kw)   # returns after the stream is drained
        lat.append(time.perf_counter() - ts)
    eng.synchronize()
    t1 = time.perf_counter()
    prof = eng.profile_read()
    eng.profile(False)
    mem = eng.memory()
    eng.close()
    return t1 - t0, lat, prof, out, mem


def roofline_of(prof):
    """Dominant hand-written kernel class by device time (library solver calls excluded)."""
    own = {k: v for k, v in prof.items() if k not in ('eigensolver', 'cholesky')}
    if not own:
        return None, None
    name = max(own, key=lambda k: own[k]['ms'])
    v = own[name]
    t_avg = v['ms'] / v['launches'] * 1e-3
    if name == 'sigma_poles':
        achieved = v['flops'] / v['launches'] / t_avg / 1e12
        r = {'bound': 'valu', 'achieved': achieved, 'peak': VALU_F64_PEAK_TFLOPS, 'unit': 'TFLOP/s'}
    else:
        achieved = v['flops'] / v['launches'] / t_avg / 1e12
        r = {'bound': 'mfma', 'achieved': achieved, 'peak': MFMA_F64_PEAK_TFLOPS, 'unit': 'TFLOP/s'}
    r['frac'] = r['achieved'] / r['peak']
    r['traffic'] = _measured_traffic(name)
    r['kernel'] = name
    r['avg_launch_ms'] = v['ms'] / v['launches']
    r['launches'] = v['launches']
    return name, r


def _measured_traffic(kclass):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, 'profiles', 'traffic.json')
    try:
        with open(path) as fh:
            return json.load(fh).get(kclass, {}).get('bytes_per_launch')
    except (OSError, ValueError):
        return None


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--workload', default='octane', choices=sorted(WORKLOADS))
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--seed', type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py: this engine runs one G0W0 solve per GPU on a single MI355X; multi-GPU runs are '
              'not part of this build (each solve is independent: shard requests across GPUs).',
              file=sys.stderr)
        sys.exit(2)
    from molgw_amd.workloads import synthetic_system, synthetic_system_device
    nbf, naux, nocc, technique, desc = WORKLOADS[args.workload]
    t0 = time.perf_counter()
    if naux * nbf * nbf > 1e9:
        system = synthetic_system_device(nbf, naux, nocc, seed=args.seed, name=args.workload)
    else:
        system = synthetic_system(nbf, naux, nocc, seed=args.seed, name=args.workload)
    print(f'# generated {args.workload}: nbf={nbf} naux={naux} nocc={nocc} in {time.perf_counter() - t0:.1f}s',
          file=sys.stderr, flush=True)
    kw = {}
    if technique == 'imaginary_pade':
        from molgw_amd.gw import selfenergy_set_state_range
        rng = selfenergy_set_state_range(system.occupation)
        rng.nsemin, rng.nsemax = rng.nhomo_G - 7, rng.nhomo_G + 8
        kw = dict(rng=rng, nomega_chi_imag=32, nomega_sigma_calc=16)
    elapsed, lat, prof, out, mem = gpu_bench(system, technique, args.steps, args.warmup, kw)
    print(f'# gpu: {args.steps} solves in {elapsed:.3f}s; per-class device ms: '
          + ', '.join(f'{k}={v["ms"]:.1f}' for k, v in prof.items()), file=sys.stderr, flush=True)
    kname, roof = roofline_of(prof)
    cpu = None
    if not args.no_cpu_baseline and technique == 'one_shot':
        print('# cpu baseline (stage-sampled oracle) ...', file=sys.stderr, flush=True)
        cpu = cpu_baseline(system)
        print(f'# cpu: {cpu["seconds_per_solve"]:.1f} s per solve, stages {cpu["stage_seconds"]}',
              file=sys.stderr, flush=True)
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
        'dtype': 'f64',
        'data': 'synthetic (seeded RI tensor, orthonormal C and energies at the molecule\'s shape; '
                'no integral engine in the image)',
        'config': {'workload': desc, 'technique': technique, 'precision': 'fp64',
                   'nbf': nbf, 'naux': naux, 'nocc': nocc, 'nstate': nbf},
        'latency_ms': {'p50': 1e3 * _percentile(lat, 50), 'p99': 1e3 * _percentile(lat, 99)},
        'roofline': roof,
        'stages_ms_per_solve': {k: v['ms'] / args.steps for k, v in prof.items()},
        'hbm_peak_bytes': mem[1],
        'cpu_baseline': cpu,
        'result_check': {'homo_qp_ha': float(out.energy_qp[nocc - 1]), 'lumo_qp_ha': float(out.energy_qp[nocc]),
                         'en_rpa': out.en_rpa, 'en_gw': out.en_gw},
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
