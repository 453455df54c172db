"""Continuous-batching G0W0 server on one GPU.

Requests (one molecule each) arrive over time. The server holds ``nslots``
libmgw contexts; each context has its own HIP stream and solver handle, so
solves admitted to different slots run concurrently on the device. A request
is admitted when it has arrived, a slot is free and its HBM footprint fits
the budget next to what the other slots hold (``mgw_footprint_bytes``,
seeded by the reference's evaluate_memory formulas, src/m_io.f90:3843-3936).
When the head of the queue does not fit, smaller requests behind it may go
first; idle slots give back their cached workspaces when memory is short.

The RI tensor of every request stays resident in HBM from submission (the
caller's device buffer); a slot unpacks it into its own working copy when
the request starts. Results carry arrival, start and end times; latency is
end - arrival (queueing included).
"""
from __future__ import annotations

import os
import queue
import threading
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .gw import PRECISION, Engine, selfenergy_set_state_range

MGW_TECH = {'one_shot': _lib.MGW_TECH_ONE_SHOT, 'imaginary_pade': _lib.MGW_TECH_IMAGINARY_PADE}


@dataclass
class Request:
    rid: int
    system: object                 # workloads.System (host) or workloads.DeviceSystem
    arrival: float = 0.0           # seconds after the start of the run
    technique: str = 'imaginary_pade'
    precision: str = 'fp16x3'
    nomega_chi_imag: int = 32
    nomega_sigma_calc: int = 16
    window: int = 8                # states HOMO-window+1 .. LUMO+window-1 (nse = 2 window)
    footprint: int = 0


@dataclass
class Result:
    rid: int
    energy_qp: np.ndarray
    zz: np.ndarray
    en_rpa: float
    en_gw: float
    arrival: float
    start: float
    end: float
    slot: int
    nsemin: int = 0
    nsemax: int = 0
    ingest: float = 0.0            # seconds of the solve spent loading its RI tensor and MO data

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def _se_range(s, window: int) -> tuple[int, int]:
    """The self-energy states (1-based, inclusive) a request's solve computes:
    HOMO-window+1 .. LUMO+window-1 inside the default range."""
    rng = selfenergy_set_state_range(s.occupation)
    return max(rng.nsemin, rng.nhomo_G - window + 1), min(rng.nsemax, rng.nhomo_G + window)


def footprint_bytes(req: Request) -> int:
    """HBM of the request's solve (mgw_footprint_bytes_nse on its
    self-energy window)."""
    s = req.system
    lib = _lib.load()
    lo, hi = _se_range(s, req.window)
    return int(lib.mgw_footprint_bytes_nse(s.nbf, s.naux, s.nstate, s.nocc, hi - lo + 1, MGW_TECH[req.technique],
                                           req.nomega_chi_imag, PRECISION[req.precision]))


def device_memory_bytes(device: int = 0) -> int:
    from .gw import device_memory
    return device_memory(device)[1]


class Server:
    def __init__(self, nslots: int = 8, device: int = 0, hbm_budget: int | None = None, graphs: bool = True):
        self.device = device
        self.slots = [Engine(device) for _ in range(nslots)]
        for e in self.slots:
            e.graphs(graphs)
        self.budget = hbm_budget if hbm_budget is not None else int(0.80 * device_memory_bytes(device))
        # the device pool that every slot allocates from gives memory back
        # above the budget at each synchronisation
        self.slots[0].pool_limit(self.budget)
        self.trims = 0       # cached-buffer releases forced by admission
        self._jobs = [queue.Queue() for _ in range(nslots)]
        self._done = queue.Queue()
        # slot threads with 64 MiB stacks: the library's calls (and a profiler's
        # callbacks on these threads) run on them
        old_stack = threading.stack_size(64 << 20)
        try:
            self._threads = [threading.Thread(target=self._worker, args=(k,), daemon=True) for k in range(nslots)]
        finally:
            threading.stack_size(old_stack)
        for t in self._threads:
            t.start()
        self._reserved = [0] * nslots
        self._busy = [False] * nslots

    # ------------------------------------------------------------ slots
    def _worker(self, k: int) -> None:
        eng = self.slots[k]
        while True:
            job = self._jobs[k].get()
            if job is None:
                return
            req, t0 = job
            try:
                res = self._solve(eng, req, k, t0)
                self._done.put((k, res, None))
            except Exception as exc:          # reported to the caller of run()
                self._done.put((k, req, exc))

    def _solve(self, eng: Engine, req: Request, k: int, t0: float) -> Result:
        s = req.system
        start = time.perf_counter() - t0
        if hasattr(s, 'eri_3center_dev'):
            eng.set_eri_3center_device(s.eri_3center_dev.data_ptr(), s.npair, s.naux, s.index_basis, s.nbf)
        else:
            eng.set_eri_3center(s.eri_3center, s.index_basis, s.nbf)
        eng.set_mo(s.c_matrix, s.energy, s.occupation)
        ingest = time.perf_counter() - t0 - start
        rng = selfenergy_set_state_range(s.occupation)
        rng.nsemin, rng.nsemax = _se_range(s, req.window)
        out = eng.g0w0(rng, technique=req.technique, precision=req.precision,
                       nomega_chi_imag=req.nomega_chi_imag, nomega_sigma_calc=req.nomega_sigma_calc,
                       exchange_m_vxc=getattr(s, 'exchange_m_vxc', None), return_sigma=False)
        end = time.perf_counter() - t0
        return Result(req.rid, out.energy_qp, out.zz, out.en_rpa, out.en_gw, req.arrival, start, end, k,
                      rng.nsemin, rng.nsemax, ingest)

    def _slot_bytes(self, k: int) -> int:
        cur, _ = self.slots[k].memory()
        return max(cur, self._reserved[k])

    def _used(self) -> int:
        return sum(self._slot_bytes(k) for k in range(len(self.slots)))

    def _others(self, k: int) -> int:
        """HBM the other slots hold or have been promised: their ledgers, or
        the footprints of the requests they run, whichever is larger. Freed
        buffers that the device pool keeps for reuse are not counted: any slot
        can allocate from them, and the pool's release threshold (set to the
        budget) hands what exceeds the budget back to the device at each
        synchronisation. Memory of the process outside this server (other
        contexts, other libraries) is outside the budget."""
        return sum(self._slot_bytes(j) for j in range(len(self.slots)) if j != k)

    def _trim_until(self, k: int, req) -> None:
        """Free the cached buffers of idle slots, largest first, until ``req``
        fits on slot ``k``. Freeing device memory waits for the whole device,
        so every in-flight solve stalls on it: free no more than needed."""
        def fits():
            return self._others(k) + max(req.footprint, self.slots[k].memory()[0]) <= self.budget
        idle = sorted((j for j in range(len(self.slots)) if not self._busy[j] and j != k),
                      key=lambda j: -self.slots[j].memory()[0])
        for j in [k] + idle:
            if fits():
                return
            if self.slots[j].memory()[0] > 0:
                self.slots[j].trim()
                self.trims += 1

    # ------------------------------------------------------------ run
    def run(self, requests: list[Request], realtime: bool = True, order: str | None = None) -> list[Result]:
        """Serve ``requests`` (arrival times in seconds). With ``realtime``
        False every request is available at t = 0 (a batch). ``order`` of the
        waiting queue: 'arrival' (FIFO) or 'largest_first' (by HBM footprint,
        a proxy of the work: the longest solves start first, so the end of a
        batch is not one large solve running alone). Default: 'arrival' for
        realtime streams, the MGW_BATCH_ORDER environment variable or
        'arrival' for batches."""
        if order is None:
            order = 'arrival' if realtime else os.environ.get('MGW_BATCH_ORDER', 'arrival')
        if order not in ('arrival', 'largest_first'):
            raise ValueError(f'unknown order {order!r}')
        for r in requests:
            r.footprint = footprint_bytes(r)
            if r.footprint > self.budget:
                raise ValueError(f'request {r.rid}: footprint {r.footprint} exceeds the HBM budget {self.budget}')
        if order == 'largest_first':
            pending = sorted(requests, key=lambda r: (r.arrival if realtime else 0.0, -r.footprint))
        else:
            pending = sorted(requests, key=lambda r: r.arrival if realtime else 0.0)
        waiting: list[Request] = []
        results: list[Result] = []
        t0 = time.perf_counter()
        inflight = 0
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                waiting.append(pending.pop(0))
            # admission: FIFO, smaller requests may pass a head that does not fit
            for req in list(waiting):
                free = [k for k in range(len(self.slots)) if not self._busy[k]]
                if not free:
                    break
                # the free slot whose cached buffers fit the request best: the
                # least memory after it grows to the request's footprint, then
                # the most buffers to reuse
                cache = {j: self.slots[j].memory()[0] for j in free}
                k = min(free, key=lambda j: (max(req.footprint, cache[j]), -cache[j]))
                others = self._others(k)
                need = max(req.footprint, cache[k])
                if others + need > self.budget:
                    # freeing idle slots' caches waits for the whole device:
                    # only when that makes room next to the running solves
                    busy = sum(self._slot_bytes(j) for j in range(len(self.slots)) if self._busy[j])
                    if busy + req.footprint > self.budget:
                        continue
                    self._trim_until(k, req)
                    others = self._others(k)
                    need = max(req.footprint, self.slots[k].memory()[0])
                if others + need > self.budget:
                    continue
                waiting.remove(req)
                self._busy[k] = True
                self._reserved[k] = req.footprint
                inflight += 1
                self._jobs[k].put((req, t0))
            # wait for a completion or the next arrival
            timeout = None
            if pending and realtime:
                timeout = max(0.0, pending[0].arrival - (time.perf_counter() - t0))
            if inflight == 0 and timeout is None and waiting:
                raise RuntimeError('requests cannot be admitted within the HBM budget')
            try:
                k, res, exc = self._done.get(timeout=timeout)
            except queue.Empty:
                continue
            self._busy[k] = False
            self._reserved[k] = 0
            inflight -= 1
            if exc is not None:
                raise RuntimeError(f'request {res.rid} failed on slot {k}') from exc
            results.append(res)
        return sorted(results, key=lambda r: r.rid)

    # ------------------------------------------------------------ profiling
    def profile(self, enable: bool = True) -> None:
        """Per-kernel-class device timing on every slot (HIP events on the
        slot's own stream); enabling resets the counters."""
        for e in self.slots:
            e.profile(enable)

    def profile_read(self) -> dict:
        """Per-class totals over all slots. Slots run concurrently, so the
        summed device time can exceed the wall time."""
        out: dict = {}
        for e in self.slots:
            for name, v in e.profile_read().items():
                t = out.setdefault(name, {'launches': 0, 'ms': 0.0, 'flops': 0.0, 'bytes': 0.0})
                for k in t:
                    t[k] += v[k]
        return out

    def escalations(self) -> int:
        return sum(e.precision_escalations() for e in self.slots)

    def close(self) -> None:
        for q in self._jobs:
            q.put(None)
        for t in self._threads:
            t.join()
        for e in self.slots:
            e.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def summarize(results: list[Result]) -> dict:
    lat = np.array([r.latency for r in results])
    t_first = min(r.arrival for r in results)
    t_last = max(r.end for r in results)
    wall = t_last - t_first
    return {'solves': len(results), 'wall_s': wall,
            'solves_per_s': len(results) / wall,
            'p50_ms': float(np.percentile(lat, 50) * 1e3), 'p99_ms': float(np.percentile(lat, 99) * 1e3),
            'mean_ms': float(lat.mean() * 1e3),
            # slots solving on average (time-weighted), and the share of that
            # time spent loading RI tensors and MO data
            'busy_slots_mean': float(sum(r.end - r.start for r in results) / wall),
            'ingest_share': float(sum(r.ingest for r in results) / max(1e-12, sum(r.end - r.start for r in results)))}


# ---------------------------------------------------------------- streams
def poisson_stream(count: int, rate: float, seed: int = 0, alkane_n=(2, 24), water_m=(1, 32)):
    """(arrival_s, kind, size) of a seeded Poisson stream: each request is an
    alkane CnH2n+2 (n uniform in alkane_n) or a water cluster (H2O)m (m
    uniform in water_m), with equal probability."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, count))
    out = []
    for k in range(count):
        if rng.random() < 0.5:
            out.append((float(t[k]), 'alkane', int(rng.integers(alkane_n[0], alkane_n[1] + 1))))
        else:
            out.append((float(t[k]), 'water', int(rng.integers(water_m[0], water_m[1] + 1))))
    return out


def alkane_batch(count: int = 64, seed: int = 0, n_range=(1, 16)):
    """Sizes n of a seeded batch of alkanes CnH2n+2 (n uniform in n_range)."""
    rng = np.random.default_rng(seed)
    return [int(v) for v in rng.integers(n_range[0], n_range[1] + 1, count)]
