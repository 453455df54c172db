"""Host-side mirror of MOLGW's RI G0W0 interfaces, executed on the GPU.

``Engine`` holds one libmgw context (one GPU) and exposes the reference's
subroutines under their own names with the same argument meaning:

================================  ==============================================
reference (src/ of MOLGW 3.4)     here
================================  ==============================================
calculate_eri_ri result           ``Engine.set_eri_3center``
  (m_eri.f90:36-62)
calculate_eri_3center_mo          ``Engine.calculate_eri_3center_mo``
  (m_eri_ao_mo.f90:416-582)
polarizability (RPA, Casida)      ``Engine.polarizability``
  (m_linear_response.f90:33-498)
polarizability_grid_scalapack     ``Engine.polarizability_grid``
  (m_gw_selfenergy_grid.f90:33-209)
gw_selfenergy                     ``Engine.gw_selfenergy``
  (m_gw_selfenergy_analytic.f90:29-196)
gw_selfenergy_imag_scalapack      ``Engine.gw_selfenergy_imag``
  (m_gw_selfenergy_grid.f90:213-338)
se_pade_fit (m_selfenergy_tools.  ``Engine.pade_fit``
  f90:965-999)
find_qp_energy_graphical /        ``Engine.find_qp_energy``
  _linearization (:202-420)
selfenergy_evaluation one-shot    ``Engine.g0w0``
  (m_selfenergy_evaluation.f90)
selfenergy_set_state_range        ``selfenergy_set_state_range``
  (m_selfenergy_tools.f90:67-122)
================================  ==============================================

Arrays use the reference's column-major meaning: ``eri_3center`` is
(npair, nauxil), ``c_matrix`` is (nbf, nstate), ``eri_3center_mo`` is
(nauxil, m, n) and ``sigma`` is (nomega, nse) complex. Errors raise
``MgwError`` (the reference calls ``die``).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import G0W0Params, Ranges, check

COMPLETELY_EMPTY = 1.0e-5   # src/m_definitions.f90:66
HA_EV = 27.21138505         # src/m_definitions.f90:43

PRECISION = {'fp64': _lib.MGW_FP64, 'fp32': _lib.MGW_FP32, 'bf16x3': _lib.MGW_BF16X3, 'bf16x6': _lib.MGW_BF16X6,
             'fp16x3': _lib.MGW_FP16X3}


def _f64(a) -> np.ndarray:
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class DeviceBuffer:
    """Device memory from libmgw (mgw_device_alloc), for arrays the host side
    hands to the entry points with where = device. ``data_ptr()`` is the
    device address; the buffer is freed by ``free()`` or when collected."""

    def __init__(self, nbytes: int, device: int = 0):
        self._lib = _lib.load()
        p = C.c_void_p()
        check(self._lib.mgw_device_alloc(int(device), int(nbytes), C.byref(p)))
        self.ptr = p.value or 0
        self.nbytes = int(nbytes)
        self.device = device

    def data_ptr(self, offset: int = 0) -> int:
        if not self.ptr:
            raise ValueError('device buffer already freed')
        if not 0 <= offset <= self.nbytes:
            raise ValueError(f'offset {offset} outside the buffer of {self.nbytes} bytes')
        return self.ptr + offset

    def upload(self, host: np.ndarray, offset: int = 0) -> None:
        """Copy a contiguous host array to byte ``offset`` of the buffer."""
        a = np.ascontiguousarray(host) if not (host.flags.c_contiguous or host.flags.f_contiguous) else host
        if offset + a.nbytes > self.nbytes:
            raise ValueError('upload past the end of the device buffer')
        check(self._lib.mgw_device_copy(C.c_void_p(self.data_ptr(offset)), C.c_void_p(a.ctypes.data), a.nbytes, 0))

    def download(self, out: np.ndarray, offset: int = 0) -> np.ndarray:
        """Copy bytes from ``offset`` into the contiguous host array ``out``."""
        if not (out.flags.c_contiguous or out.flags.f_contiguous):
            raise ValueError('download needs a contiguous host array')
        if offset + out.nbytes > self.nbytes:
            raise ValueError('download past the end of the device buffer')
        check(self._lib.mgw_device_copy(C.c_void_p(out.ctypes.data), C.c_void_p(self.data_ptr(offset)), out.nbytes, 1))
        return out

    def free(self) -> None:
        if self.ptr:
            check(self._lib.mgw_device_free(C.c_void_p(self.ptr)))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_memory(device: int = 0) -> tuple[int, int]:
    """(free, total) bytes of the device's memory."""
    lib = _lib.load()
    f, t = C.c_int64(), C.c_int64()
    check(lib.mgw_device_memory(int(device), C.byref(f), C.byref(t)))
    return f.value, t.value


@dataclass
class StateRanges:
    """1-based state ranges of the reference (nsemin, ncore_G ...)."""
    ncore_G: int
    nvirtual_G: int
    nhomo_G: int
    nlumo_G: int
    nsemin: int
    nsemax: int
    ncore_W: int
    nvirtual_W: int

    def to_c(self) -> Ranges:
        return Ranges(self.ncore_G, self.nvirtual_G, self.nsemin, self.nsemax, self.nhomo_G,
                      self.ncore_W, self.nvirtual_W, self.nlumo_G if self.nlumo_G <= self.nvirtual_G else 0)

    @property
    def nse(self) -> int:
        return self.nsemax - self.nsemin + 1


def selfenergy_set_state_range(occupation, ncoreg=0, nvirtualg=100000, ncorew=0, nvirtualw=100000,
                               selfenergy_state_min=1, selfenergy_state_max=100000,
                               selfenergy_state_range=100000) -> StateRanges:
    """State ranges as set by selfenergy_set_state_range
    (src/m_selfenergy_tools.f90:67-122) and sf_init (src/m_spectral_function.f90:183-184),
    for occupation (nstate,) or (nstate, nspin). ``selfenergy_state_range``
    counts from the HOMO."""
    occ = np.asarray(occupation, dtype=np.float64)
    occ = occ.reshape(len(occ), -1)                      # (nstate, nspin)
    nstate = occ.shape[0]
    ncore_G = ncoreg
    nvirtual_G = min(nvirtualg, nstate + 1)
    # HOMO: occupied in every channel; LUMO: empty in every channel
    filled = np.nonzero(~np.any(occ < COMPLETELY_EMPTY, axis=1))[0]
    nhomo_G = max(1, int(filled.max()) + 1) if len(filled) else 1
    empty = np.nonzero(~np.any(occ > COMPLETELY_EMPTY, axis=1))[0]
    nlumo_G = int(empty.min()) + 1 if len(empty) else 100000
    nsemin = max(ncore_G + 1, selfenergy_state_min, 1, nhomo_G - selfenergy_state_range)
    nsemax = min(nvirtual_G - 1, selfenergy_state_max, nstate, nhomo_G + selfenergy_state_range)
    return StateRanges(ncore_G, nvirtual_G, nhomo_G, nlumo_G, nsemin, nsemax, ncorew,
                       min(nvirtualw, nstate + 1))


def autotune(enable: bool = True) -> None:
    """Turn the GEMM tile autotuner on (tune each new shape class on first use)
    or off (use the tuned table / measured defaults). Process-wide."""
    check(_lib.load().mgw_autotune(1 if enable else 0, None))


def imaginary_quad_grid(nomega: int):
    """Gauss-Legendre grid of chi on the imaginary axis (alpha=1, beta=2):
    src/m_spectral_function.f90:286-306, nodes from numpy's leggauss."""
    x, w = np.polynomial.legendre.leggauss(nomega)
    x = 0.5 * (x + 1.0)
    w = 0.5 * w
    return 2.0 * x / (1.0 - x), 2.0 * w / (1.0 - x) ** 2


@dataclass
class G0W0Output:
    energy_qp: np.ndarray
    zz: np.ndarray
    energy_qp_lin: np.ndarray
    sigma: np.ndarray
    omega: np.ndarray
    en_rpa: float
    en_gw: float
    extra: dict = field(default_factory=dict)


class Engine:
    """One libmgw context on one GPU (the reference's module state, resident in HBM)."""

    def __init__(self, device: int = 0):
        self._lib = _lib.load()
        h = C.c_void_p()
        check(self._lib.mgw_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.device = int(device)
        self.nbf = self.naux = self.nstate = 0
        self._naux_kept = self._naux_bas = 0
        self.nspin = 1
        self._npole_w = 0
        self.energy = self.occupation = None

    # ------------------------------------------------------------ lifetime
    def close(self) -> None:
        if getattr(self, '_h', None) is not None and self._h.value:
            check(self._lib.mgw_ctx_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def synchronize(self) -> None:
        check(self._lib.mgw_ctx_synchronize(self._h))

    def trim(self, eager: bool = False) -> None:
        """Free cached temporaries (keeps the resident RI tensor and MO data)."""
        check(self._lib.mgw_ctx_trim(self._h, 1 if eager else 0))

    def profile(self, enable: bool) -> None:
        """Start (clearing earlier records) or stop live per-kernel-class timing."""
        check(self._lib.mgw_profile_enable(self._h, 1 if enable else 0))

    def profile_read(self) -> dict:
        """{class name: {launches, ms, flops, bytes}} over the recorded launches."""
        out = {}
        for k, name in enumerate(_lib.KCLASS):
            n, ms, fl, by = C.c_int64(), C.c_double(), C.c_double(), C.c_double()
            check(self._lib.mgw_profile_read(self._h, k, C.byref(n), C.byref(ms), C.byref(fl), C.byref(by)))
            if n.value:
                out[name] = {'launches': n.value, 'ms': ms.value, 'flops': fl.value, 'bytes': by.value}
        return out

    def memory(self):
        cur, peak = C.c_int64(), C.c_int64()
        check(self._lib.mgw_memory_report(self._h, C.byref(cur), C.byref(peak)))
        return cur.value, peak.value

    def pool_memory(self, reset_high: bool = False):
        """(reserved, reserved high-water) bytes of the device's memory pool,
        which every engine of the process allocates from; reset_high restarts
        the high-water mark from the current size after reading it."""
        cur, high = C.c_int64(), C.c_int64()
        check(self._lib.mgw_device_pool_report(self.device, 1 if reset_high else 0, C.byref(cur), C.byref(high)))
        return cur.value, high.value

    def pool_limit(self, nbytes: int) -> None:
        """Release threshold of the device's memory pool (shared by every
        engine of the process): above it, unused memory goes back to the
        device at each synchronisation."""
        check(self._lib.mgw_device_pool_limit(self.device, int(nbytes)))

    def set_energy_g(self, energy_g=None) -> None:
        """Energies in G for the following self-energy calls (W is kept);
        None returns to the energies of set_mo."""
        if energy_g is None:
            check(self._lib.mgw_set_energy_g(self._h, None))
            return
        e = np.asfortranarray(np.asarray(energy_g, dtype=np.float64)).reshape(-1, order='F')
        if len(e) != self.nstate * self.nspin:
            raise ValueError('energy_g must hold nstate values per spin channel')
        e = np.ascontiguousarray(e)
        check(self._lib.mgw_set_energy_g(self._h, _ptr(e)))

    def evgw(self, nstep: int, approx: str = 'GnWn', rng: StateRanges | None = None, exchange_m_vxc=None,
             eta: float = 0.001, spin_fact: float = 2.0, precision: str = 'fp64'):
        """Eigenvalue self-consistent GW with analytic W (postscf GNW0 / GNWN =
        EVGW; selfenergy_evaluation with EVSC, src/m_selfenergy_evaluation.f90:78-182,
        589-606). Step n evaluates Sigma_c at the QP energies of step n-1 (step 1:
        the set_mo energies) with G built from them; W is rebuilt from them for
        GnWn and kept from step 1 for GnW0. E(p) = e(p) + Re Sigma_c(p) +
        (Sigma_x - Vxc)(p) in [nsemin, nsemax] (find_qp_energy_linearization
        without a grid, src/m_selfenergy_tools.f90:236-240), e(p) elsewhere.
        Returns the QP energies of every step; the engine is left with the
        set_mo energies."""
        if approx not in ('GnW0', 'GnWn'):
            raise ValueError('approx must be GnW0 or GnWn')
        if rng is None:
            rng = selfenergy_set_state_range(self.occupation)
        e = self.energy.copy()
        occ = self.occupation.copy()
        exc = np.zeros(len(e)) if exchange_m_vxc is None else _f64(exchange_m_vxc).reshape(-1)
        win = slice(rng.nsemin - 1, rng.nsemax)
        eg = e.copy()
        history = []
        try:
            for step in range(nstep):
                if approx == 'GnWn':
                    self.set_mo(self._c_matrix, eg, occ)          # energies in W (and in G)
                    self.polarizability(rng, spin_fact=spin_fact, precision=precision)
                elif step == 0:
                    self.polarizability(rng, spin_fact=spin_fact, precision=precision)
                self.set_energy_g(eg)
                sig, _ = self.gw_selfenergy(rng, np.zeros(1), eg[win], eta=eta, spin_fact=spin_fact)
                new = e.copy()
                new[win] = e[win] + sig[0].real + exc[win]
                history.append(new)
                eg = new
        finally:
            self.set_mo(self._c_matrix, e, occ)
        return history

    def graphs(self, enable: bool = True) -> None:
        """Replay the batched Cholesky sequence from a captured HIP graph."""
        check(self._lib.mgw_ctx_graphs(self._h, 1 if enable else 0))

    def precision_escalations(self) -> int:
        """Frequencies whose reduced-precision 1 - chi0 was not positive
        definite and were rebuilt in FP64 (since the engine was created)."""
        n = C.c_int64()
        check(self._lib.mgw_precision_escalations(self._h, C.byref(n)))
        return n.value

    # ------------------------------------------------------------ A1 / MO
    def set_eri_3center(self, eri_3center, index_basis, nbf: int) -> None:
        """Upload eri_3center(npair, nauxil) (host, packed, diagonal pairs x1/2)."""
        b = _f64(eri_3center)
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        npair, naux = b.shape
        if ib.shape != (2, npair):
            raise ValueError(f'index_basis must be (2, {npair}), got {ib.shape}')
        check(self._lib.mgw_set_eri_3center(self._h, b.ctypes.data, npair, naux,
                                            ib.ctypes.data_as(C.POINTER(C.c_int32)), int(nbf),
                                            _lib.MGW_MEM_HOST))
        self.nbf, self.naux = int(nbf), int(naux)

    def set_eri_3center_device(self, ptr: int, npair: int, naux: int, index_basis, nbf: int) -> None:
        """Same as set_eri_3center for a device buffer (column-major npair x nauxil)."""
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        check(self._lib.mgw_set_eri_3center(self._h, C.c_void_p(ptr), int(npair), int(naux),
                                            ib.ctypes.data_as(C.POINTER(C.c_int32)), int(nbf),
                                            _lib.MGW_MEM_DEVICE))
        self.nbf, self.naux = int(nbf), int(naux)

    def inverse_sqrt_eri_2center(self, eri_2center, copy: bool = True):
        """calculate_inverse_sqrt_eri_2center_scalapack
        (src/m_eri_calculate.f90:1049-1262): FP64 eigen-decomposition of the
        2-center metric on the device, eigenvalues <= 1e-6 dropped. Returns
        (nauxil_kept, lowest eigenvalue, sqrtinv(nauxil_bas, nauxil_kept) or
        None); sqrtinv stays resident for build_eri_3center."""
        v = _f64(eri_2center)
        n = v.shape[0]
        if v.shape != (n, n):
            raise ValueError(f'eri_2center must be square, got {v.shape}')
        out = np.zeros((n, n), order='F') if copy else None
        kept = C.c_int(0)
        low = C.c_double(0.0)
        check(self._lib.mgw_inverse_sqrt_eri_2center(self._h, v.ctypes.data, n, _lib.MGW_MEM_HOST,
                                                     C.byref(kept), C.byref(low),
                                                     out.ctypes.data if copy else None))
        k = kept.value
        self._naux_kept, self._naux_bas = k, n
        return k, low.value, (np.asfortranarray(out[:, :k]) if copy else None)

    def build_eri_3center(self, eri_3center_tmp, index_basis, nbf: int, batch: int | None = None,
                          copy: bool = True):
        """B = eri_3center_tmp . sqrtinv (src/m_eri_calculate.f90:1877-1892) in
        batches of `batch` pairs (all at once by default); eri_3center_tmp is
        (npair, nauxil_bas) with the diagonal pairs already halved. The result
        becomes the resident RI tensor; returns eri_3center(npair, nauxil_kept)
        when copy is set."""
        t = _f64(eri_3center_tmp)
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        npair, nb = t.shape
        if ib.shape != (2, npair):
            raise ValueError(f'index_basis must be (2, {npair}), got {ib.shape}')
        kept = getattr(self, '_naux_kept', 0)
        if not kept:
            raise RuntimeError('no 2-center metric: call inverse_sqrt_eri_2center first')
        if nb != self._naux_bas:
            raise ValueError(f'eri_3center_tmp has {nb} auxiliary columns, the metric {self._naux_bas}')
        out = np.zeros((npair, kept), order='F') if copy else None
        step = npair if batch is None else max(1, int(batch))
        first = 0
        while first < npair:
            m = min(step, npair - first)
            chunk = np.asfortranarray(t[first:first + m])
            check(self._lib.mgw_build_eri_3center(self._h, chunk.ctypes.data, first + 1, m, npair,
                                                  ib.ctypes.data_as(C.POINTER(C.c_int32)), int(nbf),
                                                  _lib.MGW_MEM_HOST, out.ctypes.data if copy else None))
            first += m
        self.nbf, self.naux = int(nbf), kept
        return out

    def build_eri_3center_device(self, ptr: int, ipair_first: int, mpair: int, npair: int, index_basis,
                                 nbf: int, out_ptr: int | None = None) -> None:
        """build_eri_3center for one batch already in HBM: ``ptr`` is
        eri_3center_tmp(mpair, nauxil_bas) column-major on the device, pairs
        ipair_first.. (1-based); ``out_ptr`` (optional, device or host) is the
        whole eri_3center(npair, nauxil_kept)."""
        if not self._naux_kept:
            raise RuntimeError('no 2-center metric: call inverse_sqrt_eri_2center first')
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        check(self._lib.mgw_build_eri_3center(self._h, C.c_void_p(ptr), int(ipair_first), int(mpair), int(npair),
                                              ib.ctypes.data_as(C.POINTER(C.c_int32)), int(nbf),
                                              _lib.MGW_MEM_DEVICE, None if out_ptr is None else C.c_void_p(out_ptr)))
        if ipair_first - 1 + mpair == npair:
            self.nbf, self.naux = int(nbf), self._naux_kept

    # ------------------------------------------------------------ integral front end
    @staticmethod
    def _shells(bs):
        """(keep-alive arrays, mgw_shells) of a frontend BasisSet."""
        sh = bs.shells
        l = np.array([x.l for x in sh], dtype=np.int32)
        x0 = np.ascontiguousarray(np.array([x.center for x in sh], dtype=np.float64).reshape(-1))
        poff = np.cumsum([0] + [len(x.exps) for x in sh]).astype(np.int32)
        al = np.ascontiguousarray(np.concatenate([x.exps for x in sh]), dtype=np.float64)
        co = np.ascontiguousarray(np.concatenate([x.coefs for x in sh]), dtype=np.float64)
        tm = [np.ascontiguousarray(x.transform, dtype=np.float64).reshape(-1) for x in sh]
        toff = np.cumsum([0] + [len(t) for t in tm[:-1]]).astype(np.int32)
        tmat = np.ascontiguousarray(np.concatenate(tm))
        foff = np.asarray(bs.offsets, dtype=np.int32)
        keep = (l, x0, poff, al, co, toff, tmat, foff)
        ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        st = _lib.Shells(len(sh), l.ctypes.data_as(ip), x0.ctypes.data_as(dp), poff.ctypes.data_as(ip),
                         al.ctypes.data_as(dp), co.ctypes.data_as(dp), toff.ctypes.data_as(ip),
                         tmat.ctypes.data_as(dp), foff.ctypes.data_as(ip))
        return keep, st

    def eri_2center_ao(self, auxil):
        """(P|Q) over the final functions of a BasisSet, computed on the GPU
        (calculate_integrals_eri_2center_scalapack, src/m_eri_calculate.f90:703-945)."""
        keep, st = self._shells(auxil)
        out = np.zeros((auxil.nbf, auxil.nbf), order='F')
        check(self._lib.mgw_eri_2center_ao(self._h, C.byref(st), _lib.MGW_MEM_HOST, out.ctypes.data))
        return out

    def eri_3center_ao(self, basis, auxil, index_basis, copy: bool = True):
        """eri_3center_tmp(npair, naux_bas) = (mu nu|P) of the pairs index_basis
        (2, npair), diagonal pairs x1/2, computed on the GPU
        (calculate_integrals_eri_3center_scalapack, src/m_eri_calculate.f90:1266-1580).
        With copy False the tensor stays resident for build_eri_3center_resident."""
        kb, sb = self._shells(basis)
        ka, sa = self._shells(auxil)
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        npair = ib.shape[1]
        out = np.zeros((npair, auxil.nbf), order='F') if copy else None
        check(self._lib.mgw_eri_3center_ao(self._h, C.byref(sb), C.byref(sa), npair,
                                           ib.ctypes.data_as(C.POINTER(C.c_int32)), _lib.MGW_MEM_HOST,
                                           out.ctypes.data if copy else None))
        return out

    def build_eri_3center_resident(self, index_basis, nbf: int, out_ptr: int | None = None,
                                   copy: bool = False):
        """build_eri_3center on the 3-center integrals that eri_3center_ao left
        resident: B becomes the resident RI tensor. ``out_ptr`` (device or
        host address) receives the packed eri_3center(npair, nauxil_kept);
        with copy set a host array is returned instead."""
        if not self._naux_kept:
            raise RuntimeError('no 2-center metric: call inverse_sqrt_eri_2center first')
        ib = np.asfortranarray(np.asarray(index_basis, dtype=np.int32))
        npair = ib.shape[1]
        out = np.zeros((npair, self._naux_kept), order='F') if copy else None
        dst = out.ctypes.data if copy else (None if out_ptr is None else C.c_void_p(out_ptr))
        check(self._lib.mgw_build_eri_3center(self._h, None, 1, npair, npair, ib.ctypes.data_as(C.POINTER(C.c_int32)),
                                              int(nbf), _lib.MGW_MEM_HOST, dst))
        self.nbf, self.naux = int(nbf), self._naux_kept
        return out

    def hartree_exchange_ri(self, x, spin_fact: float = 2.0):
        """(hartree_ao, exchange_ao) of the density spin_fact x x^T through the
        resident RI tensor (setup_hartree_ri / setup_exchange_ri,
        src/m_hamiltonian_twobodies.f90:280-381, 707-786); x (nbf, nx) are the
        occupied orbitals times sqrt(occupation / spin_fact)."""
        xf = _f64(x)
        nbf, nx = xf.shape
        j = np.zeros((nbf, nbf), order='F')
        k = np.zeros((nbf, nbf), order='F')
        check(self._lib.mgw_hartree_exchange_ri(self._h, _ptr(xf), nbf, nx, float(spin_fact), _ptr(j), _ptr(k)))
        return j, k

    def exchange_m_vxc(self, hamiltonian_fock):
        """setup_exchange_m_vxc (src/m_selfenergy_tools.f90:720-802): (C^T F C -
        diag(energy)) of every spin channel for the resident MO data; F is
        (nbf, nbf) or (nbf, nbf, nspin). Returns (matrix (nstate, nstate[, nspin]),
        diagonal (nstate[, nspin]))."""
        f = _f64(np.asarray(hamiltonian_fock, dtype=np.float64).reshape(self.nbf, self.nbf, -1))
        if f.shape[2] != self.nspin:
            raise ValueError('hamiltonian_fock must hold one nbf x nbf matrix per spin channel')
        mat = np.zeros((self.nstate, self.nstate, self.nspin), order='F')
        dg = np.zeros((self.nstate, self.nspin), order='F')
        check(self._lib.mgw_exchange_m_vxc(self._h, _ptr(f), _ptr(mat), _ptr(dg)))
        if self.nspin == 1:
            return mat[:, :, 0], dg[:, 0]
        return mat, dg

    def set_mo(self, c_matrix, energy, occupation) -> None:
        """MO coefficients c_matrix (nbf, nstate) or (nbf, nstate, nspin), energies
        and occupations (nstate,) or (nstate, nspin). With two channels W sums
        the transitions of both and the self-energy / QP calls work on the
        channel chosen by set_spin (g0w0 loops over both); pass spin_fact = 1."""
        c = np.asarray(c_matrix, dtype=np.float64)
        e = np.asarray(energy, dtype=np.float64)
        o = np.asarray(occupation, dtype=np.float64)
        nspin = c.shape[2] if c.ndim == 3 else 1
        nbf, nstate = c.shape[:2]
        e = e.reshape(nstate, -1) if e.size == nstate * nspin else None
        o = o.reshape(nstate, -1) if o.size == nstate * nspin else None
        if e is None or o is None or e.shape[1] != nspin or o.shape[1] != nspin:
            raise ValueError('energy/occupation must be (nstate,) or (nstate, nspin) matching c_matrix')
        cf = np.asfortranarray(c.reshape(nbf, nstate, nspin))
        ef, of = np.asfortranarray(e), np.asfortranarray(o)
        check(self._lib.mgw_set_mo_spin(self._h, cf.ctypes.data, nbf, nstate, nspin, ef.ctypes.data,
                                        of.ctypes.data, _lib.MGW_MEM_HOST))
        self.nstate, self.nspin = nstate, nspin
        if nspin == 1:
            self.energy, self.occupation = ef[:, 0].copy(), of[:, 0].copy()
            self._c_matrix = np.asfortranarray(cf[:, :, 0])
        else:
            self.energy, self.occupation = ef.copy(), of.copy()
            self._c_matrix = cf

    def set_spin(self, ispin: int) -> None:
        """Channel (1-based) of the AO->MO, self-energy and QP calls."""
        check(self._lib.mgw_set_spin(self._h, int(ispin)))

    # ------------------------------------------------------------ A2
    def calculate_eri_3center_mo(self, mstate_min, mstate_max, nstate_min, nstate_max, copy=True):
        out = None
        if copy:
            out = np.zeros((self.naux, mstate_max - mstate_min + 1, nstate_max - nstate_min + 1), order='F')
        check(self._lib.mgw_calculate_eri_3center_mo(self._h, mstate_min, mstate_max, nstate_min, nstate_max,
                                                     _ptr(out) if copy else None))
        return out

    def destroy_eri_3center_mo(self) -> None:
        check(self._lib.mgw_destroy_eri_3center_mo(self._h))

    def calculate_eri_3center_mo_noalloc(self, mstate_min, mstate_max, nstate_min, nstate_max,
                                         out: DeviceBuffer) -> None:
        """The same transform into the device buffer ``out`` (nauxil x m x n,
        column-major) with no device allocation (mgw_calculate_eri_3center_mo_noalloc);
        queued on the context's stream: call synchronize() before reading ``out``."""
        need = 8 * self.naux * (mstate_max - mstate_min + 1) * (nstate_max - nstate_min + 1)
        if out.nbytes < need:
            raise ValueError(f'output buffer holds {out.nbytes} bytes, the block needs {need}')
        check(self._lib.mgw_calculate_eri_3center_mo_noalloc(self._h, mstate_min, mstate_max, nstate_min,
                                                             nstate_max, C.c_void_p(out.data_ptr())))

    # ------------------------------------------------------------ A4
    def npole(self, rng: StateRanges) -> int:
        n = C.c_int()
        r = rng.to_c()
        check(self._lib.mgw_npole(self._h, C.byref(r), C.byref(n)))
        return n.value

    def polarizability(self, rng: StateRanges, spin_fact=2.0, return_ws=False, precision='fp64'):
        """Returns (en_rpa, en_gw, pole, w_s or None)."""
        r = rng.to_c()
        npole = self.npole(rng)
        pole = np.zeros(npole)
        ws = np.zeros((self.naux, npole), order='F') if return_ws else None
        erpa, egw = C.c_double(), C.c_double()
        check(self._lib.mgw_polarizability(self._h, C.byref(r), spin_fact, PRECISION[precision], C.byref(erpa),
                                           C.byref(egw), _ptr(pole), _ptr(ws) if return_ws else None))
        self._npole_w = npole
        return erpa.value, egw.value, pole, ws

    def set_spectral_function(self, pole, w_s) -> None:
        """read_spectral_function: make W resident from its poles and residues
        w_s(nauxil, npole) (e.g. from formats.read_screened_coulomb)."""
        p = _f64(pole).reshape(-1)
        w = _f64(w_s)
        if w.shape != (self.naux, len(p)):
            raise ValueError(f'w_s must be ({self.naux}, {len(p)}), got {w.shape}')
        check(self._lib.mgw_set_spectral_function(self._h, len(p), self.naux, p.ctypes.data, w.ctypes.data,
                                                  _lib.MGW_MEM_HOST))
        self._npole_w = len(p)

    def gw_ingredients(self, rng: StateRanges, with_w: bool = True):
        """dump_gw_ingredients arrays of the active spin channel: (w(npole, ng, ng)
        or None, v(nauxil, ng, ng)) for states ncore_G+1 .. nvirtual_G-1."""
        ng = rng.nvirtual_G - 1 - rng.ncore_G
        v = np.zeros((self.naux, ng, ng), order='F')
        npole = self._npole_w or self.npole(rng)
        w = np.zeros((npole, ng, ng), order='F') if with_w else None
        r = rng.to_c()
        check(self._lib.mgw_gw_ingredients(self._h, C.byref(r), _ptr(w) if with_w else None, _ptr(v)))
        return w, v

    # ------------------------------------------------------------ A5
    def polarizability_grid(self, rng: StateRanges, omega_im, weight, spin_fact=2.0, return_chi=False,
                            precision='fp64'):
        """Returns (erpa, egw, chi[naux, naux, nomega] or None)."""
        r = rng.to_c()
        om = _f64(omega_im)
        wt = _f64(weight)
        n = len(om)
        chi = np.zeros((self.naux, self.naux, n), order='F') if return_chi else None
        erpa, egw = C.c_double(), C.c_double()
        check(self._lib.mgw_polarizability_grid(self._h, C.byref(r), spin_fact, PRECISION[precision], n, _ptr(om),
                                                _ptr(wt), C.byref(erpa), C.byref(egw),
                                                _ptr(chi) if return_chi else None))
        return erpa.value, egw.value, chi

    # ------------------------------------------------------------ A6
    def gw_selfenergy(self, rng: StateRanges, omega, se_energy0, eta=0.001, spin_fact=2.0):
        """Returns (sigma[nomega, nse] complex, energy_gw)."""
        r = rng.to_c()
        om = _f64(omega)
        e0 = _f64(se_energy0)
        sig = np.zeros((len(om), rng.nse), dtype=np.complex128, order='F')
        egw = C.c_double()
        check(self._lib.mgw_gw_selfenergy(self._h, C.byref(r), spin_fact, eta, len(om), _ptr(om), _ptr(e0),
                                          sig.ctypes.data_as(C.POINTER(C.c_double)), C.byref(egw)))
        return sig, egw.value

    def gw_selfenergy_qs(self, rng: StateRanges, eta=0.001, spin_fact=2.0, s_matrix=None):
        """QSGW self-energy matrix of the active channel on the resident Casida W:
        (Sigma_MO[nstate, nstate] symmetrised, Sigma_AO[nbf, nbf] = S C Sigma C^T S
        or None without s_matrix)."""
        r = rng.to_c()
        mo = np.zeros((self.nstate, self.nstate), order='F')
        ao = np.zeros((self.nbf, self.nbf), order='F') if s_matrix is not None else None
        sm = _f64(s_matrix) if s_matrix is not None else None
        check(self._lib.mgw_gw_selfenergy_qs(self._h, C.byref(r), spin_fact, eta,
                                             _ptr(sm) if sm is not None else None, _ptr(mo),
                                             _ptr(ao) if ao is not None else None))
        return mo, ao

    # ------------------------------------------------------------ A7
    def gw_selfenergy_imag(self, rng: StateRanges, omega_calc):
        r = rng.to_c()
        wc = np.asfortranarray(np.asarray(omega_calc, dtype=np.complex128))
        out = np.zeros((len(wc), rng.nse), dtype=np.complex128, order='F')
        check(self._lib.mgw_gw_selfenergy_imag(self._h, C.byref(r), len(wc),
                                               wc.ctypes.data_as(C.POINTER(C.c_double)),
                                               out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # ------------------------------------------------------------ contour deformation
    def gw_selfenergy_contour(self, rng: StateRanges, omega, se_energy0, nomega_chi_imag: int,
                              nomega_chi_real: int = 2, eta=0.001, spin_fact=None):
        """gw_selfenergy_contour for every channel: returns (sigma[nomega, nse]
        (or [nomega, nse, nspin]) complex with zero imaginary part, de_max).
        ``se_energy0`` is (nse,) or (nse, nspin)."""
        r = rng.to_c()
        ns = getattr(self, 'nspin', 1)
        if spin_fact is None:
            spin_fact = 2.0 if ns == 1 else 1.0
        om = _f64(omega)
        e0 = np.asfortranarray(np.asarray(se_energy0, dtype=np.float64).reshape(rng.nse, ns, order='F'))
        shape = (len(om), rng.nse) if ns == 1 else (len(om), rng.nse, ns)
        sig = np.zeros(shape, dtype=np.complex128, order='F')
        de_max = C.c_double()
        check(self._lib.mgw_gw_selfenergy_contour(self._h, C.byref(r), spin_fact, nomega_chi_imag,
                                                  nomega_chi_real, eta, len(om), _ptr(om), _ptr(e0),
                                                  sig.ctypes.data_as(C.POINTER(C.c_double)), C.byref(de_max)))
        return sig, de_max.value

    # ------------------------------------------------------------ A8
    def pade_fit(self, omega_calc, sigma_calc, omega, eta=0.001):
        wc = np.asfortranarray(np.asarray(omega_calc, dtype=np.complex128))
        sc = np.asfortranarray(np.asarray(sigma_calc, dtype=np.complex128))
        om = _f64(omega)
        nse = sc.shape[1]
        out = np.zeros((len(om), nse), dtype=np.complex128, order='F')
        check(self._lib.mgw_pade_fit(self._h, nse, len(wc), wc.ctypes.data_as(C.POINTER(C.c_double)),
                                     sc.ctypes.data_as(C.POINTER(C.c_double)), len(om), _ptr(om), eta,
                                     out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    # ------------------------------------------------------------ A9
    def find_qp_energy(self, rng: StateRanges, omega, sigma, se_energy0, exchange_m_vxc, energy0):
        """Returns (energy_qp_graphical, zz, energy_qp_linearized), each (nstate,)."""
        r = rng.to_c()
        om = _f64(omega)
        sg = np.asfortranarray(np.asarray(sigma, dtype=np.complex128))
        e0s = _f64(se_energy0)
        exc = _f64(exchange_m_vxc)
        e0 = _f64(energy0)
        n = len(e0)
        eqp, zz, elin = np.zeros(n), np.zeros(n), np.zeros(n)
        check(self._lib.mgw_find_qp_energy(self._h, C.byref(r), len(om), _ptr(om),
                                           sg.ctypes.data_as(C.POINTER(C.c_double)), _ptr(e0s), _ptr(exc),
                                           _ptr(e0), n, _ptr(eqp), _ptr(zz), _ptr(elin)))
        return eqp, zz, elin

    # ------------------------------------------------------------ A11
    def g0w0(self, rng: StateRanges | None = None, technique='one_shot', nomega_sigma=51,
             step_sigma=0.01, eta=0.001, nomega_chi_imag=0, nomega_sigma_calc=1,
             step_sigma_calc=0.03, exchange_m_vxc=None, spin_fact=2.0, return_sigma=True,
             precision='fp64', nomega_chi_real=2) -> G0W0Output:
        """One-shot G0W0 (postscf='GW' for technique='one_shot', 'G0W0_PADE' for
        technique='imaginary_pade', 'G0W0_CONTOUR' for technique='contour', FP64
        only, with nomega_chi_imag and nomega_chi_real) on the resident tensors. ``precision``
        ('fp64', 'fp32', 'bf16x3', 'bf16x6') applies to the AO->MO transforms and the
        chi0 / Casida-matrix build; the eigensolve, the (1 - chi0) inversion,
        the self-energy products and the QP solve are FP64."""
        if rng is None:
            rng = selfenergy_set_state_range(self.occupation)
        tech = {'one_shot': _lib.MGW_TECH_ONE_SHOT, 'imaginary_pade': _lib.MGW_TECH_IMAGINARY_PADE,
                'contour': _lib.MGW_TECH_CONTOUR}[technique]
        prm = G0W0Params(tech, nomega_sigma, step_sigma, eta, nomega_chi_imag, nomega_sigma_calc,
                         step_sigma_calc, spin_fact, PRECISION[precision], nomega_chi_real)
        r = rng.to_c()
        n, ns = self.nstate, getattr(self, 'nspin', 1)
        nom = nomega_sigma // 2
        shape = (n,) if ns == 1 else (n, ns)
        eqp, zz, elin = (np.zeros(shape, order='F') for _ in range(3))
        sshape = (2 * nom + 1, rng.nse) if ns == 1 else (2 * nom + 1, rng.nse, ns)
        sig = np.zeros(sshape, dtype=np.complex128, order='F') if return_sigma else None
        en = np.zeros(2)
        exc = None if exchange_m_vxc is None else np.asfortranarray(np.asarray(exchange_m_vxc, dtype=np.float64))
        check(self._lib.mgw_g0w0(self._h, C.byref(prm), C.byref(r), _ptr(exc) if exc is not None else None,
                                 _ptr(eqp), _ptr(zz), _ptr(elin),
                                 sig.ctypes.data_as(C.POINTER(C.c_double)) if return_sigma else None,
                                 _ptr(en)))
        if technique in ('one_shot', 'contour'):
            omega = step_sigma * np.arange(-nom, nom + 1)
        else:
            e2 = self.energy.reshape(n, -1)
            efermi = 0.5 * (e2[rng.nhomo_G - 1].max() + e2[rng.nhomo_G].min())
            omega = efermi + step_sigma * np.arange(-nom, nom + 1)
        return G0W0Output(eqp, zz, elin, sig, omega, float(en[0]), float(en[1]))
