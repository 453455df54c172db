"""CPU FP64 restatement of MOLGW's RI G0W0 hot path: the oracle.

TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or, for ``cpu_baseline``, as the timed CPU port). The product path
(``molgw_amd``) never imports it and has no CPU fallback.

Every function follows one reference routine; the docstring cites the
reference file:line it restates (files under ``src/`` of MOLGW 3.4).
Indices of states are 1-based in the arguments that mirror the reference's
module variables (``nsemin``, ``ncore_G`` ...) and 0-based in numpy arrays.
Arrays keep the reference's column-major meaning: ``eri_mo[P, m, n]`` is
``eri_3center_mo(P, m, n, spin)``.

Pinning (what checks each function against something other than itself)
lives in ``tests/test_oracle.py``:
  * ``inverse_sqrt_eri_2center`` -- sqrtinv sqrtinv^T = V^-1 (``numpy.linalg.inv``)
    for a full-rank metric; planted spectra fix the kept count and the
    projected inverse; sqrtinv^T V sqrtinv = 1
  * ``build_eri_3center`` -- from (mu nu|Q) = B V^1/2 (``scipy.linalg.sqrtm``)
    the build returns B up to a rotation (B B^T), and the G0W0 QP energies
    of the built tensor equal those of B
  * ``setup_exchange_m_vxc`` -- zero at a converged RI-HF of the front end
    (F C = S C e); a planted Fock matrix F = S C diag(e') C^T S gives
    diag(e' - e) and nothing off the diagonal
  * ``coeffs_gausslegint``  -- numpy ``leggauss``; integral of 2a/(w^2+a^2)
  * ``calculate_eri_3center_mo`` -- brute-force dense einsum; CC4S (11|11)
  * ``g0w0_casida`` end to end -- the reference's lih_gw.in golden (G0W0@HF
    HOMO-LUMO gap of LiH, cc-pVQZ, 8.053158 eV) on the integrals and RI-HF
    orbitals of molgw_amd.frontend (tests/test_frontend.py)
  * ``polarizability_casida`` -- lih_bulk08 RPA correlation golden; closed
    form for one transition; Tr[Sigma_c G] = E_GM identity
  * ``polarizability_grid`` -- chi(iw) from Casida poles (analytic identity);
    E_RPA(imag) -> E_RPA(Casida)
  * ``gw_selfenergy`` -- E_GW (1/2 Tr[Sigma_c G]) equals E_GM of the Casida
    step; analytic continuation equals the imaginary-axis quadrature
  * ``gw_selfenergy_imag`` -- analytic pole sum at complex frequency
  * ``pade`` -- exact on rational functions
  * ``find_fixed_point`` / ``find_qp_energy_*`` -- hand-built piecewise-linear
    cases with known roots, Z factors and fallbacks
  * ``vsqrt_chi_vsqrt_low_rank`` -- U diag(lambda/(1-lambda)) U^T equals the
    pinned polarizability_grid chi (imaginary axis) and a dense solve of
    (1 - chi0) chi = chi0 (real axis); the kept count equals the number of
    eigenvalues of that chi above the threshold
  * ``gw_selfenergy_contour`` -- with W at exactly |de| and nothing dropped,
    equals Re Sigma_c of the pinned pole self-energy (eta -> 0); on a grid
    holding every |de| the nearest-point rule is exact; out-of-grid error;
    closed shell as two channels equals one channel
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla

HA_EV = 27.21138505              # src/m_definitions.f90:43
PI = 3.14159265358979323         # src/m_definitions.f90:50
COMPLETELY_EMPTY = 1.0e-5        # src/m_definitions.f90:66


# ---------------------------------------------------------------------------
# Quadrature and grids
# ---------------------------------------------------------------------------
def coeffs_gausslegint(xmin: float, xmax: float, n: int):
    """Gauss-Legendre nodes/weights on [xmin, xmax].

    Restates src/m_numerical_tools.f90:35-82: Newton iteration on the Legendre
    three-term recurrence from the Chebyshev-like initial guess, tolerance
    1e-13, nodes returned in ascending order.
    """
    x = np.zeros(n)
    w = np.zeros(n)
    half = 0.5 * (xmax - xmin)
    mid = 0.5 * (xmax + xmin)
    for i in range(1, (n + 1) // 2 + 1):
        z = math.cos(PI * (i - 0.25) / (n + 0.5))
        while True:
            p_cur, p_prev = 1.0, 0.0
            for j in range(1, n + 1):
                p_prev2 = p_prev
                p_prev = p_cur
                p_cur = ((2.0 * j - 1.0) * z * p_prev - (j - 1.0) * p_prev2) / j
            dp = n * (p_prev - z * p_cur) / (1.0 - z * z)
            z_old = z
            z = z_old - p_cur / dp
            if abs(z - z_old) < 1.0e-13:
                break
        x[i - 1] = mid - half * z
        x[n - i] = mid + half * z
        w[i - 1] = 2.0 * half / ((1.0 - z * z) * dp * dp)
        w[n - i] = w[i - 1]
    return x, w


def imaginary_quad_grid(nomega: int):
    """Imaginary-axis grid of chi: src/m_spectral_function.f90:286-306.

    Gauss-Legendre on [0,1] mapped with alpha=1, beta=2:
    omega = 2x/(1-x) (imaginary part), weight = 2 w / (1-x)^2.
    Returns (omega_im, weight).
    """
    alpha, beta = 1.0, 2.0
    xg, wg = coeffs_gausslegint(0.0, 1.0, nomega)
    weight = wg / (2.0 ** alpha - 1.0) * alpha * (1.0 - xg) ** (-alpha - 1.0) * beta
    omega = 1.0 / (2.0 ** alpha - 1.0) * (1.0 / (1.0 - xg) ** alpha - 1.0) * beta
    return omega, weight


# ---------------------------------------------------------------------------
# State ranges and the transition table
# ---------------------------------------------------------------------------
@dataclass
class StateRanges:
    """1-based state ranges as held in the reference's module variables."""
    ncore_G: int
    nvirtual_G: int
    nhomo_G: int
    nlumo_G: int
    nsemin: int
    nsemax: int
    ncore_W: int
    nvirtual_W: int


def selfenergy_set_state_range(occupation, ncoreg=0, nvirtualg=100000, ncorew=0,
                               nvirtualw=100000, selfenergy_state_min=1,
                               selfenergy_state_max=100000,
                               selfenergy_state_range=100000) -> StateRanges:
    """src/m_selfenergy_tools.f90:67-122 and src/m_spectral_function.f90:183-184.

    ``occupation`` has shape (nstate, nspin). Frozen-core keywords other than
    explicit ncoreg/ncorew are out of scope (callers pass them explicitly).
    """
    occupation = np.asarray(occupation, dtype=np.float64)
    if occupation.ndim == 1:
        occupation = occupation[:, None]
    nstate = occupation.shape[0]
    ncore_G = ncoreg
    nvirtual_G = min(nvirtualg, nstate + 1)
    nhomo_G = 1
    for p in range(1, nstate + 1):
        if not np.any(occupation[p - 1, :] < COMPLETELY_EMPTY):
            nhomo_G = max(nhomo_G, p)
    nlumo_G = 100000
    for p in range(1, nstate + 1):
        if not np.any(occupation[p - 1, :] > COMPLETELY_EMPTY):
            nlumo_G = min(nlumo_G, p)
    nsemin = max(ncore_G + 1, selfenergy_state_min, 1, nhomo_G - selfenergy_state_range)
    nsemax = min(nvirtual_G - 1, selfenergy_state_max, nstate, nhomo_G + selfenergy_state_range)
    return StateRanges(ncore_G, nvirtual_G, nhomo_G, nlumo_G, nsemin, nsemax,
                       ncorew, min(nvirtualw, nstate + 1))


def transition_table(occupation, spin_fact: float, ncore_W: int, nvirtual_W: int):
    """Resonant transition table: src/m_spectral_function.f90:239-266.

    Loop order: spin, then istate (outer), then jstate (inner); a transition
    (i -> j) is kept when ``skip_transition`` (:359-378) is false and
    occ(j) - occ(i) <= 0. Returns int array (npole, 3) of 1-based
    (istate, jstate, spin).
    """
    occupation = np.asarray(occupation, dtype=np.float64)
    if occupation.ndim == 1:
        occupation = occupation[:, None]
    nstate, nspin = occupation.shape

    def skip(ib1, ib2, occ1, occ2):
        if ib1 <= ncore_W or ib2 <= ncore_W:
            return True
        if ib1 >= nvirtual_W or ib2 >= nvirtual_W:
            return True
        if occ1 < COMPLETELY_EMPTY and occ2 < COMPLETELY_EMPTY:
            return True
        if occ1 > spin_fact - COMPLETELY_EMPTY and occ2 > spin_fact - COMPLETELY_EMPTY:
            return True
        return False

    table = []
    for s in range(nspin):
        for i in range(1, nstate + 1):
            for j in range(1, nstate + 1):
                oj, oi = occupation[j - 1, s], occupation[i - 1, s]
                if skip(j, i, oj, oi):
                    continue
                if oj - oi > 0.0:
                    continue
                table.append((i, j, s + 1))
    return np.array(table, dtype=np.int64).reshape(-1, 3)


# ---------------------------------------------------------------------------
# A1/A2: RI tensor and AO -> MO transform
# ---------------------------------------------------------------------------
TOO_LOW_EIGENVAL = 1.0e-6        # src/m_eri.f90:27


def inverse_sqrt_eri_2center(eri_2center: np.ndarray):
    """Kept inverse square root of the 2-center metric (P|1/r12|Q).

    Restates the full-range branch of calculate_inverse_sqrt_eri_2center_scalapack
    (src/m_eri_calculate.f90:1049-1262): diagonalise the lower triangle
    (:1106-1108, eigenvalues ascending), keep the eigenvalues above
    TOO_LOW_EIGENVAL (:1117, :1122), and scale the kept eigenvectors,
    sqrtinv(:, j) = eigvec(:, j + nneglect) / sqrt(eigval(j + nneglect))
    (:1206-1211). Returns (sqrtinv[nauxil_bas, nauxil_kept], eigval).
    """
    w, u = np.linalg.eigh(np.asarray(eri_2center, dtype=np.float64), UPLO='L')
    kept = int(np.count_nonzero(w > TOO_LOW_EIGENVAL))
    neglect = len(w) - kept
    return u[:, neglect:] / np.sqrt(w[neglect:]), w


def build_eri_3center(eri_3center_tmp: np.ndarray, sqrtinv: np.ndarray) -> np.ndarray:
    """eri_3center(pair, P) = sum_Q eri_3center_tmp(pair, Q) sqrtinv(Q, P)
    (src/m_eri_calculate.f90:1877-1892); the diagonal pairs of
    eri_3center_tmp already carry the factor 1/2 (:1840-1841)."""
    return np.asarray(eri_3center_tmp, dtype=np.float64) @ sqrtinv


def pack_eri_3center(b_full: np.ndarray, index_basis: np.ndarray) -> np.ndarray:
    """Pack a dense symmetric B(P, mu, nu) into the reference layout.

    Layout of ``eri_3center(npair, nauxil)`` (src/m_eri.f90:38,44) with the
    diagonal pairs stored times 1/2 (src/m_eri_calculate.f90:1840-1841).
    ``index_basis`` is (2, npair), 1-based.
    """
    k = index_basis[0] - 1
    l = index_basis[1] - 1
    packed = b_full[:, k, l].T.copy()
    packed[k == l, :] *= 0.5
    return packed


def calculate_eri_3center_mo(eri_3center: np.ndarray, index_basis: np.ndarray,
                             c_matrix: np.ndarray, mstate_min: int, mstate_max: int,
                             nstate_min: int, nstate_max: int, spin: int = 0) -> np.ndarray:
    """AO -> MO transform of the 3-centre tensor: src/m_eri_ao_mo.f90:416-582.

    For each auxiliary P the reference scatters every stored pair (k,l) into
    tmp1(:,k) += C(l,:) B_kl and tmp1(:,l) += C(k,:) B_kl (:517-522), which
    for a diagonal pair adds twice the stored 1/2 value; then
    tmp2 = C_m^T tmp1^T (:527-530) and eri_3center_mo(P,m,n) = tmp2(m,n).
    The scatter is restated as the dense symmetric matrix it builds.
    Returns eri_mo[P, m - mstate_min, n - nstate_min] (0-based slices).
    """
    npair, naux = eri_3center.shape
    nbf = c_matrix.shape[0]
    k = index_basis[0] - 1
    l = index_basis[1] - 1
    b_full = np.zeros((nbf, nbf, naux))
    if len(np.unique(k * nbf + l)) == npair and not np.any(
            np.isin(l[k != l] * nbf + k[k != l], k * nbf + l)):
        # every unordered pair stored once (setup_basispair, src/m_eri.f90:252-320):
        # the two scatter branches write distinct elements, the diagonal twice
        b_full[k, l, :] = eri_3center
        b_full[l, k, :] = eri_3center
        d = k == l
        b_full[k[d], l[d], :] = 2.0 * eri_3center[d]
    else:
        np.add.at(b_full, (k, l), eri_3center)
        np.add.at(b_full, (l, k), eri_3center)
    c = c_matrix[:, :, spin] if c_matrix.ndim == 3 else c_matrix
    cm = np.ascontiguousarray(c[:, mstate_min - 1:mstate_max])
    cn = np.ascontiguousarray(c[:, nstate_min - 1:nstate_max])
    # tmp1(n, k, P) = sum_l C(l, n) B(k, l, P)
    tmp1 = np.tensordot(cn, b_full, axes=([0], [1]))
    del b_full
    out = np.tensordot(cm, tmp1, axes=([0], [1]))          # (m, n, P)
    return np.ascontiguousarray(out.transpose(2, 0, 1))


def setup_exchange_m_vxc(c_matrix: np.ndarray, hamiltonian_fock: np.ndarray, energy: np.ndarray) -> np.ndarray:
    """setup_exchange_m_vxc without a DFT core-valence split
    (src/m_selfenergy_tools.f90:786-797), one spin channel: <p|Sigma_x - Vxc|q>
    = <p|F - H|q> = (C^T F C)(p, q) - delta_pq e_p (h_ao_to_mo, then the
    diagonal minus the energies)."""
    m = c_matrix.T @ hamiltonian_fock @ c_matrix
    m[np.diag_indices_from(m)] -= energy
    return m


# ---------------------------------------------------------------------------
# A4: analytic Casida RPA
# ---------------------------------------------------------------------------
@dataclass
class SpectralFunction:
    """Subset of ``type(spectral_function)`` (src/m_spectral_function.f90:48-91)."""
    transition_table: np.ndarray
    pole: np.ndarray | None = None
    w_s: np.ndarray | None = None
    chi: np.ndarray | None = None
    omega_im: np.ndarray | None = None
    weight_quad: np.ndarray | None = None


def casida_ia_matrix(eri_mo_w: np.ndarray, ttable: np.ndarray, wlo: int) -> np.ndarray:
    """X(P, t) = eri_3center_mo(P, i_t, a_t): the gather of
    src/m_linear_response.f90:837-842 (and m_build_bse.f90:468-473).
    ``eri_mo_w`` covers states [wlo, ...] in both state indices."""
    i = ttable[:, 0] - wlo
    a = ttable[:, 1] - wlo
    return eri_mo_w[:, i, a]


def polarizability_casida(x_ia: np.ndarray, energy: np.ndarray, ttable: np.ndarray,
                          spin_fact: float = 2.0):
    """RPA branch of ``polarizability`` (src/m_linear_response.f90:33-498).

    Steps restated:
      * A-B diagonal and A+B diagonal = e_b - e_j (m_build_bse.f90:218-251)
      * Hartree term 2*spin_fact*sum_P B_ia B_jb (m_build_bse.f90:318-491)
      * E_RPA starts at -1/4 Tr(A+B) - 1/4 Tr(A-B) (m_build_bse.f90:287-314)
      * M = (A-B)^1/2 (A+B) (A-B)^1/2 = Z Omega^2 Z^T with DSYEV,
        X+Y = (A-B)^1/2 Z Omega^-1/2 (m_block_diago.f90:142-204)
      * E_RPA += 1/2 sum |Omega| (m_linear_response.f90:418)
      * w_s = sqrt(spin_fact) B (X+Y) and
        E_GM = 1/2 (sum w_s^2 - spin_fact sum B^2) (m_linear_response.f90:837-849)
    ``energy`` is (nstate,) for one channel or (nstate, nspin) with the
    transitions of ``ttable`` spin outer (x_ia columns in the same order).
    Returns (pole, w_s, en_rpa, en_gm).
    """
    i = ttable[:, 0] - 1
    a = ttable[:, 1] - 1
    if energy.ndim == 2:                 # (nstate, nspin): each transition in its channel
        sp = ttable[:, 2] - 1
        amb_diag = energy[a, sp] - energy[i, sp]
    else:
        amb_diag = energy[a] - energy[i]
    apb = 2.0 * spin_fact * (x_ia.T @ x_ia)
    apb[np.diag_indices_from(apb)] += amb_diag
    en_rpa = -0.25 * np.trace(apb) - 0.25 * np.sum(amb_diag)
    sq = np.sqrt(amb_diag)
    m = sq[:, None] * apb * sq[None, :]
    omega2, z = sla.eigh(m, driver='ev')
    bigomega = np.sqrt(omega2)
    xpy = sq[:, None] * (z * (1.0 / np.sqrt(bigomega))[None, :])
    en_rpa += 0.5 * np.sum(np.abs(bigomega))
    w_s = math.sqrt(spin_fact) * (x_ia @ xpy)
    en_gm = 0.5 * (np.sum(w_s ** 2) - spin_fact * np.sum(x_ia ** 2))
    return bigomega, w_s, en_rpa, en_gm


# ---------------------------------------------------------------------------
# A5: chi0 on the imaginary axis and chi = (1 - chi0)^-1 chi0
# ---------------------------------------------------------------------------
def polarizability_grid(eri_ai: np.ndarray, energy: np.ndarray, occupation: np.ndarray,
                        ttable: np.ndarray, omega_im: np.ndarray, weight: np.ndarray,
                        vlo: int, olo: int):
    """``polarizability_grid_scalapack`` (src/m_gw_selfenergy_grid.f90:33-209).

    ``eri_ai[P, a - vlo, i - olo]`` is the (virtual, occupied) AO->MO block
    computed at :92. Per imaginary frequency (:118-191):
      x_ia = (P|ai) sqrt(2 docc de / (w^2 + de^2))        (:126-137)
      chi0 = -x x^T                                       (:145-148)
      lambda = eig(1 - chi0)  (DSYEV 'N'), E_RPA and E_GW (:171-173)
      chi = (1 - chi0)^-1 chi0 via LU inverse and DGEMM  (:175-187)
    Returns (chi[iw, P, Q], erpa, egw).
    """
    i = ttable[:, 0] - 1
    a = ttable[:, 1] - 1
    x = eri_ai[:, a - (vlo - 1), i - (olo - 1)]
    docc = occupation[i] - occupation[a]
    de = energy[a] - energy[i]
    return polarizability_grid_x(x, docc, de, omega_im, weight)


def polarizability_grid_x(x: np.ndarray, docc: np.ndarray, de: np.ndarray, omega_im: np.ndarray,
                          weight: np.ndarray):
    """The frequency loop of ``polarizability_grid_scalapack``
    (src/m_gw_selfenergy_grid.f90:118-191) on X(P, t) = (P|a_t i_t) with the
    occupation and energy differences of each transition."""
    naux = x.shape[0]
    chi = np.zeros((len(omega_im), naux, naux))
    erpa = 0.0
    egw = 0.0
    eye = np.eye(naux)
    for iw, (om, wq) in enumerate(zip(omega_im, weight)):
        factor = np.sqrt(2.0 * docc * de / (om ** 2 + de ** 2))
        xt = x * factor[None, :]
        chi0 = -(xt @ xt.T)
        one_m = eye - chi0
        lam = sla.eigvalsh(one_m, driver='ev')
        erpa += np.sum(np.log(lam) + 1.0 - lam) / (2.0 * PI) * wq
        egw += np.sum(-(1.0 - lam) / lam + 1.0 - lam) / (2.0 * PI) * wq
        inv = sla.inv(one_m)
        chi[iw] = inv @ chi0
    return chi, erpa, egw


# ---------------------------------------------------------------------------
# A6: analytic-pole G0W0 self-energy
# ---------------------------------------------------------------------------
def gw_selfenergy(eri_pi: np.ndarray, w_s: np.ndarray, pole: np.ndarray,
                  energy: np.ndarray, occupation: np.ndarray, rng: StateRanges,
                  omega: np.ndarray, energy0: np.ndarray, eta: float,
                  spin_fact: float = 2.0, s_chunk: int = 4096):
    """``gw_selfenergy`` GW branch (src/m_gw_selfenergy_analytic.f90:29-196).

    ``eri_pi[P, p - nsemin, i - (ncore_G+1)]`` from the AO->MO call at :70.
    For each i in [ncore_G+1, nvirtual_G-1] (:80):
      w(s, p) = sum_P w_s(P, s) (P|p i)                     (:101-104)
      sigma(w, p) += w^2 [f_i / (E0_p + w - e_i + Omega_s - i eta)
                         + (1-f_i) / (E0_p + w - e_i - Omega_s + i eta)]  (:126-133)
      E_GW += (1-f_i) occ_p w^2 / (e_p - e_i - Omega_s) for filled p (:135-138)
    ``omega`` is the real grid se%omega(-n:n), ``energy0`` se%energy0(p).
    Returns (sigma[iomega, p] complex, energy_gw).
    """
    nse = rng.nsemax - rng.nsemin + 1
    p_idx = np.arange(rng.nsemin - 1, rng.nsemax)
    sigma = np.zeros((len(omega), nse), dtype=np.complex128)
    energy_gw = 0.0
    ieta = 1j * eta
    e0 = energy0[None, :] + omega[:, None]                   # (nomega, nse)
    p_full = (spin_fact - occupation[p_idx]) / spin_fact < COMPLETELY_EMPTY
    for ii, istate in enumerate(range(rng.ncore_G + 1, rng.nvirtual_G)):
        w = w_s.T @ np.ascontiguousarray(eri_pi[:, :, ii])  # (npole, nse), BLAS on a contiguous block
        w2 = w * w
        f_full = occupation[istate - 1] / spin_fact
        f_empty = (spin_fact - occupation[istate - 1]) / spin_fact
        ei = energy[istate - 1]
        for s0 in range(0, len(pole), s_chunk):
            sl = slice(s0, s0 + s_chunk)
            om = pole[sl]
            base = (e0 - ei)[:, :, None]                     # (nomega, nse, 1)
            term = np.zeros(base.shape[:2] + (len(om),), dtype=np.complex128)
            if f_full != 0.0:
                term += f_full / (base + om[None, None, :] - ieta)
            if f_empty != 0.0:
                term += f_empty / (base - om[None, None, :] + ieta)
            sigma += np.einsum('wps,sp->wp', term, w2[sl], optimize=True)
            if np.any(p_full):
                den = (energy[p_idx][p_full][None, :] - ei - om[:, None])
                energy_gw += np.sum(f_empty * occupation[p_idx][p_full][None, :]
                                    * w2[sl][:, p_full] / den)
    return sigma, energy_gw


def gw_selfenergy_qs(eri_pi: np.ndarray, w_s: np.ndarray, pole: np.ndarray, energy: np.ndarray,
                     occupation: np.ndarray, rng: StateRanges, eta: float = 0.001, spin_fact: float = 2.0):
    """``gw_selfenergy_qs`` GW branch (src/m_gw_selfenergy_analytic.f90:653-797)
    and the hermitianization of apply_qs_approximation
    (src/m_selfenergy_tools.f90:821-823), one spin channel.

    ``eri_pi[P, p - nsemin, i - (ncore_G+1)]`` as for gw_selfenergy. For each
    G state i (:699): w(s, p) = sum_P w_s(P, s) (P|p i) (:711-715), and
    Sigma(p, q) += w(s, p) w(s, q) Re[f_i / (e_q - i eta - e_i + Omega_s)
    + (1 - f_i) / (e_q + i eta - e_i - Omega_s)] (:736-744). Returns
    (Sigma + Sigma^T) / 2 over the window (nse x nse).
    """
    nse = rng.nsemax - rng.nsemin + 1
    eq = energy[rng.nsemin - 1:rng.nsemax]
    sig = np.zeros((nse, nse))
    for k, istate in enumerate(range(rng.ncore_G + 1, rng.nvirtual_G)):
        w = w_s.T @ eri_pi[:, :, k]                        # (npole, nse)
        f_full = occupation[istate - 1] / spin_fact
        f_empty = (spin_fact - occupation[istate - 1]) / spin_fact
        d = eq[None, :] - energy[istate - 1]
        a = d + pole[:, None]
        c = d - pole[:, None]
        g = f_full * a / (a * a + eta * eta) + f_empty * c / (c * c + eta * eta)
        sig += w.T @ (w * g)
    return 0.5 * (sig + sig.T)


def gw_selfenergy_c(eri_pi: np.ndarray, w_s: np.ndarray, pole: np.ndarray,
                    energy: np.ndarray, occupation: np.ndarray, rng: StateRanges,
                    omega: np.ndarray, energy0: np.ndarray, eta: float, spin_fact: float = 2.0):
    """``gw_selfenergy`` with the pole sum in C/OpenMP (oracle/c/pole_sum.c):
    the same arithmetic as ``gw_selfenergy``, used for the timed CPU baseline."""
    import ctypes as C
    import os
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), '_build', 'libpolesum.so'))
    fn = lib.pole_sum_one_i
    fn.restype = C.c_double
    dp = C.POINTER(C.c_double)
    fn.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, C.c_double, C.c_double,
                   C.c_double, C.c_double, dp]
    nse = rng.nsemax - rng.nsemin + 1
    p_idx = np.arange(rng.nsemin - 1, rng.nsemax)
    sigma = np.zeros((nse, len(omega)), dtype=np.complex128)
    pole = np.ascontiguousarray(pole, dtype=np.float64)
    om = np.ascontiguousarray(omega, dtype=np.float64)
    e0 = np.ascontiguousarray(energy0, dtype=np.float64)
    eps = np.ascontiguousarray(energy[p_idx], dtype=np.float64)
    occp = np.ascontiguousarray(occupation[p_idx], dtype=np.float64)
    ptr = lambda a: a.ctypes.data_as(dp)
    energy_gw = 0.0
    for ii, istate in enumerate(range(rng.ncore_G + 1, rng.nvirtual_G)):
        w = w_s.T @ np.ascontiguousarray(eri_pi[:, :, ii])
        w2 = np.asfortranarray(w * w)                          # (npole, nse)
        energy_gw += fn(len(pole), nse, len(om), ptr(w2), ptr(pole), ptr(om), ptr(e0), ptr(eps), ptr(occp),
                        float(energy[istate - 1]), float(occupation[istate - 1]), spin_fact, eta,
                        sigma.ctypes.data_as(dp))
    return np.ascontiguousarray(sigma.T), energy_gw


# ---------------------------------------------------------------------------
# A7: imaginary-axis self-energy
# ---------------------------------------------------------------------------
def gw_selfenergy_imag(eri_pm: np.ndarray, chi: np.ndarray, omega_im: np.ndarray,
                       weight: np.ndarray, omega_calc: np.ndarray, energy: np.ndarray,
                       rng: StateRanges):
    """``gw_selfenergy_imag_scalapack`` (src/m_gw_selfenergy_grid.f90:213-338).

    ``eri_pm[P, p - (ncore_G+1), m - nsemin]`` from the AO->MO call at :253.
    sigma(w_c, m) = - sum_w' weight_w' / (2 pi) sum_p
        [1/((w_c - e_p) + i w') + 1/((w_c - e_p) - i w')] (B_pm^T chi(iw') B_pm)
    (:297-319). Returns sigma_calc[iomega_calc, m] complex.
    """
    nse = rng.nsemax - rng.nsemin + 1
    p_idx = np.arange(rng.ncore_G, rng.nvirtual_G - 1)
    sigma = np.zeros((len(omega_calc), nse), dtype=np.complex128)
    for mm in range(nse):
        b = np.ascontiguousarray(eri_pm[:, :, mm])          # a strided slice makes BLAS crawl
        for iw in range(len(omega_im)):
            chib = chi[iw] @ b
            v = np.sum(b * chib, axis=0)                     # v_chi_v_p
            d = omega_calc[:, None] - energy[p_idx][None, :]
            g = 1.0 / (d + 1j * omega_im[iw]) + 1.0 / (d - 1j * omega_im[iw])
            sigma[:, mm] -= weight[iw] * (g @ v) / (2.0 * PI)
    return sigma


# ---------------------------------------------------------------------------
# A8: Pade analytic continuation
# ---------------------------------------------------------------------------
def calculate_pade_a(z_in: np.ndarray, f_in: np.ndarray) -> np.ndarray:
    """Thiele continued-fraction coefficients: src/m_numerical_tools.f90:125-145."""
    n = len(z_in)
    g = np.zeros((n, n), dtype=np.complex128)
    g[0, :] = f_in
    for ip in range(1, n):
        for jp in range(ip, n):
            g[ip, jp] = (g[ip - 1, ip - 1] - g[ip - 1, jp]) / ((z_in[jp] - z_in[ip - 1]) * g[ip - 1, jp])
    return np.diag(g).copy()


def pade(z_in: np.ndarray, f_in: np.ndarray, z_out: complex) -> complex:
    """Evaluate the Thiele continued fraction: src/m_numerical_tools.f90:86-120."""
    n = len(z_in)
    aa = calculate_pade_a(z_in, f_in)
    az = np.zeros(n + 1, dtype=np.complex128)
    bz = np.zeros(n + 1, dtype=np.complex128)
    az[1] = aa[0]
    bz[0] = 1.0
    bz[1] = 1.0
    for ip in range(1, n):
        az[ip + 1] = az[ip] + (z_out - z_in[ip - 1]) * aa[ip] * az[ip - 1]
        bz[ip + 1] = bz[ip] + (z_out - z_in[ip - 1]) * aa[ip] * bz[ip - 1]
    return az[n] / bz[n]


def se_pade_fit(omega_calc: np.ndarray, sigma_calc: np.ndarray, omega: np.ndarray,
                eta: float) -> np.ndarray:
    """``se_pade_fit`` (src/m_selfenergy_tools.f90:965-999).

    Symmetrise with sigma(-iw) = sigma(iw)^*, then evaluate the Pade
    approximant at omega + i eta * sign, sign = -sign(Re(omega) - Re(w_c(1))).
    ``sigma_calc`` is (nomega_calc, nse); returns sigma (nomega, nse).
    """
    nc = len(omega_calc)
    nse = sigma_calc.shape[1]
    out = np.zeros((len(omega), nse), dtype=np.complex128)
    z_sym = np.zeros(2 * nc - 1, dtype=np.complex128)
    z_sym[nc - 1] = omega_calc[0]
    z_sym[nc:] = omega_calc[1:]
    z_sym[:nc - 1] = np.conj(omega_calc[1:])[::-1]
    for p in range(nse):
        f_sym = np.zeros(2 * nc - 1, dtype=np.complex128)
        f_sym[nc - 1] = sigma_calc[0, p]
        f_sym[nc:] = sigma_calc[1:, p]
        f_sym[:nc - 1] = np.conj(sigma_calc[1:, p])[::-1]
        for k, om in enumerate(omega):
            # Fortran SIGN(1, x) is +1 for x == 0
            sign_eta = -1.0 if (om.real - omega_calc[0].real) >= 0.0 else 1.0
            out[k, p] = pade(z_sym, f_sym, om + 1j * eta * sign_eta)
    return out


# ---------------------------------------------------------------------------
# A9: quasiparticle equation
# ---------------------------------------------------------------------------
def find_fixed_point(xx: np.ndarray, fx: np.ndarray, nfpmx: int = 4):
    """``find_fixed_point`` (src/m_selfenergy_tools.f90:373-420).

    g = f - x; at each sign change g(ix) g(ix+1) < 0 the linear-interpolated
    root and Z = 1 / (1 - df/dx) are inserted into a list sorted by
    decreasing Z (up to ``nfpmx`` entries; Z < 1e-5 rejected; a new Z equal
    to a stored one goes after it). Unused entries keep Z = -1, E = 0.
    """
    z_weight = -np.ones(nfpmx)
    e_fixed = np.zeros(nfpmx)
    gx = fx - xx
    for ix in range(len(xx) - 1):
        if gx[ix] * gx[ix + 1] < 0.0:
            z_zero = 1.0 / (1.0 - (fx[ix + 1] - fx[ix]) / (xx[ix + 1] - xx[ix]))
            if z_zero < 0.00001:
                continue
            if z_zero > z_weight[nfpmx - 1]:
                jfixed = int(np.argmax(z_zero > z_weight))
                z_weight[jfixed + 1:] = z_weight[jfixed:nfpmx - 1].copy()
                e_fixed[jfixed + 1:] = e_fixed[jfixed:nfpmx - 1].copy()
                z_weight[jfixed] = z_zero
                gpxi = (gx[ix + 1] - gx[ix]) / (xx[ix + 1] - xx[ix])
                e_fixed[jfixed] = xx[ix] - gx[ix] / gpxi
    return e_fixed, z_weight


def find_qp_energy_graphical(sigma: np.ndarray, omega: np.ndarray, se_energy0: np.ndarray,
                             exchange_m_vxc: np.ndarray, energy0: np.ndarray,
                             rng: StateRanges):
    """``find_qp_energy_graphical`` (src/m_selfenergy_tools.f90:253-355).

    ``sigma`` (nomega, nse) complex on se%omega (real parts ``omega``);
    ``se_energy0`` (nse,) grid centres; ``exchange_m_vxc``/``energy0`` (nstate,).
    QP = fixed point with the largest Z (the first of the sorted list); when
    no fixed point exists: e0 + Re sigma(0) + (Sigma_x - Vxc) (:331-335).
    States outside [nsemin, nsemax] keep energy0. Returns (energy_qp, zz).
    """
    nstate = len(energy0)
    energy_qp = np.array(energy0, dtype=np.float64).copy()
    zz = np.ones(nstate)
    n0 = len(omega) // 2
    for jj, p in enumerate(range(rng.nsemin, rng.nsemax + 1)):
        lhs = omega.real + se_energy0[jj]
        rhs = sigma[:, jj].real + exchange_m_vxc[p - 1] + energy0[p - 1]
        e_fix, z_w = find_fixed_point(lhs, rhs)
        if np.max(z_w) > 1.0e-6:
            iqp = int(np.argmax(z_w))
            energy_qp[p - 1] = e_fix[iqp]
            zz[p - 1] = z_w[iqp]
        else:
            energy_qp[p - 1] = energy0[p - 1]
            zz[p - 1] = 1.0
        if np.count_nonzero(z_w > 0.0) == 0:
            energy_qp[p - 1] = energy0[p - 1] + sigma[n0, jj].real + exchange_m_vxc[p - 1]
    return energy_qp, zz


def find_qp_energy_linearization(sigma: np.ndarray, omega: np.ndarray, se_energy0: np.ndarray,
                                 exchange_m_vxc: np.ndarray, energy0: np.ndarray,
                                 rng: StateRanges):
    """``find_qp_energy_linearization`` (src/m_selfenergy_tools.f90:202-249).

    Z from the finite difference of Re sigma at omega(+-1), clipped to [0,1];
    E = E0 + Z (e - E0 + Re sigma(0) + Sigma_x - Vxc). Returns (energy_qp, zz).
    """
    nstate = len(energy0)
    energy_qp = np.array(energy0, dtype=np.float64).copy()
    zz = np.ones(nstate)
    n0 = len(omega) // 2
    for jj, p in enumerate(range(rng.nsemin, rng.nsemax + 1)):
        if n0 > 0:
            z = (sigma[n0 + 1, jj].real - sigma[n0 - 1, jj].real) / (omega[n0 + 1].real - omega[n0 - 1].real)
            z = min(max(1.0 / (1.0 - z), 0.0), 1.0)
            zz[p - 1] = z
            energy_qp[p - 1] = se_energy0[jj] + z * (energy0[p - 1] - se_energy0[jj]
                                                     + sigma[n0, jj].real + exchange_m_vxc[p - 1])
        else:
            energy_qp[p - 1] = energy0[p - 1] + sigma[n0, jj].real + exchange_m_vxc[p - 1]
    return energy_qp, zz


# ---------------------------------------------------------------------------
# Drivers: the two flows of selfenergy_evaluation for one-shot G0W0
# ---------------------------------------------------------------------------
@dataclass
class G0W0Result:
    energy_qp: np.ndarray
    zz: np.ndarray
    energy_qp_lin: np.ndarray
    sigma: np.ndarray
    omega: np.ndarray
    en_rpa: float
    en_gw: float
    en_gm: float = 0.0
    pole: np.ndarray | None = None
    extra: dict = field(default_factory=dict)


def fermi_level(energy: np.ndarray, rng: StateRanges) -> float:
    """Fermi level between HOMO and LUMO (src/m_selfenergy_tools.f90:588-595)."""
    return 0.5 * (energy[rng.nhomo_G - 1] + energy[rng.nhomo_G])


def g0w0_casida(eri_3center, index_basis, c_matrix, energy, occupation,
                exchange_m_vxc=None, nomega_sigma=51, step_sigma=0.01, eta=0.001,
                spin_fact=2.0, ranges: StateRanges | None = None) -> G0W0Result:
    """One-shot G0W0 with analytic Casida W (postscf='GW'):
    src/m_selfenergy_evaluation.f90:135-261,589-590.

    Restricted closed shell (nspin=1): ``energy``/``occupation``/
    ``exchange_m_vxc`` are (nstate,), ``c_matrix`` is (nbf, nstate).
    """
    energy = np.asarray(energy, dtype=np.float64)
    occupation = np.asarray(occupation, dtype=np.float64)
    nstate = len(energy)
    if exchange_m_vxc is None:
        exchange_m_vxc = np.zeros(nstate)
    rng = ranges or selfenergy_set_state_range(occupation[:, None])
    # se_init(one_shot): omega_k = step * k, centred on energy0 (:658-665, :682)
    nom = nomega_sigma // 2
    omega = step_sigma * np.arange(-nom, nom + 1, dtype=np.float64)
    se_energy0 = energy[rng.nsemin - 1:rng.nsemax]
    # W: polarizability with the full W range of states (m_linear_response.f90:115)
    ttable = transition_table(occupation[:, None], spin_fact, rng.ncore_W, rng.nvirtual_W)
    wlo, whi = rng.ncore_W + 1, rng.nvirtual_W - 1
    eri_w = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, wlo, whi, wlo, whi)
    x_ia = casida_ia_matrix(eri_w, ttable, wlo)
    pole, w_s, en_rpa, en_gm = polarizability_casida(x_ia, energy, ttable, spin_fact)
    del eri_w
    # Sigma (m_gw_selfenergy_analytic.f90:70)
    eri_pi = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.nsemin, rng.nsemax,
                                      rng.ncore_G + 1, rng.nvirtual_G - 1)
    sigma, en_gw = gw_selfenergy(eri_pi, w_s, pole, energy, occupation, rng, omega, se_energy0,
                                 eta, spin_fact)
    e_lin, _ = find_qp_energy_linearization(sigma, omega, se_energy0, exchange_m_vxc, energy, rng)
    e_qp, zz = find_qp_energy_graphical(sigma, omega, se_energy0, exchange_m_vxc, energy, rng)
    return G0W0Result(e_qp, zz, e_lin, sigma, omega, en_rpa, en_gw, en_gm, pole)


def g0w0_pade(eri_3center, index_basis, c_matrix, energy, occupation, nomega_chi_imag,
              exchange_m_vxc=None, nomega_sigma=51, step_sigma=0.01, nomega_sigma_calc=1,
              step_sigma_calc=0.03, eta=0.001, spin_fact=2.0,
              ranges: StateRanges | None = None) -> G0W0Result:
    """One-shot G0W0 on the imaginary axis with Pade continuation
    (postscf='G0W0_PADE'): src/m_selfenergy_evaluation.f90:215-243 with
    se_init(imaginary_axis_pade) (src/m_selfenergy_tools.f90:598-614):
    the real grid is centred on the Fermi level and se%energy0 = 0.
    """
    energy = np.asarray(energy, dtype=np.float64)
    occupation = np.asarray(occupation, dtype=np.float64)
    nstate = len(energy)
    if exchange_m_vxc is None:
        exchange_m_vxc = np.zeros(nstate)
    rng = ranges or selfenergy_set_state_range(occupation[:, None])
    efermi = fermi_level(energy, rng)
    nom = nomega_sigma // 2
    omega = efermi + step_sigma * np.arange(-nom, nom + 1, dtype=np.float64)
    omega_calc = efermi + step_sigma_calc * np.arange(nomega_sigma_calc) * 1j
    se_energy0 = np.zeros(rng.nsemax - rng.nsemin + 1)
    # chi on the imaginary axis (m_gw_selfenergy_grid.f90:92: virtual first)
    ttable = transition_table(occupation[:, None], spin_fact, rng.ncore_W, rng.nvirtual_W)
    nhomo_W = int(np.max(ttable[:, 0]))
    nlumo_W = int(np.min(ttable[:, 1]))
    eri_ai = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, nlumo_W,
                                      rng.nvirtual_W - 1, rng.ncore_W + 1, nhomo_W)
    omega_im, weight = imaginary_quad_grid(nomega_chi_imag)
    chi, en_rpa, en_gw = polarizability_grid(eri_ai, energy, occupation, ttable, omega_im,
                                             weight, nlumo_W, rng.ncore_W + 1)
    del eri_ai
    eri_pm = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.ncore_G + 1,
                                      rng.nvirtual_G - 1, rng.nsemin, rng.nsemax)
    sigma_calc = gw_selfenergy_imag(eri_pm, chi, omega_im, weight, omega_calc, energy, rng)
    sigma = se_pade_fit(omega_calc, sigma_calc, omega, eta)
    e_lin, _ = find_qp_energy_linearization(sigma, omega, se_energy0, exchange_m_vxc, energy, rng)
    e_qp, zz = find_qp_energy_graphical(sigma, omega, se_energy0, exchange_m_vxc, energy, rng)
    res = G0W0Result(e_qp, zz, e_lin, sigma, omega, en_rpa, en_gw)
    res.extra = {'sigma_calc': sigma_calc, 'omega_calc': omega_calc, 'chi': chi,
                 'omega_im': omega_im, 'weight': weight}
    return res


# ---------------------------------------------------------------------------
# eigenvalue self-consistent GW (postscf GNW0 / GNWN = EVGW)
# ---------------------------------------------------------------------------
def evgw(eri_3center, index_basis, c_matrix, energy, occupation, nstep: int, approx: str = 'GnWn',
         exchange_m_vxc=None, eta: float = 0.001, spin_fact: float = 2.0,
         ranges: StateRanges | None = None):
    """``selfenergy_evaluation`` with selfenergy_technique = EVSC
    (src/m_selfenergy_evaluation.f90:78-182, 589-606; postscf keywords at
    src/m_inputparam.f90:182-189), analytic W.

    Step n (1-based): energy_g = the QP energies of step n-1 (``read_energy_qp``;
    step 1: ``energy``, :164-170); energy_w = energy_g for GnWn, ``energy``
    for GnW0 (:174-178), and GnW0 keeps the W of step 1 (:200-202).
    se_init(EVSC) puts one point omega = 0 on top of se%energy0 = energy_g
    (src/m_selfenergy_tools.f90:594-597, 683-684), so Sigma_c is evaluated at
    energy_g(p); find_qp_energy_linearization without a grid (:236-240) gives
    E(p) = energy(p) + Re Sigma_c(p) + (Sigma_x - Vxc)(p) in [nsemin, nsemax]
    and energy(p) elsewhere (:216-217).

    Returns (list of the QP energies of each step, list of the W poles used at
    each step)."""
    energy = np.asarray(energy, dtype=np.float64)
    occupation = np.asarray(occupation, dtype=np.float64)
    nstate = len(energy)
    if exchange_m_vxc is None:
        exchange_m_vxc = np.zeros(nstate)
    if approx not in ('GnW0', 'GnWn'):
        raise ValueError('approx must be GnW0 or GnWn')
    rng = ranges or selfenergy_set_state_range(occupation[:, None])
    ttable = transition_table(occupation[:, None], spin_fact, rng.ncore_W, rng.nvirtual_W)
    wlo, whi = rng.ncore_W + 1, rng.nvirtual_W - 1
    eri_w = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, wlo, whi, wlo, whi)
    x_ia = casida_ia_matrix(eri_w, ttable, wlo)
    del eri_w
    eri_pi = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.nsemin, rng.nsemax,
                                      rng.ncore_G + 1, rng.nvirtual_G - 1)
    win = slice(rng.nsemin - 1, rng.nsemax)
    energy_g = energy.copy()
    history, poles = [], []
    pole = w_s = None
    for step in range(nstep):
        if approx == 'GnWn' or step == 0:
            energy_w = energy_g if approx == 'GnWn' else energy
            pole, w_s, _, _ = polarizability_casida(x_ia, energy_w, ttable, spin_fact)
        sigma, _ = gw_selfenergy(eri_pi, w_s, pole, energy_g, occupation, rng, np.zeros(1),
                                 energy_g[win], eta, spin_fact)
        new = energy.copy()
        new[win] = energy[win] + sigma[0].real + exchange_m_vxc[win]
        history.append(new)
        poles.append(pole)
        energy_g = new
    return history, poles


# ---------------------------------------------------------------------------
# Spin-polarised one-shot G0W0 (nspin = 2)
# ---------------------------------------------------------------------------
def transition_columns(eri_3center, index_basis, c_matrix, ttable, wlo: int, whi: int) -> np.ndarray:
    """X(P, t) = eri_3center_mo(P, a_t, i_t, s_t) for a spin-outer transition
    table, each column from its own channel's AO->MO transform
    (src/m_gw_selfenergy_grid.f90:126-137, src/m_linear_response.f90:837-842).
    ``c_matrix`` is (nbf, nstate, nspin)."""
    cols = []
    for s in range(c_matrix.shape[2]):
        sel = ttable[:, 2] == s + 1
        if not sel.any():
            continue
        eri_s = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, wlo, whi, wlo, whi, spin=s)
        cols.append(eri_s[:, ttable[sel, 1] - wlo, ttable[sel, 0] - wlo])
    return np.concatenate(cols, axis=1)


def g0w0_spin(eri_3center, index_basis, c_matrix, energy, occupation, technique: str = 'one_shot',
              nomega_chi_imag: int = 0, exchange_m_vxc=None, nomega_sigma: int = 51, step_sigma: float = 0.01,
              nomega_sigma_calc: int = 1, step_sigma_calc: float = 0.03, eta: float = 0.001,
              spin_fact: float = 1.0, ranges: StateRanges | None = None):
    """One-shot G0W0 with spin channels: ``c_matrix`` (nbf, nstate, nspin),
    ``energy``/``occupation``/``exchange_m_vxc`` (nstate, nspin).

    W from the transitions of every channel (spin-outer table,
    src/m_spectral_function.f90:239-266; chi0 at
    src/m_gw_selfenergy_grid.f90:118-148 or the Casida matrix of
    src/m_linear_response.f90:33-498), then per channel the self-energy with
    that channel's integrals and energies (the ispin loops of
    src/m_gw_selfenergy_analytic.f90:80-196 and
    src/m_gw_selfenergy_grid.f90:279-320) and its QP equation. For
    'imaginary_pade' the real grid is centred midway between the highest HOMO
    and the lowest LUMO over the channels (src/m_selfenergy_tools.f90:586-590).
    Returns a dict with energy_qp, zz, energy_qp_lin (nstate, nspin), sigma
    (nomega, nse, nspin), omega, en_rpa and en_gw (summed over the channels
    for the pole self-energy)."""
    c_matrix = np.asarray(c_matrix, dtype=np.float64)
    energy = np.asarray(energy, dtype=np.float64)
    occupation = np.asarray(occupation, dtype=np.float64)
    nstate, nspin = energy.shape
    exc = np.zeros((nstate, nspin)) if exchange_m_vxc is None else np.asarray(exchange_m_vxc, dtype=np.float64)
    rng = ranges or selfenergy_set_state_range(occupation)
    ttable = transition_table(occupation, spin_fact, rng.ncore_W, rng.nvirtual_W)
    wlo, whi = rng.ncore_W + 1, rng.nvirtual_W - 1
    x = transition_columns(eri_3center, index_basis, c_matrix, ttable, wlo, whi)
    i, a, sp = ttable[:, 0] - 1, ttable[:, 1] - 1, ttable[:, 2] - 1
    nom = nomega_sigma // 2
    nse = rng.nsemax - rng.nsemin + 1
    out = {k: np.zeros((nstate, nspin)) for k in ('energy_qp', 'zz', 'energy_qp_lin')}
    sigma = np.zeros((2 * nom + 1, nse, nspin), dtype=np.complex128)
    en_gw = 0.0
    if technique == 'one_shot':
        omega = step_sigma * np.arange(-nom, nom + 1, dtype=np.float64)
        pole, w_s, en_rpa, _ = polarizability_casida(x, energy, ttable, spin_fact)
        for s in range(nspin):
            se_energy0 = energy[rng.nsemin - 1:rng.nsemax, s]
            eri_pi = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.nsemin, rng.nsemax,
                                              rng.ncore_G + 1, rng.nvirtual_G - 1, spin=s)
            sig, eg = gw_selfenergy(eri_pi, w_s, pole, energy[:, s], occupation[:, s], rng, omega, se_energy0,
                                    eta, spin_fact)
            en_gw += eg
            sigma[:, :, s] = sig
    else:
        efermi = 0.5 * (energy[rng.nhomo_G - 1].max() + energy[rng.nhomo_G].min())
        omega = efermi + step_sigma * np.arange(-nom, nom + 1, dtype=np.float64)
        omega_calc = efermi + step_sigma_calc * np.arange(nomega_sigma_calc) * 1j
        omega_im, weight = imaginary_quad_grid(nomega_chi_imag)
        chi, en_rpa, en_gw = polarizability_grid_x(x, occupation[i, sp] - occupation[a, sp],
                                                   energy[a, sp] - energy[i, sp], omega_im, weight)
        for s in range(nspin):
            eri_pm = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.ncore_G + 1,
                                              rng.nvirtual_G - 1, rng.nsemin, rng.nsemax, spin=s)
            sc = gw_selfenergy_imag(eri_pm, chi, omega_im, weight, omega_calc, energy[:, s], rng)
            sigma[:, :, s] = se_pade_fit(omega_calc, sc, omega, eta)
    for s in range(nspin):
        se_energy0 = (energy[rng.nsemin - 1:rng.nsemax, s] if technique == 'one_shot' else np.zeros(nse))
        e_lin, _ = find_qp_energy_linearization(sigma[:, :, s], omega, se_energy0, exc[:, s], energy[:, s], rng)
        e_qp, zz = find_qp_energy_graphical(sigma[:, :, s], omega, se_energy0, exc[:, s], energy[:, s], rng)
        out['energy_qp'][:, s], out['zz'][:, s], out['energy_qp_lin'][:, s] = e_qp, zz, e_lin
    out.update(sigma=sigma, omega=omega, en_rpa=en_rpa, en_gw=en_gw)
    return out


# ---------------------------------------------------------------------------
# Contour deformation (postscf='G0W0_CONTOUR'): low-rank W on the imaginary
# and the real axis, imaginary-axis integral plus residues
# ---------------------------------------------------------------------------
TOL_LOW_EIGVAL_CHI = 1.0e-3      # src/m_spectral_function.f90:775


def vsqrt_chi_vsqrt_low_rank(x: np.ndarray, docc: np.ndarray, de: np.ndarray, omega2: np.ndarray,
                             tol: float = TOL_LOW_EIGVAL_CHI):
    """``sf_vsqrt_chi_vsqrt_rpa`` with low_rank (src/m_spectral_function.f90:768-949)
    on X(P, t) = (P|a_t i_t).

    Per frequency, with ``omega2[k]`` = omega^2 (-w^2 on the imaginary axis,
    w^2 on the real axis):
      chi0 = X diag(2 docc de / (omega^2 - de^2)) X^T          (:857-877)
      lambda, U = eig(chi0); eta = lambda / (1 - lambda)        (:892-895)
      keep the eigenpairs with |eta| > tol                      (:896-913)
    Returns a list of (eigvec[naux, neig], eta[neig])."""
    out = []
    for w2 in np.asarray(omega2, dtype=np.float64):
        factor = 2.0 * docc * de / (w2 - de ** 2)
        chi0 = (x * factor[None, :]) @ x.T
        lam, u = np.linalg.eigh(chi0)
        eta = lam / (1.0 - lam)
        keep = np.abs(eta) > tol
        out.append((u[:, keep], eta[keep]))
    return out


def contour_de_max(energy: np.ndarray, se_energy0: np.ndarray, omega: np.ndarray, rng: StateRanges,
                   eta: float) -> float:
    """Largest real frequency the residues need (src/m_gw_selfenergy_grid.f90:385-408):
    over the channels, the states m of [nsemin, nsemax] and the grid omega_k,
    e_p - (E0_m + omega_k) for occupied p in [ncore_G+1, nhomo_G] and
    E0_m + omega_k - e_p for empty p in [nlumo_G, nvirtual_G-1], counting only
    values >= -eta; 0 if there are none. ``energy`` (nstate, nspin),
    ``se_energy0`` (nse, nspin)."""
    de_max = 0.0
    nse = rng.nsemax - rng.nsemin + 1
    for s in range(energy.shape[1]):
        for m in range(nse):
            w = se_energy0[m, s] + omega
            for p in range(rng.ncore_G + 1, rng.nhomo_G + 1):
                de = energy[p - 1, s] - w
                de = de[de >= -eta]
                if de.size:
                    de_max = max(de_max, float(de.max()))
            for p in range(rng.nlumo_G, rng.nvirtual_G):
                de = w - energy[p - 1, s]
                de = de[de >= -eta]
                if de.size:
                    de_max = max(de_max, float(de.max()))
    return de_max


def real_linear_grid(nomega: int, omega_max: float) -> np.ndarray:
    """REAL_LINEAR grid of sf_init (src/m_spectral_function.f90:308-315):
    omega_j = (j - 1) / (n - 1) omega_max, j = 1..n."""
    if nomega < 2:
        raise ValueError('nomega_chi_real must be at least 2')
    return np.arange(nomega, dtype=np.float64) / (nomega - 1) * omega_max


def _v_chi_v(w, b):
    u, eta = w
    if eta.size == 0:
        return np.zeros(b.shape[1])
    t = u.T @ b
    return np.sum(t ** 2 * eta[:, None], axis=0)


def gw_selfenergy_contour(eri_pm: np.ndarray, w_imag, omega_im: np.ndarray, weight: np.ndarray, w_real,
                          omega_re: np.ndarray, energy: np.ndarray, rng: StateRanges, omega: np.ndarray,
                          se_energy0: np.ndarray, eta: float, w_real_at=None) -> np.ndarray:
    """The self-energy loop of ``gw_selfenergy_contour``
    (src/m_gw_selfenergy_grid.f90:429-530) for one channel.

    ``eri_pm[P, p - (ncore_G+1), m - nsemin]``; ``w_imag`` / ``w_real`` the
    low-rank W of ``vsqrt_chi_vsqrt_low_rank`` on the imaginary grid
    (``omega_im``, ``weight``) and on the real grid ``omega_re``; ``energy``
    this channel's energies (nstate,); ``se_energy0`` (nse,). For
    x = E0_m + omega_k - e_p and v = B_pm^T W B_pm:
      imaginary axis (:436-469), where |x| > eta:
        sigma -= weight / (2 pi) [1/(x + i w) + 1/(x - i w)] v(iw)
      occupied p (p < nlumo_G), de = e_p - (E0_m + omega_k) >= -eta (:475-500):
        sigma -= v(|de|) (x 1/2 when |de| < eta)
      empty p (p > nhomo_G), de = E0_m + omega_k - e_p >= -eta (:505-530):
        sigma += v(|de|) (x 1/2 when |de| < eta)
    v(|de|) takes the real-grid point nearest to |de|, the first one on a tie
    (sf_interpolate_vsqrt_chi_vsqrt, src/m_spectral_function.f90:953-981);
    a |de| outside the grid is an error, as there. ``w_real_at(w)``, when
    given, returns the low-rank W at exactly w instead (tests only).
    Returns sigma[iomega, m] complex (zero imaginary part)."""
    nse = rng.nsemax - rng.nsemin + 1
    p_idx = np.arange(rng.ncore_G, rng.nvirtual_G - 1)
    sigma = np.zeros((len(omega), nse), dtype=np.complex128)
    lo, hi = float(np.min(omega_re)), float(np.max(omega_re))

    def residue(b_col, w):
        if w_real_at is not None:
            return float(_v_chi_v(w_real_at(w), b_col[:, None])[0])
        if w > hi or w < lo:
            raise ValueError(f'requested frequency {w} out of range')
        j = int(np.argmin(np.abs(omega_re - w)))
        return float(_v_chi_v(w_real[j], b_col[:, None])[0])

    for mm in range(nse):
        b = np.ascontiguousarray(eri_pm[:, :, mm])
        x = se_energy0[mm] + omega[:, None] - energy[p_idx][None, :]          # (nomega, np)
        for iw in range(len(omega_im)):
            if w_imag[iw][1].size == 0:
                continue
            v = _v_chi_v(w_imag[iw], b)
            g = 1.0 / (x + 1j * omega_im[iw]) + 1.0 / (x - 1j * omega_im[iw])
            g[np.abs(x) <= eta] = 0.0
            sigma[:, mm] -= weight[iw] * (g @ v) / (2.0 * PI)
        for pl, p in enumerate(p_idx):
            if p + 1 < rng.nlumo_G:
                for k in range(len(omega)):
                    de = energy[p] - (se_energy0[mm] + omega[k])
                    if de < -eta:
                        continue
                    factor = 0.5 if abs(de) < eta else 1.0
                    sigma[k, mm] -= residue(b[:, pl], abs(de)) * factor
            if p + 1 > rng.nhomo_G:
                for k in range(len(omega)):
                    de = se_energy0[mm] + omega[k] - energy[p]
                    if de < -eta:
                        continue
                    factor = 0.5 if abs(de) < eta else 1.0
                    sigma[k, mm] += residue(b[:, pl], abs(de)) * factor
    return sigma


def g0w0_contour(eri_3center, index_basis, c_matrix, energy, occupation, nomega_chi_imag: int,
                 nomega_chi_real: int, exchange_m_vxc=None, nomega_sigma: int = 51, step_sigma: float = 0.01,
                 eta: float = 0.001, spin_fact: float | None = None, ranges: StateRanges | None = None,
                 tol: float = TOL_LOW_EIGVAL_CHI):
    """One-shot G0W0 by contour deformation (postscf='G0W0_CONTOUR'):
    src/m_selfenergy_evaluation.f90:218-240 with se_init(contour_deformation)
    (src/m_selfenergy_tools.f90:658-665, :679-680: omega_k = step k around
    E0 = the input energies) and gw_selfenergy_contour
    (src/m_gw_selfenergy_grid.f90:342-551): W from the transitions of every
    channel on both grids, the self-energy and the graphical QP solve per
    channel. ``energy``/``occupation`` (nstate,) or (nstate, nspin),
    ``c_matrix`` (nbf, nstate) or (nbf, nstate, nspin). Returns a dict with
    energy_qp, zz, energy_qp_lin (nstate, nspin), sigma (nomega, nse, nspin),
    omega, omega_re (the real grid) and de_max."""
    energy = np.asarray(energy, dtype=np.float64)
    occupation = np.asarray(occupation, dtype=np.float64)
    c_matrix = np.asarray(c_matrix, dtype=np.float64)
    if energy.ndim == 1:
        energy, occupation, c_matrix = energy[:, None], occupation[:, None], c_matrix[:, :, None]
    nstate, nspin = energy.shape
    if spin_fact is None:
        spin_fact = 2.0 if nspin == 1 else 1.0
    exc = np.zeros((nstate, nspin)) if exchange_m_vxc is None else np.asarray(exchange_m_vxc, dtype=np.float64)
    if exc.ndim == 1:
        exc = exc[:, None]
    rng = ranges or selfenergy_set_state_range(occupation)
    nom = nomega_sigma // 2
    omega = step_sigma * np.arange(-nom, nom + 1, dtype=np.float64)
    nse = rng.nsemax - rng.nsemin + 1
    se_energy0 = energy[rng.nsemin - 1:rng.nsemax, :]
    ttable = transition_table(occupation, spin_fact, rng.ncore_W, rng.nvirtual_W)
    wlo, whi = rng.ncore_W + 1, rng.nvirtual_W - 1
    x = transition_columns(eri_3center, index_basis, c_matrix, ttable, wlo, whi)
    i, a, sp = ttable[:, 0] - 1, ttable[:, 1] - 1, ttable[:, 2] - 1
    docc = occupation[i, sp] - occupation[a, sp]
    de = energy[a, sp] - energy[i, sp]
    omega_im, weight = imaginary_quad_grid(nomega_chi_imag)
    w_imag = vsqrt_chi_vsqrt_low_rank(x, docc, de, -omega_im ** 2, tol)
    de_max = contour_de_max(energy, se_energy0, omega, rng, eta)
    omega_re = real_linear_grid(nomega_chi_real, de_max)
    w_real = vsqrt_chi_vsqrt_low_rank(x, docc, de, omega_re ** 2, tol)
    out = {k: np.zeros((nstate, nspin)) for k in ('energy_qp', 'zz', 'energy_qp_lin')}
    sigma = np.zeros((len(omega), nse, nspin), dtype=np.complex128)
    for s in range(nspin):
        eri_pm = calculate_eri_3center_mo(eri_3center, index_basis, c_matrix, rng.ncore_G + 1,
                                          rng.nvirtual_G - 1, rng.nsemin, rng.nsemax, spin=s)
        sigma[:, :, s] = gw_selfenergy_contour(eri_pm, w_imag, omega_im, weight, w_real, omega_re, energy[:, s],
                                               rng, omega, se_energy0[:, s], eta)
        e_lin, _ = find_qp_energy_linearization(sigma[:, :, s], omega, se_energy0[:, s], exc[:, s],
                                                energy[:, s], rng)
        e_qp, zz = find_qp_energy_graphical(sigma[:, :, s], omega, se_energy0[:, s], exc[:, s], energy[:, s], rng)
        out['energy_qp'][:, s], out['zz'][:, s], out['energy_qp_lin'][:, s] = e_qp, zz, e_lin
    out.update(sigma=sigma, omega=omega, omega_re=omega_re, de_max=de_max)
    return out
