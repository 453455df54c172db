/*
 * mgw.h -- C-ABI of the MI355X-native RI G0W0 engine (libmgw.so).
 *
 * Every entry point replaces one Fortran subroutine of MOLGW's RI G0W0 hot
 * path (cited below as file:line under src/ of MOLGW 3.4). The Fortran
 * module state those subroutines read implicitly (eri_3center, c_matrix,
 * eri_3center_mo, wpol, se ...) lives in an opaque context that keeps the
 * tensors resident in HBM between calls.
 *
 * Conventions
 *  - Arrays are column-major FP64, exactly as the Fortran callers hold them.
 *  - State indices and ranges are 1-based, as in the Fortran module
 *    variables (nsemin, ncore_G ...).
 *  - Complex arrays are interleaved (re, im) pairs, i.e. complex(dp).
 *  - Every function returns 0 (MGW_OK) on success; on failure a non-zero
 *    code and mgw_last_error() describes it (the Fortran shim calls die()).
 *  - Output pointers marked "nullable" may be NULL: the result then stays
 *    resident in the context for the next call.
 *  - A context is not re-entrant; calls on one context are ordered on the
 *    context's HIP stream. Several contexts (one per concurrent solve) may
 *    share a GPU and run concurrently from different host threads.
 *  - Restricted (nspin = 1, spin_fact = 2, mgw_set_mo) or spin-polarised
 *    (nspin = 2, spin_fact = 1, mgw_set_mo_spin) MO data.
 */
#ifndef MGW_H
#define MGW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mgw_ctx mgw_ctx;

enum {
  MGW_OK = 0,
  MGW_ERR_ARG = 1,     /* invalid argument or inconsistent sizes        */
  MGW_ERR_HIP = 2,     /* HIP runtime error                             */
  MGW_ERR_SOLVER = 3,  /* eigensolver / factorisation failure           */
  MGW_ERR_STATE = 4,   /* required resident data not set                */
  MGW_ERR_NOMEM = 5    /* device allocation failed                      */
};

enum { MGW_MEM_HOST = 0, MGW_MEM_DEVICE = 1 };

/* Arithmetic of the transforms, the chi0 / Casida-matrix build and the
 * imaginary-axis self-energy product (the metric, the eigensolve and the QP
 * solve are FP64 in every mode; the (1 - chi0) factorisation is FP64 except
 * as stated for MGW_FP16X3):
 *   MGW_FP32   FP32 products on MFMA, FP64 accumulation
 *   MGW_BF16X3 two-way bf16 split, 3 products (~2^-17 relative)
 *   MGW_BF16X6 three-way bf16 split for chi0 and the self-energy product,
 *              6 products (~2^-24, FP32 level); the transforms stay FP64
 *   MGW_FP16X3 two-way fp16 split of power-of-two-scaled operands for chi0,
 *              the self-energy product, the AO->MO transforms (FP32
 *              accumulation over K = nbf) and the products of the batched
 *              (1 - chi0) factorisation at levels >= 128 wide (FP64
 *              64 x 64 leaves, FP64 sums across products), 3 products
 *              (~2^-22); the environment variables MGW_AO2MO_FP16X3=0 and
 *              MGW_CHOL_F16_MIN=0 keep the transforms and the factorisation
 *              FP64                                                       */
enum { MGW_FP64 = 0, MGW_FP32 = 1, MGW_BF16X3 = 2, MGW_BF16X6 = 3, MGW_FP16X3 = 4 };

enum { MGW_TECH_ONE_SHOT = 0, MGW_TECH_IMAGINARY_PADE = 1, MGW_TECH_CONTOUR = 2 };

/* State ranges held as module variables in the reference:
 * ncore_G, nvirtual_G, nsemin, nsemax, nhomo_G, nlumo_G (src/m_selfenergy_tools.f90:30-41,
 * :97-110) and ncore_W, nvirtual_W (src/m_spectral_function.f90:95-107).
 * nlumo_G = 0 means nhomo_G + 1 (integer occupations); only the contour
 * deformation self-energy reads it. */
typedef struct mgw_ranges {
  int32_t ncore_G;
  int32_t nvirtual_G;
  int32_t nsemin;
  int32_t nsemax;
  int32_t nhomo_G;
  int32_t ncore_W;
  int32_t nvirtual_W;
  int32_t nlumo_G;
} mgw_ranges;

/* One-shot G0W0 parameters (input keywords, src/input_variables.yaml:
 * nomega_sigma :1164, step_sigma :1384, eta :1051, nomega_chi_imag :1148, nomega_chi_real :1156,
 * nomega_sigma_calc :1173, step_sigma_calc :1392). */
typedef struct mgw_g0w0_params {
  int32_t technique;          /* MGW_TECH_*                                   */
  int32_t nomega_sigma;       /* real grid points (odd), default 51           */
  double  step_sigma;         /* Ha, default 0.01                             */
  double  eta;                /* Ha, default 0.001                            */
  int32_t nomega_chi_imag;    /* imaginary chi grid (Pade technique only)     */
  int32_t nomega_sigma_calc;  /* imaginary sigma points (Pade technique only) */
  double  step_sigma_calc;    /* Ha (Pade technique only)                     */
  double  spin_fact;          /* 2 for restricted closed shell                */
  int32_t precision;          /* MGW_FP64 | MGW_FP32 | MGW_BF16X3 | MGW_BF16X6
                                 | MGW_FP16X3 (transforms, chi0, the
                                 imaginary-axis self-energy product and, for
                                 MGW_FP16X3, the large products of the
                                 (1 - chi0) factorisation; metric, eigensolve
                                 and the QP solve are always FP64)            */
  int32_t nomega_chi_real;    /* real-axis W grid, >= 2 (contour technique;
                                 input keyword nomega_chi_real, default 2)    */
} mgw_g0w0_params;

/* ---------------------------------------------------------------- context */
int32_t     mgw_abi_version(void);
const char* mgw_last_error(void);
int mgw_ctx_create(int device, mgw_ctx** ctx);
int mgw_ctx_destroy(mgw_ctx* ctx);
int mgw_ctx_synchronize(mgw_ctx* ctx);
/* Free the context's cached temporaries (everything but the resident RI
 * tensor and MO data); with eager != 0, later stages also free their
 * temporaries as they finish (the reference's destroy_* calls after each
 * stage). Temporaries are cached by default because freeing device memory
 * synchronises the whole GPU. No reference counterpart beyond
 * clean_deallocate (src/m_memory.f90:238-274). */
int mgw_ctx_trim(mgw_ctx* ctx, int eager);
/* With enable != 0, the batched Cholesky sequence of 1 - chi0 (about
 * 5 nauxil / 64 kernel launches per solve) is captured into a HIP graph once
 * per shape and buffer set and replayed on later solves. No reference
 * counterpart (launch batching for the serving loop). */
int mgw_ctx_graphs(mgw_ctx* ctx, int enable);
/* Kernel autotuner (process-wide, no reference counterpart): mode 1 times
 * the candidates (GEMM tile configurations, the chi0 SYRK variants, the W.B
 * dot pipelines; all give identical results) on the first launch of each
 * new shape class and keeps the fastest; 0 uses the table or the measured defaults;
 * -1 only queries. The table is read from and appended to the file named by
 * the environment variable MGW_TUNE_FILE, if set. `previous` (nullable)
 * receives the mode before the call. */
int mgw_autotune(int mode, int* previous);
/* HBM ledger: replaces the clean_allocate bookkeeping (src/m_memory.f90:19-80). */
int mgw_memory_report(mgw_ctx* ctx, int64_t* current_bytes, int64_t* peak_bytes);
/* Bytes the device's default memory pool (from which every context
 * allocates, stream-ordered) holds now and at most since the process began
 * or the last reset (nullable outputs); reset_high != 0 restarts the
 * high-water mark after reading it. mgw_ctx_trim hands the pool's unused
 * memory back to the device. No reference counterpart (instrumentation of
 * the ledger). */
int mgw_device_pool_report(int device, int reset_high, int64_t* reserved_bytes, int64_t* reserved_high_bytes);
/* Release threshold of the device's default memory pool: at each
 * synchronisation the pool hands memory that no allocation uses back to the
 * device while it holds more than `bytes` (a serving loop passes its HBM
 * budget). Contexts start with no limit (the pool keeps freed memory for
 * the next allocation). No reference counterpart. */
int mgw_device_pool_limit(int device, int64_t bytes);
/* Caller-side device memory, for callers without a HIP runtime of their own
 * (the Python host side): a plain device allocation, its release, copies
 * (kind 0 host to device, 1 device to host, 2 device to device; dst and src
 * at byte offsets the caller computed) and the device's free and total
 * memory. Buffers from here may be passed wherever an entry point takes
 * `where` = MGW_MEM_DEVICE. The Fortran caller's counterpart is its own
 * device allocation. No reference counterpart. */
int mgw_device_alloc(int device, int64_t bytes, void** ptr);
int mgw_device_free(void* ptr);
int mgw_device_copy(void* dst, const void* src, int64_t bytes, int kind);
int mgw_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes);
/* Frequencies whose reduced-precision 1 - chi0 lost positive definiteness
 * in the Cholesky factorisation and were rebuilt in FP64 on the device,
 * counted since the context was created. No reference counterpart: the
 * reference has no reduced-precision build. */
int mgw_precision_escalations(mgw_ctx* ctx, int64_t* count);
/* Bytes of HBM one solve of this shape needs (resident RI tensor included);
 * seeded by the reference's evaluate_memory formulas (src/m_io.f90:3843-3936).
 * For MGW_TECH_CONTOUR pass nomega_chi_imag + nomega_chi_real as
 * nomega_chi_imag. */
int64_t mgw_footprint_bytes(int nbf, int nauxil, int nstate, int nocc, int technique,
                            int nomega_chi_imag, int precision);
/* The same for a self-energy range of nse states (nsemax - nsemin + 1): the
 * buffers of the self-energy stage scale with it, so a solve restricted to a
 * window around the gap needs far less than the whole range (nse = nstate,
 * as mgw_footprint_bytes). No reference counterpart (serving admission). */
int64_t mgw_footprint_bytes_nse(int nbf, int nauxil, int nstate, int nocc, int nse, int technique,
                                int nomega_chi_imag, int precision);

/* ---------------------------------- integral front end (upstream of A1) */
/* A set of contracted Gaussian shells: the reference's basis%shell(:)
 * (am, ng, alpha, coeff, x0, istart; src/m_basis_set.f90:90-364) with the
 * map from Cartesian components to the final (pure or Cartesian)
 * functions (src/m_cart_to_pure.f90). Host arrays over nshell shells:
 *   am[s]                angular momentum
 *   x0[3 s + d]          centre (bohr)
 *   prim_offset[s]       primitives of shell s: prim_offset[s] .. prim_offset[s+1]-1
 *                        (nshell + 1 entries) into
 *   alpha[], coeff[]     exponents and weights of the unnormalised Cartesian
 *                        primitives x^i y^j z^k exp(-alpha r^2) (normalisation
 *                        of the contraction included)
 *   transform_offset[s]  start in transform[] of shell s's (nfunc x ncart)
 *                        row-major map; Cartesian components ordered
 *                        (l - i, i - k, k) for i = 0..l, k = 0..i
 *   function_offset[s]   first final function (0-based; nshell + 1 entries)
 */
typedef struct mgw_shells {
  int32_t nshell;
  const int32_t* am;
  const double* x0;
  const int32_t* prim_offset;
  const double* alpha;
  const double* coeff;
  const int32_t* transform_offset;
  const double* transform;
  const int32_t* function_offset;
} mgw_shells;

/* 2-centre Coulomb integrals (P|1/r12|Q) over the final auxiliary functions:
 * the integral part of calculate_integrals_eri_2center_scalapack
 * (src/m_eri_calculate.f90:703-945), which the reference takes from libcint
 * (cint2c2e, src/m_libcint_tools.f90). eri_2center (naux x naux,
 * column-major; `where` host or device) receives both triangles; it is the
 * input of mgw_inverse_sqrt_eri_2center. Computed on the GPU (orbital
 * shells l <= 3, auxiliary shells l <= 4). */
int mgw_eri_2center_ao(mgw_ctx* ctx, const mgw_shells* auxil, int where, double* eri_2center);
/* 3-centre Coulomb integrals (mu nu|P) of the pairs index_basis(2, npair)
 * (1-based, host): the integral part of calculate_integrals_eri_3center_scalapack
 * (src/m_eri_calculate.f90:1266-1580), i.e. eri_3center_tmp(npair, naux_bas)
 * with the diagonal pairs x1/2 (:1840-1841), column-major. eri_3center_tmp
 * (`where` host or device) is nullable: the result then stays resident in
 * the context and mgw_build_eri_3center called with eri_3center_tmp = NULL
 * turns it into the RI tensor without a copy. Computed on the GPU. */
int mgw_eri_3center_ao(mgw_ctx* ctx, const mgw_shells* basis, const mgw_shells* auxil, int64_t npair,
                       const int32_t* index_basis, int where, double* eri_3center_tmp);
/* Hartree and exchange matrices through the resident RI tensor, one spin
 * channel: setup_hartree_ri (src/m_hamiltonian_twobodies.f90:280-381) and
 * setup_exchange_ri (:707-786). x (nbf x nx, host, column-major) holds the
 * occupied orbitals scaled by sqrt(occupation / spin_fact) (the reference's
 * c_t); the density is P = spin_fact x x^T. hartree_ao(mu, nu) =
 * sum_P B(P, mu nu) sum_kl B(P, k l) P(k, l) and exchange_ao = -sum_P
 * (B_P x)(B_P x)^T (nbf x nbf, host, nullable). FP64 MFMA products. */
int mgw_hartree_exchange_ri(mgw_ctx* ctx, const double* x, int nbf, int nx, double spin_fact,
                            double* hartree_ao, double* exchange_ao);

/* ------------------------------------------------- A1: resident RI tensor */
/* Replaces the module state eri_3center(npair, nauxil) and index_basis(2, npair)
 * (src/m_eri.f90:36-62) as produced by calculate_eri_ri
 * (src/m_eri_calculate.f90:609-699): packed pairs, diagonal pairs stored x1/2.
 * index_basis is 1-based, (2, npair) column-major. `where` = MGW_MEM_HOST or
 * MGW_MEM_DEVICE for eri_3center; index_basis is always host memory. */
int mgw_set_eri_3center(mgw_ctx* ctx, const double* eri_3center, int64_t npair, int nauxil,
                        const int32_t* index_basis, int nbf, int where);
int mgw_destroy_eri_3center(mgw_ctx* ctx);

/* Inverse square root of the 2-center Coulomb metric (P|1/r12|Q):
 * calculate_inverse_sqrt_eri_2center_scalapack (src/m_eri_calculate.f90:1049-1262),
 * full-range branch. eri_2center is nauxil_bas x nauxil_bas column-major
 * (`where`: host or device), of which the lower triangle is read. The metric
 * is diagonalised in FP64 on the device; eigenvalues <= 1e-6
 * (TOO_LOW_EIGENVAL, src/m_eri.f90:27) are dropped and
 * sqrtinv(:, j) = U(:, j + nneglect) / sqrt(lambda(j + nneglect)) stays
 * resident for mgw_build_eri_3center. nauxil_kept receives the kept count
 * (the reference's nauxil_global), lowest_eigval (nullable) the lowest
 * eigenvalue, sqrtinv_out (nullable; host or device, nauxil_bas x nauxil_kept
 * column-major, so nauxil_bas^2 doubles suffice) a copy of the matrix. */
int mgw_inverse_sqrt_eri_2center(mgw_ctx* ctx, const double* eri_2center, int nauxil_bas, int where,
                                 int* nauxil_kept, double* lowest_eigval, double* sqrtinv_out);
/* Second part of calculate_eri_3center_scalapack (src/m_eri_calculate.f90:1877-1892):
 * eri_3center(ipair, P) = sum_Q eri_3center_tmp(ipair, Q) sqrtinv(Q, P) for
 * the batch of pairs ipair_first .. ipair_first + mpair - 1 (1-based) of
 * npair, FP64. eri_3center_tmp is mpair x nauxil_bas column-major (`where`:
 * host or device) and already carries the 1/2 of the diagonal pairs
 * (src/m_eri_calculate.f90:1840-1841); index_basis is the full (2, npair)
 * host list. eri_3center_tmp = NULL takes the rows from the tensor that
 * mgw_eri_3center_ao left resident (same npair and naux_bas). Batches follow each other from ipair_first = 1; after the batch
 * that ends at npair the tensor is resident exactly as after
 * mgw_set_eri_3center with nauxil = nauxil_kept. eri_3center_out (nullable;
 * host or device) is the caller's whole eri_3center(npair, nauxil_kept),
 * of which this batch's rows are written. */
int mgw_build_eri_3center(mgw_ctx* ctx, const double* eri_3center_tmp, int64_t ipair_first, int64_t mpair,
                          int64_t npair, const int32_t* index_basis, int nbf, int where,
                          double* eri_3center_out);

/* MO coefficients, energies and occupations of the (single) spin channel:
 * the c_matrix(:,:,1), energy(:,1), occupation(:,1) arguments of the
 * reference routines. */
int mgw_set_mo(mgw_ctx* ctx, const double* c_matrix, int nbf, int nstate,
               const double* energy, const double* occupation, int where);
/* nspin (1 or 2) channels: c_matrix(nbf, nstate, nspin), energy(nstate, nspin),
 * occupation(nstate, nspin), column-major like the reference's arrays. W sums
 * the transitions of every channel (transition table spin outer,
 * src/m_spectral_function.f90:239-266); the AO->MO, self-energy and QP calls
 * work on the channel chosen by mgw_set_spin (channel 1 after this call),
 * and mgw_g0w0 loops over the channels. With nspin = 2 pass spin_fact = 1
 * (the reference's spin_fact = 3 - nspin, src/m_inputparam.f90:1059). */
int mgw_set_mo_spin(mgw_ctx* ctx, const double* c_matrix, int nbf, int nstate, int nspin,
                    const double* energy, const double* occupation, int where);
/* Active channel (1-based) of the AO->MO, self-energy and QP calls: the
 * reference's spin index of eri_3center_mo(:, :, :, ispin)
 * (src/m_gw_selfenergy_grid.f90:279, src/m_gw_selfenergy_analytic.f90:80-102). */
int mgw_set_spin(mgw_ctx* ctx, int ispin);

/* Energies in G for the following self-energy calls (energy_g of
 * selfenergy_evaluation, src/m_selfenergy_evaluation.f90:164-183: the QP
 * energies of the previous step of a GnW0 / GnWn = evGW run), nstate values,
 * host. The resident W and the energies in W (those of mgw_set_mo) are kept;
 * NULL returns to the energies of mgw_set_mo, as does mgw_set_mo itself. */
int mgw_set_energy_g(mgw_ctx* ctx, const double* energy_g);

/* ------------------------------------------------ A10: Sigma_x - Vxc */
/* setup_exchange_m_vxc (src/m_selfenergy_tools.f90:720-802), the branch
 * without a DFT core-valence split (:786-797): exchange_m_vxc(p, q, s) =
 * (C^T F C)(p, q) - delta_pq energy(p, s) over the resident MO data of every
 * spin channel, with hamiltonian_fock (nbf x nbf x nspin, host,
 * column-major) the Fock matrix of the starting point (for HF: h + J +
 * exchange at the final density). exchange_m_vxc (nstate x nstate x nspin)
 * and diag (nstate x nspin: its diagonal, the exchange_m_vxc argument of
 * mgw_g0w0 and mgw_find_qp_energy) are host arrays, both nullable. FP64. */
int mgw_exchange_m_vxc(mgw_ctx* ctx, const double* hamiltonian_fock, double* exchange_m_vxc, double* diag);

/* ------------------------------------------------ A2: AO -> MO transform */
/* calculate_eri_3center_mo(c_matrix, mstate_min, mstate_max, nstate_min,
 * nstate_max) (src/m_eri_ao_mo.f90:416-582). Result eri_3center_mo(P, m, n)
 * stays resident; it is also copied to `eri_3center_mo` (nullable, host,
 * nauxil x (mmax-mmin+1) x (nmax-nmin+1)). Like the reference, returns
 * immediately when a block is already resident (src/m_eri_ao_mo.f90:444-446). */
int mgw_calculate_eri_3center_mo(mgw_ctx* ctx, int mstate_min, int mstate_max,
                                 int nstate_min, int nstate_max, double* eri_3center_mo);
/* destroy_eri_3center_mo (src/m_eri_ao_mo.f90:1018). */
int mgw_destroy_eri_3center_mo(mgw_ctx* ctx);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the transform of mgw_calculate_eri_3center_mo for the
 * active spin channel, FP64, written straight into the caller's device
 * buffer eri_3center_mo_dev (nauxil x (mmax-mmin+1) x (nmax-nmin+1),
 * column-major, P fastest). Nothing is allocated and nothing is left
 * resident: the half-transformed tensor (nauxil x nbf x min(m range,
 * n range) doubles) goes to a fixed device scratch array of
 * MGW_NOALLOC_SCRATCH_DOUBLES doubles (64 MiB) per GPU, held until the
 * process ends; larger blocks are refused with MGW_ERR_ARG. The launches
 * are queued on the context's stream and the call returns without waiting
 * (mgw_ctx_synchronize waits). Every context on a GPU shares its scratch
 * array: calls that overlap on one GPU race. Needs the resident RI tensor
 * and MO data (mgw_set_eri_3center or mgw_build_eri_3center, mgw_set_mo). */
#define MGW_NOALLOC_SCRATCH_DOUBLES (8 * 1024 * 1024)
int mgw_calculate_eri_3center_mo_noalloc(mgw_ctx* ctx, int mstate_min, int mstate_max, int nstate_min,
                                         int nstate_max, double* eri_3center_mo_dev);

/* -------------------------------------------- A3/A4: Casida RPA (analytic) */
/* Number of resonant transitions of the transition table
 * (src/m_spectral_function.f90:224-266). */
int mgw_npole(mgw_ctx* ctx, const mgw_ranges* ranges, int* npole);
/* RPA branch of polarizability(enforce_rpa, calculate_w=.TRUE., ...)
 * (src/m_linear_response.f90:33-498) with build_amb_apb_diag_auxil,
 * build_apb_hartree_auxil[_scalapack], get_rpa_correlation
 * (src/m_build_bse.f90:218-491), diago_4blocks_rpa_sca
 * (src/m_block_diago.f90:142-204) and chi_to_sqrtvchisqrtv_auxil
 * (src/m_linear_response.f90:802-922). Fills wpol%pole and wpol%w_s
 * (resident; copied to pole_out[npole] and w_s_out[nauxil x npole] when
 * non-NULL), en_rpa and en_gw (Galitskii-Migdal). */
int mgw_polarizability(mgw_ctx* ctx, const mgw_ranges* ranges, double spin_fact, int precision,
                       double* en_rpa, double* en_gw, double* pole_out, double* w_s_out);
/* read_spectral_function (src/m_spectral_function.f90:508-620): W made
 * resident from its poles pole[npole] and residues w_s(nauxil, npole)
 * column-major (the arrays of a SCREENED_COULOMB file), in place of
 * mgw_polarizability; nauxil must match the resident RI tensor. */
int mgw_set_spectral_function(mgw_ctx* ctx, int npole, int nauxil, const double* pole, const double* w_s,
                              int where);
/* dump_gw_ingredients (src/m_gw_selfenergy_analytic.f90:801-881) for the
 * active spin channel, states p, q in ncore_G+1 .. nvirtual_G-1 (ng of
 * them): v_out(nauxil, ng, ng) = eri_3center_mo(P, p, q) (the v.bin
 * records) and w_out(npole, ng, ng) = sum_P w_s(P, s) v(P, p, q) (the w.bin
 * records). Both nullable; w_out needs the resident spectral function. */
int mgw_gw_ingredients(mgw_ctx* ctx, const mgw_ranges* ranges, double* w_out, double* v_out);


/* ----------------------------------------- A5: chi0 on the imaginary axis */
/* polarizability_grid_scalapack (src/m_gw_selfenergy_grid.f90:33-209):
 * chi(iw) = (1 - chi0(iw))^-1 chi0(iw) for the nomega points omega_im with
 * weights weight_quad (from sf_init, src/m_spectral_function.f90:286-306).
 * chi stays resident (wpol%chi); chi_out nullable (nauxil x nauxil x nomega). */
int mgw_polarizability_grid(mgw_ctx* ctx, const mgw_ranges* ranges, double spin_fact, int precision,
                            int nomega, const double* omega_im, const double* weight_quad,
                            double* erpa, double* egw, double* chi_out);

/* ------------------------------------------ A6: analytic-pole self-energy */
/* gw_selfenergy(GW, ...) (src/m_gw_selfenergy_analytic.f90:29-196) using the
 * resident wpol from mgw_polarizability. omega[nomega] is se%omega (real),
 * se_energy0[nse] is se%energy0. sigma (nomega x nse, complex) stays resident;
 * sigma_out nullable. energy_gw = 1/2 Tr[Sig_c G] (:140-143). */
int mgw_gw_selfenergy(mgw_ctx* ctx, const mgw_ranges* ranges, double spin_fact, double eta,
                      int nomega, const double* omega, const double* se_energy0,
                      double* sigma_out, double* energy_gw);
/* gw_selfenergy_qs, GW branch (src/m_gw_selfenergy_analytic.f90:653-797),
 * with apply_qs_approximation (src/m_selfenergy_tools.f90:809-828), for the
 * active spin channel on the resident Casida W: Sigma(p, q) = sum_{i,s}
 * w(s,p,i) w(s,q,i) Re[f_i / (e_q - i eta - e_i + W_s) + (1 - f_i) /
 * (e_q + i eta - e_i - W_s)] over p, q in nsemin..nsemax, symmetrised
 * (Sigma + Sigma^T) / 2. selfenergy_mo (nstate x nstate, zero outside the
 * window) and selfenergy_ao = S C Sigma C^T S (nbf x nbf, needs s_matrix,
 * nbf x nbf) are nullable. */
int mgw_gw_selfenergy_qs(mgw_ctx* ctx, const mgw_ranges* ranges, double spin_fact, double eta,
                         const double* s_matrix, double* selfenergy_mo, double* selfenergy_ao);

/* ------------------------------------------ A7: imaginary-axis self-energy */
/* gw_selfenergy_imag_scalapack (src/m_gw_selfenergy_grid.f90:213-338) using
 * the resident chi. omega_calc[nomega_calc] complex (se%omega_calc);
 * sigma_calc (nomega_calc x nse complex) resident, sigma_calc_out nullable. */
int mgw_gw_selfenergy_imag(mgw_ctx* ctx, const mgw_ranges* ranges, int nomega_calc,
                           const double* omega_calc, double* sigma_calc_out);

/* ------------------------------- contour deformation self-energy (f)-3 */
/* gw_selfenergy_contour (src/m_gw_selfenergy_grid.f90:342-551) for every spin
 * channel, with W in the low-rank form of sf_vsqrt_chi_vsqrt_rpa(low_rank)
 * (src/m_spectral_function.f90:768-949): chi0 of all channels' transitions
 * (energies in G, as the reference passes energy_g) on the imaginary
 * Gauss-Legendre grid (nomega_chi_imag points) and on the real grid of
 * nomega_chi_real >= 2 points over [0, de_max] (:308-315), where de_max is the
 * largest residue frequency of the reference's scan (:385-408); eigenpairs
 * with |lambda / (1 - lambda)| <= 1e-3 dropped. omega[nomega] real (the grid
 * se%omega around se_energy0), se_energy0 [nse x nspin]. sigma_out
 * [nomega x nse x nspin complex, zero imaginary part] nullable; the active
 * channel's sigma stays resident for mgw_find_qp_energy; de_max_out
 * nullable. A residue frequency outside the real grid is an error, as in
 * sf_interpolate_vsqrt_chi_vsqrt (:970-973). FP64. */
int mgw_gw_selfenergy_contour(mgw_ctx* ctx, const mgw_ranges* ranges, double spin_fact, int nomega_chi_imag,
                              int nomega_chi_real, double eta, int nomega, const double* omega,
                              const double* se_energy0, double* sigma_out, double* de_max_out);

/* ------------------------------------------------ A8: Pade continuation */
/* se_pade_fit (src/m_selfenergy_tools.f90:965-999) with pade/calculate_pade_a
 * (src/m_numerical_tools.f90:86-145) on the resident sigma_calc (or on
 * sigma_calc when non-NULL). omega[nomega] real: the real grid se%omega.
 * Result sigma (nomega x nse complex) resident; sigma_out nullable. */
int mgw_pade_fit(mgw_ctx* ctx, int nse, int nomega_calc, const double* omega_calc,
                 const double* sigma_calc, int nomega, const double* omega, double eta,
                 double* sigma_out);

/* ---------------------------------------------------- A9: QP equation */
/* find_qp_energy_graphical + find_fixed_point (src/m_selfenergy_tools.f90:253-420)
 * and find_qp_energy_linearization (:202-249) on the resident sigma (or on
 * `sigma` when non-NULL, nomega x nse complex). energy_qp/zz/energy_qp_lin
 * have nstate entries; states outside [nsemin, nsemax] get energy0. */
int mgw_find_qp_energy(mgw_ctx* ctx, const mgw_ranges* ranges, int nomega, const double* omega,
                       const double* sigma, const double* se_energy0,
                       const double* exchange_m_vxc, const double* energy0, int nstate,
                       double* energy_qp, double* zz, double* energy_qp_lin);

/* ---------------------------------------------- A11: one G0W0 solve */
/* The one-shot G0W0 flow of selfenergy_evaluation
 * (src/m_selfenergy_evaluation.f90:135-261, 589-590) on the resident RI
 * tensor and MO data. exchange_m_vxc[nstate] is the diagonal of
 * Sigma_x - Vxc (setup_exchange_m_vxc, src/m_selfenergy_tools.f90:720-802);
 * NULL means zero (converged HF). Outputs (all nullable): energy_qp, zz,
 * energy_qp_lin [nstate]; sigma_out (nomega_sigma x nse complex);
 * energies[2] = {E_RPA, E_GW}. With two spin channels (mgw_set_mo_spin)
 * exchange_m_vxc and the state outputs are [nstate x nspin], sigma_out is
 * [nomega_sigma x nse x nspin], and E_GW sums the channels.
 * MGW_TECH_CONTOUR runs gw_selfenergy_contour with the one-shot grid
 * (src/m_selfenergy_tools.f90:658-665, :679-680); its energies[] are 0, as
 * the reference computes no RPA energy on that route. */
int mgw_g0w0(mgw_ctx* ctx, const mgw_g0w0_params* params, const mgw_ranges* ranges,
             const double* exchange_m_vxc, double* energy_qp, double* zz,
             double* energy_qp_lin, double* sigma_out, double* energies);

/* ---------------------------------------- instrumentation (no reference counterpart) */
/* Live timing of the engine's kernel classes with HIP events on the context's
 * stream, for roofline reporting; the reference has only CPU timers
 * (src/m_timing.f90:378-424). mgw_profile_enable(ctx, 1) clears the records.
 * mgw_profile_read sums over the recorded launches of one class: launch
 * count, device time in ms, and algorithmic flops and HBM bytes. */
enum {
  MGW_KCLASS_GEMM_AO2MO = 0,   /* (P|mu nu) C contractions                 */
  MGW_KCLASS_GEMM_CASIDA = 1,  /* Casida SYRK and w_s = B (X+Y)            */
  MGW_KCLASS_GEMM_SIGMA = 2,   /* w_s^T (P|p i) of the pole self-energy    */
  MGW_KCLASS_SIGMA_POLES = 3,  /* pole sum (FP64 VALU)                     */
  MGW_KCLASS_EIGEN = 4,        /* eigensolvers (Casida, 2-center metric)   */
  MGW_KCLASS_GEMM_CHI0 = 5,    /* 1 - chi0(iw), all frequencies            */
  MGW_KCLASS_CHOLESKY = 6,     /* (1 - chi0)^-1 by Cholesky                */
  MGW_KCLASS_GEMM_WB = 7,      /* chi(iw) (P|p m) with fused dot           */
  MGW_KCLASS_GEMM_OTHER = 8,
  MGW_KCLASS_OTHER = 9,
  MGW_KCLASS_COUNT = 10
};
int mgw_profile_enable(mgw_ctx* ctx, int enable);
int mgw_profile_read(mgw_ctx* ctx, int kernel_class, int64_t* launches, double* total_ms,
                     double* algorithmic_flops, double* algorithmic_bytes);

#ifdef __cplusplus
}
#endif

#endif /* MGW_H */
