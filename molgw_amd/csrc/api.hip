// C-ABI of libmgw: context, resident tensors and the orchestration of the
// RI G0W0 hot path. The math of every step is documented at the kernel that
// does it; this file owns data movement, sizes and the call order of
// src/m_selfenergy_evaluation.f90 (MOLGW 3.4).
#include "mgw_internal.h"

#include <mutex>
#include <set>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <tuple>
#include <cstring>
#include <numeric>

namespace mgw {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

hipEvent_t Profiler::next_event() {
  if (used == pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    pool.push_back(e);
  }
  return pool[used++];
}

void Profiler::reset() {
  recs.clear();
  used = 0;
}

Profiler::~Profiler() {
  for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

ProfScope::ProfScope(Profiler* prof, int klass, double flops, double bytes, hipStream_t stream)
    : p(prof && prof->on ? prof : nullptr), idx(0), s(stream) {
  if (!p) return;
  hipEvent_t a = p->next_event(), b = p->next_event();
  if (!a || !b) { p = nullptr; return; }
  (void)hipEventRecord(a, s);
  p->recs.push_back({klass, flops, bytes, a, b});
  idx = p->recs.size() - 1;
}

ProfScope::~ProfScope() {
  if (p) (void)hipEventRecord(p->recs[idx].b, s);
}

}  // namespace mgw

using namespace mgw;

static_assert(sizeof(mgw_ranges) == 32, "mgw_ranges layout is part of the ABI");
static_assert(sizeof(mgw_g0w0_params) == 56, "mgw_g0w0_params layout is part of the ABI");

// A library call: its status, then any launch error its kernels left, so
// that the error is reported at this call and not at the next launch.
#define MGW_BLAS(expr)                                                                  \
  do {                                                                                  \
    rocblas_status st_ = (expr);                                                        \
    if (st_ != rocblas_status_success) {                                                \
      set_error(std::string(#expr) + ": " + rocblas_status_to_string(st_));             \
      (void)hipGetLastError();                                                          \
      return MGW_ERR_SOLVER;                                                            \
    }                                                                                   \
    hipError_t he_ = hipGetLastError();                                                 \
    if (he_ != hipSuccess) {                                                            \
      set_error(std::string(#expr) + ": launch error " + hipGetErrorString(he_) + " (" +  \
                __FILE__ + ":" + std::to_string(__LINE__) + ")");                      \
      return MGW_ERR_SOLVER;                                                            \
    }                                                                                   \
  } while (0)

#define MGW_CTX(ctx)                                                           \
  do {                                                                         \
    if (!(ctx)) {                                                              \
      set_error("null context");                                               \
      return MGW_ERR_ARG;                                                      \
    }                                                                          \
    MGW_HIP(hipSetDevice((ctx)->device));                                      \
  } while (0)

struct mgw_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  rocblas_handle blas = nullptr;
  Ledger ledger;
  Profiler prof;
  bool trim_eagerly = false;
  // A1: resident RI tensor, dense B(P, k, l)
  bool has_ri = false;
  int nbf = 0, naux = 0;
  // A1 build: kept inverse square root of the 2-center metric ("ri_sqrtinv",
  // naux_bas x naux_kept) and the next pair of a batched build
  bool has_sqrtinv = false;
  int naux_bas = 0, naux_kept = 0;
  // integral front end: eri_3center_tmp left resident by mgw_eri_3center_ao
  // ("ri_tmp3", npair x naux_bas) for mgw_build_eri_3center
  bool has_ri_tmp3 = false;
  int64_t ri_tmp3_npair = 0;
  int ri_tmp3_naux = 0;
  int64_t build_next = 0, build_npair = 0;
  int build_nbf = 0;
  // MO data of nspin channels, each vector spin-major (nstate per channel);
  // energy_g (empty: energy) are the energies in G of the self-energy, energy
  // those of W. The self-energy, AO->MO and QP calls work on channel `spin`;
  // W sums the transitions of all channels.
  bool has_mo = false;
  int nstate = 0, nspin = 1, spin = 0;
  int mo_nbf = 0;   // rows of c_matrix (must equal nbf of the resident RI tensor)
  std::vector<double> energy, occupation, energy_g;
  const std::vector<double>& energy_in_g() const { return energy_g.empty() ? energy : energy_g; }
  // offset of the active channel in the spin-major vectors
  int so() const { return spin * nstate; }
  // the half-transformed tensor H(P k, f) = B(P k, l) C(l, f) in buffer
  // "ao2mo_half" for states f = half_lo .. half_hi of channel half_spin, made
  // with arithmetic half_prec (and, in FP16x3, its maximum in "ao2mo_max"):
  // a later transform whose first states lie in that range skips the first
  // product
  bool half_valid = false;
  int half_lo = 0, half_hi = 0, half_spin = 0, half_prec = 0;
  // A2 module state eri_3center_mo(P, m, n)
  bool has_eri_mo = false;
  int mlo = 0, mhi = 0, nlo = 0, nhi = 0;
  // spectral function
  bool has_wpol = false;
  int npole = 0;
  bool has_chi = false;
  int nomega_chi = 0;
  std::vector<double> chi_omega, chi_weight;
  // contour deformation: low-rank W ("cd_u" eigenvectors, "cd_w" weights) on
  // cd_om_im (imaginary, weights cd_wt_im) then cd_om_re (real)
  bool has_cd = false;
  std::vector<double> cd_om_im, cd_wt_im, cd_om_re;
  // self-energy
  bool has_sigma = false;
  int sigma_nomega = 0, sigma_nse = 0;
  bool has_sigma_calc = false;
  int sigma_ncalc = 0, sigma_calc_nse = 0;
  // frequencies of a reduced-precision chi0 rebuilt in FP64
  int64_t escalations = 0;
  // HIP graphs of the batched Cholesky sequence, one per shape and buffer set
  struct GraphKey {
    int n, nbatch;
    const void *a, *logdet, *info, *work, *bounds, *planes;
    int f16_min;
    bool operator<(const GraphKey& o) const {
      return std::tie(n, nbatch, a, logdet, info, work, bounds, planes, f16_min) <
             std::tie(o.n, o.nbatch, o.a, o.logdet, o.info, o.work, o.bounds, o.planes, o.f16_min);
    }
  };
  bool use_graphs = false;
  std::map<GraphKey, hipGraphExec_t> graphs;
};

namespace {

// modes whose chi0 and imaginary-axis self-energy product run on split operands
bool split_precision(int p) { return p == MGW_BF16X3 || p == MGW_BF16X6 || p == MGW_FP16X3; }

// Everything computed from the RI tensor and the MO data: new inputs make it stale.
void invalidate_derived(mgw_ctx* c) {
  c->half_valid = false;
  c->has_eri_mo = false;
  c->has_wpol = false;
  c->has_chi = false;
  c->has_cd = false;
  c->has_sigma = false;
  c->has_sigma_calc = false;
}

// The MO coefficients must have as many rows as the resident RI tensor has
// basis functions (a new RI tensor of another molecule needs new MO data).
int check_mo_matches_ri(const mgw_ctx* c) {
  MGW_REQUIRE(c->has_mo, MGW_ERR_STATE, "MO data not set (mgw_set_mo)");
  MGW_REQUIRE(c->has_ri, MGW_ERR_STATE, "RI tensor not set (mgw_set_eri_3center)");
  MGW_REQUIRE(c->mo_nbf == c->nbf, MGW_ERR_STATE,
              "MO data (nbf " + std::to_string(c->mo_nbf) + ") do not match the resident RI tensor (nbf " +
                  std::to_string(c->nbf) + "): call mgw_set_mo again");
  return MGW_OK;
}

// ------------------------------------------------------------------ buffers
int ensure(mgw_ctx* c, const std::string& name, size_t bytes, void** out) {
  DevBuf& b = c->ledger.bufs[name];
  if (bytes == 0) bytes = 8;
  if (b.bytes < bytes) {
    // stream-ordered: the old buffer returns to the device's pool once this
    // stream's earlier work is done, without waiting for the other streams
    // of the device (hipFree would wait for all of them)
    if (b.ptr) {
      MGW_HIP(hipFreeAsync(b.ptr, c->stream));
      c->ledger.current -= (int64_t)b.bytes;
      b.ptr = nullptr;
      b.bytes = 0;
    }
    hipError_t e = hipMallocAsync(&b.ptr, bytes, c->stream);
    if (e != hipSuccess) {
      b.ptr = nullptr;
      set_error("hipMallocAsync(" + name + ", " + std::to_string(bytes) + " bytes): " + hipGetErrorString(e));
      (void)hipGetLastError();
      return MGW_ERR_NOMEM;
    }
    b.bytes = bytes;
    c->ledger.current += (int64_t)bytes;
    c->ledger.peak = std::max(c->ledger.peak, c->ledger.current);
  }
  *out = b.ptr;
  return MGW_OK;
}

template <class T>
int dbuf(mgw_ctx* c, const std::string& name, size_t count, T** out) {
  void* p = nullptr;
  MGW_TRY(ensure(c, name, count * sizeof(T), &p));
  *out = static_cast<T*>(p);
  return MGW_OK;
}

double* find_buf(mgw_ctx* c, const std::string& name) {
  auto it = c->ledger.bufs.find(name);
  return it == c->ledger.bufs.end() ? nullptr : static_cast<double*>(it->second.ptr);
}

int release(mgw_ctx* c, const std::string& name) {
  auto it = c->ledger.bufs.find(name);
  if (it == c->ledger.bufs.end()) return MGW_OK;
  if (it->second.ptr) {
    MGW_HIP(hipFreeAsync(it->second.ptr, c->stream));
    c->ledger.current -= (int64_t)it->second.bytes;
  }
  c->ledger.bufs.erase(it);
  return MGW_OK;
}

int scratch_done(mgw_ctx* c, const std::string& name) {
  return c->trim_eagerly ? release(c, name) : MGW_OK;
}

// Buffers that hold the caller's resident inputs survive a trim.
bool is_resident_input(const std::string& name) {
  return name == "ri_dense" || name == "ri_absmax" || name == "ri_sqrtinv" || name == "c_matrix" || name == "energy" ||
         name == "occupation";
}

// End of a stage's use of a temporary. Temporaries stay cached for the next
// solve; mgw_ctx_trim returns them to the device's pool (stream-ordered frees,
// so other contexts sharing the GPU keep running). When the context is set to
// trim eagerly, they are freed here.
int scratch_done(mgw_ctx* c, const std::string& name);

template <class T>
int upload(mgw_ctx* c, const std::string& name, const T* host, size_t count, T** out) {
  MGW_TRY(dbuf(c, name, count, out));
  if (count) MGW_HIP(hipMemcpyAsync(*out, host, count * sizeof(T), hipMemcpyHostToDevice, c->stream));
  return MGW_OK;
}

template <class T>
int download(mgw_ctx* c, T* host, const T* dev, size_t count) {
  if (count) MGW_HIP(hipMemcpyAsync(host, dev, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

// Sum of a device vector (optionally of squares): deterministic block partials
// summed on the host in a fixed order.
int dsum(mgw_ctx* c, const double* x, int64_t n, int square, double* result) {
  constexpr int NB = 256;
  double* part = nullptr;
  MGW_TRY(dbuf(c, "reduce_partials", NB, &part));
  MGW_TRY(launch_vsum(x, n, square, part, NB, c->stream));
  double h[NB];
  MGW_TRY(download(c, h, part, NB));
  double s = 0.0;
  for (int i = 0; i < NB; ++i) s += h[i];
  *result = s;
  return MGW_OK;
}

// ------------------------------------------------------------------ ranges
int check_ranges(mgw_ctx* c, const mgw_ranges* r) {
  MGW_REQUIRE(r, MGW_ERR_ARG, "null ranges");
  MGW_REQUIRE(c->has_mo, MGW_ERR_STATE, "MO data not set (mgw_set_mo)");
  if (c->has_ri) MGW_TRY(check_mo_matches_ri(c));
  const int ns = c->nstate;
  MGW_REQUIRE(r->ncore_G >= 0 && r->nvirtual_G <= ns + 1 && r->ncore_G + 1 <= r->nvirtual_G - 1,
              MGW_ERR_ARG, "invalid ncore_G / nvirtual_G");
  MGW_REQUIRE(r->nsemin >= 1 && r->nsemax <= ns && r->nsemin <= r->nsemax, MGW_ERR_ARG,
              "invalid nsemin / nsemax");
  MGW_REQUIRE(r->ncore_W >= 0 && r->nvirtual_W <= ns + 1, MGW_ERR_ARG, "invalid ncore_W / nvirtual_W");
  return MGW_OK;
}

// Transition table (src/m_spectral_function.f90:224-266, skip_transition :359-378),
// spin outer. Entries are 0-based indices (i, a) into the spin-major state
// vectors: state i of spin s is s * nstate + i.
void transition_table(const mgw_ctx* c, const mgw_ranges* r, double spin_fact, std::vector<int32_t>& ia) {
  ia.clear();
  const double ce = 1.0e-5;
  const int ns = c->nstate;
  auto skip = [&](int ib1, int ib2, double o1, double o2) {
    if (ib1 <= r->ncore_W || ib2 <= r->ncore_W) return true;
    if (ib1 >= r->nvirtual_W || ib2 >= r->nvirtual_W) return true;
    if (o1 < ce && o2 < ce) return true;
    if (o1 > spin_fact - ce && o2 > spin_fact - ce) return true;
    return false;
  };
  for (int s = 0; s < c->nspin; ++s) {
    const double* occ = c->occupation.data() + (size_t)s * ns;
    for (int i = 1; i <= ns; ++i)
      for (int j = 1; j <= ns; ++j) {
        const double oj = occ[j - 1], oi = occ[i - 1];
        if (skip(j, i, oj, oi)) continue;
        if (oj - oi > 0.0) continue;
        ia.push_back(s * ns + i - 1);
        ia.push_back(s * ns + j - 1);
      }
  }
}

// Is the table the full product i in [i0, i1] (outer) x a in [a0, a1] (inner)?
bool rectangular(const std::vector<int32_t>& ia, int& i0, int& i1, int& a0, int& a1) {
  const int np = (int)ia.size() / 2;
  if (np == 0) return false;
  i0 = ia[0]; a0 = ia[1];
  i1 = ia[2 * (np - 1)]; a1 = ia[2 * (np - 1) + 1];
  const int na = a1 - a0 + 1, ni = i1 - i0 + 1;
  if (na <= 0 || ni <= 0 || (int64_t)na * ni != np) return false;
  for (int t = 0; t < np; ++t)
    if (ia[2 * t] != i0 + t / na || ia[2 * t + 1] != a0 + t % na) return false;
  return true;
}

// ------------------------------------------------------------------ A2
// eri_mo(P, m, n) = sum_kl C(k, m) B(P, k, l) C(l, n) for m in [mlo, mhi],
// n in [nlo, nhi] (1-based), into `out` (naux x nm x nn). The smaller range
// is contracted first: H(P, k, f) = B(P k, l) C(l, f) is one GEMM over the
// dense resident tensor, then a GEMM batched over f contracts k and writes
// straight into the (P, m, n) layout of the reference.
// FP16x3 AO->MO transforms in MGW_FP16X3 solves (default); MGW_AO2MO_FP16X3
// = 0 (read at every transform) keeps the transforms of that mode FP64.
// Measured on MI355X over the serving stream (real molecules, one solve at a
// time): AO->MO 4.5 -> 3.5 ms per solve, 42.0 -> 43.3-44.1 solves/s, the
// largest QP error against FP64 0.0013 -> 0.0014 meV.
// FP16x3 transforms on pre-split planes (default: both operands of the
// second product, and the columns of C in the first) or splitting on every
// tile load: MGW_AO2MO_PLANES = 1 / 0, read at every transform.
bool ao2mo_planes() {
  const char* e = std::getenv("MGW_AO2MO_PLANES");
  return e ? std::atoi(e) != 0 : true;
}

bool ao2mo_fp16x3() {
  const char* e = std::getenv("MGW_AO2MO_FP16X3");
  return e ? std::atoi(e) != 0 : true;
}

// H(P k, f) = sum_l B(P k, l) C(l, f) for the nf1 states from flo (the first
// product of the transform).
int ao2mo_first(mgw_ctx* c, const double* B, const double* C, int flo, int nf1, double* H, int precision, bool f16,
                double* mx, bool no_tune) {
  const int nbf = c->nbf, naux = c->naux;
  GemmF64 g1;
  g1.no_tune = no_tune;
  if (f16) {
    g1.amax = find_buf(c, "ri_absmax");
    g1.bmax = mx;
    g1.cmax = mx + 1;
  }
  g1.M = naux * nbf; g1.N = nf1; g1.K = nbf;
  g1.A = B; g1.lda = (int64_t)naux * nbf;
  g1.B = C + (int64_t)(flo - 1) * nbf; g1.ldb = nbf;
  g1.C = H; g1.ldc = (int64_t)naux * nbf;
  g1.klass = MGW_KCLASS_GEMM_AO2MO;
  g1.precision = precision;
  ProfScope ps(&c->prof, MGW_KCLASS_GEMM_AO2MO, 2.0 * g1.M * (double)g1.N * g1.K,
               8.0 * ((double)g1.M * g1.K + (double)g1.K * g1.N + (double)g1.M * g1.N), c->stream);
  if (f16 && !no_tune && ao2mo_planes()) {
    // the columns of C split once (scale of max |C|, the kernel's own split
    // of B) instead of by every row tile of the large operand: the same
    // planes, so the same products
    const Fp16Planes dc = fp16_planes(nf1, nbf);
    _Float16* pc = nullptr;   // dbuf counts elements of the pointer's type: fp16 here
    MGW_TRY(dbuf(c, "ao2mo_cplanes", (size_t)dc.per_f, &pc));
    MGW_TRY(fp16_split_planes(g1.B, nbf, 0, 1, dc, 1, mx, pc, MGW_KCLASS_GEMM_AO2MO, c->stream));
    g1.bplanes = pc;
    g1.bplane_q = dc.per_q;
    g1.bplane_rpad = dc.rpad;
  }
  return gemm_f64(g1, c->stream);
}

// `half` (nullable): caller-provided space for the half-transformed tensor
// (naux x nbf x min(nm, nn) doubles); null takes a context buffer. With it,
// the call allocates nothing, times no GEMM candidates and leaves nothing
// cached. `extend_hi` (context buffer only): the first product also covers
// the first-index states up to extend_hi, so that a later transform of those
// (the self-energy window after the W transform) reuses it.
int ao2mo(mgw_ctx* c, int mlo, int mhi, int nlo, int nhi, double* out, int precision = MGW_FP64, int spin = -1,
          double* half = nullptr, int extend_hi = 0) {
  MGW_TRY(check_mo_matches_ri(c));
  MGW_REQUIRE(mlo >= 1 && mhi <= c->nstate && mlo <= mhi && nlo >= 1 && nhi <= c->nstate && nlo <= nhi,
              MGW_ERR_ARG, "ao2mo: invalid state range");
  const int nbf = c->nbf, naux = c->naux;
  const int nm = mhi - mlo + 1, nn = nhi - nlo + 1;
  const double* B = find_buf(c, "ri_dense");
  // MO coefficients of the requested (default: active) spin channel
  const double* C = find_buf(c, "c_matrix") + (int64_t)(spin < 0 ? c->spin : spin) * nbf * c->nstate;
  const bool first_n = nn <= nm;
  const int nf = first_n ? nn : nm;
  const int flo = first_n ? nlo : mlo;
  const int no = first_n ? nm : nn;
  const int olo = first_n ? mlo : nlo;
  // FP16x3 transforms: operand maxima on the device (max |B| from the unpack,
  // max |C| of this channel, max |H| from the first product's epilogue)
  const bool f16 = precision == MGW_FP16X3 && ao2mo_fp16x3() && find_buf(c, "ri_absmax");
  const int sp = spin < 0 ? c->spin : spin;
  const int prec_key = f16 ? -1 : precision;
  const int fhi = flo + nf - 1;
  const bool reuse = !half && c->half_valid && c->half_spin == sp && c->half_prec == prec_key && flo >= c->half_lo &&
                     fhi <= c->half_hi;
  double* H = half;
  double* mx = nullptr;
  if (reuse) {
    H = find_buf(c, "ao2mo_half") + (int64_t)(flo - c->half_lo) * naux * nbf;
    if (f16) mx = find_buf(c, "ao2mo_max");
  } else {
    // the first product's states: the requested ones, extended up to extend_hi
    const int f1hi = (!half && extend_hi > fhi) ? std::min(extend_hi, c->nstate) : fhi;
    const int nf1 = f1hi - flo + 1;
    c->half_valid = false;
    if (!H) MGW_TRY(dbuf(c, "ao2mo_half", (size_t)naux * nbf * nf1, &H));
    if (f16) {
      double* part = nullptr;
      MGW_TRY(dbuf(c, "ao2mo_max", 2, &mx));
      MGW_TRY(dbuf(c, "ao2mo_maxpart", ABSMAX_BLOCKS, &part));
      MGW_TRY(launch_absmax_scalar(C, (int64_t)nbf * c->nstate, part, mx, c->stream));
      MGW_TRY(launch_fill(mx + 1, 1, 0.0, c->stream));
    }
    MGW_TRY(ao2mo_first(c, B, C, flo, nf1, H, precision, f16, mx, half != nullptr));
    if (!half) {
      c->half_valid = true;
      c->half_lo = flo;
      c->half_hi = f1hi;
      c->half_spin = sp;
      c->half_prec = prec_key;
    }
  }
  if (f16 && !half && ao2mo_planes()) {
    // the second product on planes: H split once per matrix f, the
    // transform's columns of C once for all f, instead of splitting both on
    // every tile load (the same arithmetic in the same order)
    const Fp16Planes dh = fp16_planes(naux, nbf), dc = fp16_planes(no, nbf);
    _Float16* pl = nullptr;   // dbuf counts elements of the pointer's type: fp16 here
    MGW_TRY(dbuf(c, "ao2mo_planes", (size_t)(dh.per_f * nf + dc.per_f), &pl));
    _Float16* pc = pl + dh.per_f * nf;
    Fp16PlaneProduct p;
    p.M = naux; p.N = no; p.nbatch = nf;
    p.a_f = dh.per_f; p.b_f = 0;
    p.C = out;
    if (first_n) { p.ldc = naux; p.strideC = (int64_t)naux * nm; }
    else { p.ldc = (int64_t)naux * nm; p.strideC = naux; }
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_AO2MO, 2.0 * naux * (double)no * nbf * nf,
                 8.0 * ((double)naux * nbf * nf + (double)nbf * no + (double)naux * no * nf), c->stream);
    constexpr int K = MGW_KCLASS_GEMM_AO2MO;
    MGW_TRY(fp16_split_planes(H, naux, (int64_t)naux * nbf, 0, dh, nf, mx + 1, pl, K, c->stream));
    MGW_TRY(fp16_split_planes(C + (int64_t)(olo - 1) * nbf, nbf, 0, 1, dc, 1, mx, pc, K, c->stream));
    MGW_TRY(fp16_plane_gemm(dh, pl, mx + 1, dc, pc, mx, p, K, c->stream));
    return scratch_done(c, "ao2mo_planes");
  }
  GemmF64 g2;
  g2.no_tune = half != nullptr;
  if (f16) {
    g2.amax = mx + 1;
    g2.bmax = mx;
  }
  g2.M = naux; g2.N = no; g2.K = nbf;
  g2.A = H; g2.lda = naux; g2.strideA = (int64_t)naux * nbf;
  g2.B = C + (int64_t)(olo - 1) * nbf; g2.ldb = nbf; g2.strideB = 0;
  g2.batch = nf;
  if (first_n) {
    g2.C = out; g2.ldc = naux; g2.strideC = (int64_t)naux * nm;
  } else {
    g2.C = out; g2.ldc = (int64_t)naux * nm; g2.strideC = naux;
  }
  g2.klass = MGW_KCLASS_GEMM_AO2MO;
  g2.precision = precision;
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_AO2MO, 2.0 * g2.M * (double)g2.N * g2.K * nf,
                 8.0 * ((double)g2.M * g2.K * nf + (double)g2.K * g2.N + (double)g2.M * g2.N * nf), c->stream);
    MGW_TRY(gemm_f64(g2, c->stream));
  }
  return MGW_OK;
}

// X(P, t) = (P| a_t i_t) for the W transition table, into buffer "x_ia";
// the transitions of each spin channel are contiguous (spin outer) and come
// from that channel's AO->MO transform.
int build_x_ia(mgw_ctx* c, const mgw_ranges* r, double spin_fact, std::vector<int32_t>& ia,
               double** x_out, int precision) {
  transition_table(c, r, spin_fact, ia);
  const int npole = (int)ia.size() / 2;
  MGW_REQUIRE(npole > 0, MGW_ERR_ARG, "empty transition space");
  double* x = nullptr;
  MGW_TRY(dbuf(c, "x_ia", (size_t)c->naux * npole, &x));
  const int ns = c->nstate;
  int t0 = 0;
  for (int s = 0; s < c->nspin && t0 < npole; ++s) {
    int t1 = t0;
    while (t1 < npole && ia[2 * t1] / ns == s) ++t1;
    if (t1 == t0) continue;
    // this channel's table with state indices relative to the channel
    std::vector<int32_t> sub(ia.begin() + 2 * t0, ia.begin() + 2 * t1);
    for (auto& v : sub) v -= s * ns;
    double* xs = x + (int64_t)c->naux * t0;
    int i0, i1, a0, a1;
    if (rectangular(sub, i0, i1, a0, a1)) {
      // (P, a, i) with a inner is exactly X(P, t), t = (i outer, a inner)
      // the first product also covers the self-energy window (nsemax), which
      // the self-energy's transform of this channel then reuses
      MGW_TRY(ao2mo(c, a0 + 1, a1 + 1, i0 + 1, i1 + 1, xs, precision, s, nullptr, r->nsemax));
    } else {
      const int wlo = r->ncore_W + 1, whi = r->nvirtual_W - 1;
      const int nw = whi - wlo + 1;
      double* e = nullptr;
      MGW_TRY(dbuf(c, "ao2mo_w", (size_t)c->naux * nw * nw, &e));
      MGW_TRY(ao2mo(c, wlo, whi, wlo, whi, e, precision, s));
      for (auto& v : sub) v -= (wlo - 1);
      int32_t* ia_d = nullptr;
      MGW_TRY(upload(c, "ia_rel", sub.data(), sub.size(), &ia_d));
      MGW_TRY(launch_gather_ia(e, c->naux, nw, nw, ia_d, t1 - t0, xs, c->stream));
      MGW_TRY(scratch_done(c, "ao2mo_w"));
    }
    t0 = t1;
  }
  *x_out = x;
  return MGW_OK;
}

// ------------------------------------------------------------------ A4
int polarizability_casida(mgw_ctx* c, const mgw_ranges* r, double spin_fact, double* en_rpa,
                          double* en_gm, int precision) {
  std::vector<int32_t> ia;
  double* x = nullptr;
  MGW_TRY(build_x_ia(c, r, spin_fact, ia, &x, precision));
  const int npole = (int)ia.size() / 2;
  const int naux = c->naux;
  int32_t* ia_d = nullptr;
  MGW_TRY(upload(c, "ia", ia.data(), ia.size(), &ia_d));
  double *de, *sq, *de2, *n2, *cs, *eig, *ework, *pole;
  MGW_TRY(dbuf(c, "cas_de", npole, &de));
  MGW_TRY(dbuf(c, "cas_sq", npole, &sq));
  MGW_TRY(dbuf(c, "cas_de2", npole, &de2));
  MGW_TRY(dbuf(c, "cas_n2", npole, &n2));
  MGW_TRY(dbuf(c, "cas_cs", npole, &cs));
  MGW_TRY(dbuf(c, "cas_eig", npole, &eig));
  MGW_TRY(dbuf(c, "cas_ework", npole, &ework));
  MGW_TRY(dbuf(c, "wpol_pole", npole, &pole));
  double* energy_d = find_buf(c, "energy");
  MGW_TRY(launch_casida_diag(ia_d, npole, energy_d, de, sq, de2, c->stream));
  MGW_TRY(launch_colnorm2(x, naux, naux, npole, n2, c->stream));
  // M = (A-B)^1/2 (A+B) (A-B)^1/2 = diag(de^2) + 2 spin_fact sq X^T X sq
  double* m = nullptr;
  MGW_TRY(dbuf(c, "cas_m", (size_t)npole * npole, &m));
  GemmF64 g;
  g.M = npole; g.N = npole; g.K = naux;
  g.transA = true; g.A = x; g.lda = naux;
  g.B = x; g.ldb = naux;
  g.C = m; g.ldc = npole;
  g.alpha = 2.0 * spin_fact;
  g.rowscale = sq; g.colscale = sq; g.diag_add = de2;
  g.lower_only = true;
  g.klass = MGW_KCLASS_GEMM_CASIDA;
  g.precision = precision;
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_CASIDA, (double)npole * (npole + 1.0) * naux,
                 8.0 * ((double)naux * npole + 0.5 * (double)npole * npole), c->stream);
    MGW_TRY(gemm_f64(g, c->stream));
  }
  // eigen-decomposition (src/m_block_diago.f90:185)
  int32_t* info_d = nullptr;
  MGW_TRY(dbuf(c, "solver_info", 64, &info_d));
  {
    const double n3 = (double)npole * npole * npole;
    ProfScope ps(&c->prof, MGW_KCLASS_EIGEN, (4.0 / 3.0 + 2.0) * n3, 16.0 * (double)npole * npole, c->stream);
    MGW_BLAS(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_lower, npole, m, npole, eig,
                              ework, info_d));
  }
  int32_t info = 0;
  MGW_TRY(download(c, &info, info_d, 1));
  MGW_REQUIRE(info == 0, MGW_ERR_SOLVER, "rocsolver_dsyevd failed, info=" + std::to_string(info));
  MGW_TRY(launch_casida_post(eig, npole, pole, cs, c->stream));
  // w_s = sqrt(spin_fact) X diag(sq) Z diag(Omega^-1/2)  (src/m_linear_response.f90:845-849)
  double* ws = nullptr;
  MGW_TRY(dbuf(c, "wpol_ws", (size_t)naux * npole, &ws));
  GemmF64 h;
  h.M = naux; h.N = npole; h.K = npole;
  h.A = x; h.lda = naux;
  h.B = m; h.ldb = npole;
  h.C = ws; h.ldc = naux;
  h.kscale = sq; h.colscale = cs;
  h.alpha = std::sqrt(spin_fact);
  h.klass = MGW_KCLASS_GEMM_CASIDA;
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_CASIDA, 2.0 * naux * (double)npole * npole,
                 8.0 * (2.0 * (double)naux * npole + (double)npole * npole), c->stream);
    MGW_TRY(gemm_f64(h, c->stream));
  }
  // energies (src/m_build_bse.f90:306-307, src/m_linear_response.f90:418, :849)
  double sum_de, sum_n2, sum_pole, sum_ws2;
  MGW_TRY(dsum(c, de, npole, 0, &sum_de));
  MGW_TRY(dsum(c, n2, npole, 0, &sum_n2));
  MGW_TRY(dsum(c, pole, npole, 0, &sum_pole));
  MGW_TRY(dsum(c, ws, (int64_t)naux * npole, 1, &sum_ws2));
  if (en_rpa) *en_rpa = -0.25 * (sum_de + 2.0 * spin_fact * sum_n2) - 0.25 * sum_de + 0.5 * sum_pole;
  if (en_gm) *en_gm = 0.5 * (sum_ws2 - spin_fact * sum_n2);
  c->npole = npole;
  c->has_wpol = true;
  MGW_TRY(scratch_done(c, "cas_m"));
  return MGW_OK;
}

// ------------------------------------------------------------------ A6
int gw_selfenergy(mgw_ctx* c, const mgw_ranges* r, double spin_fact, double eta, int nomega,
                  const double* omega, const double* se_energy0, double* energy_gw, int precision) {
  MGW_REQUIRE(c->has_wpol, MGW_ERR_STATE, "no spectral function: call mgw_polarizability first");
  const int naux = c->naux, npole = c->npole;
  const int nse = r->nsemax - r->nsemin + 1;
  const int glo = r->ncore_G + 1, ghi = r->nvirtual_G - 1;
  const int ng = ghi - glo + 1;
  double* epi = nullptr;
  MGW_TRY(dbuf(c, "eri_pi", (size_t)naux * nse * ng, &epi));
  MGW_TRY(ao2mo(c, r->nsemin, r->nsemax, glo, ghi, epi, precision));
  // small per-state vectors
  std::vector<double> e0p(se_energy0, se_energy0 + nse), ei(ng), fi(ng), ep(nse), occp(nse);
  // energies and occupations of the active spin channel
  const double* eg = c->energy_in_g().data() + c->so();
  const double* oc = c->occupation.data() + c->so();
  for (int k = 0; k < ng; ++k) { ei[k] = eg[glo - 1 + k]; fi[k] = oc[glo - 1 + k]; }
  for (int k = 0; k < nse; ++k) { ep[k] = eg[r->nsemin - 1 + k]; occp[k] = oc[r->nsemin - 1 + k]; }
  double *om_d, *e0p_d, *ei_d, *fi_d, *ep_d, *occp_d;
  MGW_TRY(upload(c, "se_omega", omega, nomega, &om_d));
  MGW_TRY(upload(c, "se_e0p", e0p.data(), nse, &e0p_d));
  MGW_TRY(upload(c, "se_ei", ei.data(), ng, &ei_d));
  MGW_TRY(upload(c, "se_fi", fi.data(), ng, &fi_d));
  MGW_TRY(upload(c, "se_ep", ep.data(), nse, &ep_d));
  MGW_TRY(upload(c, "se_occp", occp.data(), nse, &occp_d));
  double *sigma, *egw_p, *part;
  MGW_TRY(dbuf(c, "sigma", (size_t)2 * nomega * nse, &sigma));
  MGW_TRY(dbuf(c, "sigma_egw", nse, &egw_p));
  MGW_TRY(dbuf(c, "sigma_part", (size_t)sigma_part_doubles(npole, nse), &part));
  // chunk over i so that w(s, p, i) stays within ~1 GiB
  const size_t per_i = (size_t)npole * nse * sizeof(double);
  int ic = (int)std::max<size_t>(1, std::min<size_t>((size_t)ng, (size_t(1) << 30) / per_i));
  double* w = nullptr;
  MGW_TRY(dbuf(c, "sigma_w", (size_t)npole * nse * ic, &w));
  const double* ws = find_buf(c, "wpol_ws");
  const double* pole = find_buf(c, "wpol_pole");
  for (int i0 = 0; i0 < ng; i0 += ic) {
    const int cnt = std::min(ic, ng - i0);
    // w(s, p, i) = sum_P w_s(P, s) (P | p i)   (src/m_gw_selfenergy_analytic.f90:101-104)
    GemmF64 g;
    g.M = npole; g.N = nse * cnt; g.K = naux;
    g.transA = true; g.A = ws; g.lda = naux;
    g.B = epi + (int64_t)naux * nse * i0; g.ldb = naux;
    g.C = w; g.ldc = npole;
    g.klass = MGW_KCLASS_GEMM_SIGMA;
    {
      ProfScope ps(&c->prof, MGW_KCLASS_GEMM_SIGMA, 2.0 * g.M * (double)g.N * g.K,
                   8.0 * ((double)g.M * g.K + (double)g.K * g.N + (double)g.M * g.N), c->stream);
      MGW_TRY(gemm_f64(g, c->stream));
    }
    {
      const double terms = (double)npole * nse * cnt * nomega;
      ProfScope ps(&c->prof, MGW_KCLASS_SIGMA_POLES, 8.0 * terms, 8.0 * (double)npole * nse * cnt, c->stream);
      MGW_TRY(launch_sigma_poles(w, npole, nse, cnt, pole, om_d, nomega, e0p_d, ei_d + i0, fi_d + i0,
                                 ep_d, occp_d, spin_fact, eta, part, i0 > 0 ? 1 : 0, sigma, egw_p,
                                 c->stream));
    }
  }
  std::vector<double> egw(nse);
  MGW_TRY(download(c, egw.data(), egw_p, nse));
  if (energy_gw) {
    double s = 0.0;
    for (int k = 0; k < nse; ++k) s += egw[k];
    *energy_gw = s;
  }
  c->has_sigma = true;
  c->sigma_nomega = nomega;
  c->sigma_nse = nse;
  MGW_TRY(scratch_done(c, "sigma_w"));
  MGW_TRY(scratch_done(c, "eri_pi"));
  return MGW_OK;
}

// The batched Cholesky sequence (6 n / 64 launches, no host synchronisation)
// is captured once per shape and buffer set into a HIP graph and replayed
// when the context has graphs enabled; the autotuner's timing mode launches
// eagerly because it synchronises.
void clear_graphs(mgw_ctx* c) {
  for (auto& kv : c->graphs) (void)hipGraphExecDestroy(kv.second);
  c->graphs.clear();
}

int run_chol(mgw_ctx* c, double* A, int n, int nbatch, double* logdet, int32_t* info, double* work,
             double* bounds, void* planes) {
  if (!c->use_graphs || gemm_autotune_mode(-1) == 1)
    return chol_inverse_batched(A, n, nbatch, logdet, info, work, c->stream, bounds, planes);
  const mgw_ctx::GraphKey key{n, nbatch, A, logdet, info, work, bounds, planes,
                             bounds ? chol_f16_min() * 2 + (chol_epi_planes() ? 1 : 0) : 0};
  auto it = c->graphs.find(key);
  if (it == c->graphs.end()) {
    MGW_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = chol_inverse_batched(A, n, nbatch, logdet, info, work, c->stream, bounds, planes);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc != MGW_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    MGW_HIP(e);
    hipGraphExec_t exec = nullptr;
    const hipError_t e2 = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MGW_HIP(e2);
    it = c->graphs.emplace(key, exec).first;
  }
  MGW_HIP(hipGraphLaunch(it->second, c->stream));
  return MGW_OK;
}

// ------------------------------------------------------------------ A5
int polarizability_grid(mgw_ctx* c, const mgw_ranges* r, double spin_fact, int nomega,
                        const double* omega_im, const double* weight, double* erpa, double* egw,
                        int precision) {
  MGW_REQUIRE(nomega >= 1 && nomega <= 65535, MGW_ERR_ARG, "nomega_chi_imag out of range");
  std::vector<int32_t> ia;
  double* x = nullptr;
  MGW_TRY(build_x_ia(c, r, spin_fact, ia, &x, precision));
  const int npole = (int)ia.size() / 2;
  const int naux = c->naux;
  int32_t* ia_d = nullptr;
  MGW_TRY(upload(c, "ia", ia.data(), ia.size(), &ia_d));
  double *om_d, *sw, *n2, *ones, *tr0, *chi;
  MGW_TRY(upload(c, "chi_omega", omega_im, nomega, &om_d));
  MGW_TRY(dbuf(c, "chi_sw", (size_t)npole * nomega, &sw));
  MGW_TRY(dbuf(c, "cas_n2", npole, &n2));
  MGW_TRY(dbuf(c, "ones", naux, &ones));
  MGW_TRY(dbuf(c, "chi_tr0", nomega, &tr0));
  // 1 - chi0, then its inverse Cholesky factor, resident for the self-energy
  MGW_TRY(dbuf(c, "wpol_linv", (size_t)naux * naux * nomega, &chi));
  MGW_TRY(launch_fill(ones, naux, 1.0, c->stream));
  MGW_TRY(launch_chi_weights(ia_d, npole, find_buf(c, "energy"), find_buf(c, "occupation"), om_d,
                             nomega, sw, c->stream));
  MGW_TRY(launch_colnorm2(x, naux, naux, npole, n2, c->stream));
  MGW_TRY(launch_chi_trace(sw, n2, npole, nullptr, naux, 0, nomega, tr0, c->stream));
  // 1 - chi0(iw) = 1 + X diag(s(iw)) X^T for all frequencies in one launch
  GemmF64 g;
  g.M = naux; g.N = naux; g.K = npole;
  g.A = x; g.lda = naux; g.strideA = 0;
  g.B = x; g.ldb = naux; g.transB = true; g.strideB = 0;
  g.C = chi; g.ldc = naux; g.strideC = (int64_t)naux * naux;
  g.kscale = sw; g.strideK = npole;
  g.diag_add = ones; g.strideDiag = 0;
  g.batch = nomega;
  g.lower_only = true;
  g.upper_zero = true;   // the Cholesky wants the strict upper triangle zero
  g.klass = MGW_KCLASS_GEMM_CHI0;
  g.precision = precision;
  // the split-operand kernels factor s = sqrt(s) sqrt(s): every weight must be >= 0
  bool nonneg = true;
  for (int t = 0; t < npole; ++t) {
    const int i = ia[2 * t], a = ia[2 * t + 1];
    if ((c->occupation[i] - c->occupation[a]) * (c->energy[a] - c->energy[i]) < 0.0) nonneg = false;
  }
  if (split_precision(precision) && !nonneg)
    g.precision = MGW_FP32;   // per-element split kernels take signed weights at FP32
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_CHI0, (double)naux * (naux + 1.0) * npole * nomega,
                 8.0 * ((double)naux * npole + 0.5 * (double)naux * naux * nomega), c->stream);
    if (split_precision(precision) && nonneg) {
      const size_t bytes = chi0_split_workspace(naux, npole, nomega, precision);
      double* work = nullptr;
      MGW_TRY(dbuf(c, "chi_split", (bytes + 7) / 8, &work));
      MGW_TRY(chi0_split_syrk(x, naux, npole, sw, nomega, precision, work, bytes, chi, c->stream));
    } else {
      MGW_TRY(gemm_f64(g, c->stream));
    }
  }
  if (precision != MGW_FP64) MGW_TRY(scratch_done(c, "chi_split"));
  // 1 - chi0 = L L^T for every frequency, then L^-1 in place (chol.hip)
  int32_t* info_d = nullptr;
  double *logdet, *work, *frob;
  MGW_TRY(dbuf(c, "solver_info", std::max(64, nomega), &info_d));
  MGW_TRY(dbuf(c, "chol_logdet", nomega, &logdet));
  MGW_TRY(dbuf(c, "chol_work", chol_work_doubles(naux) * nomega, &work));
  MGW_TRY(dbuf(c, "chol_frob", (size_t)FROB_BLOCKS * nomega, &frob));
  // FP16x3 solves factor with FP16x3 products at the large levels: 1 - chi0
  // >= I holds there (every weight s >= 0), which the operand bounds need
  double* bounds = nullptr;
  double* planes = nullptr;
  if (precision == MGW_FP16X3 && nonneg && chol_f16_min() > 0) {
    MGW_TRY(dbuf(c, "chol_bounds", 4 + (size_t)nomega, &bounds));
    MGW_TRY(dbuf(c, "chol_planes", std::max<size_t>(1, (chol_plane_bytes(naux, nomega) + 7) / 8), &planes));
  }
  const double chol_flops = 2.0 / 3.0 * (double)naux * naux * naux;
  {
    ProfScope chol(&c->prof, MGW_KCLASS_CHOLESKY, chol_flops * nomega, 16.0 * (double)naux * naux * nomega,
                   c->stream);
    MGW_TRY(run_chol(c, chi, naux, nomega, logdet, info_d, work, bounds, planes));
  }
  if (planes) MGW_TRY(scratch_done(c, "chol_planes"));
  std::vector<int32_t> info(nomega);
  MGW_TRY(download(c, info.data(), info_d, nomega));
  bool checked_x = false;
  for (int k = 0; k < nomega; ++k) {
    if (info[k] == 0) continue;
    MGW_REQUIRE(precision != MGW_FP64, MGW_ERR_SOLVER,
                "Cholesky of 1-chi0 failed at frequency " + std::to_string(k) + " (column " +
                    std::to_string(info[k]) + ")");
    if (!checked_x) {
      // 1 + X diag(s) X^T with s >= 0 is positive definite for any finite X
      std::vector<double> h(npole);
      MGW_TRY(download(c, h.data(), n2, npole));
      for (int t = 0; t < npole; ++t)
        MGW_REQUIRE(std::isfinite(h[t]), MGW_ERR_SOLVER,
                    "non-finite (P|ia) column " + std::to_string(t) + " before the chi0 build");
      checked_x = true;
    }
    // a reduced-precision 1 - chi0 that is not positive definite: rebuild
    // this frequency in FP64 and factor it again
    GemmF64 g1 = g;
    g1.precision = MGW_FP64;
    g1.batch = 1;
    g1.kscale = sw + (int64_t)npole * k;
    g1.C = chi + (int64_t)naux * naux * k;
    MGW_TRY(gemm_f64(g1, c->stream));
    MGW_TRY(chol_inverse_batched(g1.C, naux, 1, logdet + k, info_d + k, work, c->stream));
    int32_t one = 0;
    MGW_TRY(download(c, &one, info_d + k, 1));
    MGW_REQUIRE(one == 0, MGW_ERR_SOLVER,
                "Cholesky of the FP64 rebuild of 1-chi0 failed at frequency " + std::to_string(k) + " (column " +
                    std::to_string(one) + "; reduced-precision column " + std::to_string(info[k]) + ")");
    ++c->escalations;
  }
  MGW_TRY(launch_lower_frob2(chi, naux, nomega, frob, c->stream));
  // D = L^-1 - 1 stays resident: chi = L^-T L^-1 - 1 = D^T D + D + D^T, and
  // the self-energy product works on the screening part D alone
  MGW_TRY(launch_diag_add(chi, naux, nomega, -1.0, c->stream));
  std::vector<double> h0(nomega), hl(nomega), fr((size_t)FROB_BLOCKS * nomega);
  MGW_TRY(download(c, h0.data(), tr0, nomega));
  MGW_TRY(download(c, hl.data(), logdet, nomega));
  MGW_TRY(download(c, fr.data(), frob, (size_t)FROB_BLOCKS * nomega));
  // sum ln(lambda) + 1 - lambda = ln det(1-chi0) + Tr chi0 ;
  // sum -(1-lambda)/lambda + 1 - lambda = n - Tr (1-chi0)^-1 + Tr chi0
  // (src/m_gw_selfenergy_grid.f90:171-173); Tr (1-chi0)^-1 = ||L^-1||_F^2
  const double twopi = 2.0 * 3.14159265358979323;
  double er = 0.0, eg = 0.0;
  for (int k = 0; k < nomega; ++k) {
    double tri = 0.0;
    for (int b = 0; b < FROB_BLOCKS; ++b) tri += fr[(size_t)k * FROB_BLOCKS + b];
    er += (2.0 * hl[k] - h0[k]) / twopi * weight[k];
    eg += ((double)naux - tri - h0[k]) / twopi * weight[k];
  }
  if (erpa) *erpa = er;
  if (egw) *egw = eg;
  c->has_chi = true;
  c->nomega_chi = nomega;
  c->chi_omega.assign(omega_im, omega_im + nomega);
  c->chi_weight.assign(weight, weight + nomega);
  MGW_TRY(scratch_done(c, "chi_sw"));
  return MGW_OK;
}

// ------------------------------------------------------------------ A7
int gw_selfenergy_imag(mgw_ctx* c, const mgw_ranges* r, int ncalc, const double* omega_calc, int precision) {
  MGW_REQUIRE(c->has_chi, MGW_ERR_STATE, "no chi: call mgw_polarizability_grid first");
  const int naux = c->naux, nomega = c->nomega_chi;
  const int nse = r->nsemax - r->nsemin + 1;
  const int glo = r->ncore_G + 1, ghi = r->nvirtual_G - 1;
  const int np = ghi - glo + 1;
  const int64_t ncol = (int64_t)np * nse;
  MGW_REQUIRE(ncol < (1ll << 31), MGW_ERR_ARG, "sigma_imag: too many columns");
  double* epm = nullptr;
  MGW_TRY(dbuf(c, "eri_pm", (size_t)naux * ncol, &epm));
  MGW_TRY(ao2mo(c, glo, ghi, r->nsemin, r->nsemax, epm, precision));
  const int nrb = gemm_f64_rowblocks(naux);
  double *part, *v, *om_d, *wt_d, *ep_d, *wc_d, *sc;
  MGW_TRY(dbuf(c, "dot_part", (size_t)nrb * ncol * nomega, &part));
  MGW_TRY(dbuf(c, "dot_v", (size_t)ncol * nomega, &v));
  // diag(B_m^T chi(iw) B_m) for every (p, m) and iw (src/m_gw_selfenergy_grid.f90:297-310).
  // With chi = L^-T L^-1 - 1 and D = L^-1 - 1 (lower triangular, resident):
  // B^T chi B = ||D B||^2 + 2 B^T D B, a triangular product whose epilogue
  // sums C (C + 2 B) over the rows; no cancellation against ||B||^2
  GemmF64 g;
  g.M = naux; g.N = (int)ncol; g.K = naux;
  g.A = find_buf(c, "wpol_linv"); g.lda = naux; g.strideA = (int64_t)naux * naux; g.tri_a = true;
  g.B = epm; g.ldb = naux; g.strideB = 0;
  g.dotB = epm; g.lddot = naux; g.dot_cc = 1.0; g.dot_cd = 2.0;
  g.dotpart = part; g.strideDot = (int64_t)nrb * ncol;
  g.batch = nomega;
  g.batch_outer = true;   // one L^-1 stays in cache while B streams past it
  const bool split = split_precision(precision);
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_WB, (double)naux * (naux + 1.0) * ncol * nomega,
                 8.0 * (0.5 * (double)naux * naux * nomega + (double)naux * ncol), c->stream);
    if (split) {
      const size_t bytes = wb_split_workspace(naux, ncol, nomega, precision);
      double* work = nullptr;
      MGW_TRY(dbuf(c, "wb_split", (bytes + 7) / 8, &work));
      MGW_TRY(wb_split_dot(g.A, naux, nomega, epm, ncol, precision, work, bytes, part, c->stream));
    } else {
      MGW_TRY(gemm_f64(g, c->stream));
    }
  }
  if (split) MGW_TRY(scratch_done(c, "wb_split"));
  MGW_TRY(launch_dot_reduce(part, nrb, ncol, nomega, nullptr, v, c->stream));
  std::vector<double> ep(np);
  for (int k = 0; k < np; ++k) ep[k] = c->energy_in_g()[c->so() + glo - 1 + k];
  MGW_TRY(upload(c, "sgi_omega", c->chi_omega.data(), nomega, &om_d));
  MGW_TRY(upload(c, "sgi_weight", c->chi_weight.data(), nomega, &wt_d));
  MGW_TRY(upload(c, "sgi_ep", ep.data(), np, &ep_d));
  MGW_TRY(upload(c, "sgi_wc", omega_calc, (size_t)2 * ncalc, &wc_d));
  MGW_TRY(dbuf(c, "sigma_calc", (size_t)2 * ncalc * nse, &sc));
  MGW_TRY(launch_sigma_imag_sum(v, np, nse, nomega, om_d, wt_d, ep_d, wc_d, ncalc, sc, c->stream));
  c->has_sigma_calc = true;
  c->sigma_ncalc = ncalc;
  c->sigma_calc_nse = nse;
  MGW_TRY(scratch_done(c, "dot_part"));
  MGW_TRY(scratch_done(c, "eri_pm"));
  return MGW_OK;
}

// ------------------------------------------------------------------ A8
int pade_fit(mgw_ctx* c, int nse, int ncalc, const double* omega_calc, const double* sigma_calc_host,
             int nomega, const double* omega, double eta) {
  const double* sc = nullptr;
  if (sigma_calc_host) {
    double* d = nullptr;
    MGW_TRY(upload(c, "pade_sc_in", sigma_calc_host, (size_t)2 * ncalc * nse, &d));
    sc = d;
  } else {
    MGW_REQUIRE(c->has_sigma_calc && c->sigma_ncalc == ncalc && c->sigma_calc_nse == nse,
                MGW_ERR_STATE, "no resident sigma_calc of this shape");
    sc = find_buf(c, "sigma_calc");
  }
  double *wc_d, *om_d, *work, *sigma;
  MGW_TRY(upload(c, "pade_wc", omega_calc, (size_t)2 * ncalc, &wc_d));
  MGW_TRY(upload(c, "pade_omega", omega, nomega, &om_d));
  MGW_TRY(dbuf(c, "pade_work", (size_t)pade_work_doubles(ncalc, nse), &work));
  MGW_TRY(dbuf(c, "sigma", (size_t)2 * nomega * nse, &sigma));
  MGW_TRY(launch_pade(wc_d, sc, ncalc, nse, om_d, nomega, eta, work, sigma, c->stream));
  c->has_sigma = true;
  c->sigma_nomega = nomega;
  c->sigma_nse = nse;
  return MGW_OK;
}

// ------------------------------------------------------------------ A9
int find_qp(mgw_ctx* c, const mgw_ranges* r, int nomega, const double* omega, const double* sigma_host,
            const double* se_energy0, const double* exc, const double* energy0, int nstate,
            double* energy_qp, double* zz, double* energy_qp_lin) {
  const int nse = r->nsemax - r->nsemin + 1;
  MGW_REQUIRE(r->nsemin >= 1 && r->nsemax <= nstate, MGW_ERR_ARG, "find_qp: state range exceeds nstate");
  const double* sg = nullptr;
  if (sigma_host) {
    double* d = nullptr;
    MGW_TRY(upload(c, "qp_sigma_in", sigma_host, (size_t)2 * nomega * nse, &d));
    sg = d;
  } else {
    MGW_REQUIRE(c->has_sigma && c->sigma_nomega == nomega && c->sigma_nse == nse, MGW_ERR_STATE,
                "no resident sigma of this shape");
    sg = find_buf(c, "sigma");
  }
  std::vector<double> e0s(nse), excs(nse);
  for (int k = 0; k < nse; ++k) {
    e0s[k] = energy0[r->nsemin - 1 + k];
    excs[k] = exc ? exc[r->nsemin - 1 + k] : 0.0;
  }
  double *om_d, *se0_d, *exc_d, *e0_d, *out;
  MGW_TRY(upload(c, "qp_omega", omega, nomega, &om_d));
  MGW_TRY(upload(c, "qp_se0", se_energy0, nse, &se0_d));
  MGW_TRY(upload(c, "qp_exc", excs.data(), nse, &exc_d));
  MGW_TRY(upload(c, "qp_e0", e0s.data(), nse, &e0_d));
  MGW_TRY(dbuf(c, "qp_out", (size_t)3 * nse, &out));
  MGW_TRY(launch_qp(sg, nomega, nse, om_d, se0_d, exc_d, e0_d, out, out + nse, out + 2 * nse, c->stream));
  std::vector<double> h(3 * nse);
  MGW_TRY(download(c, h.data(), out, 3 * nse));
  for (int p = 0; p < nstate; ++p) {
    const bool in = p >= r->nsemin - 1 && p <= r->nsemax - 1;
    const int k = p - (r->nsemin - 1);
    if (energy_qp) energy_qp[p] = in ? h[k] : energy0[p];
    if (zz) zz[p] = in ? h[nse + k] : 1.0;
    if (energy_qp_lin) energy_qp_lin[p] = in ? h[2 * nse + k] : energy0[p];
  }
  return MGW_OK;
}

// Gauss-Legendre on [0,1] mapped to the imaginary axis
// (src/m_numerical_tools.f90:35-82, src/m_spectral_function.f90:286-306).
void imaginary_grid(int n, std::vector<double>& om, std::vector<double>& wt) {
  const double pi = 3.14159265358979323;
  std::vector<double> x(n), w(n);
  for (int i = 1; i <= (n + 1) / 2; ++i) {
    double z = std::cos(pi * (i - 0.25) / (n + 0.5)), dp = 0.0;
    for (;;) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      dp = n * (p2 - z * p1) / (1.0 - z * z);
      const double z1 = z;
      z = z1 - p1 / dp;
      if (std::fabs(z - z1) < 1.0e-13) break;
    }
    x[i - 1] = 0.5 - 0.5 * z;
    x[n - i] = 0.5 + 0.5 * z;
    w[i - 1] = 1.0 / ((1.0 - z * z) * dp * dp);
    w[n - i] = w[i - 1];
  }
  om.resize(n);
  wt.resize(n);
  for (int k = 0; k < n; ++k) {
    wt[k] = w[k] * std::pow(1.0 - x[k], -2.0) * 2.0;
    om[k] = (1.0 / (1.0 - x[k]) - 1.0) * 2.0;
  }
}

// ------------------------------------------------------------------ contour deformation
int nlumo_of(const mgw_ranges* r) { return r->nlumo_G > 0 ? r->nlumo_G : r->nhomo_G + 1; }

// Largest residue frequency (src/m_gw_selfenergy_grid.f90:385-408): over the
// channels, the states m and the grid, e_p - (E0_m + omega_k) for occupied p
// in [ncore_G+1, nhomo_G] and E0_m + omega_k - e_p for empty p in
// [nlumo_G, nvirtual_G-1], values >= -eta only; 0 if none.
double contour_de_max(const mgw_ctx* c, const mgw_ranges* r, double eta, int nomega, const double* omega,
                      const double* se_energy0) {
  const int nse = r->nsemax - r->nsemin + 1, ns = c->nstate;
  const std::vector<double>& e = c->energy_in_g();
  double de_max = 0.0;
  for (int s = 0; s < c->nspin; ++s)
    for (int m = 0; m < nse; ++m)
      for (int k = 0; k < nomega; ++k) {
        const double w = se_energy0[m + (int64_t)nse * s] + omega[k];
        for (int p = r->ncore_G + 1; p <= r->nhomo_G; ++p) {
          const double de = e[(size_t)s * ns + p - 1] - w;
          if (de >= -eta) de_max = std::max(de_max, de);
        }
        for (int p = nlumo_of(r); p <= r->nvirtual_G - 1; ++p) {
          const double de = w - e[(size_t)s * ns + p - 1];
          if (de >= -eta) de_max = std::max(de_max, de);
        }
      }
  return de_max;
}

// W of sf_vsqrt_chi_vsqrt_rpa(low_rank) (src/m_spectral_function.f90:768-949)
// on nimag imaginary and nreal real frequencies at once: chi0_f = -X diag(s_f) X^T
// (one batched MFMA SYRK), the eigenpairs of every chi0_f (rocSOLVER dsyevd,
// strided batched), and the weights lambda / (1 - lambda) where they exceed
// 1e-3 in magnitude, 0 for the dropped eigenpairs (contour_eta_kernel).
int contour_w(mgw_ctx* c, const mgw_ranges* r, double spin_fact, int nimag, int nreal, double de_max) {
  MGW_REQUIRE(nimag >= 1 && nreal >= 2 && nimag + nreal <= 65535, MGW_ERR_ARG,
              "contour deformation needs nomega_chi_imag >= 1 and nomega_chi_real >= 2");
  std::vector<int32_t> ia;
  double* x = nullptr;
  MGW_TRY(build_x_ia(c, r, spin_fact, ia, &x, MGW_FP64));
  const int npole = (int)ia.size() / 2, naux = c->naux, nf = nimag + nreal;
  int32_t* ia_d = nullptr;
  MGW_TRY(upload(c, "ia", ia.data(), ia.size(), &ia_d));
  std::vector<double> om_im, wt_im, om_re(nreal);
  imaginary_grid(nimag, om_im, wt_im);
  // REAL_LINEAR grid (src/m_spectral_function.f90:308-315)
  for (int j = 0; j < nreal; ++j) om_re[j] = (double)j / (double)(nreal - 1) * de_max;
  double *om_d, *eg_d, *sw, *u, *lam, *ework, *w;
  std::vector<double> om_all(om_im);
  om_all.insert(om_all.end(), om_re.begin(), om_re.end());
  MGW_TRY(upload(c, "cd_omega", om_all.data(), nf, &om_d));
  // the reference builds this W from the energies in G (energy_g)
  MGW_TRY(upload(c, "cd_energy", c->energy_in_g().data(), c->energy_in_g().size(), &eg_d));
  MGW_TRY(dbuf(c, "chi_sw", (size_t)npole * nf, &sw));
  MGW_TRY(launch_chi_weights(ia_d, npole, eg_d, find_buf(c, "occupation"), om_d, nimag, sw, c->stream, 1.0));
  MGW_TRY(launch_chi_weights(ia_d, npole, eg_d, find_buf(c, "occupation"), om_d + nimag, nreal,
                             sw + (int64_t)npole * nimag, c->stream, -1.0));
  MGW_TRY(dbuf(c, "cd_u", (size_t)naux * naux * nf, &u));
  GemmF64 g;
  g.M = naux; g.N = naux; g.K = npole;
  g.A = x; g.lda = naux; g.strideA = 0;
  g.B = x; g.ldb = naux; g.transB = true; g.strideB = 0;
  g.C = u; g.ldc = naux; g.strideC = (int64_t)naux * naux;
  g.kscale = sw; g.strideK = npole;
  g.alpha = -1.0;                // chi0 itself
  g.batch = nf;
  g.lower_only = true; g.no_mirror = true;   // dsyevd reads the lower triangle
  g.klass = MGW_KCLASS_GEMM_CHI0;
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_CHI0, (double)naux * (naux + 1.0) * npole * nf,
                 8.0 * ((double)naux * npole + 0.5 * (double)naux * naux * nf), c->stream);
    MGW_TRY(gemm_f64(g, c->stream));
  }
  MGW_TRY(scratch_done(c, "chi_sw"));
  int32_t* info_d = nullptr;
  MGW_TRY(dbuf(c, "solver_info", std::max(64, nf), &info_d));
  MGW_TRY(dbuf(c, "cd_lambda", (size_t)naux * nf, &lam));
  MGW_TRY(dbuf(c, "cd_ework", (size_t)naux * nf, &ework));
  {
    const double n3 = (double)naux * naux * naux;
    ProfScope ps(&c->prof, MGW_KCLASS_EIGEN, (4.0 / 3.0 + 2.0) * n3 * nf, 16.0 * (double)naux * naux * nf,
                 c->stream);
    MGW_BLAS(rocsolver_dsyevd_strided_batched(c->blas, rocblas_evect_original, rocblas_fill_lower, naux, u, naux,
                                              (rocblas_stride)naux * naux, lam, naux, ework, naux, info_d, nf));
  }
  std::vector<int32_t> info(nf);
  MGW_TRY(download(c, info.data(), info_d, nf));
  for (int f = 0; f < nf; ++f)
    MGW_REQUIRE(info[f] == 0, MGW_ERR_SOLVER,
                "rocsolver_dsyevd (contour W) failed at frequency " + std::to_string(f) + ", info=" +
                    std::to_string(info[f]));
  MGW_TRY(dbuf(c, "cd_w", (size_t)naux * nf, &w));
  MGW_TRY(launch_contour_eta(lam, (int64_t)naux * nf, 1.0e-3, w, c->stream));
  MGW_TRY(scratch_done(c, "cd_ework"));
  c->has_cd = true;
  c->cd_om_im = om_im;
  c->cd_wt_im = wt_im;
  c->cd_om_re = om_re;
  return MGW_OK;
}

// The self-energy of gw_selfenergy_contour (src/m_gw_selfenergy_grid.f90:429-530)
// for the active channel: v(p, m, f) = B_pm^T W_f B_pm on every frequency of
// both grids in one batched MFMA product with a row-weighted column-dot
// epilogue (C = U_f^T B, v = sum_e w_e C_e^2), then sigma_contour_kernel.
int contour_sigma(mgw_ctx* c, const mgw_ranges* r, double eta, int nomega, const double* omega,
                  const double* se_energy0) {
  MGW_REQUIRE(c->has_cd, MGW_ERR_STATE, "no contour W");
  const int naux = c->naux, nimag = (int)c->cd_om_im.size(), nreal = (int)c->cd_om_re.size();
  const int nf = nimag + nreal;
  const int nse = r->nsemax - r->nsemin + 1;
  const int glo = r->ncore_G + 1, ghi = r->nvirtual_G - 1;
  const int np = ghi - glo + 1;
  const int64_t ncol = (int64_t)np * nse;
  MGW_REQUIRE(ncol < (1ll << 31), MGW_ERR_ARG, "sigma_contour: too many columns");
  double* epm = nullptr;
  MGW_TRY(dbuf(c, "eri_pm", (size_t)naux * ncol, &epm));
  MGW_TRY(ao2mo(c, glo, ghi, r->nsemin, r->nsemax, epm, MGW_FP64));
  const int nrb = gemm_f64_rowblocks(naux);
  double *part, *v;
  MGW_TRY(dbuf(c, "dot_part", (size_t)nrb * ncol * nf, &part));
  MGW_TRY(dbuf(c, "dot_v", (size_t)ncol * nf, &v));
  GemmF64 g;
  g.M = naux; g.N = (int)ncol; g.K = naux;
  g.transA = true; g.A = find_buf(c, "cd_u"); g.lda = naux; g.strideA = (int64_t)naux * naux;
  g.B = epm; g.ldb = naux; g.strideB = 0;
  g.dot_w = find_buf(c, "cd_w"); g.strideDotW = naux;
  g.dotpart = part; g.strideDot = (int64_t)nrb * ncol;
  g.batch = nf;
  g.batch_outer = true;
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_WB, 2.0 * naux * (double)naux * ncol * nf,
                 8.0 * ((double)naux * naux * nf + (double)naux * ncol), c->stream);
    MGW_TRY(gemm_f64(g, c->stream));
  }
  MGW_TRY(launch_dot_reduce(part, nrb, ncol, nf, nullptr, v, c->stream));
  std::vector<double> ep(np);
  std::vector<int32_t> ps(np);
  for (int k = 0; k < np; ++k) {
    ep[k] = c->energy_in_g()[c->so() + glo - 1 + k];
    ps[k] = glo + k;
  }
  double *omi_d, *wti_d, *omr_d, *ep_d, *e0_d, *om_d, *sigma;
  int32_t *ps_d, *err_d;
  MGW_TRY(upload(c, "cd_om_im", c->cd_om_im.data(), nimag, &omi_d));
  MGW_TRY(upload(c, "cd_wt_im", c->cd_wt_im.data(), nimag, &wti_d));
  MGW_TRY(upload(c, "cd_om_re", c->cd_om_re.data(), nreal, &omr_d));
  MGW_TRY(upload(c, "cd_ep", ep.data(), np, &ep_d));
  MGW_TRY(upload(c, "cd_pstate", ps.data(), np, &ps_d));
  MGW_TRY(upload(c, "se_e0p", se_energy0, nse, &e0_d));
  MGW_TRY(upload(c, "se_omega", omega, nomega, &om_d));
  MGW_TRY(dbuf(c, "sigma", (size_t)2 * nomega * nse, &sigma));
  const int32_t zero = 0;
  MGW_TRY(upload(c, "cd_err", &zero, 1, &err_d));
  MGW_TRY(launch_sigma_contour(v, np, nse, nimag, nreal, omi_d, wti_d, omr_d, ep_d, ps_d, r->nhomo_G, nlumo_of(r),
                               e0_d, om_d, nomega, eta, sigma, err_d, c->stream));
  int32_t err = 0;
  MGW_TRY(download(c, &err, err_d, 1));
  MGW_REQUIRE(err == 0, MGW_ERR_ARG, "sigma_contour: a residue frequency lies outside the real W grid");
  c->has_sigma = true;
  c->sigma_nomega = nomega;
  c->sigma_nse = nse;
  MGW_TRY(scratch_done(c, "dot_part"));
  MGW_TRY(scratch_done(c, "eri_pm"));
  return MGW_OK;
}

}  // namespace

// ======================================================================
//                               C-ABI
// ======================================================================
extern "C" {

int32_t mgw_abi_version(void) { return 1; }

const char* mgw_last_error(void) { return last_error().c_str(); }

int mgw_ctx_create(int device, mgw_ctx** out) {
  MGW_REQUIRE(out, MGW_ERR_ARG, "null output pointer");
  int ndev = 0;
  MGW_HIP(hipGetDeviceCount(&ndev));
  MGW_REQUIRE(device >= 0 && device < ndev, MGW_ERR_ARG, "invalid device " + std::to_string(device));
  MGW_HIP(hipSetDevice(device));
  // freed buffers stay in the device's pool for the next allocation of any
  // context of this process instead of going back to the driver at each
  // synchronisation
  {
    // once per device and process, so that a limit set by mgw_device_pool_limit stays
    static std::mutex m;
    static std::set<int> done;
    std::lock_guard<std::mutex> lock(m);
    hipMemPool_t pool;
    if (!done.count(device) && hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t keep = UINT64_MAX;
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
      done.insert(device);
    }
  }
  mgw_ctx* c = new mgw_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return MGW_ERR_HIP;
  }
  if (rocblas_create_handle(&c->blas) != rocblas_status_success ||
      rocblas_set_stream(c->blas, c->stream) != rocblas_status_success) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    set_error("rocblas handle creation failed");
    return MGW_ERR_SOLVER;
  }
  *out = c;
  return MGW_OK;
}

int mgw_ctx_destroy(mgw_ctx* c) {
  if (!c) return MGW_OK;
  MGW_HIP(hipSetDevice(c->device));
  MGW_HIP(hipStreamSynchronize(c->stream));
  clear_graphs(c);
  for (auto& kv : c->ledger.bufs)
    if (kv.second.ptr) MGW_HIP(hipFreeAsync(kv.second.ptr, c->stream));
  MGW_HIP(hipStreamSynchronize(c->stream));
  rocblas_destroy_handle(c->blas);
  MGW_HIP(hipStreamDestroy(c->stream));
  delete c;
  return MGW_OK;
}

int mgw_set_energy_g(mgw_ctx* c, const double* energy_g) {
  MGW_CTX(c);
  MGW_REQUIRE(c->has_mo, MGW_ERR_STATE, "MO data not set (mgw_set_mo)");
  if (energy_g) c->energy_g.assign(energy_g, energy_g + (size_t)c->nstate * c->nspin);
  else c->energy_g.clear();
  c->has_sigma = false;
  c->has_sigma_calc = false;
  return MGW_OK;
}

int mgw_ctx_graphs(mgw_ctx* c, int enable) {
  MGW_CTX(c);
  MGW_HIP(hipStreamSynchronize(c->stream));
  if (!enable) clear_graphs(c);
  c->use_graphs = enable != 0;
  return MGW_OK;
}

int mgw_ctx_synchronize(mgw_ctx* c) {
  MGW_CTX(c);
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_autotune(int mode, int* previous) {
  MGW_REQUIRE(mode >= -1 && mode <= 1, MGW_ERR_ARG, "mgw_autotune: mode must be -1, 0 or 1");
  const int old = gemm_autotune_mode(mode);
  if (previous) *previous = old;
  return MGW_OK;
}

int mgw_ctx_trim(mgw_ctx* c, int eager) {
  MGW_CTX(c);
  MGW_HIP(hipStreamSynchronize(c->stream));
  std::vector<std::string> names;
  for (auto& kv : c->ledger.bufs)
    if (!is_resident_input(kv.first)) names.push_back(kv.first);
  for (auto& n : names) MGW_TRY(release(c, n));
  clear_graphs(c);
  invalidate_derived(c);
  c->trim_eagerly = eager != 0;
  // the device pool keeps freed memory (release threshold at the maximum, so
  // that later allocations do not go back to the driver): hand what no
  // context holds back to the device, so that the pool's reserved size stays
  // within what the contexts' ledgers account for
  MGW_HIP(hipStreamSynchronize(c->stream));
  hipMemPool_t pool;
  MGW_HIP(hipDeviceGetDefaultMemPool(&pool, c->device));
  MGW_HIP(hipMemPoolTrimTo(pool, 0));
  return MGW_OK;
}

int mgw_profile_enable(mgw_ctx* c, int enable) {
  MGW_CTX(c);
  MGW_HIP(hipStreamSynchronize(c->stream));
  c->prof.reset();
  c->prof.on = enable != 0;
  return MGW_OK;
}

int mgw_profile_read(mgw_ctx* c, int kernel_class, int64_t* launches, double* total_ms,
                     double* algorithmic_flops, double* algorithmic_bytes) {
  MGW_CTX(c);
  MGW_REQUIRE(kernel_class >= 0 && kernel_class < MGW_KCLASS_COUNT, MGW_ERR_ARG, "invalid kernel class");
  MGW_HIP(hipStreamSynchronize(c->stream));
  int64_t n = 0;
  double ms = 0.0, fl = 0.0, by = 0.0;
  for (const auto& r : c->prof.recs) {
    if (r.klass != kernel_class) continue;
    float t = 0.0f;
    MGW_HIP(hipEventElapsedTime(&t, r.a, r.b));
    ++n;
    ms += t;
    fl += r.flops;
    by += r.bytes;
  }
  if (launches) *launches = n;
  if (total_ms) *total_ms = ms;
  if (algorithmic_flops) *algorithmic_flops = fl;
  if (algorithmic_bytes) *algorithmic_bytes = by;
  return MGW_OK;
}

int mgw_precision_escalations(mgw_ctx* c, int64_t* count) {
  MGW_REQUIRE(c && count, MGW_ERR_ARG, "mgw_precision_escalations: null argument");
  *count = c->escalations;
  return MGW_OK;
}

int mgw_device_pool_report(int device, int reset_high, int64_t* reserved_bytes, int64_t* reserved_high_bytes) {
  int ndev = 0;
  MGW_HIP(hipGetDeviceCount(&ndev));
  MGW_REQUIRE(device >= 0 && device < ndev, MGW_ERR_ARG, "invalid device " + std::to_string(device));
  MGW_HIP(hipSetDevice(device));
  hipMemPool_t pool;
  MGW_HIP(hipDeviceGetDefaultMemPool(&pool, device));
  uint64_t cur = 0, high = 0;
  MGW_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &cur));
  MGW_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemHigh, &high));
  if (reset_high) {
    uint64_t zero = 0;   // the high-water mark restarts from the current size
    MGW_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReservedMemHigh, &zero));
  }
  if (reserved_bytes) *reserved_bytes = (int64_t)cur;
  if (reserved_high_bytes) *reserved_high_bytes = (int64_t)high;
  return MGW_OK;
}

int mgw_device_pool_limit(int device, int64_t bytes) {
  MGW_REQUIRE(bytes >= 0, MGW_ERR_ARG, "mgw_device_pool_limit: negative limit");
  int ndev = 0;
  MGW_HIP(hipGetDeviceCount(&ndev));
  MGW_REQUIRE(device >= 0 && device < ndev, MGW_ERR_ARG, "invalid device " + std::to_string(device));
  MGW_HIP(hipSetDevice(device));
  hipMemPool_t pool;
  MGW_HIP(hipDeviceGetDefaultMemPool(&pool, device));
  uint64_t v = (uint64_t)bytes;
  MGW_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &v));
  return MGW_OK;
}

int mgw_device_alloc(int device, int64_t bytes, void** ptr) {
  MGW_REQUIRE(ptr && bytes >= 0, MGW_ERR_ARG, "mgw_device_alloc: invalid arguments");
  int ndev = 0;
  MGW_HIP(hipGetDeviceCount(&ndev));
  MGW_REQUIRE(device >= 0 && device < ndev, MGW_ERR_ARG, "invalid device " + std::to_string(device));
  MGW_HIP(hipSetDevice(device));
  *ptr = nullptr;
  if (bytes == 0) return MGW_OK;
  const hipError_t e = hipMalloc(ptr, (size_t)bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    set_error("mgw_device_alloc: " + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? MGW_ERR_NOMEM : MGW_ERR_HIP;
  }
  return MGW_OK;
}

int mgw_device_free(void* ptr) {
  if (ptr) MGW_HIP(hipFree(ptr));
  return MGW_OK;
}

int mgw_device_copy(void* dst, const void* src, int64_t bytes, int kind) {
  MGW_REQUIRE(bytes >= 0 && kind >= 0 && kind <= 2, MGW_ERR_ARG, "mgw_device_copy: invalid arguments");
  if (bytes == 0) return MGW_OK;
  MGW_REQUIRE(dst && src, MGW_ERR_ARG, "mgw_device_copy: null pointer");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
  MGW_HIP(hipMemcpy(dst, src, (size_t)bytes, k));
  return MGW_OK;
}

int mgw_device_memory(int device, int64_t* free_bytes, int64_t* total_bytes) {
  int ndev = 0;
  MGW_HIP(hipGetDeviceCount(&ndev));
  MGW_REQUIRE(device >= 0 && device < ndev, MGW_ERR_ARG, "invalid device " + std::to_string(device));
  MGW_HIP(hipSetDevice(device));
  size_t f = 0, t = 0;
  MGW_HIP(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = (int64_t)f;
  if (total_bytes) *total_bytes = (int64_t)t;
  return MGW_OK;
}

int mgw_memory_report(mgw_ctx* c, int64_t* current_bytes, int64_t* peak_bytes) {
  MGW_REQUIRE(c, MGW_ERR_ARG, "null context");
  if (current_bytes) *current_bytes = c->ledger.current;
  if (peak_bytes) *peak_bytes = c->ledger.peak;
  return MGW_OK;
}

int64_t mgw_footprint_bytes(int nbf, int nauxil, int nstate, int nocc, int technique,
                            int nomega_chi_imag, int precision) {
  return mgw_footprint_bytes_nse(nbf, nauxil, nstate, nocc, nstate, technique, nomega_chi_imag, precision);
}

int64_t mgw_footprint_bytes_nse(int nbf, int nauxil, int nstate, int nocc, int nse, int technique,
                                int nomega_chi_imag, int precision) {
  const int64_t nb = nbf, na = nauxil, ns = nstate, no = nocc, nv = nstate - nocc;
  const int64_t ne = std::max(1, std::min(nse, nstate));   // self-energy states
  const int64_t npole = no * nv;
  int64_t d = na * nb * nb;                       // resident dense RI tensor
  d += nb * ns;                                   // C
  int64_t stage_w;
  if (technique == MGW_TECH_ONE_SHOT) {
    stage_w = na * nb * no + na * npole           // half transform + (P|ia)
              + 2 * npole * npole                 // Casida matrix + eigen workspace
              + na * npole;                       // w_s
  } else if (technique == MGW_TECH_CONTOUR) {
    // eigenvectors of chi0 on both grids, eigenvalues, weights
    stage_w = na * nb * no + na * npole + na * na * (int64_t)nomega_chi_imag + npole * nomega_chi_imag +
              3 * na * (int64_t)nomega_chi_imag;
  } else {
    stage_w = na * nb * no + na * npole + na * na * (int64_t)nomega_chi_imag + npole * nomega_chi_imag +
              (int64_t)mgw::chol_work_doubles((int)na) * nomega_chi_imag + 4 + nomega_chi_imag;
    if (precision == MGW_FP16X3)
      stage_w += (int64_t)((mgw::chol_plane_bytes((int)na, nomega_chi_imag) + 7) / 8);
    if (split_precision(precision))
      stage_w += (int64_t)(mgw::chi0_split_workspace((int)na, (int)npole, nomega_chi_imag, precision) / 8);
  }
  // FP16x3 transforms: the first product's states run up to the top of the
  // self-energy range (its half transform is reused there), and the fp16
  // hi/lo planes of the half-transformed tensor (4 bytes per element, rows
  // padded to 128, k to 32)
  if (precision == MGW_FP16X3) stage_w += na * nb * ne + (na + 128) * (nb + 32) * (no + ne + 32) / 2;
  // self-energy stage: half transform and (P|pm) for the ne states p, all m
  int64_t stage_s = na * nb * ne + na * ne * ns + na * npole + (int64_t(1) << 27);
  if (technique != MGW_TECH_ONE_SHOT)   // column-dot partials and v on every frequency
    stage_s += (int64_t)(mgw::gemm_f64_rowblocks((int)na) + 1) * ns * ne * nomega_chi_imag;
  if (technique != MGW_TECH_ONE_SHOT && split_precision(precision))
    stage_s += (int64_t)((mgw::wb_split_workspace((int)na, ns * ne, nomega_chi_imag, precision) + 7) / 8);
  // the temporaries of the W stage stay cached while the self-energy runs
  return 8 * (d + stage_w + stage_s);
}

// Does the 0-based pair list (k, l) hold every unordered pair of nbf
// functions exactly once?
bool all_pairs(const std::vector<int32_t>& kl, int nbf) {
  const int64_t npair = (int64_t)kl.size() / 2;
  if (npair != (int64_t)nbf * (nbf + 1) / 2) return false;
  std::vector<char> seen((size_t)nbf * nbf, 0);
  for (int64_t p = 0; p < npair; ++p) {
    const int a = std::min(kl[2 * p], kl[2 * p + 1]), b = std::max(kl[2 * p], kl[2 * p + 1]);
    char& f = seen[(size_t)a * nbf + b];
    if (f) return false;
    f = 1;
  }
  return true;
}

int mgw_set_eri_3center(mgw_ctx* c, const double* eri, int64_t npair, int nauxil,
                        const int32_t* index_basis, int nbf, int where) {
  MGW_CTX(c);
  MGW_REQUIRE(eri && index_basis && npair > 0 && nauxil > 0 && nbf > 0, MGW_ERR_ARG,
              "mgw_set_eri_3center: invalid arguments");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  MGW_REQUIRE((int64_t)nauxil * nbf * (int64_t)nbf < (int64_t)1 << 40, MGW_ERR_ARG, "RI tensor too large");
  std::vector<int32_t> kl(2 * npair);
  for (int64_t p = 0; p < npair; ++p) {
    const int32_t k = index_basis[2 * p] - 1, l = index_basis[2 * p + 1] - 1;
    MGW_REQUIRE(k >= 0 && k < nbf && l >= 0 && l < nbf, MGW_ERR_ARG,
                "index_basis entry out of range at pair " + std::to_string(p + 1));
    kl[2 * p] = k;
    kl[2 * p + 1] = l;
  }
  // a list of every unordered pair writes every element of the dense tensor:
  // the zero fill is needed only when pairs are screened out
  const bool complete = all_pairs(kl, nbf);
  const double* packed = eri;
  if (where == MGW_MEM_HOST) {
    double* d = nullptr;
    MGW_TRY(upload(c, "ri_packed", eri, (size_t)npair * nauxil, &d));
    packed = d;
  }
  int32_t* kl_d = nullptr;
  MGW_TRY(upload(c, "ri_kl", kl.data(), kl.size(), &kl_d));
  double* dense = nullptr;
  MGW_TRY(dbuf(c, "ri_dense", (size_t)nauxil * nbf * nbf, &dense));
  double* amax = nullptr;                     // max |B|: the scale of reduced-precision transforms
  MGW_TRY(dbuf(c, "ri_absmax", 1, &amax));
  MGW_TRY(launch_fill(amax, 1, 0.0, c->stream));
  MGW_TRY(launch_unpack_ri(packed, npair, nauxil, kl_d, nbf, dense, c->stream, !complete, amax));
  MGW_HIP(hipStreamSynchronize(c->stream));
  MGW_TRY(release(c, "ri_packed"));
  MGW_TRY(release(c, "ri_kl"));
  c->nbf = nbf;
  c->naux = nauxil;
  c->has_ri = true;
  invalidate_derived(c);
  return MGW_OK;
}

// Threshold below which an eigenvalue of the 2-center metric is dropped
// (TOO_LOW_EIGENVAL, src/m_eri.f90:27).
constexpr double TOO_LOW_EIGENVAL = 1.0e-6;

int mgw_inverse_sqrt_eri_2center(mgw_ctx* c, const double* eri_2center, int nauxil_bas, int where,
                                 int* nauxil_kept, double* lowest_eigval, double* sqrtinv_out) {
  MGW_CTX(c);
  MGW_REQUIRE(eri_2center && nauxil_bas > 0 && nauxil_kept, MGW_ERR_ARG,
              "mgw_inverse_sqrt_eri_2center: invalid arguments");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  const int n = nauxil_bas;
  c->has_sqrtinv = false;
  double *u = nullptr, *eig = nullptr, *ework = nullptr;
  int32_t* info_d = nullptr;
  MGW_TRY(dbuf(c, "metric_u", (size_t)n * n, &u));
  MGW_TRY(dbuf(c, "metric_eig", n, &eig));
  MGW_TRY(dbuf(c, "metric_ework", n, &ework));
  MGW_TRY(dbuf(c, "solver_info", 64, &info_d));
  MGW_HIP(hipMemcpyAsync(u, eri_2center, (size_t)n * n * sizeof(double),
                         where == MGW_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
  // the lower triangle is read, as by the reference's diagonalization
  // (src/m_eri_calculate.f90:1106-1108)
  {
    const double n3 = (double)n * n * n;
    ProfScope ps(&c->prof, MGW_KCLASS_EIGEN, (4.0 / 3.0 + 2.0) * n3, 16.0 * (double)n * n, c->stream);
    MGW_BLAS(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_lower, n, u, n, eig, ework, info_d));
  }
  int32_t info = 0;
  MGW_TRY(download(c, &info, info_d, 1));
  MGW_REQUIRE(info == 0, MGW_ERR_SOLVER, "rocsolver_dsyevd (2-center metric) failed, info=" + std::to_string(info));
  std::vector<double> ev(n);
  MGW_TRY(download(c, ev.data(), eig, n));
  // src/m_eri_calculate.f90:1117 and :1122
  int kept = 0;
  for (double v : ev) kept += v > TOO_LOW_EIGENVAL;
  MGW_REQUIRE(kept > 0, MGW_ERR_SOLVER, "2-center metric has no eigenvalue above 1e-6");
  const int neglect = n - kept;
  double* s = nullptr;
  MGW_TRY(dbuf(c, "ri_sqrtinv", (size_t)n * kept, &s));
  MGW_TRY(launch_invsqrt_select(u, n, eig, neglect, kept, s, c->stream));
  if (sqrtinv_out) {
    MGW_HIP(hipMemcpyAsync(sqrtinv_out, s, (size_t)n * kept * sizeof(double), hipMemcpyDefault, c->stream));
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  MGW_TRY(release(c, "metric_u"));
  MGW_TRY(release(c, "metric_eig"));
  MGW_TRY(release(c, "metric_ework"));
  c->naux_bas = n;
  c->naux_kept = kept;
  c->has_sqrtinv = true;
  c->build_next = 0;
  *nauxil_kept = kept;
  if (lowest_eigval) *lowest_eigval = ev[0];
  return MGW_OK;
}

int mgw_build_eri_3center(mgw_ctx* c, const double* eri_3center_tmp, int64_t ipair_first, int64_t mpair,
                          int64_t npair, const int32_t* index_basis, int nbf, int where, double* eri_3center_out) {
  MGW_CTX(c);
  MGW_REQUIRE(c->has_sqrtinv, MGW_ERR_STATE, "no 2-center metric: call mgw_inverse_sqrt_eri_2center first");
  MGW_REQUIRE(index_basis && npair > 0 && mpair > 0 && nbf > 0 && ipair_first >= 1 &&
                  ipair_first - 1 + mpair <= npair,
              MGW_ERR_ARG, "mgw_build_eri_3center: invalid arguments");
  const bool resident_tmp = eri_3center_tmp == nullptr;
  if (resident_tmp) {
    MGW_REQUIRE(c->has_ri_tmp3 && c->ri_tmp3_npair == npair && c->ri_tmp3_naux == c->naux_bas, MGW_ERR_STATE,
                "mgw_build_eri_3center: no resident 3-center integrals of this shape (mgw_eri_3center_ao)");
  }
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  const int nb = c->naux_bas, nk = c->naux_kept;
  MGW_REQUIRE((int64_t)nk * nbf * (int64_t)nbf < (int64_t)1 << 40, MGW_ERR_ARG, "RI tensor too large");
  if (ipair_first > 1) {
    MGW_REQUIRE(ipair_first == c->build_next && npair == c->build_npair && nbf == c->build_nbf, MGW_ERR_STATE,
                "mgw_build_eri_3center: batches must follow each other from ipair_first = 1 (expected " +
                    std::to_string(c->build_next) + ")");
  }
  std::vector<int32_t> kl(2 * mpair);
  for (int64_t p = 0; p < mpair; ++p) {
    const int64_t q = ipair_first - 1 + p;
    const int32_t k = index_basis[2 * q] - 1, l = index_basis[2 * q + 1] - 1;
    MGW_REQUIRE(k >= 0 && k < nbf && l >= 0 && l < nbf, MGW_ERR_ARG,
                "index_basis entry out of range at pair " + std::to_string(q + 1));
    kl[2 * p] = k;
    kl[2 * p + 1] = l;
  }
  double* dense = nullptr;
  if (ipair_first == 1) {
    c->has_ri = false;
    c->has_eri_mo = false;
    MGW_TRY(dbuf(c, "ri_dense", (size_t)nk * nbf * nbf, &dense));
    MGW_TRY(launch_fill(dense, (int64_t)nk * nbf * nbf, 0.0, c->stream));
    double* amax = nullptr;
    MGW_TRY(dbuf(c, "ri_absmax", 1, &amax));
    MGW_TRY(launch_fill(amax, 1, 0.0, c->stream));
    c->build_npair = npair;
    c->build_nbf = nbf;
  } else {
    dense = find_buf(c, "ri_dense");
    MGW_REQUIRE(dense, MGW_ERR_STATE, "mgw_build_eri_3center: the tensor of the first batch is gone");
  }
  const double* tmp = eri_3center_tmp;
  int64_t ld_tmp = mpair;
  if (resident_tmp) {
    tmp = find_buf(c, "ri_tmp3") + (ipair_first - 1);
    ld_tmp = npair;
  } else if (where == MGW_MEM_HOST) {
    double* d = nullptr;
    MGW_TRY(upload(c, "ri_tmp", eri_3center_tmp, (size_t)mpair * nb, &d));
    tmp = d;
  }
  double* packed = nullptr;
  MGW_TRY(dbuf(c, "ri_packed", (size_t)mpair * nk, &packed));
  // eri_3center(ipair, P) = sum_Q eri_3center_tmp(ipair, Q) sqrtinv(Q, P)
  // (src/m_eri_calculate.f90:1889-1892); FP64
  GemmF64 g;
  g.M = (int)mpair; g.N = nk; g.K = nb;
  g.A = tmp; g.lda = ld_tmp;
  g.B = find_buf(c, "ri_sqrtinv"); g.ldb = nb;
  g.C = packed; g.ldc = mpair;
  g.klass = MGW_KCLASS_GEMM_OTHER;
  MGW_REQUIRE(mpair < ((int64_t)1 << 31), MGW_ERR_ARG, "mgw_build_eri_3center: batch too large");
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_OTHER, 2.0 * (double)mpair * nk * nb,
                 8.0 * ((double)mpair * nb + (double)nb * nk + (double)mpair * nk), c->stream);
    MGW_TRY(gemm_f64(g, c->stream));
  }
  int32_t* kl_d = nullptr;
  MGW_TRY(upload(c, "ri_kl", kl.data(), kl.size(), &kl_d));
  MGW_TRY(launch_unpack_ri(packed, mpair, nk, kl_d, nbf, dense, c->stream, false, find_buf(c, "ri_absmax")));
  if (eri_3center_out) {
    // rows ipair_first.. of the caller's eri_3center(npair, nauxil_kept)
    MGW_HIP(hipMemcpy2DAsync(eri_3center_out + (ipair_first - 1), (size_t)npair * sizeof(double), packed,
                             (size_t)mpair * sizeof(double), (size_t)mpair * sizeof(double), nk, hipMemcpyDefault,
                             c->stream));
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  MGW_TRY(release(c, "ri_tmp"));
  MGW_TRY(release(c, "ri_packed"));
  MGW_TRY(release(c, "ri_kl"));
  c->build_next = ipair_first + mpair;
  if (c->build_next == npair + 1) {
    c->nbf = nbf;
    c->naux = nk;
    c->has_ri = true;
    c->build_next = 0;
    invalidate_derived(c);
    if (resident_tmp) {
      MGW_TRY(release(c, "ri_tmp3"));
      c->has_ri_tmp3 = false;
    }
  }
  return MGW_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ integral front end
namespace {

int ncart_of(int l) { return (l + 1) * (l + 2) / 2; }

// Checks a caller's shell set and copies it to the device under name prefix.
int upload_shells(mgw_ctx* c, const std::string& prefix, const mgw_shells* sh, EriShells& out, int& nfunc) {
  MGW_REQUIRE(sh && sh->nshell > 0 && sh->am && sh->x0 && sh->prim_offset && sh->alpha && sh->coeff &&
                  sh->transform_offset && sh->transform && sh->function_offset,
              MGW_ERR_ARG, "mgw_shells: null field or no shells");
  const int n = sh->nshell;
  MGW_REQUIRE(sh->prim_offset[0] == 0 && sh->function_offset[0] == 0, MGW_ERR_ARG,
              "mgw_shells: offsets must start at 0");
  int64_t ntmat = 0;
  for (int k = 0; k < n; ++k) {
    const int l = sh->am[k];
    MGW_REQUIRE(l >= 0 && l <= 6, MGW_ERR_ARG, "mgw_shells: angular momentum out of range at shell " + std::to_string(k));
    MGW_REQUIRE(sh->prim_offset[k + 1] > sh->prim_offset[k], MGW_ERR_ARG,
                "mgw_shells: shell " + std::to_string(k) + " has no primitive");
    const int nf = sh->function_offset[k + 1] - sh->function_offset[k];
    MGW_REQUIRE(nf >= 1 && nf <= ncart_of(l), MGW_ERR_ARG,
                "mgw_shells: function count of shell " + std::to_string(k) + " out of range");
    MGW_REQUIRE(sh->transform_offset[k] >= 0, MGW_ERR_ARG, "mgw_shells: negative transform offset");
    ntmat = std::max<int64_t>(ntmat, (int64_t)sh->transform_offset[k] + (int64_t)nf * ncart_of(l));
  }
  const int nprim = sh->prim_offset[n];
  nfunc = sh->function_offset[n];
  int32_t *l_d, *poff_d, *toff_d, *foff_d;
  double *cen_d, *exp_d, *coef_d, *tm_d;
  MGW_TRY(upload(c, prefix + "_l", sh->am, n, &l_d));
  MGW_TRY(upload(c, prefix + "_x0", sh->x0, 3 * (size_t)n, &cen_d));
  MGW_TRY(upload(c, prefix + "_poff", sh->prim_offset, n + 1, &poff_d));
  MGW_TRY(upload(c, prefix + "_alpha", sh->alpha, nprim, &exp_d));
  MGW_TRY(upload(c, prefix + "_coeff", sh->coeff, nprim, &coef_d));
  MGW_TRY(upload(c, prefix + "_toff", sh->transform_offset, n, &toff_d));
  MGW_TRY(upload(c, prefix + "_tmat", sh->transform, (size_t)ntmat, &tm_d));
  MGW_TRY(upload(c, prefix + "_foff", sh->function_offset, n + 1, &foff_d));
  out.l = l_d; out.center = cen_d; out.poff = poff_d; out.exps = exp_d; out.coefs = coef_d;
  out.toff = toff_d; out.tmat = tm_d; out.foff = foff_d;
  return MGW_OK;
}

int release_shells(mgw_ctx* c, const std::string& prefix) {
  for (const char* f : {"_l", "_x0", "_poff", "_alpha", "_coeff", "_toff", "_tmat", "_foff"})
    MGW_TRY(release(c, prefix + f));
  return MGW_OK;
}

int boys_table(mgw_ctx* c, const double** out) {
  double* d = find_buf(c, "boys_table");
  if (!d) {
    std::vector<double> h(boys_table_doubles());
    boys_table_fill(h.data());
    MGW_TRY(upload(c, "boys_table", h.data(), h.size(), &d));
  }
  *out = d;
  return MGW_OK;
}

// ket functions (shell, function) grouped by angular momentum: kfun holds
// them class by class; first[l] .. first[l + 1] is class l
void ket_functions(const mgw_shells* sh, std::vector<int32_t>& kfun, std::vector<int>& first) {
  first.assign(8, 0);
  kfun.clear();
  for (int l = 0; l <= 6; ++l) {
    first[l] = (int)kfun.size() / 2;
    for (int k = 0; k < sh->nshell; ++k) {
      if (sh->am[k] != l) continue;
      const int nf = sh->function_offset[k + 1] - sh->function_offset[k];
      for (int f = 0; f < nf; ++f) {
        kfun.push_back(k);
        kfun.push_back(f);
      }
    }
  }
  first[7] = (int)kfun.size() / 2;
}

}  // namespace

extern "C" {

int mgw_eri_2center_ao(mgw_ctx* c, const mgw_shells* auxil, int where, double* eri_2center) {
  MGW_CTX(c);
  MGW_REQUIRE(eri_2center, MGW_ERR_ARG, "mgw_eri_2center_ao: null output");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  EriShells A;
  int naux = 0;
  MGW_TRY(upload_shells(c, "int_aux", auxil, A, naux));
  const int n = auxil->nshell;
  for (int k = 0; k < n; ++k)
    MGW_REQUIRE(eri_class_supported(auxil->am[k], 0, auxil->am[k]), MGW_ERR_ARG,
                "mgw_eri_2center_ao: auxiliary angular momentum above 4 is not supported on the GPU");
  // the bra's second shell: a unit s function (exponent 0) at each shell's centre
  std::vector<int32_t> ul(n, 0), upoff(n + 1), utoff(n), ufoff(n + 1), pairs;
  std::vector<double> uexp(n, 0.0), ucoef(n, 1.0), utm(n, 1.0);
  for (int k = 0; k <= n; ++k) upoff[k] = ufoff[k] = k;
  for (int k = 0; k < n; ++k) utoff[k] = k;
  const mgw_shells unit{n, ul.data(), auxil->x0, upoff.data(), uexp.data(), ucoef.data(), utoff.data(), utm.data(),
                        ufoff.data()};
  EriShells U;
  int nunit = 0;
  MGW_TRY(upload_shells(c, "int_unit", &unit, U, nunit));
  std::vector<int32_t> kfun;
  std::vector<int> first;
  ket_functions(auxil, kfun, first);
  // bra pairs (P shell, its unit shell) grouped by l_P, in the same order
  std::vector<int32_t> kshell_pairs;
  std::vector<int> pfirst(8, 0);
  for (int l = 0; l <= 6; ++l) {
    pfirst[l] = (int)kshell_pairs.size() / 2;
    for (int k = 0; k < n; ++k)
      if (auxil->am[k] == l) {
        kshell_pairs.push_back(k);
        kshell_pairs.push_back(k);
      }
  }
  pfirst[7] = (int)kshell_pairs.size() / 2;
  int32_t *kf_d, *pr_d;
  MGW_TRY(upload(c, "int_kfun", kfun.data(), kfun.size(), &kf_d));
  MGW_TRY(upload(c, "int_pairs", kshell_pairs.data(), kshell_pairs.size(), &pr_d));
  const double* boys = nullptr;
  MGW_TRY(boys_table(c, &boys));
  double* out = eri_2center;
  if (where == MGW_MEM_HOST) MGW_TRY(dbuf(c, "int_out2", (size_t)naux * naux, &out));
  for (int la = 0; la <= 4; ++la)
    for (int lc = 0; lc <= 4; ++lc) {
      EriLaunch e;
      e.la = la; e.lb = 0; e.lc = lc;
      e.A = A; e.B = U; e.C = A;
      e.pairs = pr_d + 2 * (int64_t)pfirst[la];
      e.npairs = pfirst[la + 1] - pfirst[la];
      e.kfun = kf_d + 2 * (int64_t)first[lc];
      e.nkfun = first[lc + 1] - first[lc];
      e.boys = boys;
      e.metric = 1;
      e.naux = naux;
      e.out = out;
      MGW_TRY(launch_eri_class(e, c->stream));
    }
  if (where == MGW_MEM_HOST) MGW_TRY(download(c, eri_2center, out, (size_t)naux * naux));
  MGW_HIP(hipStreamSynchronize(c->stream));
  for (const char* b : {"int_kfun", "int_pairs", "int_out2"}) MGW_TRY(release(c, b));
  MGW_TRY(release_shells(c, "int_aux"));
  MGW_TRY(release_shells(c, "int_unit"));
  return MGW_OK;
}

int mgw_eri_3center_ao(mgw_ctx* c, const mgw_shells* basis, const mgw_shells* auxil, int64_t npair,
                       const int32_t* index_basis, int where, double* eri_3center_tmp) {
  MGW_CTX(c);
  MGW_REQUIRE(index_basis && npair > 0, MGW_ERR_ARG, "mgw_eri_3center_ao: invalid pair list");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  EriShells S, A;
  int nbf = 0, naux = 0;
  MGW_TRY(upload_shells(c, "int_bas", basis, S, nbf));
  MGW_TRY(upload_shells(c, "int_aux", auxil, A, naux));
  MGW_REQUIRE(npair <= (int64_t)nbf * (nbf + 1) / 2, MGW_ERR_ARG, "mgw_eri_3center_ao: more pairs than nbf(nbf+1)/2");
  // pair index of every (mu, nu), both orders
  std::vector<int32_t> pmap((size_t)nbf * nbf, -1);
  for (int64_t p = 0; p < npair; ++p) {
    const int k = index_basis[2 * p] - 1, l = index_basis[2 * p + 1] - 1;
    MGW_REQUIRE(k >= 0 && k < nbf && l >= 0 && l < nbf, MGW_ERR_ARG,
                "index_basis entry out of range at pair " + std::to_string(p + 1));
    int32_t& a = pmap[(size_t)k * nbf + l];
    MGW_REQUIRE(a < 0, MGW_ERR_ARG, "index_basis lists a pair twice at pair " + std::to_string(p + 1));
    a = (int32_t)p;
    pmap[(size_t)l * nbf + k] = (int32_t)p;
  }
  // shell pairs (s1, s2) with l1 >= l2, grouped by (l1, l2)
  const int n = basis->nshell;
  std::vector<std::vector<int32_t>> cls(64);
  for (int a = 0; a < n; ++a)
    for (int b = 0; b <= a; ++b) {
      int s1 = a, s2 = b;
      if (basis->am[s1] < basis->am[s2]) std::swap(s1, s2);
      const int l1 = basis->am[s1], l2 = basis->am[s2];
      MGW_REQUIRE(l1 <= 3, MGW_ERR_ARG, "mgw_eri_3center_ao: orbital angular momentum above 3 is not supported on the GPU");
      cls[l1 * 8 + l2].push_back(s1);
      cls[l1 * 8 + l2].push_back(s2);
    }
  for (int k = 0; k < auxil->nshell; ++k)
    MGW_REQUIRE(auxil->am[k] <= 4, MGW_ERR_ARG,
                "mgw_eri_3center_ao: auxiliary angular momentum above 4 is not supported on the GPU");
  std::vector<int32_t> pairs;
  std::vector<int64_t> cfirst(65, 0);
  for (int k = 0; k < 64; ++k) {
    cfirst[k] = (int64_t)pairs.size() / 2;
    pairs.insert(pairs.end(), cls[k].begin(), cls[k].end());
  }
  cfirst[64] = (int64_t)pairs.size() / 2;
  std::vector<int32_t> kfun;
  std::vector<int> first;
  ket_functions(auxil, kfun, first);
  int32_t *pm_d, *pr_d, *kf_d;
  MGW_TRY(upload(c, "int_pairmap", pmap.data(), pmap.size(), &pm_d));
  MGW_TRY(upload(c, "int_pairs", pairs.data(), pairs.size(), &pr_d));
  MGW_TRY(upload(c, "int_kfun", kfun.data(), kfun.size(), &kf_d));
  const double* boys = nullptr;
  MGW_TRY(boys_table(c, &boys));
  double* out = eri_3center_tmp;
  if (!eri_3center_tmp) {
    MGW_TRY(dbuf(c, "ri_tmp3", (size_t)npair * naux, &out));
  } else if (where == MGW_MEM_HOST) {
    MGW_TRY(dbuf(c, "int_out3", (size_t)npair * naux, &out));
  }
  {
    ProfScope ps(&c->prof, MGW_KCLASS_OTHER, 0.0, 8.0 * (double)npair * naux, c->stream);
    for (int l1 = 0; l1 <= 3; ++l1)
      for (int l2 = 0; l2 <= l1; ++l2)
        for (int lc = 0; lc <= 4; ++lc) {
          EriLaunch e;
          e.la = l1; e.lb = l2; e.lc = lc;
          e.A = S; e.B = S; e.C = A;
          e.pairs = pr_d + 2 * cfirst[l1 * 8 + l2];
          e.npairs = cfirst[l1 * 8 + l2 + 1] - cfirst[l1 * 8 + l2];
          e.kfun = kf_d + 2 * (int64_t)first[lc];
          e.nkfun = first[lc + 1] - first[lc];
          e.boys = boys;
          e.metric = 0;
          e.pairmap = pm_d;
          e.nbf = nbf;
          e.npair = npair;
          e.out = out;
          MGW_TRY(launch_eri_class(e, c->stream));
        }
  }
  if (eri_3center_tmp && where == MGW_MEM_HOST) MGW_TRY(download(c, eri_3center_tmp, out, (size_t)npair * naux));
  MGW_HIP(hipStreamSynchronize(c->stream));
  for (const char* b : {"int_pairmap", "int_pairs", "int_kfun", "int_out3"}) MGW_TRY(release(c, b));
  MGW_TRY(release_shells(c, "int_bas"));
  MGW_TRY(release_shells(c, "int_aux"));
  if (!eri_3center_tmp) {
    c->has_ri_tmp3 = true;
    c->ri_tmp3_npair = npair;
    c->ri_tmp3_naux = naux;
  }
  return MGW_OK;
}

int mgw_exchange_m_vxc(mgw_ctx* c, const double* fock, double* exchange_m_vxc, double* diag) {
  MGW_CTX(c);
  MGW_REQUIRE(fock, MGW_ERR_ARG, "mgw_exchange_m_vxc: null hamiltonian_fock");
  MGW_REQUIRE(c->has_mo, MGW_ERR_STATE, "MO data not set (mgw_set_mo)");
  const int nbf = c->mo_nbf, ns = c->nstate;
  double *f, *t, *m;
  MGW_TRY(dbuf(c, "xc_fock", (size_t)nbf * nbf, &f));
  MGW_TRY(dbuf(c, "xc_fc", (size_t)nbf * ns, &t));
  MGW_TRY(dbuf(c, "xc_mo", (size_t)ns * ns, &m));
  std::vector<double> h((size_t)ns * ns);
  for (int sp = 0; sp < c->nspin; ++sp) {
    const double* C = find_buf(c, "c_matrix") + (int64_t)sp * nbf * ns;
    MGW_HIP(hipMemcpyAsync(f, fock + (size_t)sp * nbf * nbf, sizeof(double) * nbf * nbf, hipMemcpyHostToDevice,
                           c->stream));
    // (F C), then C^T (F C)
    GemmF64 a;
    a.M = nbf; a.N = ns; a.K = nbf;
    a.A = f; a.lda = nbf;
    a.B = C; a.ldb = nbf;
    a.C = t; a.ldc = nbf;
    MGW_TRY(gemm_f64(a, c->stream));
    GemmF64 b;
    b.M = ns; b.N = ns; b.K = nbf;
    b.transA = true;
    b.A = C; b.lda = nbf;
    b.B = t; b.ldb = nbf;
    b.C = m; b.ldc = ns;
    MGW_TRY(gemm_f64(b, c->stream));
    MGW_TRY(download(c, h.data(), m, h.size()));
    for (int p = 0; p < ns; ++p) h[(size_t)p * ns + p] -= c->energy[(size_t)sp * ns + p];
    if (exchange_m_vxc) std::memcpy(exchange_m_vxc + (size_t)sp * ns * ns, h.data(), sizeof(double) * h.size());
    if (diag)
      for (int p = 0; p < ns; ++p) diag[(size_t)sp * ns + p] = h[(size_t)p * ns + p];
  }
  for (const char* b : {"xc_fock", "xc_fc", "xc_mo"}) MGW_TRY(scratch_done(c, b));
  return MGW_OK;
}

int mgw_hartree_exchange_ri(mgw_ctx* c, const double* x, int nbf, int nx, double spin_fact, double* hartree_ao,
                            double* exchange_ao) {
  MGW_CTX(c);
  MGW_REQUIRE(c->has_ri, MGW_ERR_STATE, "RI tensor not set (mgw_set_eri_3center / mgw_build_eri_3center)");
  MGW_REQUIRE(x && nbf == c->nbf && nx >= 1, MGW_ERR_ARG, "mgw_hartree_exchange_ri: x must be nbf x nx with nbf of the RI tensor");
  const int naux = c->naux;
  MGW_REQUIRE((int64_t)naux * nx < ((int64_t)1 << 31) && (int64_t)nbf * nbf < ((int64_t)1 << 31), MGW_ERR_ARG,
              "mgw_hartree_exchange_ri: problem too large");
  const double* B = find_buf(c, "ri_dense");
  double *xd, *xt, *h, *gam, *jv, *kv;
  MGW_TRY(upload(c, "hf_x", x, (size_t)nbf * nx, &xd));
  std::vector<double> xth((size_t)nx * nbf);   // xt[i + nx k] = x(k, i)
  for (int k = 0; k < nbf; ++k)
    for (int i = 0; i < nx; ++i) xth[(size_t)i + (size_t)nx * k] = x[(size_t)k + (size_t)nbf * i];
  MGW_TRY(upload(c, "hf_xt", xth.data(), xth.size(), &xt));
  MGW_TRY(dbuf(c, "hf_h", (size_t)naux * nx * nbf, &h));
  MGW_TRY(dbuf(c, "hf_gamma", naux, &gam));
  MGW_TRY(dbuf(c, "hf_j", (size_t)nbf * nbf, &jv));
  MGW_TRY(dbuf(c, "hf_k", (size_t)nbf * nbf, &kv));
  {
    ProfScope ps(&c->prof, MGW_KCLASS_GEMM_OTHER,
                 2.0 * naux * (double)nbf * nbf * nx * 2.0 + 4.0 * naux * (double)nbf * nbf,
                 8.0 * ((double)naux * nbf * nbf * 2.0 + 2.0 * naux * nx * nbf), c->stream);
    // H(P, i, k) = sum_l B(P, k, l) x(l, i): one product per k (B(:, :, k) is contiguous)
    GemmF64 g;
    g.M = naux; g.N = nx; g.K = nbf;
    g.A = B; g.lda = naux; g.strideA = (int64_t)naux * nbf;
    g.B = xd; g.ldb = nbf; g.strideB = 0;
    g.C = h; g.ldc = naux; g.strideC = (int64_t)naux * nx;
    g.batch = nbf;
    MGW_TRY(gemm_f64(g, c->stream));
    // exchange = -H^T H over the rows (P, i)
    GemmF64 k;
    k.M = nbf; k.N = nbf; k.K = naux * nx;
    k.transA = true;
    k.A = h; k.lda = (int64_t)naux * nx;
    k.B = h; k.ldb = (int64_t)naux * nx;
    k.C = kv; k.ldc = nbf;
    k.alpha = -1.0;
    k.lower_only = true;
    MGW_TRY(gemm_f64(k, c->stream));
    // gamma(P) = spin_fact sum_{i,k} H(P, i, k) x(k, i)
    GemmF64 y;
    y.M = naux; y.N = 1; y.K = nx * nbf;
    y.A = h; y.lda = naux;
    y.B = xt; y.ldb = (int64_t)nx * nbf;
    y.C = gam; y.ldc = naux;
    y.alpha = spin_fact;
    MGW_TRY(gemm_f64(y, c->stream));
    // hartree(k, l) = sum_P B(P, k, l) gamma(P)
    GemmF64 jj;
    jj.M = nbf * nbf; jj.N = 1; jj.K = naux;
    jj.transA = true;
    jj.A = B; jj.lda = naux;
    jj.B = gam; jj.ldb = naux;
    jj.C = jv; jj.ldc = (int64_t)nbf * nbf;
    MGW_TRY(gemm_f64(jj, c->stream));
  }
  if (hartree_ao) MGW_TRY(download(c, hartree_ao, jv, (size_t)nbf * nbf));
  if (exchange_ao) MGW_TRY(download(c, exchange_ao, kv, (size_t)nbf * nbf));
  MGW_HIP(hipStreamSynchronize(c->stream));
  for (const char* b : {"hf_x", "hf_xt", "hf_gamma", "hf_j", "hf_k"}) MGW_TRY(scratch_done(c, b));
  MGW_TRY(scratch_done(c, "hf_h"));
  return MGW_OK;
}

int mgw_destroy_eri_3center(mgw_ctx* c) {
  MGW_CTX(c);
  MGW_TRY(release(c, "ri_dense"));
  c->has_ri = false;
  invalidate_derived(c);
  return MGW_OK;
}

int mgw_set_mo_spin(mgw_ctx* c, const double* c_matrix, int nbf, int nstate, int nspin, const double* energy,
                    const double* occupation, int where) {
  MGW_CTX(c);
  MGW_REQUIRE(c_matrix && energy && occupation && nbf > 0 && nstate > 0 && (nspin == 1 || nspin == 2),
              MGW_ERR_ARG, "mgw_set_mo_spin: invalid arguments");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  MGW_REQUIRE(!c->has_ri || nbf == c->nbf, MGW_ERR_ARG, "mgw_set_mo: nbf differs from the RI tensor");
  const size_t nc = (size_t)nbf * nstate * nspin, ne = (size_t)nstate * nspin;
  double* cd = nullptr;
  MGW_TRY(dbuf(c, "c_matrix", nc, &cd));
  c->energy.resize(ne);
  c->occupation.resize(ne);
  if (where == MGW_MEM_DEVICE) {
    MGW_HIP(hipMemcpyAsync(cd, c_matrix, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
    MGW_HIP(hipMemcpyAsync(c->energy.data(), energy, sizeof(double) * ne, hipMemcpyDeviceToHost, c->stream));
    MGW_HIP(hipMemcpyAsync(c->occupation.data(), occupation, sizeof(double) * ne, hipMemcpyDeviceToHost, c->stream));
    MGW_HIP(hipStreamSynchronize(c->stream));
  } else {
    MGW_HIP(hipMemcpyAsync(cd, c_matrix, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream));
    std::memcpy(c->energy.data(), energy, sizeof(double) * ne);
    std::memcpy(c->occupation.data(), occupation, sizeof(double) * ne);
  }
  double *ed, *od;
  MGW_TRY(upload(c, "energy", c->energy.data(), ne, &ed));
  MGW_TRY(upload(c, "occupation", c->occupation.data(), ne, &od));
  MGW_HIP(hipStreamSynchronize(c->stream));
  c->nstate = nstate;
  c->nspin = nspin;
  c->spin = 0;
  c->energy_g.clear();
  c->has_mo = true;
  c->mo_nbf = nbf;
  invalidate_derived(c);
  return MGW_OK;
}

int mgw_set_mo(mgw_ctx* c, const double* c_matrix, int nbf, int nstate, const double* energy,
               const double* occupation, int where) {
  return mgw_set_mo_spin(c, c_matrix, nbf, nstate, 1, energy, occupation, where);
}

int mgw_set_spin(mgw_ctx* c, int ispin) {
  MGW_CTX(c);
  MGW_REQUIRE(c->has_mo, MGW_ERR_STATE, "MO data not set (mgw_set_mo)");
  MGW_REQUIRE(ispin >= 1 && ispin <= c->nspin, MGW_ERR_ARG, "mgw_set_spin: spin channel out of range");
  if (c->spin != ispin - 1) {
    c->spin = ispin - 1;
    c->has_eri_mo = false;
    c->has_sigma = false;
    c->has_sigma_calc = false;
  }
  return MGW_OK;
}

int mgw_calculate_eri_3center_mo(mgw_ctx* c, int mstate_min, int mstate_max, int nstate_min,
                                 int nstate_max, double* out) {
  MGW_CTX(c);
  const int nm = mstate_max - mstate_min + 1, nn = nstate_max - nstate_min + 1;
  if (!c->has_eri_mo) {
    double* e = nullptr;
    MGW_REQUIRE(nm > 0 && nn > 0, MGW_ERR_ARG, "empty state range");
    MGW_TRY(dbuf(c, "eri_3center_mo", (size_t)c->naux * nm * nn, &e));
    MGW_TRY(ao2mo(c, mstate_min, mstate_max, nstate_min, nstate_max, e));
    c->has_eri_mo = true;
    c->mlo = mstate_min; c->mhi = mstate_max; c->nlo = nstate_min; c->nhi = nstate_max;
  }
  if (out) {
    MGW_REQUIRE(c->mlo == mstate_min && c->mhi == mstate_max && c->nlo == nstate_min && c->nhi == nstate_max,
                MGW_ERR_STATE, "a different eri_3center_mo block is resident; destroy it first");
    MGW_TRY(download(c, out, find_buf(c, "eri_3center_mo"), (size_t)c->naux * nm * nn));
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

// ---------------------------------------------------------------- no-allocation AO -> MO
// The half-transformed tensor of mgw_calculate_eri_3center_mo_noalloc lives
// in this fixed array (one copy per GPU, held until the process ends). Host
// code never passes `noalloc_scratch` itself to a kernel: its device address
// is looked up per device with hipGetSymbolAddress on the first call there.
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = MGW_NOALLOC_SCRATCH_DOUBLES;
static __device__ __align__(256) double noalloc_scratch[NOALLOC_SCRATCH_DOUBLES];

int noalloc_scratch_ptr(int device, double** out) {
  constexpr int MAXDEV = 64;
  static std::mutex m;
  static double* addr[MAXDEV] = {};
  MGW_REQUIRE(device >= 0 && device < MAXDEV, MGW_ERR_ARG, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lock(m);
  if (!addr[device]) {
    MGW_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      set_error(std::string("no-allocation scratch: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
      return MGW_ERR_HIP;
    }
    addr[device] = static_cast<double*>(p);
  }
  *out = addr[device];
  return MGW_OK;
}

int mgw_calculate_eri_3center_mo_noalloc(mgw_ctx* c, int mstate_min, int mstate_max, int nstate_min,
                                         int nstate_max, double* eri_3center_mo_dev) {
  MGW_CTX(c);
  MGW_REQUIRE(eri_3center_mo_dev, MGW_ERR_ARG, "mgw_calculate_eri_3center_mo_noalloc: null output");
  MGW_REQUIRE(c->has_ri && c->has_mo, MGW_ERR_STATE,
              "mgw_calculate_eri_3center_mo_noalloc: set the RI tensor and the MO data first");
  MGW_REQUIRE(mstate_min >= 1 && mstate_max <= c->nstate && mstate_min <= mstate_max && nstate_min >= 1 &&
                  nstate_max <= c->nstate && nstate_min <= nstate_max,
              MGW_ERR_ARG, "mgw_calculate_eri_3center_mo_noalloc: invalid state range");
  const int64_t nf = std::min(mstate_max - mstate_min + 1, nstate_max - nstate_min + 1);
  const int64_t need = (int64_t)c->naux * c->nbf * nf;
  MGW_REQUIRE(need <= NOALLOC_SCRATCH_DOUBLES, MGW_ERR_ARG,
              "mgw_calculate_eri_3center_mo_noalloc: nauxil x nbf x min(m range, n range) = " +
                  std::to_string(need) + " doubles exceeds the scratch of " +
                  std::to_string(NOALLOC_SCRATCH_DOUBLES));
  double* scratch = nullptr;
  MGW_TRY(noalloc_scratch_ptr(c->device, &scratch));
  return ao2mo(c, mstate_min, mstate_max, nstate_min, nstate_max, eri_3center_mo_dev, MGW_FP64, -1, scratch);
}

int mgw_destroy_eri_3center_mo(mgw_ctx* c) {
  MGW_CTX(c);
  MGW_TRY(release(c, "eri_3center_mo"));
  c->has_eri_mo = false;
  return MGW_OK;
}

int mgw_npole(mgw_ctx* c, const mgw_ranges* r, int* npole) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  std::vector<int32_t> ia;
  transition_table(c, r, 2.0, ia);
  if (npole) *npole = (int)ia.size() / 2;
  return MGW_OK;
}

int mgw_polarizability(mgw_ctx* c, const mgw_ranges* r, double spin_fact, int precision,
                       double* en_rpa, double* en_gw, double* pole_out, double* w_s_out) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(precision >= MGW_FP64 && precision <= MGW_FP16X3, MGW_ERR_ARG,
              "mgw_polarizability: unknown precision");
  MGW_TRY(polarizability_casida(c, r, spin_fact, en_rpa, en_gw, precision));
  if (pole_out) MGW_TRY(download(c, pole_out, find_buf(c, "wpol_pole"), c->npole));
  if (w_s_out) MGW_TRY(download(c, w_s_out, find_buf(c, "wpol_ws"), (size_t)c->naux * c->npole));
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_set_spectral_function(mgw_ctx* c, int npole, int nauxil, const double* pole, const double* w_s,
                              int where) {
  MGW_CTX(c);
  MGW_REQUIRE(pole && w_s && npole > 0 && nauxil > 0, MGW_ERR_ARG, "mgw_set_spectral_function: invalid arguments");
  MGW_REQUIRE(where == MGW_MEM_HOST || where == MGW_MEM_DEVICE, MGW_ERR_ARG, "invalid memory kind");
  MGW_REQUIRE(c->has_ri && nauxil == c->naux, MGW_ERR_ARG,
              "mgw_set_spectral_function: nauxil differs from the resident RI tensor");
  double *pd = nullptr, *wd = nullptr;
  MGW_TRY(dbuf(c, "wpol_pole", npole, &pd));
  MGW_TRY(dbuf(c, "wpol_ws", (size_t)nauxil * npole, &wd));
  const hipMemcpyKind k = where == MGW_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  MGW_HIP(hipMemcpyAsync(pd, pole, sizeof(double) * npole, k, c->stream));
  MGW_HIP(hipMemcpyAsync(wd, w_s, sizeof(double) * nauxil * (size_t)npole, k, c->stream));
  MGW_HIP(hipStreamSynchronize(c->stream));
  c->npole = npole;
  c->has_wpol = true;
  c->has_sigma = false;
  return MGW_OK;
}

int mgw_gw_ingredients(mgw_ctx* c, const mgw_ranges* r, double* w_out, double* v_out) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(!w_out || c->has_wpol, MGW_ERR_STATE, "no spectral function: call mgw_polarizability first");
  const int glo = r->ncore_G + 1, ghi = r->nvirtual_G - 1;
  const int ng = ghi - glo + 1, naux = c->naux;
  MGW_REQUIRE(ng > 0, MGW_ERR_ARG, "mgw_gw_ingredients: empty state range");
  double* v = nullptr;
  MGW_TRY(dbuf(c, "ingr_v", (size_t)naux * ng * ng, &v));
  MGW_TRY(ao2mo(c, glo, ghi, glo, ghi, v));
  if (w_out) {
    // w(s, p, q) = sum_P w_s(P, s) (P|p q)  (src/m_gw_selfenergy_analytic.f90:851-857)
    const int npole = c->npole;
    double* w = nullptr;
    MGW_TRY(dbuf(c, "ingr_w", (size_t)npole * ng * ng, &w));
    GemmF64 g;
    g.M = npole; g.N = ng * ng; g.K = naux;
    g.transA = true; g.A = find_buf(c, "wpol_ws"); g.lda = naux;
    g.B = v; g.ldb = naux;
    g.C = w; g.ldc = npole;
    MGW_REQUIRE((int64_t)ng * ng < (1ll << 31), MGW_ERR_ARG, "mgw_gw_ingredients: state range too large");
    MGW_TRY(gemm_f64(g, c->stream));
    MGW_TRY(download(c, w_out, w, (size_t)npole * ng * ng));
  }
  if (v_out) MGW_TRY(download(c, v_out, v, (size_t)naux * ng * ng));
  MGW_HIP(hipStreamSynchronize(c->stream));
  MGW_TRY(release(c, "ingr_w"));
  MGW_TRY(release(c, "ingr_v"));
  return MGW_OK;
}

int mgw_polarizability_grid(mgw_ctx* c, const mgw_ranges* r, double spin_fact, int precision,
                            int nomega, const double* omega_im, const double* weight_quad,
                            double* erpa, double* egw, double* chi_out) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(precision >= MGW_FP64 && precision <= MGW_FP16X3, MGW_ERR_ARG,
              "mgw_polarizability_grid: unknown precision");
  MGW_REQUIRE(omega_im && weight_quad, MGW_ERR_ARG, "null frequency grid");
  MGW_TRY(polarizability_grid(c, r, spin_fact, nomega, omega_im, weight_quad, erpa, egw, precision));
  if (chi_out) {
    // chi = (L L^T)^-1 - 1 = L^-T L^-1 - 1, formed only for callers that ask
    // for it, from L^-1 = D + 1 (D is restored afterwards)
    const int n = c->naux;
    double* linv = find_buf(c, "wpol_linv");
    double *full, *mones;
    MGW_TRY(dbuf(c, "chi_full", (size_t)n * n * nomega, &full));
    MGW_TRY(dbuf(c, "minus_ones", n, &mones));
    MGW_TRY(launch_fill(mones, n, -1.0, c->stream));
    MGW_TRY(launch_diag_add(linv, n, nomega, 1.0, c->stream));
    GemmF64 g;
    g.M = n; g.N = n; g.K = n;
    g.A = linv; g.lda = n; g.strideA = (int64_t)n * n; g.transA = true;
    g.B = linv; g.ldb = n; g.strideB = (int64_t)n * n;
    g.C = full; g.ldc = n; g.strideC = (int64_t)n * n;
    g.diag_add = mones;
    g.lower_only = true;
    g.batch = nomega;
    g.batch_outer = true;
    MGW_TRY(gemm_f64(g, c->stream));
    MGW_TRY(launch_diag_add(linv, n, nomega, -1.0, c->stream));
    MGW_TRY(download(c, chi_out, full, (size_t)n * n * nomega));
    MGW_TRY(release(c, "chi_full"));
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_gw_selfenergy_qs(mgw_ctx* c, const mgw_ranges* r, double spin_fact, double eta, const double* s_matrix,
                         double* selfenergy_mo, double* selfenergy_ao) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(c->has_wpol, MGW_ERR_STATE, "no spectral function: call mgw_polarizability first");
  MGW_REQUIRE(!selfenergy_ao || s_matrix, MGW_ERR_ARG, "mgw_gw_selfenergy_qs: the AO matrix needs s_matrix");
  const int naux = c->naux, npole = c->npole, nbf = c->nbf, ns = c->nstate;
  const int nse = r->nsemax - r->nsemin + 1;
  const int glo = r->ncore_G + 1, ghi = r->nvirtual_G - 1;
  const int ng = ghi - glo + 1;
  double* epi = nullptr;
  MGW_TRY(dbuf(c, "eri_pi", (size_t)naux * nse * ng, &epi));
  MGW_TRY(ao2mo(c, r->nsemin, r->nsemax, glo, ghi, epi));
  const double* eg = c->energy_in_g().data() + c->so();
  const double* oc = c->occupation.data() + c->so();
  std::vector<double> ei(ng), fi(ng), eq(nse);
  for (int k = 0; k < ng; ++k) { ei[k] = eg[glo - 1 + k]; fi[k] = oc[glo - 1 + k]; }
  for (int k = 0; k < nse; ++k) eq[k] = eg[r->nsemin - 1 + k];
  double *ei_d, *fi_d, *eq_d, *sig;
  MGW_TRY(upload(c, "qs_ei", ei.data(), ng, &ei_d));
  MGW_TRY(upload(c, "qs_fi", fi.data(), ng, &fi_d));
  MGW_TRY(upload(c, "qs_eq", eq.data(), nse, &eq_d));
  MGW_TRY(dbuf(c, "qs_sigma", (size_t)nse * nse, &sig));
  // chunk over i so that w(s, p, i) and its scaled copy stay within ~1 GiB each
  const size_t per_i = (size_t)npole * nse * sizeof(double);
  const int ic = (int)std::max<size_t>(1, std::min<size_t>((size_t)ng, (size_t(1) << 30) / per_i));
  double *w, *b;
  MGW_TRY(dbuf(c, "sigma_w", (size_t)npole * nse * ic, &w));
  MGW_TRY(dbuf(c, "qs_b", (size_t)npole * nse * ic, &b));
  const double* ws = find_buf(c, "wpol_ws");
  const double* pole = find_buf(c, "wpol_pole");
  for (int i0 = 0; i0 < ng; i0 += ic) {
    const int cnt = std::min(ic, ng - i0);
    // w(s, p, i) = sum_P w_s(P, s) (P | p i)   (src/m_gw_selfenergy_analytic.f90:711-715)
    GemmF64 g;
    g.M = npole; g.N = nse * cnt; g.K = naux;
    g.transA = true; g.A = ws; g.lda = naux;
    g.B = epi + (int64_t)naux * nse * i0; g.ldb = naux;
    g.C = w; g.ldc = npole;
    g.klass = MGW_KCLASS_GEMM_SIGMA;
    MGW_TRY(gemm_f64(g, c->stream));
    MGW_TRY(launch_qs_scale(w, b, npole, nse, cnt, pole, eq_d, ei_d + i0, fi_d + i0, spin_fact, eta, c->stream));
    // Sigma(p, q) += sum_s w(s, p, i) b(s, q, i)   (:736-744)
    for (int k = 0; k < cnt; ++k) {
      GemmF64 h;
      h.M = nse; h.N = nse; h.K = npole;
      h.transA = true; h.A = w + (int64_t)npole * nse * k; h.lda = npole;
      h.B = b + (int64_t)npole * nse * k; h.ldb = npole;
      h.C = sig; h.ldc = nse;
      h.beta = (i0 + k == 0) ? 0.0 : 1.0;
      h.klass = MGW_KCLASS_GEMM_SIGMA;
      MGW_TRY(gemm_f64(h, c->stream));
    }
  }
  MGW_TRY(launch_symmetrize(sig, nse, c->stream));
  if (selfenergy_mo) {
    std::vector<double> h((size_t)nse * nse);
    MGW_TRY(download(c, h.data(), sig, h.size()));
    std::fill(selfenergy_mo, selfenergy_mo + (size_t)ns * ns, 0.0);
    for (int q = 0; q < nse; ++q)
      for (int p = 0; p < nse; ++p)
        selfenergy_mo[(r->nsemin - 1 + p) + (size_t)ns * (r->nsemin - 1 + q)] = h[p + (size_t)nse * q];
  }
  if (selfenergy_ao) {
    // h_mo_to_ao (src/m_selfenergy_tools.f90:826): S C Sigma C^T S over the window's MOs
    double *sd, *t1, *t2;
    MGW_TRY(upload(c, "qs_s", s_matrix, (size_t)nbf * nbf, &sd));
    MGW_TRY(dbuf(c, "qs_t1", (size_t)nbf * nse, &t1));
    MGW_TRY(dbuf(c, "qs_t2", (size_t)2 * nbf * nbf, &t2));
    const double* cw = find_buf(c, "c_matrix") + (int64_t)c->spin * nbf * ns + (int64_t)(r->nsemin - 1) * nbf;
    GemmF64 g1;   // T1 = C_w Sigma
    g1.M = nbf; g1.N = nse; g1.K = nse;
    g1.A = cw; g1.lda = nbf; g1.B = sig; g1.ldb = nse; g1.C = t1; g1.ldc = nbf;
    MGW_TRY(gemm_f64(g1, c->stream));
    GemmF64 g2;   // T2 = T1 C_w^T
    g2.M = nbf; g2.N = nbf; g2.K = nse;
    g2.A = t1; g2.lda = nbf; g2.B = cw; g2.ldb = nbf; g2.transB = true; g2.C = t2; g2.ldc = nbf;
    MGW_TRY(gemm_f64(g2, c->stream));
    double* t3 = t2 + (size_t)nbf * nbf;
    GemmF64 g3;   // T3 = S T2
    g3.M = nbf; g3.N = nbf; g3.K = nbf;
    g3.A = sd; g3.lda = nbf; g3.B = t2; g3.ldb = nbf; g3.C = t3; g3.ldc = nbf;
    MGW_TRY(gemm_f64(g3, c->stream));
    GemmF64 g4;   // Sigma_AO = T3 S
    g4.M = nbf; g4.N = nbf; g4.K = nbf;
    g4.A = t3; g4.lda = nbf; g4.B = sd; g4.ldb = nbf; g4.C = t2; g4.ldc = nbf;
    MGW_TRY(gemm_f64(g4, c->stream));
    MGW_TRY(download(c, selfenergy_ao, t2, (size_t)nbf * nbf));
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  MGW_TRY(scratch_done(c, "sigma_w"));
  MGW_TRY(scratch_done(c, "qs_b"));
  MGW_TRY(scratch_done(c, "eri_pi"));
  return MGW_OK;
}

int mgw_gw_selfenergy(mgw_ctx* c, const mgw_ranges* r, double spin_fact, double eta, int nomega,
                      const double* omega, const double* se_energy0, double* sigma_out,
                      double* energy_gw) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(omega && se_energy0 && nomega >= 1, MGW_ERR_ARG, "invalid frequency grid");
  MGW_TRY(gw_selfenergy(c, r, spin_fact, eta, nomega, omega, se_energy0, energy_gw, MGW_FP64));
  const int nse = r->nsemax - r->nsemin + 1;
  if (sigma_out) MGW_TRY(download(c, sigma_out, find_buf(c, "sigma"), (size_t)2 * nomega * nse));
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_gw_selfenergy_imag(mgw_ctx* c, const mgw_ranges* r, int nomega_calc, const double* omega_calc,
                           double* sigma_calc_out) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(omega_calc && nomega_calc >= 1 && nomega_calc <= 65535, MGW_ERR_ARG, "invalid omega_calc");
  MGW_TRY(gw_selfenergy_imag(c, r, nomega_calc, omega_calc, MGW_FP64));
  const int nse = r->nsemax - r->nsemin + 1;
  if (sigma_calc_out) MGW_TRY(download(c, sigma_calc_out, find_buf(c, "sigma_calc"), (size_t)2 * nomega_calc * nse));
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_gw_selfenergy_contour(mgw_ctx* c, const mgw_ranges* r, double spin_fact, int nomega_chi_imag,
                              int nomega_chi_real, double eta, int nomega, const double* omega,
                              const double* se_energy0, double* sigma_out, double* de_max_out) {
  MGW_CTX(c);
  MGW_TRY(check_ranges(c, r));
  MGW_REQUIRE(omega && se_energy0 && nomega >= 1, MGW_ERR_ARG, "invalid frequency grid");
  const int nse = r->nsemax - r->nsemin + 1;
  const double de_max = contour_de_max(c, r, eta, nomega, omega, se_energy0);
  MGW_TRY(contour_w(c, r, spin_fact, nomega_chi_imag, nomega_chi_real, de_max));
  // every channel, the active one last so that its sigma stays resident
  const int active = c->spin;
  for (int k = 0; k < c->nspin; ++k) {
    const int s = k == c->nspin - 1 ? active : (k < active ? k : k + 1);
    c->spin = s;
    const int rc = contour_sigma(c, r, eta, nomega, omega, se_energy0 + (int64_t)nse * s);
    if (rc == MGW_OK && sigma_out)
      MGW_TRY(download(c, sigma_out + (int64_t)2 * nomega * nse * s, find_buf(c, "sigma"), (size_t)2 * nomega * nse));
    c->spin = active;
    MGW_TRY(rc);
  }
  if (de_max_out) *de_max_out = de_max;
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_pade_fit(mgw_ctx* c, int nse, int nomega_calc, const double* omega_calc, const double* sigma_calc,
                 int nomega, const double* omega, double eta, double* sigma_out) {
  MGW_CTX(c);
  MGW_REQUIRE(nse >= 1 && nomega_calc >= 1 && nomega >= 1 && omega_calc && omega, MGW_ERR_ARG,
              "mgw_pade_fit: invalid arguments");
  MGW_TRY(pade_fit(c, nse, nomega_calc, omega_calc, sigma_calc, nomega, omega, eta));
  if (sigma_out) MGW_TRY(download(c, sigma_out, find_buf(c, "sigma"), (size_t)2 * nomega * nse));
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

int mgw_find_qp_energy(mgw_ctx* c, const mgw_ranges* r, int nomega, const double* omega,
                       const double* sigma, const double* se_energy0, const double* exchange_m_vxc,
                       const double* energy0, int nstate, double* energy_qp, double* zz,
                       double* energy_qp_lin) {
  MGW_CTX(c);
  MGW_REQUIRE(r && omega && se_energy0 && energy0 && nomega >= 1 && nstate >= 1, MGW_ERR_ARG,
              "mgw_find_qp_energy: invalid arguments");
  return find_qp(c, r, nomega, omega, sigma, se_energy0, exchange_m_vxc, energy0, nstate, energy_qp, zz,
                 energy_qp_lin);
}

int mgw_g0w0(mgw_ctx* c, const mgw_g0w0_params* prm, const mgw_ranges* r, const double* exc,
             double* energy_qp, double* zz, double* energy_qp_lin, double* sigma_out, double* energies) {
  MGW_CTX(c);
  MGW_REQUIRE(prm, MGW_ERR_ARG, "null parameters");
  MGW_TRY(check_ranges(c, r));
  const int prec = prm->precision;
  MGW_REQUIRE(prec >= MGW_FP64 && prec <= MGW_FP16X3, MGW_ERR_ARG, "mgw_g0w0: unknown precision");
  MGW_REQUIRE(prm->nomega_sigma >= 1, MGW_ERR_ARG, "nomega_sigma must be positive");
  const int nstate = c->nstate, nspin = c->nspin;
  const int nse = r->nsemax - r->nsemin + 1;
  const int nom = prm->nomega_sigma / 2;
  const int nomega = 2 * nom + 1;
  std::vector<double> omega(nomega), se0(nse);
  double en_rpa = 0.0, en_gw = 0.0;
  struct SpinGuard {     // the active channel is restored on every return
    mgw_ctx* c;
    int s;
    ~SpinGuard() { c->spin = s; }
  } guard{c, c->spin};
  // W from the transitions of every channel, then Sigma and the QP equation
  // channel by channel (the reference's ispin loops); outputs are spin-major
  auto per_spin = [&](int s) {
    const int64_t o = (int64_t)s * nstate;
    return find_qp(c, r, nomega, omega.data(), nullptr, se0.data(), exc ? exc + o : nullptr,
                   c->energy.data() + o, nstate, energy_qp ? energy_qp + o : nullptr, zz ? zz + o : nullptr,
                   energy_qp_lin ? energy_qp_lin + o : nullptr);
  };
  auto sigma_copy = [&](int s) -> int {
    if (sigma_out)
      MGW_TRY(download(c, sigma_out + (int64_t)s * 2 * nomega * nse, find_buf(c, "sigma"), (size_t)2 * nomega * nse));
    return MGW_OK;
  };
  if (prm->technique == MGW_TECH_ONE_SHOT) {
    // se_init(one_shot): grid centred on each energy0 (src/m_selfenergy_tools.f90:682-690)
    for (int k = 0; k < nomega; ++k) omega[k] = prm->step_sigma * (k - nom);
    MGW_TRY(polarizability_casida(c, r, prm->spin_fact, &en_rpa, nullptr, prec));
    for (int s = 0; s < nspin; ++s) {
      c->spin = s;
      for (int k = 0; k < nse; ++k) se0[k] = c->energy_in_g()[(size_t)s * nstate + r->nsemin - 1 + k];
      double eg = 0.0;
      MGW_TRY(gw_selfenergy(c, r, prm->spin_fact, prm->eta, nomega, omega.data(), se0.data(), &eg, prec));
      en_gw += eg;
      MGW_TRY(per_spin(s));
      MGW_TRY(sigma_copy(s));
    }
  } else if (prm->technique == MGW_TECH_IMAGINARY_PADE) {
    MGW_REQUIRE(prm->nomega_chi_imag >= 1 && prm->nomega_sigma_calc >= 1, MGW_ERR_ARG,
                "imaginary-axis technique needs nomega_chi_imag and nomega_sigma_calc");
    MGW_REQUIRE(r->nhomo_G >= 1 && r->nhomo_G < nstate, MGW_ERR_ARG, "nhomo_G out of range");
    // se_init(imaginary_axis_pade) (src/m_selfenergy_tools.f90:586-614):
    // midway between the highest HOMO and the lowest LUMO over the channels
    double ehomo = -1.0e300, elumo = 1.0e300;
    for (int s = 0; s < nspin; ++s) {
      ehomo = std::max(ehomo, c->energy_in_g()[(size_t)s * nstate + r->nhomo_G - 1]);
      elumo = std::min(elumo, c->energy_in_g()[(size_t)s * nstate + r->nhomo_G]);
    }
    const double efermi = 0.5 * (ehomo + elumo);
    for (int k = 0; k < nomega; ++k) omega[k] = efermi + prm->step_sigma * (k - nom);
    std::fill(se0.begin(), se0.end(), 0.0);
    std::vector<double> wc(2 * prm->nomega_sigma_calc);
    for (int k = 0; k < prm->nomega_sigma_calc; ++k) {
      wc[2 * k] = efermi;
      wc[2 * k + 1] = prm->step_sigma_calc * k;
    }
    std::vector<double> om, wt;
    imaginary_grid(prm->nomega_chi_imag, om, wt);
    MGW_TRY(polarizability_grid(c, r, prm->spin_fact, prm->nomega_chi_imag, om.data(), wt.data(), &en_rpa, &en_gw,
                                prec));
    for (int s = 0; s < nspin; ++s) {
      c->spin = s;
      MGW_TRY(gw_selfenergy_imag(c, r, prm->nomega_sigma_calc, wc.data(), prec));
      MGW_TRY(pade_fit(c, nse, prm->nomega_sigma_calc, wc.data(), nullptr, nomega, omega.data(), prm->eta));
      MGW_TRY(per_spin(s));
      MGW_TRY(sigma_copy(s));
    }
  } else if (prm->technique == MGW_TECH_CONTOUR) {
    MGW_REQUIRE(prec == MGW_FP64, MGW_ERR_ARG, "the contour deformation technique runs in FP64");
    // se_init(contour_deformation): the one-shot grid around energy0
    // (src/m_selfenergy_tools.f90:658-665, :679-680)
    for (int k = 0; k < nomega; ++k) omega[k] = prm->step_sigma * (k - nom);
    std::vector<double> se0_all((size_t)nse * nspin);
    for (int s = 0; s < nspin; ++s)
      for (int k = 0; k < nse; ++k) se0_all[k + (size_t)nse * s] = c->energy_in_g()[(size_t)s * nstate + r->nsemin - 1 + k];
    const double de_max = contour_de_max(c, r, prm->eta, nomega, omega.data(), se0_all.data());
    MGW_TRY(contour_w(c, r, prm->spin_fact, prm->nomega_chi_imag, prm->nomega_chi_real, de_max));
    for (int s = 0; s < nspin; ++s) {
      c->spin = s;
      std::copy(se0_all.begin() + (size_t)nse * s, se0_all.begin() + (size_t)nse * (s + 1), se0.begin());
      MGW_TRY(contour_sigma(c, r, prm->eta, nomega, omega.data(), se0.data()));
      MGW_TRY(per_spin(s));
      MGW_TRY(sigma_copy(s));
    }
  } else {
    set_error("unknown technique");
    return MGW_ERR_ARG;
  }
  if (energies) {
    energies[0] = en_rpa;
    energies[1] = en_gw;
  }
  MGW_HIP(hipStreamSynchronize(c->stream));
  return MGW_OK;
}

}  // extern "C"
