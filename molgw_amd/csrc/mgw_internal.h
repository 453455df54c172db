// Internal declarations of libmgw: context, device buffers, kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <map>
#include <functional>
#include <string>
#include <vector>

#include "../../include/mgw.h"

namespace mgw {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
const std::string& last_error();

struct Status {
  int code = MGW_OK;
  static Status ok() { return Status{}; }
};

#define MGW_HIP(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      ::mgw::set_error(std::string(#expr) + ": " + hipGetErrorString(e_) + " (" +  \
                       __FILE__ + ":" + std::to_string(__LINE__) + ")");           \
      return MGW_ERR_HIP;                                                          \
    }                                                                              \
  } while (0)

#define MGW_TRY(expr)             \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != MGW_OK) return rc_; \
  } while (0)

#define MGW_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      ::mgw::set_error(msg);         \
      return (code);                 \
    }                                \
  } while (0)

// ---------------------------------------------------------------- HBM ledger
// A named device buffer owned by a context. Buffers grow on demand and are
// reused across calls, so a solve loop does no hipMalloc after warm-up.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct Ledger {
  std::map<std::string, DevBuf> bufs;
  int64_t current = 0;
  int64_t peak = 0;
};

// ---------------------------------------------------------------- profiling
// Live per-kernel-class timing with HIP events recorded on the stream the
// kernels are launched on; classes are the MGW_KCLASS_* of mgw.h.
struct Profiler {
  bool on = false;
  struct Rec {
    int klass;
    double flops, bytes;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  hipEvent_t next_event();
  void reset();
  ~Profiler();
};

// RAII scope: records a start event now and an end event at destruction.
struct ProfScope {
  Profiler* p;
  size_t idx;
  hipStream_t s;
  ProfScope(Profiler* prof, int klass, double flops, double bytes, hipStream_t stream);
  ~ProfScope();
};

// ---------------------------------------------------------------- GEMM
// C(MxN) = alpha * opA(A) * diag(kscale) * opB(B) scaled by rowscale/colscale,
// plus diag_add on the diagonal, plus beta * C. Column-major; batch strides in
// elements. lower_only: M == N, only tiles on/below the diagonal are computed
// and mirrored to the upper triangle (symmetric rank-k update).
struct GemmF64 {
  int M = 0, N = 0, K = 0;
  bool transA = false, transB = false;
  const double* A = nullptr; int64_t lda = 0, strideA = 0;
  const double* B = nullptr; int64_t ldb = 0, strideB = 0;
  double* C = nullptr; int64_t ldc = 0, strideC = 0;
  const double* kscale = nullptr; int64_t strideK = 0;
  const double* rowscale = nullptr;
  const double* colscale = nullptr;
  const double* diag_add = nullptr; int64_t strideDiag = 0;
  double alpha = 1.0, beta = 0.0;
  int batch = 1;
  bool lower_only = false;
  bool no_mirror = false;     // lower_only: do not mirror into the upper triangle
  bool upper_zero = false;    // lower_only: write zeros, not the mirror, above the diagonal
  bool tri_a = false;         // A (not transposed, M == K) is lower triangular
  bool batch_outer = false;   // tiles of one batch entry before the next (no shared operands)
  // Fused column dot: instead of storing C, accumulate
  // dotpart[(rowblock * N + n) + batch * strideDot]
  //   = sum_{m in rowblock} C(m,n) (dot_cc C(m,n) + dot_cd D(m,n))
  // with D(m, n) = dotB[m + n * lddot] (batched with strideB), or
  // sum w(m) C(m,n)^2 when dotB is null, w = dot_w[m + batch * strideDotW]
  // (FP64 only) or 1.
  const double* dotB = nullptr; int64_t lddot = 0;
  double dot_cc = 0.0, dot_cd = 1.0;
  double* dotpart = nullptr; int64_t strideDot = 0;
  const double* dot_w = nullptr; int64_t strideDotW = 0;
  // kernel class: the hot classes get their own kernel symbol so that
  // profiler per-kernel statistics separate them
  int klass = MGW_KCLASS_GEMM_OTHER;
  // arithmetic of the product: MGW_FP64 (default), MGW_FP32 or MGW_BF16X3
  // (FP64 operands rounded on the way to LDS, FP64 accumulation); the
  // triangular and no-mirror options are FP64 only
  int precision = MGW_FP64;
  // use the tuned or default tiles without timing candidates (the timing
  // creates events and synchronises the stream)
  bool no_tune = false;
  // MGW_FP16X3 (gemm_lowp, plain non-transposed products): device scalars
  // max|A| and max|B| that set the power-of-two operand scales; cmax
  // (nullable, zeroed by the caller) receives max|C| of the stored results
  const double* amax = nullptr;
  const double* bmax = nullptr;
  double* cmax = nullptr;
  // MGW_FP16X3, batch 1: op(B)^T already split (fp16_split_planes: rows N,
  // k K, scale of bmax), read instead of splitting B on every tile load
  const _Float16* bplanes = nullptr;
  int64_t bplane_q = 0;   // elements of one plane
  int bplane_rpad = 0;    // padded rows of a plane
};
int gemm_f64(const GemmF64& g, hipStream_t stream);
int gemm_lowp(const GemmF64& g, int precision, hipStream_t stream);
// 1 - chi0(iw_f) = I + X diag(s_f) X^T (lower triangle, zeros above it) for all
// nomega frequencies from bf16-split operands (precision MGW_BF16X3 or
// MGW_BF16X6); needs s_f >= 0 and a workspace of chi0_split_workspace bytes.
size_t chi0_split_bytes_per_frequency(int naux, int npole, int precision, int nomega);
size_t chi0_split_workspace(int naux, int npole, int nomega, int precision);
int chi0_split_syrk(const double* x, int naux, int npole, const double* sw, int nomega, int precision,
                    void* work, size_t work_bytes, double* chi, hipStream_t stream);
int gemm_f64_rowblocks(int M);
// v(n, f) = B_n^T chi(iw_f) B_n = sum_m C (C + 2 B), C = D_f B, from bf16-split
// operands (MGW_BF16X3 / MGW_BF16X6) of D_f = L_f^-1 - 1 (lower triangle of
// naux x naux, stride naux^2) and B (naux x ncol); writes the row-block
// partials dotpart[(f * gemm_f64_rowblocks(naux) + rb) * ncol + n].
size_t wb_split_workspace(int naux, int64_t ncol, int nomega, int precision);
int wb_split_dot(const double* d, int naux, int nomega, const double* b, int64_t ncol, int precision, void* work,
                 size_t work_bytes, double* dotpart, hipStream_t stream);
// In-place inverse Cholesky factor of a batch of symmetric positive definite
// matrices: A_f (n x n, lower triangle, stride n^2, strict upper triangle
// zero on entry) becomes L_f^-1 (lower triangle; the upper stays zero). logdet[f] receives
// sum_i ln L_f(i, i); info[f] the 1-based column of the first non-positive
// pivot (0 on success). work: chol_work_doubles(n) doubles per matrix.
// bounds (nullable; 4 + nbatch doubles of device memory): when given, the matrices
// must satisfy A_f >= I (1 - chi0 with chi0 <= 0), and the products of the
// large levels of the recursion run in FP16x3 on operands split into fp16
// planes (planes: chol_plane_bytes(n, nbatch) of device memory) with
// operand bounds derived from max_i A_f(i, i), which a first kernel writes
// to bounds.
// FP16x3 operand planes (chol.hip): op(X) (rows x k) of each matrix of a
// batch split once into fp16 hi and lo of 2^e x (2^e from a device bound,
// so that the bound lands in [2^13, 2^14)), in the k-tiled layout
// plane[matrix][q][kt][row][32] (rows padded to 128, k to 32, zeros in the
// padding); trans = 0: op(X)(r, k) = X[r + k ld], trans = 1: X[k + r ld].
struct Fp16Planes {
  int rows, k;           // logical operand: rows x k
  int rpad, nkt;         // padded rows, k tiles
  int64_t per_q;         // elements of one plane of one matrix
  int64_t per_f;         // elements of both planes of one matrix
};
Fp16Planes fp16_planes(int rows, int k);
int fp16_split_planes(const double* X, int64_t ld, int64_t strideX, int trans, const Fp16Planes& d, int nbatch,
                      const double* bound, _Float16* dst, int klass, hipStream_t stream);
// C(m, n) = alpha sum_k A(m, k) B^T(n, k) + beta C(m, n) for each matrix of
// the batch from the planes of A (M rows) and B^T (N rows) with the same k:
// hi hi + hi lo + lo hi on fp16 MFMA, FP32 accumulation over the tile's K
// range, FP64 epilogue. a_f / b_f: plane elements between matrices (0 for
// an operand shared by the batch). lower: lower-triangle tiles of a square
// C, m < n not written; tri_a: A lower triangular (K stops at the tile's
// last row); kb_hi / kb_lo: B^T(n, k) zero for k > n / k < n. bm: tile, 128
// or 64. Tiles of one matrix before the next. klass (MGW_KCLASS_CHOLESKY or
// MGW_KCLASS_GEMM_AO2MO) picks the kernel symbol, so profiles separate the
// callers.
// out_planes (optional, beta = 0 only): the result written as the planes
// (layout out_dims, scale from the device bound out_bound) of op(C), C itself
// (out_trans = 0) or C^T (1), instead of C; the padding inside the tiles
// is written with zeros.
struct Fp16PlaneProduct {
  int M = 0, N = 0, nbatch = 1;
  int64_t a_f = 0, b_f = 0;
  double* C = nullptr; int64_t ldc = 0, strideC = 0;
  double alpha = 1.0, beta = 0.0;
  bool lower = false;
  int tri_a = 0, kb_hi = 0, kb_lo = 0;
  int bm = 128;
  _Float16* out_planes = nullptr;
  Fp16Planes out_dims = {};
  int out_trans = 0;
  const double* out_bound = nullptr;
};
int fp16_plane_gemm(const Fp16Planes& da, const _Float16* pa, const double* amax, const Fp16Planes& db,
                    const _Float16* pb, const double* bmax, const Fp16PlaneProduct& p, int klass,
                    hipStream_t stream);

constexpr int CHOL_NB = 64;
size_t chol_work_doubles(int n);
size_t chol_plane_bytes(int n, int nbatch);
int chol_f16_min();
bool chol_epi_planes();
int chol_inverse_batched(double* A, int n, int nbatch, double* logdet, int32_t* info, double* work,
                         hipStream_t stream, double* bounds = nullptr, void* planes = nullptr);
// tr[f] = sum of squares of the lower triangle of M_f (n x n, stride n^2).
int launch_lower_frob2(const double* m, int n, int nbatch, double* partial, hipStream_t s);
constexpr int FROB_BLOCKS = 64;
int gemm_autotune_mode(int mode); // -1 query; returns the previous mode
// The same tuner for the split-operand kernel families: launch(c) runs
// candidate c (0 <= c < ncand); every candidate must give the same result.
// Key: family, planes, fp16 flag and three log2 shape buckets. With tuning
// on, an unseen key times every candidate on the actual operands (the last
// run's output stands) and records the fastest; otherwise the recorded
// choice, or `fallback`, is launched.
int autotune_select(int family, int ns, int h, int mb, int nb, int kb, int ncand, int fallback,
                    const std::function<int(int)>& launch, hipStream_t stream);
constexpr int MGW_TUNE_CHI_SYRK = 105;   // K-split chi0 SYRK variants
constexpr int MGW_TUNE_WB_DOT = 107;     // W.B dot pipelines
inline int log2_floor(int64_t v) {
  int b = 0;
  while (v > 1) { v >>= 1; ++b; }
  return b;
}

// ---------------------------------------------------------------- integrals
// Device view of a set of contracted shells (struct of arrays): angular
// momentum, centre (3 per shell), primitive range [poff[s], poff[s+1]) into
// exps/coefs (weights of the unnormalised Cartesian primitives), the
// (nfunc x ncart) row-major map to the final functions at toff[s], and the
// first final function foff[s] (foff has nshell + 1 entries).
struct EriShells {
  const int32_t* l = nullptr;
  const double* center = nullptr;
  const int32_t* poff = nullptr;
  const double* exps = nullptr;
  const double* coefs = nullptr;
  const int32_t* toff = nullptr;
  const double* tmat = nullptr;
  const int32_t* foff = nullptr;
};
// One class (la, lb | lc) of Coulomb integrals: pairs (sa in A, sb in B) as
// int32 pairs, ket functions (sc in C, function f of sc) as int32 pairs.
// metric = 0: out(ipair, P) of packed pairs (pairmap: nbf x nbf pair index
// or -1, diagonal pairs x1/2, npair rows); metric = 1: out(P, Q) (naux x
// naux) with B the unit shells and pairs (s, s).
struct EriLaunch {
  int la = 0, lb = 0, lc = 0;
  EriShells A, B, C;
  const int32_t* pairs = nullptr;
  int64_t npairs = 0;
  const int32_t* kfun = nullptr;
  int nkfun = 0;
  const double* boys = nullptr;
  int metric = 0;
  const int32_t* pairmap = nullptr;
  int nbf = 0;
  int64_t npair = 0;
  int naux = 0;
  double* out = nullptr;
};
bool eri_class_supported(int la, int lb, int lc);
size_t boys_table_doubles();
void boys_table_fill(double* host);
int launch_eri_class(const EriLaunch& e, hipStream_t s);

// ---------------------------------------------------------------- kernels
// zero_dense false: only the pairs listed are written (a later batch of a
// tensor whose buffer was cleared before the first batch)
// amax (nullable, zeroed by the caller): atomically raised to max |dense| of
// the elements written
int launch_unpack_ri(const double* packed, int64_t npair, int naux, const int32_t* kl,
                     int nbf, double* dense, hipStream_t s, bool zero_dense = true, double* amax = nullptr);
// out(:, j) = u(:, j + neglect) / sqrt(eig[j + neglect]), j < kept (u n x n)
int launch_invsqrt_select(const double* u, int n, const double* eig, int neglect, int kept, double* out,
                          hipStream_t s);
int launch_fill(double* p, int64_t n, double v, hipStream_t s);
// b(s, q, i) = w(s, q, i) times the real part of the QS pole factor
int launch_qs_scale(const double* w, double* b, int npole, int nse, int cnt, const double* pole, const double* eq,
                    const double* ei, const double* fi, double spin_fact, double eta, hipStream_t s);
// m = (m + m^T) / 2 (n x n)
int launch_symmetrize(double* m, int n, hipStream_t s);
// part[b] = max |x_i| over the elements of block b, b < ABSMAX_BLOCKS
constexpr int ABSMAX_BLOCKS = 1024;
int launch_absmax(const double* x, int64_t n, double* part, hipStream_t s);
// out[0] = max |x_i| (part: ABSMAX_BLOCKS doubles of workspace)
int launch_absmax_scalar(const double* x, int64_t n, double* part, double* out, hipStream_t s);
// m_f(i, i) += v for the nbatch n x n matrices m_f (stride n^2)
int launch_diag_add(double* m, int n, int nbatch, double v, hipStream_t s);
int launch_gather_ia(const double* eri, int naux, int nm, int nn, const int32_t* ia_dev, int npole,
                     double* x, hipStream_t s);
int launch_colnorm2(const double* x, int64_t ld, int rows, int cols, double* out, hipStream_t s);
int launch_vsum(const double* x, int64_t n, int square, double* partial, int nblk, hipStream_t s);
int launch_casida_diag(const int32_t* ia, int npole, const double* energy, double* de, double* sq,
                       double* de2, hipStream_t s);
int launch_casida_post(const double* eig, int npole, double* pole, double* cs, hipStream_t s);
int sigma_part_doubles(int npole, int nse);
int launch_sigma_poles(const double* w, int npole, int nse, int ni, const double* pole,
                       const double* omega, int nomega, const double* e0p, const double* ei,
                       const double* fi, const double* ep, const double* occp, double spin_fact,
                       double eta, double* part, int accumulate, double* sigma, double* egw_p,
                       hipStream_t s);
// sw[t + npole f] = 2 docc de / (wsign omega_f^2 + de^2): wsign = 1 for
// imaginary frequencies i omega_f, -1 for real ones (then 1 - chi0 = 1 + X diag(sw) X^T
// on both axes)
int launch_chi_weights(const int32_t* ia, int npole, const double* energy, const double* occ,
                       const double* omega, int nomega, double* sw, hipStream_t s, double wsign = 1.0);
int launch_contour_eta(const double* lam, int64_t n, double tol, double* w, hipStream_t s);
int launch_sigma_contour(const double* v, int np, int nse, int nimag, int nreal, const double* om_im,
                         const double* wt_im, const double* om_re, const double* ep, const int32_t* pstate,
                         int nhomo, int nlumo, const double* e0, const double* omega, int nomega, double eta,
                         double* sigma, int32_t* err, hipStream_t s);
int launch_chi_trace(const double* sw, const double* n2, int npole, const double* mat, int naux,
                     int mode, int nomega, double* out, hipStream_t s);
int launch_dot_reduce(const double* part, int nrb, int64_t ncol, int nomega, const double* sub, double* v,
                      hipStream_t s);
int launch_sigma_imag_sum(const double* v, int np, int nse, int nomega, const double* omega,
                          const double* weight, const double* ep, const double* wc, int ncalc,
                          double* sigma_calc, hipStream_t s);
int pade_work_doubles(int ncalc, int nse);
int launch_pade(const double* wc, const double* sc, int ncalc, int nse, const double* omega,
                int nomega, double eta, double* work, double* sigma, hipStream_t s);
int launch_qp(const double* sigma, int nomega, int nse, const double* omega, const double* se_e0,
              const double* exc, const double* e0, double* eqp, double* zz, double* elin,
              hipStream_t s);

}  // namespace mgw
