// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), gfx950.
//
// One kernel serves every dense contraction of the RI G0W0 path:
//   AO->MO half transforms ((P|mu nu) . C), the Casida SYRK (B^T B scaled by
//   sqrt(de)), w_s = B (X+Y), the self-energy products w_s^T (P|p i), the
//   imaginary-axis chi0 = X diag(s(iw)) X^T batched over frequencies and
//   chi(iw) (P|p m) with a fused column-dot epilogue.
//
// Tiling: a workgroup of WM x WN wave64s owns a BM x BN tile of C and walks K
// in slabs of BK = 16 through double-buffered LDS (one barrier per slab, the
// next slab's global loads in flight during the MFMAs of the current one).
// Each wave owns (BM/WM) x (BN/WN) and issues (BM/WM/16) * (BN/WN/16)
// independent MFMA chains per k-step of 4, which hides the MFMA latency.
//
// Fragment layout of v_mfma_f64_16x16x4_f64 (checked on gfx950): operand "a"
// lane l = a[i = l & 15][k = l >> 4], operand "b" lane l = b[k = l >> 4][j = l & 15],
// result lane l register r = d[i = (l >> 4) + 4 r][j = l & 15]. The kernel
// feeds a = B^T and b = A^T so that d[i][j] = C(m = j, n = i): the 16 lanes of
// a row group hold 16 consecutive rows m of one column of the column-major C
// and every store is a 128-byte contiguous segment.
#include "mgw_internal.h"
#include "tile_order.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace mgw {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int PAD = 1;
constexpr int GROUP_M = 8;     // tile rows per group of the grouped ordering

struct GemmDev {
  int M, N, K;
  const double* A; int64_t lda, strideA;
  const double* B; int64_t ldb, strideB;
  double* C; int64_t ldc, strideC;
  const double* kscale; int64_t strideK;
  const double* rowscale;
  const double* colscale;
  const double* diag_add; int64_t strideDiag;
  double alpha, beta;
  int lower_only;
  int no_mirror;       // lower_only: leave the strict upper triangle untouched
  int upper_zero;      // lower_only: write zeros above the diagonal instead of the mirror
  int tri_a;           // A (not transposed) is lower triangular: A(m, k) = 0 for k > m
  int batch_outer;     // walk the tiles of one batch entry before the next
  int tiles_m, tiles_n, batch;
  int64_t nblocks;
  const double* dotB; int64_t lddot;
  double dot_cc, dot_cd;
  double* dotpart; int64_t strideDot;
  const double* dot_w; int64_t strideDotW;
};

template <int BM, int BN, int WM, int WN, int BK, bool TA, bool TB, bool DOT, int KLASS>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_f64_kernel(GemmDev g) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM;
  constexpr int TN = BN / WN;
  constexpr int FM = TM / 16;
  constexpr int FN = TN / 16;
  constexpr int LA = BM * BK / NT;   // A elements loaded per thread per slab
  constexpr int LB = BN * BK / NT;
  static_assert(BM * BK % NT == 0 && BN * BK % NT == 0, "tile/thread mismatch");
  // contiguous-dimension strides of the per-thread element walk
  static_assert(TA ? (NT % BK == 0) : (NT % BM == 0), "A walk");
  static_assert(TB ? (NT % BN == 0) : (NT % BK == 0), "B walk");

  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM;
  const int wn = wave / WM;

  // logical block: batch fastest when the batch entries share operands (the
  // frequencies of one chi0 tile read the same panels), batch slowest when
  // they do not; then the tile in grouped order
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int64_t ntile = g.nblocks / g.batch;
  const int batch = (int)(g.batch_outer ? L / ntile : L % g.batch);
  const int tile = (int)(g.batch_outer ? L % ntile : L / g.batch);
  int bm, bn;
  if (g.lower_only && BM == BN) {
    tri_grouped(tile, g.tiles_m, bm, bn);
  } else {
    const int per_group = GROUP_M * g.tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(g.tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    bm = first_m + in_group % gsize;
    bn = in_group / gsize;
  }
  const int m0 = bm * BM;
  const int n0 = bn * BN;
  // symmetric rank-k update: tiles strictly above the diagonal have nothing to do
  if (g.lower_only && m0 + BM - 1 < n0) return;

  const double* __restrict__ A = g.A + (int64_t)batch * g.strideA;
  const double* __restrict__ B = g.B + (int64_t)batch * g.strideB;
  const double* __restrict__ ks = g.kscale ? g.kscale + (int64_t)batch * g.strideK : nullptr;

  // Per-thread walk over a slab: element e = tid + i * NT.
  //  A non-trans: m = tid % BM fixed, k = tid / BM + i * (NT / BM)
  //  A trans:     k = tid % BK fixed, m = tid / BK + i * (NT / BK)
  const int a_fix = TA ? tid % BK : tid % BM;
  const int a_var = TA ? tid / BK : tid / BM;
  constexpr int a_step = TA ? NT / BK : NT / BM;
  const int b_fix = TB ? tid % BN : tid % BK;
  const int b_var = TB ? tid / BN : tid / BK;
  constexpr int b_step = TB ? NT / BN : NT / BK;

  double ra[LA], rb[LB];

  auto load_slab = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int v = a_var + i * a_step;
      const int gm = m0 + (TA ? v : a_fix);
      const int gk = k0 + (TA ? a_fix : v);
      double x = 0.0;
      if (gm < g.M && gk < g.K && (!g.tri_a || gk <= gm))
        x = TA ? A[gk + (int64_t)gm * g.lda] : A[gm + (int64_t)gk * g.lda];
      ra[i] = x;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      const int gn = n0 + (TB ? b_fix : v);
      const int gk = k0 + (TB ? v : b_fix);
      double x = 0.0;
      if (gn < g.N && gk < g.K) {
        x = TB ? B[gn + (int64_t)gk * g.ldb] : B[gk + (int64_t)gn * g.ldb];
        if (ks) x *= ks[gk];
      }
      rb[i] = x;
    }
  };
  auto store_slab = [&](int buf) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int v = a_var + i * a_step;
      if (TA) As[buf][a_fix][v] = ra[i]; else As[buf][v][a_fix] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      if (TB) Bs[buf][v][b_fix] = rb[i]; else Bs[buf][b_fix][v] = rb[i];
    }
  };

  d4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  // a lower-triangular A has no terms beyond the tile's last row
  const int kend = g.tri_a ? min(g.K, m0 + BM) : g.K;
  const int nslab = (kend + BK - 1) / BK;
  load_slab(0);
  store_slab(0);
  __syncthreads();

  const int fr = lane & 15;
  const int fk = lane >> 4;
  for (int s = 0; s < nslab; ++s) {
    const int buf = s & 1;
    if (s + 1 < nslab) load_slab((s + 1) * BK);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      double af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[buf][k4 * 4 + fk][wm * TM + i * 16 + fr];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][k4 * 4 + fk][wn * TN + j * 16 + fr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bf[j], af[i], acc[i][j], 0, 0, 0);
    }
    if (s + 1 < nslab) store_slab(buf ^ 1);
    __syncthreads();
  }

  // ------------------------------------------------------------ epilogue
  double* __restrict__ C = g.C ? g.C + (int64_t)batch * g.strideC : nullptr;
  const double* __restrict__ dg = g.diag_add ? g.diag_add + (int64_t)batch * g.strideDiag : nullptr;

  if constexpr (DOT) {
    // Column dots of this row block: sum_m C(m,n) (cc C(m,n) + cd D(m,n)),
    // or sum_m w(m) C(m,n)^2 without D (w = 1 without row weights).
    __shared__ double red[WM][BN];
    const double* __restrict__ D = g.dotB ? g.dotB + (int64_t)batch * g.strideB : nullptr;  // batched as B
    const double* __restrict__ W = g.dot_w ? g.dot_w + (int64_t)batch * g.strideDotW : nullptr;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int nl = wn * TN + j * 16 + fk + 4 * r;
        const int n = n0 + nl;
        double part = 0.0;
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int m = m0 + wm * TM + i * 16 + fr;
          if (m < g.M && n < g.N) {
            const double c = g.alpha * acc[i][j][r];
            part += D ? c * (g.dot_cc * c + g.dot_cd * D[m + (int64_t)n * g.lddot]) : (W ? W[m] : 1.0) * c * c;
          }
        }
        part += __shfl_xor(part, 8, 16);
        part += __shfl_xor(part, 4, 16);
        part += __shfl_xor(part, 2, 16);
        part += __shfl_xor(part, 1, 16);
        if (fr == 0) red[wm][nl] = part;
      }
    }
    __syncthreads();
    double* __restrict__ out = g.dotpart + (int64_t)batch * g.strideDot;
    for (int nl = tid; nl < BN; nl += NT) {
      const int n = n0 + nl;
      if (n < g.N) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < WM; ++w) sum += red[w][nl];
        out[(int64_t)bm * g.N + n] = sum;
      }
    }
    return;
  }

#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + fk + 4 * r;
        if (m >= g.M || n >= g.N) continue;
        if (g.lower_only && m < n) continue;
        double v = g.alpha * acc[i][j][r];
        if (g.rowscale) v *= g.rowscale[m];
        if (g.colscale) v *= g.colscale[n];
        if (dg && m == n) v += dg[m];
        const int64_t idx = m + (int64_t)n * g.ldc;
        if (g.beta != 0.0) v += g.beta * C[idx];
        C[idx] = v;
        if (g.lower_only && m > n && !g.no_mirror) C[n + (int64_t)m * g.ldc] = g.upper_zero ? 0.0 : v;
      }
    }
  }
}

template <int BM, int BN, int WM, int WN, int BK, bool DOT, int KLASS>
int launch_cfg(const GemmF64& h, hipStream_t stream) {
  GemmDev g;
  g.M = h.M; g.N = h.N; g.K = h.K;
  g.A = h.A; g.lda = h.lda; g.strideA = h.strideA;
  g.B = h.B; g.ldb = h.ldb; g.strideB = h.strideB;
  g.C = h.C; g.ldc = h.ldc; g.strideC = h.strideC;
  g.kscale = h.kscale; g.strideK = h.strideK;
  g.rowscale = h.rowscale; g.colscale = h.colscale;
  g.diag_add = h.diag_add; g.strideDiag = h.strideDiag;
  g.alpha = h.alpha; g.beta = h.beta;
  g.lower_only = h.lower_only ? 1 : 0;
  g.no_mirror = h.no_mirror ? 1 : 0;
  g.upper_zero = h.upper_zero ? 1 : 0;
  g.tri_a = h.tri_a ? 1 : 0;
  g.batch_outer = h.batch_outer ? 1 : 0;
  g.tiles_m = (h.M + BM - 1) / BM;
  g.tiles_n = (h.N + BN - 1) / BN;
  g.batch = h.batch;
  g.dotB = h.dotB; g.lddot = h.lddot; g.dotpart = h.dotpart; g.strideDot = h.strideDot;
  g.dot_cc = h.dot_cc; g.dot_cd = h.dot_cd;
  g.dot_w = h.dot_w; g.strideDotW = h.strideDotW;
  // square-tile symmetric updates enumerate the lower triangle of tiles only;
  // other lower-only shapes launch the full grid and skip upper tiles
  const int64_t tiles = (h.lower_only && BM == BN) ? (int64_t)g.tiles_m * (g.tiles_m + 1) / 2
                                                   : (int64_t)g.tiles_m * g.tiles_n;
  g.nblocks = tiles * h.batch;
  MGW_REQUIRE(tiles < (1ll << 31) && g.nblocks < (1ll << 31), MGW_ERR_ARG, "gemm_f64: too many tiles");
  MGW_REQUIRE(h.batch >= 1, MGW_ERR_ARG, "gemm_f64: batch out of range");
  dim3 grid((unsigned)g.nblocks, 1, 1);
  dim3 block(64 * WM * WN, 1, 1);
  const int ta = h.transA ? 1 : 0, tb = h.transB ? 1 : 0;
#define MGW_GEMM_CASE(TA_, TB_)                                                                        \
  if (ta == TA_ && tb == TB_) {                                                                        \
    hipLaunchKernelGGL((gemm_f64_kernel<BM, BN, WM, WN, BK, TA_, TB_, DOT, KLASS>), grid, block, 0,    \
                       stream, g);                                                                     \
  }
  MGW_GEMM_CASE(0, 0)
  MGW_GEMM_CASE(0, 1)
  MGW_GEMM_CASE(1, 0)
  MGW_GEMM_CASE(1, 1)
#undef MGW_GEMM_CASE
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace

int gemm_f64_rowblocks(int M) { return (M + 127) / 128; }   // row tile of the dot epilogue

// ------------------------------------------------------------------ autotuner
// Candidate tile configurations (all FP64 MFMA, BK = 16):
//   0: 256x128, 16 waves   1: 128x128, 8 waves   2: 128x128, 4 waves   3: 64x64, 4 waves
// The defaults were measured on MI355X (0 for large products, 1 for the
// long-K symmetric update). With tuning on, the first product of a new
// (class, transposes, lower, log2 M/N/K bucket) times every candidate on
// the actual operands and keeps the fastest; the table can be persisted.
namespace {

constexpr int NCFG = 4;

struct TuneKey {
  int klass, ta, tb, lower, mb, nb, kb;
  bool operator<(const TuneKey& o) const {
    return std::tie(klass, ta, tb, lower, mb, nb, kb) < std::tie(o.klass, o.ta, o.tb, o.lower, o.mb, o.nb, o.kb);
  }
};

struct Tuner {
  std::mutex mu;
  std::map<TuneKey, int> table;
  int mode = 0;           // 0 off, 1 tune unseen shapes
  std::string file;       // persisted table (appended as entries are found)
  bool loaded = false;
};

Tuner& tuner() {
  static Tuner t;
  return t;
}

int log2_bucket(int64_t v) {
  int b = 0;
  while (v > 1) { v >>= 1; ++b; }
  return b;
}

void load_table(Tuner& t) {
  if (t.loaded) return;
  t.loaded = true;
  const char* f = std::getenv("MGW_TUNE_FILE");
  if (!f || !*f) return;
  t.file = f;
  FILE* fp = std::fopen(f, "r");
  if (!fp) return;
  TuneKey k;
  int cfg;
  while (std::fscanf(fp, "%d %d %d %d %d %d %d %d", &k.klass, &k.ta, &k.tb, &k.lower, &k.mb, &k.nb, &k.kb, &cfg) == 8)
    if (cfg >= 0 && cfg < 16) t.table[k] = cfg;   // each family checks its own candidate count
  std::fclose(fp);
}

template <int KLASS>
int launch_candidate(int cfg, const GemmF64& h, hipStream_t stream) {
  switch (cfg) {
    case 0: return launch_cfg<256, 128, 4, 4, 16, false, KLASS>(h, stream);
    case 1: return launch_cfg<128, 128, 4, 2, 16, false, KLASS>(h, stream);
    case 2: return launch_cfg<128, 128, 2, 2, 16, false, KLASS>(h, stream);
    default: return launch_cfg<64, 64, 2, 2, 16, false, KLASS>(h, stream);
  }
}

int launch_by_class(int cfg, const GemmF64& h, hipStream_t stream) {
  switch (h.klass) {
    case MGW_KCLASS_GEMM_AO2MO: return launch_candidate<MGW_KCLASS_GEMM_AO2MO>(cfg, h, stream);
    case MGW_KCLASS_GEMM_CASIDA: return launch_candidate<MGW_KCLASS_GEMM_CASIDA>(cfg, h, stream);
    case MGW_KCLASS_GEMM_SIGMA: return launch_candidate<MGW_KCLASS_GEMM_SIGMA>(cfg, h, stream);
    case MGW_KCLASS_GEMM_CHI0: return launch_candidate<MGW_KCLASS_GEMM_CHI0>(cfg, h, stream);
    case MGW_KCLASS_CHOLESKY: return launch_candidate<MGW_KCLASS_CHOLESKY>(cfg, h, stream);
    default: return launch_candidate<MGW_KCLASS_GEMM_OTHER>(cfg, h, stream);
  }
}

int tune(const GemmF64& h, hipStream_t stream, int* best) {
  hipEvent_t a, b;
  MGW_HIP(hipEventCreate(&a));
  MGW_HIP(hipEventCreate(&b));
  float best_ms = 1e30f;
  *best = -1;
  for (int cfg = 0; cfg < NCFG; ++cfg) {
    MGW_HIP(hipEventRecord(a, stream));
    MGW_TRY(launch_by_class(cfg, h, stream));
    MGW_HIP(hipEventRecord(b, stream));
    MGW_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    MGW_HIP(hipEventElapsedTime(&ms, a, b));
    if (ms < best_ms) { best_ms = ms; *best = cfg; }
  }
  MGW_HIP(hipEventDestroy(a));
  MGW_HIP(hipEventDestroy(b));
  return MGW_OK;
}

}  // namespace

int gemm_autotune_mode(int mode) {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> lk(t.mu);
  load_table(t);
  const int old = t.mode;
  if (mode >= 0) t.mode = mode;
  return old;
}

int autotune_select(int family, int ns, int h, int mb, int nb, int kb, int ncand, int fallback,
                    const std::function<int(int)>& launch, hipStream_t stream) {
  Tuner& t = tuner();
  const TuneKey key{family, ns, h, 0, mb, nb, kb};
  int choice = fallback;
  bool need_tune = false;
  {
    std::lock_guard<std::mutex> lk(t.mu);
    load_table(t);
    auto it = t.table.find(key);
    if (it != t.table.end() && it->second < ncand) choice = it->second;
    else need_tune = t.mode == 1;
  }
  if (!need_tune) return launch(choice);
  hipEvent_t a, b;
  MGW_HIP(hipEventCreate(&a));
  MGW_HIP(hipEventCreate(&b));
  // two rounds over the candidates, the best time of each (other streams of
  // the device may run meanwhile); another candidate replaces the fallback
  // only if it is at least 3 % faster
  std::vector<float> ms(ncand, 1e30f);
  auto time_all = [&]() -> int {
    for (int round = 0; round < 2; ++round)
      for (int cand = 0; cand < ncand; ++cand) {
        MGW_HIP(hipEventRecord(a, stream));
        MGW_TRY(launch(cand));
        MGW_HIP(hipEventRecord(b, stream));
        MGW_HIP(hipEventSynchronize(b));
        float e = 0.0f;
        MGW_HIP(hipEventElapsedTime(&e, a, b));
        ms[cand] = std::min(ms[cand], e);
      }
    return MGW_OK;
  };
  const int rc = time_all();
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  MGW_TRY(rc);
  int best = fallback;
  for (int cand = 0; cand < ncand; ++cand)
    if (ms[cand] < ms[best] && ms[cand] < 0.97f * ms[fallback]) best = cand;
  std::lock_guard<std::mutex> lk(t.mu);
  if (!t.table.emplace(key, best).second) return MGW_OK;   // another stream tuned it meanwhile
  if (!t.file.empty()) {
    if (FILE* fp = std::fopen(t.file.c_str(), "a")) {
      std::fprintf(fp, "%d %d %d %d %d %d %d %d\n", key.klass, key.ta, key.tb, key.lower, key.mb, key.nb, key.kb,
                   best);
      std::fclose(fp);
    }
  }
  return MGW_OK;   // every candidate wrote the same result
}

int gemm_f64(const GemmF64& h, hipStream_t stream) {
  if (h.M <= 0 || h.N <= 0 || h.batch <= 0) return MGW_OK;
  // MGW_BF16X6 and MGW_FP16X3 split only the chi0 and imaginary-axis
  // self-energy operands (chi0_split.hip, wb_split.hip); every other product
  // of those modes stays FP64
  MGW_REQUIRE(h.dot_w == nullptr || h.precision == MGW_FP64, MGW_ERR_ARG, "gemm_f64: row-weighted dots are FP64 only");
  if (h.precision == MGW_FP32 || h.precision == MGW_BF16X3) return gemm_lowp(h, h.precision, stream);
  // FP16x3 products need the operand maxima; without them the product stays FP64
  if (h.precision == MGW_FP16X3 && h.amax && h.bmax) return gemm_lowp(h, h.precision, stream);
  MGW_REQUIRE(h.K >= 0, MGW_ERR_ARG, "gemm_f64: negative K");
  MGW_REQUIRE(!h.lower_only || h.M == h.N, MGW_ERR_ARG, "gemm_f64: lower_only needs M == N");
  MGW_REQUIRE(h.dotpart == nullptr || !h.lower_only, MGW_ERR_ARG, "gemm_f64: dot epilogue needs a full tile grid");
  MGW_REQUIRE(!h.tri_a || (!h.transA && h.M == h.K), MGW_ERR_ARG,
              "gemm_f64: a triangular A must be square and not transposed");
  MGW_REQUIRE(h.dotpart != nullptr || h.C != nullptr, MGW_ERR_ARG, "gemm_f64: no output");
  if (h.dotpart) return launch_cfg<128, 128, 4, 2, 16, true, MGW_KCLASS_GEMM_WB>(h, stream);
  // Skinny N: a 256 x 32 tile keeps the matrix cores busy when N is small.
  if (h.N <= 32 && !h.lower_only) return launch_cfg<256, 32, 4, 1, 16, false, MGW_KCLASS_GEMM_AO2MO>(h, stream);
  // Defaults measured on MI355X: 256x128 with 16 waves for the large products
  // (57-62 TF/s), 128x128 with 8 waves for the long-K symmetric updates;
  // 64x64 tiles when the default would leave most of the 256 CUs idle (the
  // block columns of the batched Cholesky, small molecules), also when they
  // give as many workgroups (a product that fits one 64x64 tile: the larger
  // tiles only add idle waves; the tuner picks 64x64 for every such product
  // of the Cholesky, 4.26 -> 3.91 ms per solve on the FP16x3 stream).
  int cfg = h.lower_only ? 1 : 0;
  {
    auto wgs = [&](int bm, int bn) {
      const int64_t tm = (h.M + bm - 1) / bm, tn = (h.N + bn - 1) / bn;
      return (h.lower_only && bm == bn ? tm * (tm + 1) / 2 : tm * tn) * h.batch;
    };
    const int64_t have = h.lower_only ? wgs(128, 128) : wgs(256, 128);
    if (have < 512 && wgs(64, 64) >= have) cfg = 3;
  }
  Tuner& t = tuner();
  const TuneKey key{h.klass, h.transA ? 1 : 0, h.transB ? 1 : 0, h.lower_only ? 1 : 0,
                    log2_bucket(h.M), log2_bucket(h.N), log2_bucket((int64_t)h.K * h.batch)};
  bool need_tune = false;
  {
    std::lock_guard<std::mutex> lk(t.mu);
    load_table(t);
    auto it = t.table.find(key);
    if (it != t.table.end() && it->second < NCFG) cfg = it->second;
    else need_tune = t.mode == 1 && h.beta == 0.0 && !h.no_tune;   // candidates overwrite C: only pure products
  }
  if (need_tune) {
    int best = cfg;
    MGW_TRY(tune(h, stream, &best));
    std::lock_guard<std::mutex> lk(t.mu);
    t.table[key] = best;
    cfg = best;
    if (!t.file.empty()) {
      if (FILE* fp = std::fopen(t.file.c_str(), "a")) {
        std::fprintf(fp, "%d %d %d %d %d %d %d %d\n", key.klass, key.ta, key.tb, key.lower, key.mb, key.nb, key.kb,
                     best);
        std::fclose(fp);
      }
    }
  }
  return launch_by_class(cfg, h, stream);
}

}  // namespace mgw
