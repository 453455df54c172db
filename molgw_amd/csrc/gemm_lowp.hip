// Reduced-precision products on the CDNA4 matrix cores with FP64 accumulation.
//
// Same contract as gemm_f64 (FP64 operands and results in HBM, the same
// fused scales, symmetric lower-triangle mode) for the contractions that may
// run below FP64: the AO->MO transforms and the chi0(iw) build.
//
//  * MGW_FP32   : operands rounded to FP32 on the way to LDS,
//                 v_mfma_f32_16x16x4_f32 (2x the FP64 MFMA rate).
//  * MGW_BF16X3 : each operand split x = hi + lo with hi = bf16(x),
//                 lo = bf16(x - hi); a*b ~ hi*hi + hi*lo + lo*hi on
//                 v_mfma_f32_16x16x32_bf16 (3 products, ~2^-16 relative per
//                 product; 10x the FP64 MFMA rate at equal work).
//
// Both accumulate each K slab in FP32 MFMA registers and flush into FP64
// accumulators every FLUSH slabs, so long sums (K up to ~10^5) keep FP64
// accumulation error. The optional K scale is applied in FP64 before the
// operand is rounded.
//
// Fragment layouts on gfx950 (checked on the box):
//  f32 16x16x4 : a[i = l & 15][k = l >> 4], b[k = l >> 4][j = l & 15],
//                d: lane l reg r -> d[4 (l >> 4) + r][l & 15]
//  bf16 16x16x32: a[i = l & 15][k = 8 (l >> 4) + t], b[k = 8 (l >> 4) + t][j = l & 15]
//                (t = 0..7 packed in one operand), d as for f32.
// As in gemm_f64 the kernels feed a = B^T, b = A^T, so d[i][j] = C(m = j, n = i)
// and lane l, register r holds C(m = l & 15, n = 4 (l >> 4) + r).
#include "mgw_internal.h"
#include "tile_order.h"

#include <cmath>

namespace mgw {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));


struct LowpDev {
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
  int upper_zero;      // lower_only: zeros above the diagonal instead of the mirror
  int tiles_m, tiles_n, batch;
  int64_t nblocks;
};

__device__ inline void tile_of(const LowpDev& g, int tile, int BMv, int BNv, int& bm, int& bn) {
  if (g.lower_only && BMv == BNv) {
    tri_grouped(tile, g.tiles_m, bm, bn);
  } else {
    constexpr int GROUP_M = 8;
    const int per_group = GROUP_M * g.tiles_n;
    const int group = tile / per_group;
    const int first_m = group * GROUP_M;
    const int gsize = min(g.tiles_m - first_m, GROUP_M);
    const int in_group = tile - group * per_group;
    bm = first_m + in_group % gsize;
    bn = in_group / gsize;
  }
}

// Shared epilogue: lane l, register r of tile (i, j) holds C(m, n) with
// m = m_base + (l & 15), n = n_base + 4 (l >> 4) + r.
template <int FM, int FN, int TM, int TN>
__device__ inline void store_tile(const LowpDev& g, double (&acc)[FM][FN][4], int batch, int m0, int n0,
                                  int wm, int wn, int lane) {
  double* __restrict__ C = g.C + (int64_t)batch * g.strideC;
  const double* __restrict__ dg = g.diag_add ? g.diag_add + (int64_t)batch * g.strideDiag : nullptr;
  const int fr = lane & 15, fq = lane >> 4;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
        if (m >= g.M || n >= g.N) continue;
        if (g.lower_only && m < n) continue;
        double v = g.alpha * acc[i][j][r];
        if (g.rowscale) v *= g.rowscale[m];
        if (g.colscale) v *= g.colscale[n];
        if (dg && m == n) v += dg[m];
        const int64_t idx = m + (int64_t)n * g.ldc;
        if (g.beta != 0.0) v += g.beta * C[idx];
        C[idx] = v;
        if (g.lower_only && m > n) C[n + (int64_t)m * g.ldc] = g.upper_zero ? 0.0 : v;
      }
    }
  }
}

// ----------------------------------------------------------------- BF16x3
// WG tile BM x BN, WM x WN waves, K slab of 32 (one bf16 MFMA step). LDS
// holds hi and lo parts k-contiguous per row: X[row][k], so a lane's 8
// consecutive k are one 16-byte read.
template <int BM, int BN, int WM, int WN, int FLUSH, bool TA, bool TB, int KLASS>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_bf16x3_kernel(LowpDev g) {
  constexpr int BK = 32;
  constexpr int KP = BK + 8;            // padded row (bf16 elements): 80 bytes
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int LA = BM * BK / NT, LB = BN * BK / NT;
  static_assert(BM * BK % NT == 0 && BN * BK % NT == 0, "tile/thread mismatch");
  static_assert(TA ? (NT % BK == 0) : (NT % BM == 0), "A walk");
  static_assert(TB ? (NT % BN == 0) : (NT % BK == 0), "B walk");

  __shared__ __bf16 Ah[BM][KP], Al[BM][KP], Bh[BN][KP], Bl[BN][KP];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int batch = (int)(L % g.batch);
  int bm, bn;
  tile_of(g, (int)(L / g.batch), BM, BN, bm, bn);
  const int m0 = bm * BM, n0 = bn * BN;
  if (g.lower_only && m0 + BM - 1 < n0) return;

  const double* __restrict__ A = g.A + (int64_t)batch * g.strideA;
  const double* __restrict__ B = g.B + (int64_t)batch * g.strideB;
  const double* __restrict__ ks = g.kscale ? g.kscale + (int64_t)batch * g.strideK : nullptr;

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
      const int gm = m0 + (TA ? v : a_fix), gk = k0 + (TA ? a_fix : v);
      double x = 0.0;
      if (gm < g.M && gk < g.K) x = TA ? A[gk + (int64_t)gm * g.lda] : A[gm + (int64_t)gk * g.lda];
      ra[i] = x;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      const int gn = n0 + (TB ? b_fix : v), gk = k0 + (TB ? v : b_fix);
      double x = 0.0;
      if (gn < g.N && gk < g.K) {
        x = TB ? B[gn + (int64_t)gk * g.ldb] : B[gk + (int64_t)gn * g.ldb];
        if (ks) x *= ks[gk];
      }
      rb[i] = x;
    }
  };
  auto split_store = [&]() {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int v = a_var + i * a_step;
      const int m = TA ? v : a_fix, k = TA ? a_fix : v;
      const __bf16 hi = (__bf16)(float)ra[i];
      Ah[m][k] = hi;
      Al[m][k] = (__bf16)(float)(ra[i] - (double)(float)hi);
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      const int n = TB ? b_fix : v, k = TB ? v : b_fix;
      const __bf16 hi = (__bf16)(float)rb[i];
      Bh[n][k] = hi;
      Bl[n][k] = (__bf16)(float)(rb[i] - (double)(float)hi);
    }
  };

  double acc[FM][FN][4];
  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }

  const int fr = lane & 15, fq = lane >> 4;
  const int nslab = (g.K + BK - 1) / BK;
  load_slab(0);
  for (int s = 0; s < nslab; ++s) {
    __syncthreads();               // previous slab's fragments consumed
    split_store();
    __syncthreads();
    if (s + 1 < nslab) load_slab((s + 1) * BK);
    bf8 ah[FM], al[FM], bh[FN], bl[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int row = wm * TM + i * 16 + fr;
      ah[i] = *reinterpret_cast<const bf8*>(&Ah[row][8 * fq]);
      al[i] = *reinterpret_cast<const bf8*>(&Al[row][8 * fq]);
    }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int row = wn * TN + j * 16 + fr;
      bh[j] = *reinterpret_cast<const bf8*>(&Bh[row][8 * fq]);
      bl[j] = *reinterpret_cast<const bf8*>(&Bl[row][8 * fq]);
    }
    // product-major: consecutive MFMAs write different accumulators
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl[j], ah[i], part[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], al[i], part[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh[j], ah[i], part[i][j], 0, 0, 0);
    if ((s + 1) % FLUSH == 0 || s + 1 == nslab) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)part[i][j][r];
          part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
  store_tile<FM, FN, TM, TN>(g, acc, batch, m0, n0, wm, wn, lane);
}

// ----------------------------------------------------------------- FP32
// Same walk as the FP64 kernel (k-major LDS rows, K slab 16), FP32 operands.
template <int BM, int BN, int WM, int WN, int FLUSH, bool TA, bool TB, int KLASS>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_f32acc64_kernel(LowpDev g) {
  constexpr int BK = 16;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int LA = BM * BK / NT, LB = BN * BK / NT;
  static_assert(BM * BK % NT == 0 && BN * BK % NT == 0, "tile/thread mismatch");
  static_assert(TA ? (NT % BK == 0) : (NT % BM == 0), "A walk");
  static_assert(TB ? (NT % BN == 0) : (NT % BK == 0), "B walk");

  __shared__ float As[BK][BM + 1], Bs[BK][BN + 1];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int batch = (int)(L % g.batch);
  int bm, bn;
  tile_of(g, (int)(L / g.batch), BM, BN, bm, bn);
  const int m0 = bm * BM, n0 = bn * BN;
  if (g.lower_only && m0 + BM - 1 < n0) return;

  const double* __restrict__ A = g.A + (int64_t)batch * g.strideA;
  const double* __restrict__ B = g.B + (int64_t)batch * g.strideB;
  const double* __restrict__ ks = g.kscale ? g.kscale + (int64_t)batch * g.strideK : nullptr;

  const int a_fix = TA ? tid % BK : tid % BM;
  const int a_var = TA ? tid / BK : tid / BM;
  constexpr int a_step = TA ? NT / BK : NT / BM;
  const int b_fix = TB ? tid % BN : tid % BK;
  const int b_var = TB ? tid / BN : tid / BK;
  constexpr int b_step = TB ? NT / BN : NT / BK;

  float ra[LA], rb[LB];
  auto load_slab = [&](int k0) {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int v = a_var + i * a_step;
      const int gm = m0 + (TA ? v : a_fix), gk = k0 + (TA ? a_fix : v);
      double x = 0.0;
      if (gm < g.M && gk < g.K) x = TA ? A[gk + (int64_t)gm * g.lda] : A[gm + (int64_t)gk * g.lda];
      ra[i] = (float)x;
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      const int gn = n0 + (TB ? b_fix : v), gk = k0 + (TB ? v : b_fix);
      double x = 0.0;
      if (gn < g.N && gk < g.K) {
        x = TB ? B[gn + (int64_t)gk * g.ldb] : B[gk + (int64_t)gn * g.ldb];
        if (ks) x *= ks[gk];
      }
      rb[i] = (float)x;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < LA; ++i) {
      const int v = a_var + i * a_step;
      if (TA) As[a_fix][v] = ra[i]; else As[v][a_fix] = ra[i];
    }
#pragma unroll
    for (int i = 0; i < LB; ++i) {
      const int v = b_var + i * b_step;
      if (TB) Bs[v][b_fix] = rb[i]; else Bs[b_fix][v] = rb[i];
    }
  };

  double acc[FM][FN][4];
  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }
  const int fr = lane & 15, fk = lane >> 4;
  const int nslab = (g.K + BK - 1) / BK;
  load_slab(0);
  for (int s = 0; s < nslab; ++s) {
    __syncthreads();
    store();
    __syncthreads();
    if (s + 1 < nslab) load_slab((s + 1) * BK);
#pragma unroll
    for (int k4 = 0; k4 < BK / 4; ++k4) {
      float af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = As[k4 * 4 + fk][wm * TM + i * 16 + fr];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = Bs[k4 * 4 + fk][wn * TN + j * 16 + fr];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          part[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bf[j], af[i], part[i][j], 0, 0, 0);
    }
    if ((s + 1) % FLUSH == 0 || s + 1 == nslab) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)part[i][j][r];
          part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  }
  store_tile<FM, FN, TM, TN>(g, acc, batch, m0, n0, wm, wn, lane);
}

// ----------------------------------------------------------------- FP16x3
// Plain products C = alpha A B (+ beta C), A (M x K) and B (K x N) both
// column-major and not transposed, batched by strides: the AO->MO
// transforms. Each operand is scaled by an exact power of two that brings
// its maximum (the device scalars amax, bmax) into [2^13, 2^14) and split
// into two fp16 parts, x 2^e = hi + lo with hi = fp16(y), lo = fp16(y - hi)
// (y the FP32 value); a b ~ hi hi + hi lo + lo hi on
// v_mfma_f32_16x16x32_f16 (~2^-22 relative per product, against 2^-16 for
// bf16), accumulated in FP32 over the whole K (K = nbf <= a few thousand:
// 3 products per 32-k slab) and unscaled exactly in FP64 in the epilogue.
// The split happens as the tile goes to LDS: A, the large operand (the RI
// tensor or the half-transformed one), is split once per column tile of C,
// i.e. about once per element for the narrow N of these products.
// LDS rows hold 32 k (64 bytes) with the 16-byte chunks swizzled as in the
// chi0 kernels. cmax (nullable) receives max |C| of the stored results, the
// scale input of the next transform.
__device__ inline int lds_chunk16(int r, int j) { return j ^ ((r >> 1) & 3); }
typedef _Float16 h8v __attribute__((ext_vector_type(8)));
typedef unsigned int u4v __attribute__((ext_vector_type(4)));

struct F16Dev {
  int M, N, K;
  const double* A; int64_t lda, strideA;
  const double* B; int64_t ldb, strideB;
  double* C; int64_t ldc, strideC;
  double alpha, beta;
  const double* amax;
  const double* bmax;
  double* cmax;
  int tiles_m, tiles_n, batch;
  int64_t nblocks;
  const _Float16* bp;  // BPL: planes of op(B)^T (plane[q][kt][row n][32])
  int64_t bq;
  int brpad;
};

__device__ inline double pow2_scale(double mx) {
  return (mx > 0.0 && mx <= 1.0e300) ? ldexp(1.0, 13 - ilogb(mx)) : 1.0;
}

template <int BM, int BN, int WM, int WN, int KLASS, bool BPL>
__global__ void __launch_bounds__(64 * WM * WN)
gemm_f16x3_kernel(F16Dev g) {
  constexpr int BK = 32;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  // A: thread = (m, k octet); B: thread = (k octet, n)
  constexpr int AU = BM * (BK / 8) / NT;      // A units (8 k of one m) per thread
  constexpr int BU = (BN * (BK / 8) + NT - 1) / NT;
  static_assert(BM * (BK / 8) % NT == 0 && NT % BM == 0, "A walk");
  __shared__ __attribute__((aligned(16))) _Float16 As[2][BM][BK];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[2][BN][BK];
  __shared__ double cred[NT / 64];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int batch = (int)(L % g.batch);
  const int tile = (int)(L / g.batch);
  // groups of 8 row tiles walked column by column: concurrent workgroups share B
  constexpr int GROUP_M = 8;
  const int per_group = GROUP_M * g.tiles_n;
  const int group = tile / per_group;
  const int first_m = group * GROUP_M;
  const int gsize = min(g.tiles_m - first_m, GROUP_M);
  const int in_group = tile - group * per_group;
  const int m0 = (first_m + in_group % gsize) * BM, n0 = (in_group / gsize) * BN;

  const double* __restrict__ A = g.A + (int64_t)batch * g.strideA;
  const double* __restrict__ B = g.B + (int64_t)batch * g.strideB;
  const double sa = pow2_scale(g.amax[0]), sb = pow2_scale(g.bmax[0]);
  const double unscale = 1.0 / (sa * sb);     // exact: a power of two

  const int am = tid % BM, ako = tid / BM;    // A: row m, first k octet
  constexpr int A_KO_STEP = NT / BM;
  double ra[AU][8], rb[BPL ? 1 : BU][8];
  u4v rbp[BPL ? BU : 1][2];        // BPL: the hi and lo 16-byte chunks of B^T
  auto load = [&](int k0) __attribute__((always_inline)) {
    const int gm = m0 + am;
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int kb = k0 + 8 * (ako + u * A_KO_STEP);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int gk = kb + e;
        ra[u][e] = (gm < g.M && gk < g.K) ? A[gm + (int64_t)gk * g.lda] : 0.0;
      }
    }
    if constexpr (BPL) {
      // rows n < N of the planes (padded with zeros past N and K); rows past
      // N are not read (their columns are not stored)
      const int kt = k0 / BK;
#pragma unroll
      for (int u = 0; u < BU; ++u) {
        const int c = tid + u * NT;
        const int n = c >> 2, ko = c & 3;
        const int gn = n0 + n;
        const bool in = c < BN * 4 && gn < g.N;
        const int64_t off = ((int64_t)kt * g.brpad + gn) * BK + 8 * ko;
#pragma unroll
        for (int q = 0; q < 2; ++q)
          rbp[u][q] = in ? *reinterpret_cast<const u4v*>(g.bp + q * g.bq + off) : u4v{0u, 0u, 0u, 0u};
      }
    } else {
#pragma unroll
      for (int u = 0; u < BU; ++u) {
        const int c = tid + u * NT;             // (n, k octet) with the octet fastest
        const int n = c >> 2, ko = c & 3;
        const int gn = n0 + n, kb = k0 + 8 * ko;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int gk = kb + e;
          rb[u][e] = (c < BN * 4 && gn < g.N && gk < g.K) ? B[gk + (int64_t)gn * g.ldb] : 0.0;
        }
      }
    }
  };
  auto store = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < AU; ++u) {
      const int ko = ako + u * A_KO_STEP;
      h8v hi, lo;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float y = (float)(ra[u][e] * sa);
        const _Float16 h = (_Float16)y;
        hi[e] = h;
        lo[e] = (_Float16)(y - (float)h);
      }
      *reinterpret_cast<h8v*>(&As[0][am][8 * lds_chunk16(am, ko)]) = hi;
      *reinterpret_cast<h8v*>(&As[1][am][8 * lds_chunk16(am, ko)]) = lo;
    }
#pragma unroll
    for (int u = 0; u < BU; ++u) {
      const int c = tid + u * NT;
      if (c >= BN * 4) break;
      const int n = c >> 2, ko = c & 3;
      if constexpr (BPL) {
        *reinterpret_cast<u4v*>(&Bs[0][n][8 * lds_chunk16(n, ko)]) = rbp[u][0];
        *reinterpret_cast<u4v*>(&Bs[1][n][8 * lds_chunk16(n, ko)]) = rbp[u][1];
      } else {
        h8v hi, lo;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float y = (float)(rb[u][e] * sb);
          const _Float16 h = (_Float16)y;
          hi[e] = h;
          lo[e] = (_Float16)(y - (float)h);
        }
        *reinterpret_cast<h8v*>(&Bs[0][n][8 * lds_chunk16(n, ko)]) = hi;
        *reinterpret_cast<h8v*>(&Bs[1][n][8 * lds_chunk16(n, ko)]) = lo;
      }
    }
  };

  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fq = lane >> 4;
  const int nslab = (g.K + BK - 1) / BK;
  load(0);
  for (int s = 0; s < nslab; ++s) {
    __syncthreads();                 // previous slab's fragments consumed
    store();
    __syncthreads();
    if (s + 1 < nslab) load((s + 1) * BK);
    h8v a[2][FM], b[2][FN];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * TM + i * 16 + fr;
        a[q][i] = *reinterpret_cast<const h8v*>(&As[q][r][8 * lds_chunk16(r, fq)]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wn * TN + j * 16 + fr;
        b[q][j] = *reinterpret_cast<const h8v*>(&Bs[q][r][8 * lds_chunk16(r, fq)]);
      }
    }
    // lo hi + hi lo + hi hi, product-major (consecutive MFMAs write different accumulators)
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int qa = pr == 1 ? 1 : 0, qb = pr == 0 ? 1 : 0;
          part[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(b[qb][j], a[qa][i], part[i][j], 0, 0, 0);
        }
  }
  // epilogue: lane l, register r holds C(m = m_base + (l & 15), n = n_base + 4 (l >> 4) + r)
  double* __restrict__ C = g.C + (int64_t)batch * g.strideC;
  double cm = 0.0;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
        if (m >= g.M || n >= g.N) continue;
        const int64_t idx = m + (int64_t)n * g.ldc;
        double v = g.alpha * ((double)part[i][j][r] * unscale);
        if (g.beta != 0.0) v += g.beta * C[idx];
        C[idx] = v;
        cm = fmax(cm, fabs(v));
      }
  }
  if (g.cmax) {
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_down(cm, off, 64));
    if (lane == 0) cred[wave] = cm;
    __syncthreads();
    if (tid == 0) {
      double b = 0.0;
#pragma unroll
      for (int w = 0; w < NT / 64; ++w) b = fmax(b, cred[w]);
      atomicMax(reinterpret_cast<unsigned long long*>(g.cmax), (unsigned long long)__double_as_longlong(b));
    }
  }
}

template <int BM, int BN, int WM, int WN, int KLASS>
int launch_f16x3(const GemmF64& h, hipStream_t stream) {
  MGW_REQUIRE(!h.bplanes || (h.batch == 1 && h.bplane_rpad >= h.N), MGW_ERR_ARG,
              "gemm_lowp: B planes need one matrix and rows covering N");
  F16Dev g;
  g.M = h.M; g.N = h.N; g.K = h.K;
  g.A = h.A; g.lda = h.lda; g.strideA = h.strideA;
  g.B = h.B; g.ldb = h.ldb; g.strideB = h.strideB;
  g.C = h.C; g.ldc = h.ldc; g.strideC = h.strideC;
  g.alpha = h.alpha; g.beta = h.beta;
  g.amax = h.amax; g.bmax = h.bmax; g.cmax = h.cmax;
  g.tiles_m = (h.M + BM - 1) / BM;
  g.tiles_n = (h.N + BN - 1) / BN;
  g.batch = h.batch;
  g.nblocks = (int64_t)g.tiles_m * g.tiles_n * h.batch;
  MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "gemm_lowp: too many tiles");
  g.bp = h.bplanes; g.bq = h.bplane_q; g.brpad = h.bplane_rpad;
  if (g.bp)
    hipLaunchKernelGGL((gemm_f16x3_kernel<BM, BN, WM, WN, KLASS, true>), dim3((unsigned)g.nblocks),
                       dim3(64 * WM * WN), 0, stream, g);
  else
    hipLaunchKernelGGL((gemm_f16x3_kernel<BM, BN, WM, WN, KLASS, false>), dim3((unsigned)g.nblocks),
                       dim3(64 * WM * WN), 0, stream, g);
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

template <int BM, int BN, int WM, int WN, int FLUSH, int KLASS, bool BF16>
int launch_lowp(const GemmF64& h, hipStream_t stream) {
  LowpDev g;
  g.M = h.M; g.N = h.N; g.K = h.K;
  g.A = h.A; g.lda = h.lda; g.strideA = h.strideA;
  g.B = h.B; g.ldb = h.ldb; g.strideB = h.strideB;
  g.C = h.C; g.ldc = h.ldc; g.strideC = h.strideC;
  g.kscale = h.kscale; g.strideK = h.strideK;
  g.rowscale = h.rowscale; g.colscale = h.colscale;
  g.diag_add = h.diag_add; g.strideDiag = h.strideDiag;
  g.alpha = h.alpha; g.beta = h.beta;
  g.lower_only = h.lower_only ? 1 : 0;
  g.upper_zero = h.upper_zero ? 1 : 0;
  g.tiles_m = (h.M + BM - 1) / BM;
  g.tiles_n = (h.N + BN - 1) / BN;
  g.batch = h.batch;
  const int64_t tiles = (h.lower_only && BM == BN) ? (int64_t)g.tiles_m * (g.tiles_m + 1) / 2
                                                   : (int64_t)g.tiles_m * g.tiles_n;
  g.nblocks = tiles * h.batch;
  MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "gemm_lowp: too many tiles");
  dim3 grid((unsigned)g.nblocks), block(64 * WM * WN);
  const int ta = h.transA ? 1 : 0, tb = h.transB ? 1 : 0;
#define MGW_LOWP_CASE(TA_, TB_)                                                                          \
  if (ta == TA_ && tb == TB_) {                                                                          \
    if constexpr (BF16)                                                                                  \
      hipLaunchKernelGGL((gemm_bf16x3_kernel<BM, BN, WM, WN, FLUSH, TA_, TB_, KLASS>), grid, block, 0,   \
                         stream, g);                                                                     \
    else                                                                                                 \
      hipLaunchKernelGGL((gemm_f32acc64_kernel<BM, BN, WM, WN, FLUSH, TA_, TB_, KLASS>), grid, block, 0, \
                         stream, g);                                                                     \
  }
  MGW_LOWP_CASE(0, 0)
  MGW_LOWP_CASE(0, 1)
  MGW_LOWP_CASE(1, 0)
  MGW_LOWP_CASE(1, 1)
#undef MGW_LOWP_CASE
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace

int gemm_lowp(const GemmF64& h, int precision, hipStream_t stream) {
  if (h.M <= 0 || h.N <= 0 || h.batch <= 0) return MGW_OK;
  MGW_REQUIRE(h.dotpart == nullptr && h.C != nullptr && !h.tri_a && !h.no_mirror, MGW_ERR_ARG,
              "gemm_lowp: plain products only");
  MGW_REQUIRE(!h.lower_only || h.M == h.N, MGW_ERR_ARG, "gemm_lowp: lower_only needs M == N");
  if (precision == MGW_BF16X3) {
    if (h.klass == MGW_KCLASS_GEMM_CHI0) return launch_lowp<128, 128, 4, 2, 4, MGW_KCLASS_GEMM_CHI0, true>(h, stream);
    return launch_lowp<128, 128, 4, 2, 4, MGW_KCLASS_GEMM_AO2MO, true>(h, stream);
  }
  if (precision == MGW_FP16X3) {
    MGW_REQUIRE(!h.transA && !h.transB && !h.lower_only && !h.kscale && !h.rowscale && !h.colscale && !h.diag_add &&
                    h.amax && h.bmax,
                MGW_ERR_ARG, "gemm_lowp: FP16x3 takes plain non-transposed products with operand maxima");
    // one column tile when N (the first transform's occupied and window
    // columns) fits 64, 128 or 192, so that the large operand A is read once
    if (h.N <= 64) return launch_f16x3<128, 64, 2, 2, MGW_KCLASS_GEMM_AO2MO>(h, stream);
    if (h.N > 128 && h.N <= 192) return launch_f16x3<128, 192, 2, 2, MGW_KCLASS_GEMM_AO2MO>(h, stream);
    return launch_f16x3<128, 128, 2, 2, MGW_KCLASS_GEMM_AO2MO>(h, stream);
  }
  if (precision == MGW_FP32) {
    if (h.klass == MGW_KCLASS_GEMM_CHI0) return launch_lowp<128, 128, 4, 2, 8, MGW_KCLASS_GEMM_CHI0, false>(h, stream);
    return launch_lowp<128, 128, 4, 2, 8, MGW_KCLASS_GEMM_AO2MO, false>(h, stream);
  }
  set_error("gemm_lowp: unknown precision");
  return MGW_ERR_ARG;
}

}  // namespace mgw
