// Reduced-precision 1 - chi0(iw) from pre-split operands.
//
// For every imaginary frequency w_f of the grid
//   1 - chi0(iw_f) = I + Y_f Y_f^T,   Y_f = X diag(sqrt(s_f)),
//   s_f(t) = 2 (f_i - f_a)(e_a - e_i) / (w_f^2 + (e_a - e_i)^2)   (>= 0; the caller checks)
// (src/m_gw_selfenergy_grid.f90:110-150 forms X diag(s) X^T with DSYRK in FP64).
//
// A pre-pass writes Y_f once per frequency, split into NS bf16 planes
// y = y0 + y1 (+ y2), with y0 = bf16(y), y1 = bf16(y - y0), y2 = bf16(y - y0 - y1),
// or into two fp16 planes of 2^e_f Y_f (FP16X3): the power-of-two scale puts
// the largest element near 2^14 (max|X| sqrt(max s_f) bounds it), so the
// 11-bit hi and lo parts stay normal for every element that matters and the
// result is unscaled exactly by 2^-2e_f.
// Each plane is k-tiled, plane[f][kt][P][32] (P padded to a multiple of 128,
// t padded to a multiple of 32, zeros in the padding), so the K slab of a
// row block is one contiguous run and a lane's 8 consecutive k are one
// 16-byte load. The SYRK kernel multiplies the planes on
// v_mfma_f32_16x16x32_bf16:
//   NS = 2 (BF16x3): y0 y0 + y0 y1 + y1 y0                  (~2^-17 relative per product)
//   NS = 3 (BF16x6): + y1 y1 + y0 y2 + y2 y0                (~2^-24, FP32 level)
//   NS = 2, fp16 (FP16X3): y0 y0 + y0 y1 + y1 y0 on v_mfma_f32_16x16x32_f16 (~2^-22)
// FP32 partials are flushed into FP64 accumulators every FLUSH slabs of 32,
// so sums over K ~ 10^5 transitions keep FP64 accumulation error.
//
// Fragment layout (gfx950, v_mfma_f32_16x16x32_bf16): a[i = l & 15][k = 8 (l >> 4) + e],
// b[k = 8 (l >> 4) + e][j = l & 15], d: lane l, register r -> d[4 (l >> 4) + r][l & 15].
// The kernel feeds a = rows of the n block, b = rows of the m block, so lane
// l, register r holds C(m = l & 15, n = 4 (l >> 4) + r) of the 16 x 16 tile.
#include "mgw_internal.h"
#include "tile_order.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>

namespace mgw {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

// operand element and 8-vector of the planes: bf16, or fp16 when H
template <bool H> struct PlaneT { typedef __bf16 E; typedef bf8 V; };
template <> struct PlaneT<true> { typedef _Float16 E; typedef h8 V; };

template <bool H>
__device__ inline f4 mfma16x16x32(typename PlaneT<H>::V b, typename PlaneT<H>::V a, f4 c) {
  if constexpr (H) return __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a, c, 0, 0, 0);
  else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, c, 0, 0, 0);
}

constexpr int KT = 32;          // k per tile of the plane layout (one MFMA step)

// LDS rows of a slab are 64 bytes (32 k) without padding; 16-byte chunk j of
// row r is stored at chunk j ^ ((r >> 1) & 3). Eight consecutive rows read at
// one chunk index (the MFMA fragment loads), and the four chunks of two rows
// written by eight consecutive lanes (the slab stores), then fall on eight
// distinct 16-byte bank groups. Rows of a fragment start at multiples of 16,
// so (fr >> 1) & 3 gives the swizzle of row base + fr.
__device__ inline int lds_chunk(int r, int j) { return j ^ ((r >> 1) & 3); }
constexpr int PAD_ROWS = 128;   // P padding of the plane layout

// ---------------------------------------------------------------- split
// Block: 64 rows P x 32 transitions t (one k tile), all nf frequencies of
// the chunk. X is read once through LDS (coalesced along P) into registers;
// the weights sqrt(s_f(t)) (times the plane scale) of 32 frequencies at a
// time are staged in LDS, so the frequency loop issues no global loads and
// each thread writes 8 consecutive t of one row per plane per frequency (one
// 16-byte store).
template <int NS, bool H>
__global__ void __launch_bounds__(256)
chi_split_kernel(const double* __restrict__ x, int naux, int npole, const double* __restrict__ sw,
                 const double* __restrict__ scale, int f0, int nf, int ppad, int nkt,
                 typename PlaneT<H>::E* __restrict__ y, int64_t plane) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int FB = 32;                    // frequencies per staged weight block
  // two planes carry at most 22 bits: the product and the split run in FP32
  // (24 bits); three planes need the FP64 product
  typedef typename std::conditional<NS == 2, float, double>::type R;
  __shared__ double tile[KT][65];
  __shared__ R wt[FB][KT];
  const int tid = threadIdx.x;
  const int p0 = blockIdx.x * 64, kt = blockIdx.y, t0 = kt * KT;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tl = (tid >> 6) + 4 * j, p = tid & 63;
    const int P = p0 + p, t = t0 + tl;
    tile[tl][p] = (P < naux && t < npole) ? x[P + (int64_t)t * naux] : 0.0;
  }
  __syncthreads();
  const int p = tid >> 2, c8 = tid & 3;
  const int64_t row = (int64_t)(p0 + p) * KT + 8 * c8;
  R xv[8];
#pragma unroll
  for (int e = 0; e < 8; ++e) xv[e] = (R)tile[8 * c8 + e][p];
  for (int fb = 0; fb < nf; fb += FB) {
    const int nb = min(FB, nf - fb);
    __syncthreads();                        // previous block's weights consumed
    for (int e = tid; e < FB * KT; e += 256) {
      const int fl = e / KT, tl = e % KT, t = t0 + tl;
      double w = 0.0;
      if (fl < nb && t < npole) {
        const int f = f0 + fb + fl;
        w = sqrt(sw[(int64_t)f * npole + t]) * (scale ? scale[f] : 1.0);
      }
      wt[fl][tl] = (R)w;
    }
    __syncthreads();
    for (int fl = 0; fl < nb; ++fl) {
      V out[NS];
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        R v = xv[e] * wt[fl][8 * c8 + e];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const E h = (E)(float)v;
          out[q][e] = h;
          v -= (R)(float)h;
        }
      }
      const int64_t off = ((int64_t)(fb + fl) * nkt + kt) * (int64_t)ppad * KT + row;
#pragma unroll
      for (int q = 0; q < NS; ++q) *reinterpret_cast<V*>(y + q * plane + off) = out[q];
    }
  }
}

// The two-plane split from the k-tiled FP32 copy of X (xf[kt][P][32], made
// once per build by chi_xf32_kernel): a thread's 8 transitions of one row are
// two 16-byte loads, so the block needs no LDS transpose, and each frequency
// chunk reads 4 bytes per element of X instead of 8. Same arithmetic as
// chi_split_kernel<2, H> (FP32 product, then the split): identical planes.
template <bool H>
__global__ void __launch_bounds__(256)
chi_split_xf_kernel(const float* __restrict__ xf, int npole, const double* __restrict__ sw,
                    const double* __restrict__ scale, int f0, int nf, int ppad, int nkt,
                    typename PlaneT<H>::E* __restrict__ y, int64_t plane) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int FB = 32;
  __shared__ float wt[FB][KT];
  const int tid = threadIdx.x;
  const int p0 = blockIdx.x * 64, kt = blockIdx.y, t0 = kt * KT;
  const int p = tid >> 2, c8 = tid & 3;
  const int64_t row = (int64_t)(p0 + p) * KT + 8 * c8;
  const float* __restrict__ src = xf + (int64_t)kt * ppad * KT + row;
  const f4 x0 = *reinterpret_cast<const f4*>(src), x1 = *reinterpret_cast<const f4*>(src + 4);
  float xv[8];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    xv[e] = x0[e];
    xv[4 + e] = x1[e];
  }
  for (int fb = 0; fb < nf; fb += FB) {
    const int nb = min(FB, nf - fb);
    __syncthreads();                        // previous block's weights consumed
    for (int e = tid; e < FB * KT; e += 256) {
      const int fl = e / KT, tl = e % KT, t = t0 + tl;
      double w = 0.0;
      if (fl < nb && t < npole) {
        const int f = f0 + fb + fl;
        w = sqrt(sw[(int64_t)f * npole + t]) * (scale ? scale[f] : 1.0);
      }
      wt[fl][tl] = (float)w;
    }
    __syncthreads();
    for (int fl = 0; fl < nb; ++fl) {
      V hi, lo;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float v = xv[e] * wt[fl][8 * c8 + e];
        const E h = (E)v;
        hi[e] = h;
        lo[e] = (E)(v - (float)h);
      }
      const int64_t off = ((int64_t)(fb + fl) * nkt + kt) * (int64_t)ppad * KT + row;
      *reinterpret_cast<V*>(y + off) = hi;
      *reinterpret_cast<V*>(y + plane + off) = lo;
    }
  }
}

// ---------------------------------------------------------------- SYRK
struct SyrkArgs {
  const void* y;       // planes (bf16 or fp16)
  int64_t plane;       // elements between planes
  int64_t fstride;     // elements between frequencies
  int ppad, nkt, naux, nf, tiles, tm;
  double* C;           // chi + f0 * naux^2
  int64_t strideC;
  int64_t nblocks;
  // K split: workgroup (f, kc, tile) multiplies slabs [kc * kc_slabs, ...) and
  // writes its FP32 tile to part[((f * tiles + tile) * nkc + kc) * BM * BM]
  int kc_slabs, nkc;
  float* part;
  const double* unscale;   // per frequency of the chunk (FP16X3), else null
};

// Lower-triangle tiles BM x BM, WM x WN waves, one 32-k slab per step.
// DEPTH 1, 2: that many register sets prefetch the next slabs from global
// memory while the MFMAs consume the LDS copy of the current one (two
// barriers per slab). DEPTH 3: two LDS buffers; the next slab is written to
// the idle buffer while the MFMAs read the other (one barrier per slab).
template <int NS, bool H, int BM, int WM, int WN, int FLUSH, int DEPTH, int KS, bool KSPLIT>
__global__ void __launch_bounds__(64 * WM * WN, KSPLIT ? 2 : 1)
chi_syrk_kernel(SyrkArgs g) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int KP = KT;                  // LDS row: 64 bytes, chunks swizzled (lds_chunk)
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BM / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int CH = BM * KT / 8;         // 16-byte chunks per plane per operand slab
  constexpr int LPT = CH / NT;
  static_assert(CH % NT == 0, "chunks per thread");
  static_assert(TM % 16 == 0 && TN % 16 == 0, "wave tile");
  static_assert(DEPTH >= 1 && DEPTH <= 3, "prefetch depth");
  static_assert(KS == 1 || KS == 2, "k tiles per slab");
  constexpr int NBUF = DEPTH == 3 ? 2 : 1;

  // KS k tiles of 32 per slab (the plane layout pads the k tiles to a multiple of 2)
  __shared__ __attribute__((aligned(16))) E As[NBUF][NS][KS][BM][KP];
  __shared__ __attribute__((aligned(16))) E Bs[NBUF][NS][KS][BM][KP];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  // K split: (frequency, chunk, tile) with the tile fastest, so concurrent
  // workgroups read the same K range of their panels
  const int64_t per_f = KSPLIT ? (int64_t)g.nkc * g.tiles : g.tiles;
  const int f = (int)(L / per_f);
  const int kc = KSPLIT ? (int)((L % per_f) / g.tiles) : 0;
  const int tile = (int)(L % g.tiles);
  int bm, bn;
  tri_grouped(tile, g.tm, bm, bn);
  const int m0 = bm * BM, n0 = bn * BM;

  const E* __restrict__ Ya = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)m0 * KT;
  const E* __restrict__ Yb = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)n0 * KT;
  const int64_t slab = (int64_t)g.ppad * KT;

  // prefetch registers (native vectors, so they stay in VGPRs)
  u4 ra0[NS][KS][LPT], rb0[NS][KS][LPT], ra1[NS][KS][LPT], rb1[NS][KS][LPT];
#define MGW_SYRK_LOAD(RA, RB, S_)                                                                  \
  _Pragma("unroll") for (int q = 0; q < NS; ++q)                                                    \
    _Pragma("unroll") for (int kk = 0; kk < KS; ++kk)                                               \
      _Pragma("unroll") for (int i = 0; i < LPT; ++i) {                                             \
        const int c = tid + i * NT;                                                                 \
        const int64_t off = q * g.plane + ((int64_t)(S_) * KS + kk) * slab + 8 * c;                 \
        RA[q][kk][i] = *reinterpret_cast<const u4*>(Ya + off);                                      \
        RB[q][kk][i] = *reinterpret_cast<const u4*>(Yb + off);                                      \
      }
#define MGW_SYRK_STORE(RA, RB, BUF)                                                  \
  _Pragma("unroll") for (int q = 0; q < NS; ++q)                                      \
    _Pragma("unroll") for (int kk = 0; kk < KS; ++kk)                                 \
      _Pragma("unroll") for (int i = 0; i < LPT; ++i) {                               \
        const int c = tid + i * NT;                                                   \
        *reinterpret_cast<u4*>(&As[BUF][q][kk][c >> 2][8 * lds_chunk(c >> 2, c & 3)]) = RA[q][kk][i];  \
        *reinterpret_cast<u4*>(&Bs[BUF][q][kk][c >> 2][8 * lds_chunk(c >> 2, c & 3)]) = RB[q][kk][i];  \
      }

  double acc[FM][FN][4];
  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }

  const int fr = lane & 15, fq = lane >> 4;
  const int nslab = g.nkt / KS;
  const int s0 = KSPLIT ? kc * g.kc_slabs : 0;
  const int s1 = KSPLIT ? min(nslab, s0 + g.kc_slabs) : nslab;
  // MFMAs of the slab in LDS, then the FP64 flush every FLUSH slabs
  auto compute = [&](int s, int buf) __attribute__((always_inline)) {
    // products (a plane, b plane): BF16x6 all six, BF16x3 the last three
    constexpr int QA[6] = {NS == 3 ? 2 : 0, 0, 1, 1, 0, 0};
    constexpr int QB[6] = {0, NS == 3 ? 2 : 0, 1, 0, 1, 0};
    constexpr int P0 = NS == 3 ? 0 : 3;
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
      V a[NS][FM], b[NS][FN];
#pragma unroll
      for (int q = 0; q < NS; ++q) {
#pragma unroll
        for (int i = 0; i < FM; ++i)
          a[q][i] = *reinterpret_cast<const V*>(&As[buf][q][kk][wm * TM + i * 16 + fr][8 * lds_chunk(fr, fq)]);
#pragma unroll
        for (int j = 0; j < FN; ++j)
          b[q][j] = *reinterpret_cast<const V*>(&Bs[buf][q][kk][wn * TN + j * 16 + fr][8 * lds_chunk(fr, fq)]);
      }
      // product-major: consecutive MFMAs write different accumulators, so
      // none waits for the previous one's result (small products first)
#pragma unroll
      for (int pr = P0; pr < 6; ++pr)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            part[i][j] = mfma16x16x32<H>(b[QB[pr]][j], a[QA[pr]][i], part[i][j]);
    }
    if (!KSPLIT && ((s + 1) % FLUSH == 0 || s + 1 == nslab)) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[i][j][r] += (double)part[i][j][r];
          part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
        }
    }
  };

  if constexpr (DEPTH == 1) {
    MGW_SYRK_LOAD(ra0, rb0, s0)
    for (int s = s0; s < s1; ++s) {
      __syncthreads();                   // previous slab's fragments consumed
      MGW_SYRK_STORE(ra0, rb0, 0)
      __syncthreads();
      if (s + 1 < s1) { MGW_SYRK_LOAD(ra0, rb0, s + 1) }
      compute(s, 0);
    }
  } else if constexpr (DEPTH == 2) {
    MGW_SYRK_LOAD(ra0, rb0, s0)
    if (s0 + 1 < s1) { MGW_SYRK_LOAD(ra1, rb1, s0 + 1) }
    for (int s = s0; s < s1; s += 2) {
      __syncthreads();
      MGW_SYRK_STORE(ra0, rb0, 0)
      __syncthreads();
      if (s + 2 < s1) { MGW_SYRK_LOAD(ra0, rb0, s + 2) }
      compute(s, 0);
      if (s + 1 >= s1) break;
      __syncthreads();
      MGW_SYRK_STORE(ra1, rb1, 0)
      __syncthreads();
      if (s + 3 < s1) { MGW_SYRK_LOAD(ra1, rb1, s + 3) }
      compute(s + 1, 0);
    }
  } else {
    // buffer (s - s0) & 1 holds slab s; slab s + 1 goes to the other buffer
    // while slab s is multiplied, and the one barrier per slab orders both ways
    MGW_SYRK_LOAD(ra0, rb0, s0)
    MGW_SYRK_STORE(ra0, rb0, 0)
    if (s0 + 1 < s1) { MGW_SYRK_LOAD(ra0, rb0, s0 + 1) }
    __syncthreads();
    for (int s = s0; s < s1; ++s) {
      const int cur = (s - s0) & 1;
      compute(s, cur);
      if (s + 1 < s1) {
        MGW_SYRK_STORE(ra0, rb0, cur ^ 1)
        if (s + 2 < s1) { MGW_SYRK_LOAD(ra0, rb0, s + 2) }
      }
      __syncthreads();
    }
  }

  if constexpr (KSPLIT) {
    // FP32 partial tile, column-major BM x BM; summed in FP64 by chi_ksplit_reduce_kernel
    float* __restrict__ P = g.part + ((int64_t)(f * g.tiles + tile) * g.nkc + kc) * (BM * BM);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          P[(wm * TM + i * 16 + fr) + (wn * TN + j * 16 + 4 * fq + r) * BM] = part[i][j][r];
    return;
  }

  // epilogue: lower triangle with 1 on the diagonal, zeros above it (the
  // Cholesky that follows reads the lower triangle and wants the upper zero)
  double* __restrict__ C = g.C + (int64_t)f * g.strideC;
  const int n_ld = g.naux;
  const double us = g.unscale ? g.unscale[f] : 1.0;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
        if (m >= g.naux || n >= g.naux || m < n) continue;
        double v = acc[i][j][r] * us;
        if (m == n) v += 1.0;
        C[m + (int64_t)n * n_ld] = v;
        if (m > n) C[n + (int64_t)m * n_ld] = 0.0;
      }
  }
}

#undef MGW_SYRK_LOAD
#undef MGW_SYRK_STORE

// The K-split SYRK with the slabs copied from the planes straight into LDS
// (global_load_lds_dwordx4: each lane's 16 bytes land at the wave's LDS base
// + 16 x lane), so no prefetch registers: more waves fit a SIMD, or a wave
// can hold a 64 x 128 tile (WN = 1, a quarter fewer LDS fragment reads per
// MFMA than 64 x 64). Lane p of a 1 KB window of 16 rows writes LDS chunk
// p, i.e. row p / 4 and slot p % 4, which holds the swizzled chunk
// j = slot ^ ((row >> 1) & 3) of that row, so the lane loads that global
// chunk. NBUF: with 2 to 4 LDS buffers the copies of the next NBUF - 1
// slabs run during this slab's MFMAs (one barrier per slab); with one, each
// slab is copied, waited for and multiplied (two barriers per slab). Same
// products in the same order per accumulator as chi_syrk_kernel: identical
// results.

// s_waitcnt immediate that waits until at most v vector-memory operations of
// this wave are outstanding (expcnt and lgkmcnt not waited on): vmcnt bits
// [3:0] and [15:14], expcnt [6:4], lgkmcnt [11:8]
constexpr int waitcnt_vm(int v) { return (v & 15) | (7 << 4) | (15 << 8) | ((v >> 4) << 14); }

template <bool H, int BM, int WM, int WN, int NBUF>
__global__ void __launch_bounds__(64 * WM * WN)
chi_syrk_lds_kernel(SyrkArgs g) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int NS = 2;
  constexpr int NW = WM * WN;
  constexpr int TM = BM / WM, TN = BM / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int WINS = BM / 16;                 // 1 KB windows per plane per operand
  constexpr int LPW = 2 * NS * WINS / NW;       // copies per wave per slab
  static_assert((2 * NS * WINS) % NW == 0, "windows per wave");
  static_assert(NBUF == 1 || NBUF == 2, "LDS buffers");
  __shared__ __attribute__((aligned(16))) E As[NBUF][NS][BM][KT];
  __shared__ __attribute__((aligned(16))) E Bs[NBUF][NS][BM][KT];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int64_t per_f = (int64_t)g.nkc * g.tiles;
  const int f = (int)(L / per_f);
  const int kc = (int)((L % per_f) / g.tiles);
  const int tile = (int)(L % g.tiles);
  int bm, bn;
  tri_grouped(tile, g.tm, bm, bn);
  const int m0 = bm * BM, n0 = bn * BM;
  const E* __restrict__ Ya = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)m0 * KT;
  const E* __restrict__ Yb = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)n0 * KT;
  const int64_t slab = (int64_t)g.ppad * KT;
  // this lane's chunk within a window: row rl = lane / 4 (of 16), logical chunk j
  const int rl = lane >> 2;
  // the XOR of lds_chunk depends on the row's bits 1-2: the same for row
  // 16 w + rl as for rl, since windows start at multiples of 16
  const int jl = (lane & 3) ^ ((rl >> 1) & 3);
  auto issue = [&](int s, int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < 2 * NS * WINS / NW; ++u) {
      const int w = wave + u * NW;             // window index over (operand, plane, 16-row block)
      const int op = w / (NS * WINS), q = (w / WINS) % NS, win = w % WINS;
      const int64_t off = q * g.plane + (int64_t)s * slab + (int64_t)(16 * win + rl) * KT + 8 * jl;
      // operands cast to non-dependent types: with the element type a template
      // parameter, the host pass drops the kernel's launch stub
      if (op == 0)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned*>(Ya + off),
                                         reinterpret_cast<unsigned*>(&As[buf][q][16 * win][0]), 16, 0, 0);
      else
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned*>(Yb + off),
                                         reinterpret_cast<unsigned*>(&Bs[buf][q][16 * win][0]), 16, 0, 0);
    }
  };

  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fq = lane >> 4;
  auto compute = [&](int buf) __attribute__((always_inline)) {
    V a[NS][FM], b[NS][FN];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[q][i] = *reinterpret_cast<const V*>(&As[buf][q][wm * TM + i * 16 + fr][8 * lds_chunk(fr, fq)]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        b[q][j] = *reinterpret_cast<const V*>(&Bs[buf][q][wn * TN + j * 16 + fr][8 * lds_chunk(fr, fq)]);
    }
    constexpr int QA[3] = {1, 0, 0};
    constexpr int QB[3] = {0, 1, 0};
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) part[i][j] = mfma16x16x32<H>(b[QB[pr]][j], a[QA[pr]][i], part[i][j]);
  };

  const int nslab = g.nkt;
  const int s0 = kc * g.kc_slabs;
  const int s1 = min(nslab, s0 + g.kc_slabs);
  if constexpr (NBUF >= 2) {
    // slab s + D is copied into the buffer of slab s - 1 while slab s is
    // multiplied: D = NBUF - 1 slabs in flight
    constexpr int D = NBUF - 1;
#pragma unroll
    for (int p = 0; p < D; ++p)
      if (s0 + p < s1) issue(s0 + p, p);
    for (int s = s0; s < s1; ++s) {
      const int cur = (s - s0) % NBUF;
      // this wave's copies of slab s have landed once at most the copies of
      // the D - 1 later slabs in flight are outstanding (all, near the end)
      if (s + D - 1 < s1) __builtin_amdgcn_s_waitcnt(waitcnt_vm((D - 1) * LPW));
      else __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();                         // everyone's; and slab s - 1's buffer is free
      if (s + D < s1) issue(s + D, (s + D - s0) % NBUF);
      compute(cur);
    }
  } else {
    for (int s = s0; s < s1; ++s) {
      issue(s, 0);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      compute(0);
      __syncthreads();                         // fragments read before the next copies
    }
  }

  float* __restrict__ P = g.part + ((int64_t)(f * g.tiles + tile) * g.nkc + kc) * (BM * BM);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        P[(wm * TM + i * 16 + fr) + (wn * TN + j * 16 + 4 * fq + r) * BM] = part[i][j][r];
}

// The K-split SYRK (128 x 128 tiles, 4 waves of 64 x 64) with NBUF LDS
// buffers, each its own __shared__ array, and NBUF - 1 slabs copied ahead
// by global_load_lds. A slab is waited for by an explicit vmcnt that leaves
// the later slabs' copies in flight, and the workgroup then meets at a bare
// s_barrier: __syncthreads() would also wait for every outstanding copy,
// and so would the fragment reads of a single array that the copies may
// alias. With one array per buffer and the slab loop unrolled by NBUF, each
// fragment read names its buffer, so it waits only for that buffer's copies.
// Same products in the same order per accumulator as chi_syrk_kernel.
template <bool H, int NBUF>
__global__ void __launch_bounds__(256)
chi_syrk_pipe_kernel(SyrkArgs g) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int BM = 128, NS = 2, WM = 2, NW = 4;
  constexpr int TM = 64, TN = 64, FM = 4, FN = 4;
  constexpr int WINS = BM / 16;                 // 1 KB windows per plane per operand
  constexpr int LPW = 2 * NS * WINS / NW;       // copies per wave per slab
  constexpr int D = NBUF - 1;                   // slabs in flight
  static_assert(NBUF == 3 || NBUF == 4, "LDS buffers");
  static_assert((D - 1) * LPW <= 63, "vmcnt range");
  // buffer u: [operand][plane][row][k]
  __shared__ __attribute__((aligned(16))) E S0[2][NS][BM][KT];
  __shared__ __attribute__((aligned(16))) E S1[2][NS][BM][KT];
  __shared__ __attribute__((aligned(16))) E S2[2][NS][BM][KT];
  __shared__ __attribute__((aligned(16))) E S3[2][NS][BM][KT];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int64_t per_f = (int64_t)g.nkc * g.tiles;
  const int f = (int)(L / per_f);
  const int kc = (int)((L % per_f) / g.tiles);
  const int tile = (int)(L % g.tiles);
  int bm, bn;
  tri_grouped(tile, g.tm, bm, bn);
  const int m0 = bm * BM, n0 = bn * BM;
  const E* __restrict__ Ya = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)m0 * KT;
  const E* __restrict__ Yb = static_cast<const E*>(g.y) + (int64_t)f * g.fstride + (int64_t)n0 * KT;
  const int64_t slab = (int64_t)g.ppad * KT;
  const int rl = lane >> 2;
  const int jl = (lane & 3) ^ ((rl >> 1) & 3);
  const int fr = lane & 15, fq = lane >> 4;

  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) part[i][j] = f4{0.f, 0.f, 0.f, 0.f};

  // the buffers are named directly (no pointer chosen at run time), so
  // every LDS access carries its array's alias scope
#define MGW_PIPE_ISSUE(S_, SB)                                                                         \
  _Pragma("unroll") for (int c = 0; c < LPW; ++c) {                                                     \
    const int w = wave + c * NW;                                                                        \
    const int op = w / (NS * WINS), q = (w / WINS) % NS, win = w % WINS;                                \
    const int64_t off = q * g.plane + (int64_t)(S_) * slab + (int64_t)(16 * win + rl) * KT + 8 * jl;    \
    __builtin_amdgcn_global_load_lds(reinterpret_cast<const unsigned*>((op == 0 ? Ya : Yb) + off),      \
                                     reinterpret_cast<unsigned*>(&SB[op][q][16 * win][0]), 16, 0, 0);   \
  }
#define MGW_PIPE_COMPUTE(SB)                                                                           \
  {                                                                                                     \
    V a[NS][FM], b[NS][FN];                                                                             \
    _Pragma("unroll") for (int q = 0; q < NS; ++q) {                                                    \
      _Pragma("unroll") for (int i = 0; i < FM; ++i)                                                    \
        a[q][i] = *reinterpret_cast<const V*>(&SB[0][q][wm * TM + i * 16 + fr][8 * lds_chunk(fr, fq)]); \
      _Pragma("unroll") for (int j = 0; j < FN; ++j)                                                    \
        b[q][j] = *reinterpret_cast<const V*>(&SB[1][q][wn * TN + j * 16 + fr][8 * lds_chunk(fr, fq)]); \
    }                                                                                                   \
    constexpr int QA[3] = {1, 0, 0};                                                                    \
    constexpr int QB[3] = {0, 1, 0};                                                                    \
    _Pragma("unroll") for (int pr = 0; pr < 3; ++pr)                                                    \
      _Pragma("unroll") for (int i = 0; i < FM; ++i)                                                    \
        _Pragma("unroll") for (int j = 0; j < FN; ++j)                                                  \
          part[i][j] = mfma16x16x32<H>(b[QB[pr]][j], a[QA[pr]][i], part[i][j]);                         \
  }
  // slab ss = s + U sits in SCUR; slab ss + D goes to SNEXT, the buffer of
  // slab ss - 1. This wave's copies of slab ss have landed once at most
  // those of the D - 1 later slabs in flight are outstanding (all, near the
  // end); then the bare barrier: everyone's copies have landed and every
  // wave is done with slab ss - 1. The empty asm statements keep the
  // compiler from moving LDS accesses across it.
#define MGW_PIPE_SUB(U, SCUR, SNEXT)                                                                   \
  if (s + (U) < s1) {                                                                                   \
    if (s + (U) + D - 1 < s1) __builtin_amdgcn_s_waitcnt(waitcnt_vm((D - 1) * LPW));                    \
    else __builtin_amdgcn_s_waitcnt(0);                                                                 \
    __asm__ __volatile__("" ::: "memory");                                                              \
    __builtin_amdgcn_s_barrier();                                                                       \
    __asm__ __volatile__("" ::: "memory");                                                              \
    if (s + (U) + D < s1) { MGW_PIPE_ISSUE(s + (U) + D, SNEXT) }                                        \
    MGW_PIPE_COMPUTE(SCUR)                                                                              \
  }

  const int s0 = kc * g.kc_slabs;
  const int s1 = min(g.nkt, s0 + g.kc_slabs);
  if (s0 < s1) { MGW_PIPE_ISSUE(s0, S0) }
  if (s0 + 1 < s1) { MGW_PIPE_ISSUE(s0 + 1, S1) }
  if constexpr (NBUF == 4) {
    if (s0 + 2 < s1) { MGW_PIPE_ISSUE(s0 + 2, S2) }
  }
  for (int s = s0; s < s1; s += NBUF) {
    if constexpr (NBUF == 3) {
      MGW_PIPE_SUB(0, S0, S2)
      MGW_PIPE_SUB(1, S1, S0)
      MGW_PIPE_SUB(2, S2, S1)
    } else {
      MGW_PIPE_SUB(0, S0, S3)
      MGW_PIPE_SUB(1, S1, S0)
      MGW_PIPE_SUB(2, S2, S1)
      MGW_PIPE_SUB(3, S3, S2)
    }
  }
#undef MGW_PIPE_SUB
#undef MGW_PIPE_COMPUTE
#undef MGW_PIPE_ISSUE

  float* __restrict__ P = g.part + ((int64_t)(f * g.tiles + tile) * g.nkc + kc) * (BM * BM);
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        P[(wm * TM + i * 16 + fr) + (wn * TN + j * 16 + 4 * fq + r) * BM] = part[i][j][r];
}

// SYRK kernel of the K-split planes path: MGW_CHI_SYRK = 0 (register
// prefetch, default), 1 (LDS copies, one buffer, 4 waves), 2 (two buffers,
// 4 waves), 3 (one buffer, 2 waves of 64 x 128), 4 (two buffers, 2 waves),
// 5 (256 x 256 tiles, LDS copies, one buffer, 8 waves of 64 x 128), 6 (the
// same with two buffers), 7 and 8 (128 x 128 tiles, 4 waves, three or four
// LDS buffers: two or three slabs in flight, for the latency of panel reads
// that miss L2). A 256-wide tile reads each row and column panel half as
// often as a 128-wide one. Read at every chi0 build.
int chi_syrk_variant() {
  const char* e = std::getenv("MGW_CHI_SYRK");
  const int v = e ? std::atoi(e) : 0;
  return (v >= 0 && v <= 8) ? v : 0;
}

// Output tile of the planes path's SYRK (two planes, K split): 256 for the
// variants with 256-wide tiles, else 128.
int chi_syrk_bm(int ns, bool ksplit) {
  const int v = chi_syrk_variant();
  return (ns == 2 && ksplit && (v == 5 || v == 6)) ? 256 : 128;
}

int launch_syrk_lds(bool h, int var, dim3 grid, const SyrkArgs& g, hipStream_t stream) {
  constexpr int BM = 128;
  if (var == 5 || var == 6) {
    const bool db = var == 6;
    if (h && db) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, 256, 4, 2, 2>), grid, dim3(512), 0, stream, g);
    else if (h) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, 256, 4, 2, 1>), grid, dim3(512), 0, stream, g);
    else if (db) hipLaunchKernelGGL((chi_syrk_lds_kernel<false, 256, 4, 2, 2>), grid, dim3(512), 0, stream, g);
    else hipLaunchKernelGGL((chi_syrk_lds_kernel<false, 256, 4, 2, 1>), grid, dim3(512), 0, stream, g);
    MGW_HIP(hipGetLastError());
    return MGW_OK;
  }
  if (h) {
    if (var == 1) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, BM, 2, 2, 1>), grid, dim3(256), 0, stream, g);
    else if (var == 2) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, BM, 2, 2, 2>), grid, dim3(256), 0, stream, g);
    else if (var == 3) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, BM, 2, 1, 1>), grid, dim3(128), 0, stream, g);
    else if (var == 4) hipLaunchKernelGGL((chi_syrk_lds_kernel<true, BM, 2, 1, 2>), grid, dim3(128), 0, stream, g);
    else if (var == 7) hipLaunchKernelGGL((chi_syrk_pipe_kernel<true, 3>), grid, dim3(256), 0, stream, g);
    else hipLaunchKernelGGL((chi_syrk_pipe_kernel<true, 4>), grid, dim3(256), 0, stream, g);
  } else {
    if (var == 1) hipLaunchKernelGGL((chi_syrk_lds_kernel<false, BM, 2, 2, 1>), grid, dim3(256), 0, stream, g);
    else if (var == 2) hipLaunchKernelGGL((chi_syrk_lds_kernel<false, BM, 2, 2, 2>), grid, dim3(256), 0, stream, g);
    else if (var == 3) hipLaunchKernelGGL((chi_syrk_lds_kernel<false, BM, 2, 1, 1>), grid, dim3(128), 0, stream, g);
    else if (var == 4) hipLaunchKernelGGL((chi_syrk_lds_kernel<false, BM, 2, 1, 2>), grid, dim3(128), 0, stream, g);
    else if (var == 7) hipLaunchKernelGGL((chi_syrk_pipe_kernel<false, 3>), grid, dim3(256), 0, stream, g);
    else hipLaunchKernelGGL((chi_syrk_pipe_kernel<false, 4>), grid, dim3(256), 0, stream, g);
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

// 1 - chi0 = I + sum over the K chunks of the FP32 partial tiles, summed in
// FP64 in chunk order; lower triangle, zeros above it.
template <int BM>
__global__ void __launch_bounds__(256)
chi_ksplit_reduce_kernel(const float* __restrict__ part, int nkc, int tiles, int tm, int naux,
                         const double* __restrict__ unscale, double* __restrict__ C, int64_t strideC) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  const int tile = blockIdx.y, f = blockIdx.z;
  if (e >= BM * BM) return;
  int bm, bn;
  tri_grouped(tile, tm, bm, bn);
  const int ml = e % BM, nl = e / BM;
  const int m = bm * BM + ml, n = bn * BM + nl;
  double* __restrict__ out = C + (int64_t)f * strideC;
  // zeros above the diagonal, in the same coalesced column order: the whole
  // mirror tile (bn, bm) of an off-diagonal tile, the strict upper part of a
  // diagonal one (every upper element exactly once)
  if (bm != bn) {
    const int mu = bn * BM + ml, nu = bm * BM + nl;
    if (nu < naux) out[mu + (int64_t)nu * naux] = 0.0;
  } else if (m < n && n < naux) {
    out[m + (int64_t)n * naux] = 0.0;
  }
  if (m >= naux || n >= naux || m < n) return;
  const float* __restrict__ p = part + (int64_t)(f * tiles + tile) * nkc * (BM * BM) + e;
  // four chunks' loads in flight, summed in chunk order
  double v = 0.0;
  int c = 0;
  for (; c + 4 <= nkc; c += 4) {
    const float a0 = p[(int64_t)c * (BM * BM)], a1 = p[(int64_t)(c + 1) * (BM * BM)];
    const float a2 = p[(int64_t)(c + 2) * (BM * BM)], a3 = p[(int64_t)(c + 3) * (BM * BM)];
    v += (double)a0;
    v += (double)a1;
    v += (double)a2;
    v += (double)a3;
  }
  for (; c < nkc; ++c) v += (double)p[(int64_t)c * (BM * BM)];
  if (unscale) v *= unscale[f];
  if (m == n) v += 1.0;
  out[m + (int64_t)n * naux] = v;
}

// ---------------------------------------------------------------- fused split
// The two-plane modes (FP16X3, BF16X3) split inside the SYRK instead of in a
// pre-pass per frequency: X is converted once to FP32 in the k-tiled layout
// xf[kt][P][32], and every workgroup (frequency f, tile, K chunk) scales the
// FP32 rows of its slab by w_f(t) = sqrt(s_f(t)) 2^e_f (a per-frequency FP32
// table), splits y = x w into hi = E(y) and lo = E(y - hi) in registers and
// stores the two planes to LDS. The arithmetic is the planes path's (FP32
// product, then the split), so the results are the same bit for bit; what
// goes away is writing and reading two planes per frequency through HBM
// (4 bytes per element and frequency). Workgroups are ordered frequency
// fastest: the frequencies of one tile and K chunk read the same X slabs and
// share them in L2.
__global__ void __launch_bounds__(256)
chi_xf32_kernel(const double* __restrict__ x, int naux, int npole, int ppad, float* __restrict__ xf) {
  __shared__ double tile[KT][65];
  const int tid = threadIdx.x;
  const int p0 = blockIdx.x * 64, kt = blockIdx.y, t0 = kt * KT;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tl = (tid >> 6) + 4 * j, p = tid & 63;
    const int P = p0 + p, t = t0 + tl;
    tile[tl][p] = (P < naux && t < npole) ? x[P + (int64_t)t * naux] : 0.0;
  }
  __syncthreads();
  const int p = tid >> 2, c8 = tid & 3;
  f4 lo4, hi4;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    lo4[e] = (float)tile[8 * c8 + e][p];
    hi4[e] = (float)tile[8 * c8 + 4 + e][p];
  }
  float* __restrict__ dst = xf + ((int64_t)kt * ppad + p0 + p) * KT + 8 * c8;
  *reinterpret_cast<f4*>(dst) = lo4;
  *reinterpret_cast<f4*>(dst + 4) = hi4;
}

// wf[f][t] = sqrt(s_f(t)) scale_f in FP32 (zero for t >= npole, up to nkt * 32);
// with `whole`, s_f(t) scale_f (the one-sided split carries the whole weight)
__global__ void __launch_bounds__(256)
chi_wf_kernel(const double* __restrict__ sw, const double* __restrict__ scale, int npole, int ntpad,
              float* __restrict__ wf, int whole) {
  const int t = blockIdx.x * 256 + threadIdx.x, f = blockIdx.y;
  if (t >= ntpad) return;
  double w = 0.0;
  if (t < npole) {
    const double sv = sw[(int64_t)f * npole + t];
    w = (whole ? sv : sqrt(sv)) * (scale ? scale[f] : 1.0);
  }
  wf[(int64_t)f * ntpad + t] = (float)w;
}

// One-sided split: the frequency-independent operand 2^eA X split once into
// hi and lo planes in the k-tiled layout ap[q][kt][P][32] (zeros in the
// padding); scale (nullable) holds 2^eA.
template <bool H>
__global__ void __launch_bounds__(256)
chi_xplanes_kernel(const double* __restrict__ x, int naux, int npole, int ppad, const double* __restrict__ scale,
                   typename PlaneT<H>::E* __restrict__ ap, int64_t aplane) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  __shared__ double tile[KT][65];
  const int tid = threadIdx.x;
  const int p0 = blockIdx.x * 64, kt = blockIdx.y, t0 = kt * KT;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int tl = (tid >> 6) + 4 * j, p = tid & 63;
    const int P = p0 + p, t = t0 + tl;
    tile[tl][p] = (P < naux && t < npole) ? x[P + (int64_t)t * naux] : 0.0;
  }
  __syncthreads();
  const float sc = scale ? (float)scale[0] : 1.0f;
  const int p = tid >> 2, c8 = tid & 3;
  V hi, lo;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float y = (float)tile[8 * c8 + e][p] * sc;
    const E h = (E)y;
    hi[e] = h;
    lo[e] = (E)(y - (float)h);
  }
  E* __restrict__ dst = ap + ((int64_t)kt * ppad + p0 + p) * KT + 8 * c8;
  *reinterpret_cast<V*>(dst) = hi;
  *reinterpret_cast<V*>(dst + aplane) = lo;
}

// Scales of the one-sided split (FP16X3): 2^eA brings max|X| into
// [2^13, 2^14) (scale[nomega] holds it); 2^eB_f brings the bound
// max|X| max_t s_f(t) of X diag(s_f) there; unscale_f = 2^-(eA + eB_f) up to
// the 2^26 the two targets carry.
__global__ void __launch_bounds__(256)
chi_scale_half_kernel(const double* __restrict__ xpart, int nb, const double* __restrict__ sw, int npole,
                      int nomega, double* __restrict__ scale, double* __restrict__ unscale) {
  __shared__ double red[256];
  const int f = blockIdx.x, tid = threadIdx.x;
  const double* __restrict__ s = sw + (int64_t)f * npole;
  double m = 0.0;
  for (int t = tid; t < npole; t += 256) m = fmax(m, s[t]);
  double xm = 0.0;
  for (int b = tid; b < nb; b += 256) xm = fmax(xm, xpart[b]);
  red[tid] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  const double smax = red[0];
  __syncthreads();
  red[tid] = xm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) {
    const double xb = red[0], bb = red[0] * smax;
    const int ea = (xb > 0.0 && xb <= 1.0e300) ? ilogb(xb) : 0;
    const int eb = (bb > 0.0 && bb <= 1.0e300) ? ilogb(bb) : 0;
    scale[f] = ldexp(1.0, 13 - eb);
    unscale[f] = ldexp(1.0, ea + eb - 26);
    if (f == 0) scale[nomega] = ldexp(1.0, 13 - ea);
  }
}

struct FusedArgs {
  const float* xf;       // [kt][ppad][32]
  const float* wf;       // [f][nkt * 32]
  const void* ap;        // one-sided split: A planes [q][kt][ppad][32] (bf16 or fp16), else null
  int64_t aplane;        // elements between the A planes
  int ppad, nkt, naux, nf, tiles, tm;
  double* C;             // chi + f0 * naux^2
  int64_t strideC;
  int64_t nblocks;
  int kc_slabs, nkc;     // K split (nkc == 1: none)
  float* part;
  const double* unscale; // per frequency of the chunk (FP16X3), else null
};

// HALF: the one-sided split. The A operand (rows m) comes from the
// frequency-independent planes of 2^eA X (copied to LDS as they are); only
// the B operand (rows n) is scaled by 2^eB_f s_f(t) and split in the kernel:
// y_m y_n = x_m (x_n s_f), half the VALU work of the two-sided split and the
// same two-plane accuracy.
template <bool H, int BM, int WM, int WN, int FLUSH, bool KSPLIT, bool HALF = false>
__global__ void __launch_bounds__(64 * WM * WN, KSPLIT ? 2 : 1)
chi_fsyrk_kernel(FusedArgs g) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int NS = 2;
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BM / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int CH = BM * KT / 8;         // 8-k chunks per operand slab
  constexpr int LPT = CH / NT;
  static_assert(CH % NT == 0 && (NT % 4) == 0, "chunks per thread");
  __shared__ __attribute__((aligned(16))) E As[NS][BM][KT];
  __shared__ __attribute__((aligned(16))) E Bs[NS][BM][KT];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int f = (int)(L % g.nf);
  const int64_t rest = L / g.nf;
  const int tile = (int)(rest % g.tiles);
  const int kc = KSPLIT ? (int)(rest / g.tiles) : 0;
  int bm, bn;
  tri_grouped(tile, g.tm, bm, bn);
  const int m0 = bm * BM, n0 = bn * BM;
  const float* __restrict__ Xa = g.xf + (int64_t)m0 * KT;
  const float* __restrict__ Xb = g.xf + (int64_t)n0 * KT;
  const float* __restrict__ W = g.wf + (int64_t)f * g.nkt * KT + 8 * (tid & 3);
  const int64_t slab = (int64_t)g.ppad * KT;

  f4 ra[LPT][2], rb[LPT][2], rw[2];
  u4 rp[NS][LPT];                          // HALF: the A planes' chunks
  const E* __restrict__ Pa = static_cast<const E*>(g.ap) + (int64_t)m0 * KT;
  auto load = [&](int s) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int c = tid + i * NT;
      const int64_t off = (int64_t)s * slab + (c >> 2) * KT + 8 * (c & 3);
      if constexpr (HALF) {
#pragma unroll
        for (int q = 0; q < NS; ++q) rp[q][i] = *reinterpret_cast<const u4*>(Pa + q * g.aplane + off);
      } else {
        ra[i][0] = *reinterpret_cast<const f4*>(Xa + off);
        ra[i][1] = *reinterpret_cast<const f4*>(Xa + off + 4);
      }
      rb[i][0] = *reinterpret_cast<const f4*>(Xb + off);
      rb[i][1] = *reinterpret_cast<const f4*>(Xb + off + 4);
    }
    rw[0] = *reinterpret_cast<const f4*>(W + (int64_t)s * KT);
    rw[1] = *reinterpret_cast<const f4*>(W + (int64_t)s * KT + 4);
  };
  auto copy_store = [&](E (&S)[NS][BM][KT]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int c = tid + i * NT, row = c >> 2, j = c & 3;
#pragma unroll
      for (int q = 0; q < NS; ++q) *reinterpret_cast<u4*>(&S[q][row][8 * lds_chunk(row, j)]) = rp[q][i];
    }
  };
  auto split_store = [&](f4 (&r)[LPT][2], E (&S)[NS][BM][KT]) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < LPT; ++i) {
      const int c = tid + i * NT, row = c >> 2, j = c & 3;
      V hi, lo;
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float y = r[i][e >> 2][e & 3] * rw[e >> 2][e & 3];
        const E h = (E)y;
        hi[e] = h;
        lo[e] = (E)(y - (float)h);
      }
      *reinterpret_cast<V*>(&S[0][row][8 * lds_chunk(row, j)]) = hi;
      *reinterpret_cast<V*>(&S[1][row][8 * lds_chunk(row, j)]) = lo;
    }
  };

  double acc[FM][FN][4];
  f4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      part[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;
    }
  const int fr = lane & 15, fq = lane >> 4;
  const int nslab = g.nkt;
  const int s0 = KSPLIT ? kc * g.kc_slabs : 0;
  const int s1 = KSPLIT ? min(nslab, s0 + g.kc_slabs) : nslab;
  load(s0);
  for (int s = s0; s < s1; ++s) {
    __syncthreads();                   // previous slab's fragments consumed
    if constexpr (HALF) copy_store(As);
    else split_store(ra, As);
    split_store(rb, Bs);
    __syncthreads();
    if (s + 1 < s1) load(s + 1);
    V a[NS][FM], b[NS][FN];
#pragma unroll
    for (int q = 0; q < NS; ++q) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
        a[q][i] = *reinterpret_cast<const V*>(&As[q][wm * TM + i * 16 + fr][8 * lds_chunk(fr, fq)]);
#pragma unroll
      for (int j = 0; j < FN; ++j)
        b[q][j] = *reinterpret_cast<const V*>(&Bs[q][wn * TN + j * 16 + fr][8 * lds_chunk(fr, fq)]);
    }
    // hi lo + lo hi + hi hi, product-major (small products first)
#pragma unroll
    for (int pr = 0; pr < 3; ++pr)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int qa = pr == 0 ? 1 : 0, qb = pr == 1 ? 1 : 0;
          part[i][j] = mfma16x16x32<H>(b[qb][j], a[qa][i], part[i][j]);
        }
    if (!KSPLIT && ((s + 1) % FLUSH == 0 || s + 1 == nslab)) {
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

  if constexpr (KSPLIT) {
    float* __restrict__ P = g.part + ((int64_t)(f * g.tiles + tile) * g.nkc + kc) * (BM * BM);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          P[(wm * TM + i * 16 + fr) + (wn * TN + j * 16 + 4 * fq + r) * BM] = part[i][j][r];
    return;
  }
  double* __restrict__ C = g.C + (int64_t)f * g.strideC;
  const double us = g.unscale ? g.unscale[f] : 1.0;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
        if (m >= g.naux || n >= g.naux || m < n) continue;
        double v = acc[i][j][r] * us;
        if (m == n) v += 1.0;
        C[m + (int64_t)n * g.naux] = v;
        if (m > n) C[n + (int64_t)m * g.naux] = 0.0;
      }
  }
}

// Pipeline of the SYRK. Measured on MI355X (1200/3600/120, 32 frequencies):
// two LDS buffers for two planes (BF16x3, 731 vs 850 ms with two register
// sets), two register sets for three planes (BF16x6, 1392 vs 1419 ms).
// MGW_CHI_PREFETCH = 1 (one register set), 2 or 3 overrides the choice.
int chi_prefetch_depth(int ns) {
  static const int forced = [] {
    const char* v = std::getenv("MGW_CHI_PREFETCH");
    const int d = v ? std::atoi(v) : 0;
    return (d >= 1 && d <= 3) ? d : 0;
  }();
  return forced ? forced : (ns == 2 ? 3 : 2);
}

// k tiles of 32 per SYRK slab: MGW_CHI_KS = 1 or 2 (default 1; two tiles do
// not fit two LDS buffers).
int chi_k_tiles(int ns) {
  static const int forced = [] {
    const char* v = std::getenv("MGW_CHI_KS");
    const int k = v ? std::atoi(v) : 0;
    return (k == 1 || k == 2) ? k : 0;
  }();
  (void)ns;
  const int k = forced ? forced : 1;
  return (k == 2 && chi_prefetch_depth(ns) == 3) ? 1 : k;
}

// K split of the SYRK (FP32 partial tiles summed in FP64 by a second kernel):
// frees the FP64 accumulators, so two workgroups share a CU, and multiplies
// the workgroups of small problems. MGW_CHI_KSPLIT = 0 or 1 overrides the
// default (on).
bool chi_ksplit() {
  static const int forced = [] {
    const char* v = std::getenv("MGW_CHI_KSPLIT");
    return v ? std::atoi(v) : -1;
  }();
  return forced != 0;
}

// K chunk of the split SYRK: 64 slabs (2048 transitions). Larger chunks
// write fewer FP32 partial tiles but accumulate longer in FP32; measured at
// 1200/3600/120 (FP16x3), 256-slab chunks were slower (fewer workgroups per
// launch) and doubled the QP error. MGW_CHI_KC (slabs) overrides it.
constexpr int KC_SLABS = 64;

int chi_kc_slabs() {
  static const int forced = [] {
    const char* v = std::getenv("MGW_CHI_KC");
    const int k = v ? std::atoi(v) : 0;
    return k >= 1 ? k : 0;
  }();
  return forced ? forced : KC_SLABS;
}

struct SplitLayout {
  int bm, ppad, nkt, tm, tiles, kc_slabs, nkc;
  int64_t per_plane_f;   // elements of one plane of one frequency
  size_t planes_per_f;   // bytes of all planes of one frequency
  size_t part_per_f;     // bytes of the FP32 partial tiles of one frequency
};

// Largest K chunk (slabs) that the grid-filling rule of split_layout may
// choose: MGW_CHI_KC_CAP (read at every chi0 build), default 128 slabs
// (4096 transitions). Measured on MI355X over the serving stream (FP16x3,
// one request at a time): chi0 12.53 ms per solve with 128, against
// 12.87-12.93 with 64 and 12.72 with 256, at the same QP error.
constexpr int KC_CAP_DEFAULT = 128;
int chi_kc_cap() {
  const char* v = std::getenv("MGW_CHI_KC_CAP");
  const int k = v ? std::atoi(v) : KC_CAP_DEFAULT;
  return k >= KC_SLABS ? k : KC_SLABS;
}

SplitLayout split_layout(int naux, int npole, int ns, bool ksplit, int nomega = 1) {
  const int bm = chi_syrk_bm(ns, ksplit);
  constexpr int64_t FILL = 4096;   // workgroups that keep 256 CUs busy at 2-3 per CU, several waves deep
  SplitLayout L;
  L.bm = bm;
  // rows padded to whole tiles: the SYRK reads full BM-row panels
  const int pad = std::max(PAD_ROWS, bm);
  L.ppad = (naux + pad - 1) / pad * pad;
  L.nkt = ((npole + KT - 1) / KT + 1) / 2 * 2;   // even: slabs of one or two k tiles
  L.tm = (naux + bm - 1) / bm;
  L.tiles = L.tm * (L.tm + 1) / 2;
  L.kc_slabs = chi_kc_slabs();
  if (!std::getenv("MGW_CHI_KC")) {
    // the fewest K chunks that still fill the device: each chunk costs an
    // FP32 partial tile written and read back
    const int64_t base = (int64_t)L.tiles * std::max(1, nomega);
    const int64_t want = std::max<int64_t>(1, (FILL + base - 1) / base);
    const int kc = (int)std::max<int64_t>(KC_SLABS, (L.nkt + want - 1) / want);
    L.kc_slabs = std::min(kc, chi_kc_cap());
  }
  L.nkc = (L.nkt + L.kc_slabs - 1) / L.kc_slabs;
  L.per_plane_f = (int64_t)L.nkt * L.ppad * KT;
  L.planes_per_f = (size_t)L.per_plane_f * ns * sizeof(__bf16);
  L.part_per_f = ksplit ? (size_t)L.tiles * L.nkc * bm * bm * sizeof(float) : 0;
  return L;
}

// bytes of the FP32 k-tiled copy of X that the two-plane split reads
size_t split_xf_bytes(const SplitLayout& L, int ns) {
  return ns == 2 ? ((size_t)L.nkt * L.ppad * KT * sizeof(float) + 255) / 256 * 256 : 0;
}

// ---------------------------------------------------------------- scales
constexpr int NB_MAX = ABSMAX_BLOCKS;   // blocks of the |X| max reduction

// Power-of-two scale of the fp16 planes of Y_f = X diag(sqrt(s_f)): the bound
// max|X| sqrt(max_t s_f(t)) >= max|Y_f| is brought into [2^13, 2^14); the
// product is unscaled by the exact inverse square.
__global__ void __launch_bounds__(256)
chi_scale_kernel(const double* __restrict__ xpart, int nb, const double* __restrict__ sw, int npole,
                 double* __restrict__ scale, double* __restrict__ unscale) {
  __shared__ double red[256];
  const int f = blockIdx.x, tid = threadIdx.x;
  const double* __restrict__ s = sw + (int64_t)f * npole;
  double m = 0.0;
  for (int t = tid; t < npole; t += 256) m = fmax(m, s[t]);
  double xm = 0.0;
  for (int b = tid; b < nb; b += 256) xm = fmax(xm, xpart[b]);
  red[tid] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  const double smax = red[0];
  __syncthreads();
  red[tid] = xm;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) {
    const double bound = red[0] * sqrt(smax);
    const int e = (bound > 0.0 && bound <= 1.0e300) ? ilogb(bound) : 0;   // bound in [2^e, 2^(e+1))
    scale[f] = ldexp(1.0, 13 - e);
    unscale[f] = ldexp(1.0, 2 * (e - 13));
  }
}

// bytes in front of the planes: scale, unscale (per frequency), |X| partial maxima
size_t scale_header_bytes(int nomega) { return ((size_t)(2 * nomega + NB_MAX) * sizeof(double) + 255) / 256 * 256; }

template <int NS, bool H>
int split_and_syrk(const double* x, int naux, int npole, const double* sw, int nomega, void* work,
                   size_t work_bytes, double* chi, hipStream_t stream) {
  typedef typename PlaneT<H>::E E;
  constexpr int BM = 128;
  const bool ksplit = chi_ksplit();
  const SplitLayout L = split_layout(naux, npole, NS, ksplit, nomega);
  const size_t per_f = L.planes_per_f + L.part_per_f;
  const size_t head = scale_header_bytes(nomega);
  const size_t xf_bytes = split_xf_bytes(L, NS);
  MGW_REQUIRE(work_bytes > head + xf_bytes + 256, MGW_ERR_ARG, "chi0 split: workspace too small");
  const int nf_max = (int)std::min<size_t>((size_t)nomega, (work_bytes - head - xf_bytes - 256) / per_f);
  MGW_REQUIRE(nf_max >= 1, MGW_ERR_ARG, "chi0 split: workspace smaller than one frequency");
  double* scl = static_cast<double*>(work);
  double* uns = scl + nomega;
  double* xpart = uns + nomega;
  E* planes = reinterpret_cast<E*>(static_cast<char*>(work) + head);
  float* part = reinterpret_cast<float*>(static_cast<char*>(work) + head +
                                         (L.planes_per_f * nf_max + 255) / 256 * 256);
  // the FP32 k-tiled copy of X behind the partial tiles (two planes only)
  float* xf = reinterpret_cast<float*>(static_cast<char*>(work) + head +
                                       (L.planes_per_f * nf_max + 255) / 256 * 256 +
                                       (L.part_per_f * nf_max + 255) / 256 * 256);
  if (H) {
    MGW_TRY(launch_absmax(x, (int64_t)naux * npole, xpart, stream));
    hipLaunchKernelGGL(chi_scale_kernel, dim3(nomega), dim3(256), 0, stream, xpart, NB_MAX, sw, npole, scl, uns);
    MGW_HIP(hipGetLastError());
  }
  if (NS == 2) {
    hipLaunchKernelGGL(chi_xf32_kernel, dim3(L.ppad / 64, L.nkt), dim3(256), 0, stream, x, naux, npole, L.ppad, xf);
    MGW_HIP(hipGetLastError());
  }
  for (int f0 = 0; f0 < nomega; f0 += nf_max) {
    const int nf = std::min(nf_max, nomega - f0);
    const int64_t plane = L.per_plane_f * nf;
    if constexpr (NS == 2)
      hipLaunchKernelGGL((chi_split_xf_kernel<H>), dim3(L.ppad / 64, L.nkt), dim3(256), 0, stream, xf, npole, sw,
                         H ? scl : nullptr, f0, nf, L.ppad, L.nkt, planes, plane);
    else
      hipLaunchKernelGGL((chi_split_kernel<NS, H>), dim3(L.ppad / 64, L.nkt), dim3(256), 0, stream, x, naux, npole,
                         sw, H ? scl : nullptr, f0, nf, L.ppad, L.nkt, planes, plane);
    MGW_HIP(hipGetLastError());
    SyrkArgs g;
    g.y = planes; g.plane = plane; g.fstride = L.per_plane_f;
    g.ppad = L.ppad; g.nkt = L.nkt; g.naux = naux; g.nf = nf; g.tiles = L.tiles; g.tm = L.tm;
    g.C = chi + (int64_t)f0 * naux * naux; g.strideC = (int64_t)naux * naux;
    g.kc_slabs = L.kc_slabs; g.nkc = L.nkc; g.part = part;
    g.unscale = H ? uns + f0 : nullptr;
    const int depth = chi_prefetch_depth(NS);
    const dim3 block(256);
    if (ksplit) {
      g.nblocks = (int64_t)L.tiles * L.nkc * nf;
      MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "chi0 split: too many tiles");
      const dim3 grid((unsigned)g.nblocks);
      // one LDS buffer and one register set: 162 (two planes) / 196 (three)
      // VGPRs, two or three waves per SIMD; a second register set spills
      // (two LDS buffers with one barrier per slab, or two k tiles per slab:
      // 64 KB of LDS and two waves per SIMD, measured slower)
      const int var = NS == 2 ? chi_syrk_variant() : 0;
      MGW_REQUIRE((L.bm == 256) == (var == 5 || var == 6), MGW_ERR_ARG,
                  "chi0 split: tile size and SYRK variant disagree");
      auto launch = [&](int v) -> int {
        if (v) return launch_syrk_lds(H, v, grid, g, stream);
        hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 4, 1, 1, true>), grid, block, 0, stream, g);
        MGW_HIP(hipGetLastError());
        return MGW_OK;
      };
      if (NS == 2 && !std::getenv("MGW_CHI_SYRK")) {
        // the tuner picks among the 128-wide variants (same products in the
        // same order: identical results) per shape class
        static constexpr int CAND[7] = {0, 1, 2, 3, 4, 7, 8};
        MGW_TRY(autotune_select(MGW_TUNE_CHI_SYRK, NS, H ? 1 : 0, log2_floor(naux), log2_floor(L.nkt), log2_floor(nf),
                                7, 0, [&](int c) { return launch(CAND[c]); }, stream));
      } else {
        MGW_TRY(launch(var));
      }
      if (L.bm == 256)
        hipLaunchKernelGGL((chi_ksplit_reduce_kernel<256>), dim3(256 * 256 / 256, L.tiles, nf), block, 0, stream,
                           part, L.nkc, L.tiles, L.tm, naux, g.unscale, g.C, g.strideC);
      else
        hipLaunchKernelGGL((chi_ksplit_reduce_kernel<BM>), dim3(BM * BM / 256, L.tiles, nf), block, 0, stream, part,
                           L.nkc, L.tiles, L.tm, naux, g.unscale, g.C, g.strideC);
      MGW_HIP(hipGetLastError());
      continue;
    }
    g.nblocks = (int64_t)L.tiles * nf;
    MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "chi0 split: too many tiles");
    const int ks = chi_k_tiles(NS);
    const dim3 grid((unsigned)g.nblocks);
    if (ks == 2) {
      if (depth == 1) hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 2, 1, 2, false>), grid, block, 0, stream, g);
      else hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 2, 2, 2, false>), grid, block, 0, stream, g);
    } else {
      if (depth == 1) hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 4, 1, 1, false>), grid, block, 0, stream, g);
      else if (depth == 2) hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 4, 2, 1, false>), grid, block, 0, stream, g);
      else hipLaunchKernelGGL((chi_syrk_kernel<NS, H, BM, 2, 2, 4, 3, 1, false>), grid, block, 0, stream, g);
    }
    MGW_HIP(hipGetLastError());
  }
  return MGW_OK;
}

// Fused split (two planes): layout and K chunking. The K chunk keeps the
// FP32 accumulation short (at most KC_MAX slabs) and the grid large enough to
// fill the device; a problem with enough (tile, frequency) workgroups runs
// without a K split and with the FP64 flush every FUSED_FLUSH slabs, which
// writes no partial tiles. MGW_CHI_FUSED_KSPLIT = 0 / 1 forces one of the two.
constexpr int FUSED_FLUSH = 4;
constexpr int KC_MAX = 256;
constexpr int64_t FILL_BLOCKS = 2048;     // 256 CUs x 8 workgroups

struct FusedLayout {
  int ppad, nkt, tm, tiles, kc_slabs, nkc;
  bool ksplit, half;
  size_t head, xf_bytes, wf_bytes, ap_bytes, part_per_f;
};

// read at every call (a getenv per chi0 build), so that a process can compare the variants
int fused_ksplit_forced() {
  const char* e = std::getenv("MGW_CHI_FUSED_KSPLIT");
  return e ? std::atoi(e) : -1;
}

FusedLayout fused_layout(int naux, int npole, int nomega, bool half) {
  constexpr int bm = 128;
  FusedLayout L;
  L.half = half;
  L.head = scale_header_bytes(nomega + 1);   // scales and unscales of nomega + 1 (the one-sided A scale)
  L.ppad = (naux + PAD_ROWS - 1) / PAD_ROWS * PAD_ROWS;
  L.nkt = (npole + KT - 1) / KT;
  L.tm = (naux + bm - 1) / bm;
  L.tiles = L.tm * (L.tm + 1) / 2;
  const int64_t base = (int64_t)L.tiles * nomega;
  const int forced = fused_ksplit_forced();
  L.ksplit = forced >= 0 ? forced != 0 : base < FILL_BLOCKS;
  if (L.ksplit) {
    const int64_t want = (FILL_BLOCKS + base - 1) / base;
    int kc = (int)std::max<int64_t>(KC_SLABS, (L.nkt + want - 1) / want);
    kc = std::min(kc, KC_MAX);
    L.kc_slabs = kc;
    L.nkc = (L.nkt + kc - 1) / kc;
  } else {
    L.kc_slabs = L.nkt;
    L.nkc = 1;
  }
  L.xf_bytes = ((size_t)L.nkt * L.ppad * KT * sizeof(float) + 255) / 256 * 256;
  L.wf_bytes = ((size_t)nomega * L.nkt * KT * sizeof(float) + 255) / 256 * 256;
  L.ap_bytes = half ? ((size_t)2 * L.nkt * L.ppad * KT * sizeof(_Float16) + 255) / 256 * 256 : 0;
  L.part_per_f = L.ksplit ? (size_t)L.tiles * L.nkc * bm * bm * sizeof(float) : 0;
  return L;
}

// The split of the two-plane modes: MGW_CHI_FUSED = 0 per-frequency planes
// (pre-pass), 1 both operands split in the SYRK, 2 one-sided (A planes split
// once, B scaled and split in the SYRK). Read at every chi0 build. Measured
// on MI355X over the serving stream (real molecules, one solve at a time),
// the chi0 stage took 22.3 ms per solve with both operands split in the
// kernel against 13.5 ms with the planes: the in-kernel scale and split of
// both operands of every (tile, frequency) workgroup costs more VALU time
// than the plane traffic it saves.
int fused_mode() {
  const char* e = std::getenv("MGW_CHI_FUSED");
  const int v = e ? std::atoi(e) : 0;
  return (v >= 0 && v <= 2) ? v : 0;
}

bool fused_enabled() { return fused_mode() != 0; }

size_t fused_workspace(int naux, int npole, int nomega) {
  const FusedLayout L = fused_layout(naux, npole, nomega, fused_mode() == 2);
  size_t nf = (size_t)nomega;
  if (L.part_per_f) {
    const size_t budget = std::max<size_t>((size_t)32 * naux * npole, (size_t)2 << 30);
    nf = std::max<size_t>(1, std::min<size_t>((size_t)nomega, budget / L.part_per_f));
  }
  return L.head + L.xf_bytes + L.wf_bytes + L.ap_bytes + nf * L.part_per_f + 512;
}

template <bool H, bool HALF>
int fused_split_syrk(const double* x, int naux, int npole, const double* sw, int nomega, void* work,
                     size_t work_bytes, double* chi, hipStream_t stream) {
  typedef typename PlaneT<H>::E E;
  constexpr int BM = 128;
  const FusedLayout L = fused_layout(naux, npole, nomega, HALF);
  const size_t fixed = L.head + L.xf_bytes + L.wf_bytes + L.ap_bytes;
  MGW_REQUIRE(work_bytes >= fixed + 256, MGW_ERR_ARG, "chi0 fused: workspace too small");
  size_t nf_max = (size_t)nomega;
  if (L.part_per_f) nf_max = std::min(nf_max, (work_bytes - fixed - 256) / L.part_per_f);
  MGW_REQUIRE(nf_max >= 1, MGW_ERR_ARG, "chi0 fused: workspace smaller than one frequency");
  double* scl = static_cast<double*>(work);      // nomega + 1
  double* uns = scl + nomega + 1;
  double* xpart = uns + nomega + 1;
  char* w8 = static_cast<char*>(work);
  float* xf = reinterpret_cast<float*>(w8 + L.head);
  float* wf = reinterpret_cast<float*>(w8 + L.head + L.xf_bytes);
  E* ap = reinterpret_cast<E*>(w8 + L.head + L.xf_bytes + L.wf_bytes);
  float* part = reinterpret_cast<float*>(w8 + fixed);
  const int64_t aplane = (int64_t)L.nkt * L.ppad * KT;
  if (H) {
    MGW_TRY(launch_absmax(x, (int64_t)naux * npole, xpart, stream));
    if (HALF)
      hipLaunchKernelGGL(chi_scale_half_kernel, dim3(nomega), dim3(256), 0, stream, xpart, NB_MAX, sw, npole, nomega,
                         scl, uns);
    else
      hipLaunchKernelGGL(chi_scale_kernel, dim3(nomega), dim3(256), 0, stream, xpart, NB_MAX, sw, npole, scl, uns);
    MGW_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL(chi_xf32_kernel, dim3(L.ppad / 64, L.nkt), dim3(256), 0, stream, x, naux, npole, L.ppad, xf);
  MGW_HIP(hipGetLastError());
  if (HALF) {
    hipLaunchKernelGGL((chi_xplanes_kernel<H>), dim3(L.ppad / 64, L.nkt), dim3(256), 0, stream, x, naux, npole,
                       L.ppad, H ? scl + nomega : nullptr, ap, aplane);
    MGW_HIP(hipGetLastError());
  }
  const int ntpad = L.nkt * KT;
  hipLaunchKernelGGL(chi_wf_kernel, dim3((ntpad + 255) / 256, nomega), dim3(256), 0, stream, sw,
                     H ? scl : nullptr, npole, ntpad, wf, HALF ? 1 : 0);
  MGW_HIP(hipGetLastError());
  for (int f0 = 0; f0 < nomega; f0 += (int)nf_max) {
    const int nf = std::min((int)nf_max, nomega - f0);
    FusedArgs g;
    g.xf = xf;
    g.wf = wf + (int64_t)f0 * ntpad;
    g.ap = HALF ? ap : nullptr;
    g.aplane = aplane;
    g.ppad = L.ppad; g.nkt = L.nkt; g.naux = naux; g.nf = nf; g.tiles = L.tiles; g.tm = L.tm;
    g.C = chi + (int64_t)f0 * naux * naux; g.strideC = (int64_t)naux * naux;
    g.kc_slabs = L.kc_slabs; g.nkc = L.nkc; g.part = part;
    g.unscale = H ? uns + f0 : nullptr;
    g.nblocks = (int64_t)L.tiles * L.nkc * nf;
    MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "chi0 fused: too many tiles");
    const dim3 grid((unsigned)g.nblocks), block(256);
    if (L.ksplit) {
      hipLaunchKernelGGL((chi_fsyrk_kernel<H, BM, 2, 2, FUSED_FLUSH, true, HALF>), grid, block, 0, stream, g);
      MGW_HIP(hipGetLastError());
      hipLaunchKernelGGL((chi_ksplit_reduce_kernel<BM>), dim3(BM * BM / 256, L.tiles, nf), block, 0, stream, part,
                         L.nkc, L.tiles, L.tm, naux, g.unscale, g.C, g.strideC);
    } else {
      hipLaunchKernelGGL((chi_fsyrk_kernel<H, BM, 2, 2, FUSED_FLUSH, false, HALF>), grid, block, 0, stream, g);
    }
    MGW_HIP(hipGetLastError());
  }
  return MGW_OK;
}

}  // namespace

size_t chi0_split_bytes_per_frequency(int naux, int npole, int precision, int nomega) {
  const SplitLayout L = split_layout(naux, npole, precision == MGW_BF16X6 ? 3 : 2, chi_ksplit(), nomega);
  return L.planes_per_f + L.part_per_f;
}

// Workspace of the split planes (and K-split partial tiles): all frequencies
// when they fit in max(4 x the FP64 operand, 2 GiB), else as many as fit (at
// least one).
size_t chi0_split_workspace(int naux, int npole, int nomega, int precision) {
  if (precision != MGW_BF16X6 && fused_enabled()) return fused_workspace(naux, npole, nomega);
  const int ns = precision == MGW_BF16X6 ? 3 : 2;
  const SplitLayout L = split_layout(naux, npole, ns, chi_ksplit(), nomega);
  const size_t per_f = L.planes_per_f + L.part_per_f;
  const size_t budget = std::max<size_t>((size_t)32 * naux * npole, (size_t)2 << 30);
  const size_t nf = std::max<size_t>(1, std::min<size_t>((size_t)nomega, budget / per_f));
  return scale_header_bytes(nomega) + nf * per_f + split_xf_bytes(L, ns) + 1024;
}

int chi0_split_syrk(const double* x, int naux, int npole, const double* sw, int nomega, int precision,
                    void* work, size_t work_bytes, double* chi, hipStream_t stream) {
  if (naux <= 0 || nomega <= 0) return MGW_OK;
  MGW_REQUIRE(npole > 0, MGW_ERR_ARG, "chi0 split: empty transition space");
  const int fm = fused_mode();
  if (precision == MGW_BF16X3 && fm == 1)
    return fused_split_syrk<false, false>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_FP16X3 && fm == 1)
    return fused_split_syrk<true, false>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_BF16X3 && fm == 2)
    return fused_split_syrk<false, true>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_FP16X3 && fm == 2)
    return fused_split_syrk<true, true>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_BF16X3)
    return split_and_syrk<2, false>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_BF16X6)
    return split_and_syrk<3, false>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  if (precision == MGW_FP16X3)
    return split_and_syrk<2, true>(x, naux, npole, sw, nomega, work, work_bytes, chi, stream);
  set_error("chi0 split: precision must be BF16X3, BF16X6 or FP16X3");
  return MGW_ERR_ARG;
}

}  // namespace mgw
