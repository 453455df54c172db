// Reduced-precision self-energy product on the imaginary axis.
//
// For every frequency w_f of the chi grid and every column n = (p, m) of
// B = (P|pm) (naux x ncol, column-major) the self-energy needs
//   v(n, f) = B_n^T chi(iw_f) B_n       (src/m_gw_selfenergy_grid.f90:297-310).
// With 1 - chi0 = L L^T and D_f = L_f^-1 - 1 (lower triangular, resident),
//   chi = L^-T L^-1 - 1,  v = ||D_f B_n||^2 + 2 B_n^T D_f B_n,
// so C_f = D_f B is the only product, and v(n, f) = sum_m C (C + 2 B) over
// the rows m. Its error is relative to the screening part D_f alone (no
// cancellation against ||B_n||^2), which is what lets the product run below
// FP64.
//
// A pre-pass splits D_f and B^T into NS bf16 planes x = x0 + x1 (+ x2), or
// two fp16 planes of power-of-two-scaled operands (FP16X3: D_f times 2^13,
// since |D_f(i, j)| <= ||L_f^-1||_2 <= 1, and B times 2^(13 - e) with
// 2^e <= max|B| < 2^(e+1)),
// k-tiled as plane[kt][row][32] (rows padded to 128, k = aux index padded to
// 32, zeros in the padding and, for D_f, above the diagonal). The product
// kernel multiplies the planes on v_mfma_f32_16x16x32_bf16 / _f16 (NS = 2: three
// products, NS = 3: six, as the chi0 SYRK), stops the K loop of a row block at
// its last row (D_f is lower triangular), accumulates in FP32 MFMA registers
// (NS = 2: over the whole K, at most naux terms; NS = 3: flushed into FP64
// every FLUSH slabs), and reduces C (C + 2 B) over the rows of the block in
// its epilogue (B in FP64): out[f][row block][n], summed over the row blocks
// by dot_reduce_kernel.
//
// Fragment layout as in the chi0 SYRK: lane l, register r holds
// C(m = l & 15, n = 4 (l >> 4) + r) of each 16 x 16 tile.
#include "mgw_internal.h"
#include "tile_order.h"

#include <algorithm>
#include <cstdlib>
#include <type_traits>

namespace mgw {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));

template <bool H> struct PlaneT { typedef __bf16 E; typedef bf8 V; };
template <> struct PlaneT<true> { typedef _Float16 E; typedef h8 V; };

template <bool H>
__device__ inline f4 mfma16x16x32(typename PlaneT<H>::V b, typename PlaneT<H>::V a, f4 c) {
  if constexpr (H) return __builtin_amdgcn_mfma_f32_16x16x32_f16(b, a, c, 0, 0, 0);
  else return __builtin_amdgcn_mfma_f32_16x16x32_bf16(b, a, c, 0, 0, 0);
}

constexpr double D_SCALE = 8192.0;   // 2^13: |D_f| <= 1

constexpr int KT = 32;    // k per tile of the plane layout (one MFMA step)

// LDS rows of a slab are 64 bytes (32 k) without padding; 16-byte chunk j of
// row r is stored at chunk j ^ ((r >> 1) & 3). Eight consecutive rows read at
// one chunk index (the MFMA fragment loads), and the four chunks of two rows
// written by eight consecutive lanes (the slab stores), then fall on eight
// distinct 16-byte bank groups. Rows of a fragment start at multiples of 16,
// so (fr >> 1) & 3 gives the swizzle of row base + fr.
__device__ inline int lds_chunk(int r, int j) { return j ^ ((r >> 1) & 3); }
constexpr int PAD = 128;  // row padding of the plane layout (the tile size)

// ---------------------------------------------------------------- split
// Rows of D_f (k = column index, lower triangle only): block of 64 rows x 32
// columns of one frequency, read coalesced along the rows through LDS, each
// thread then writes 8 consecutive k of one row per plane (16 bytes).
template <int NS, bool H>
__global__ void __launch_bounds__(256)
wb_split_d_kernel(const double* __restrict__ d, int n, int64_t strideD, int rpad, int nkt,
                  typename PlaneT<H>::E* __restrict__ y, int64_t plane, int64_t fstride) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  __shared__ double tile[KT][65];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * 64, kt = blockIdx.y, k0 = kt * KT, f = blockIdx.z;
  const double* __restrict__ src = d + (int64_t)f * strideD;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int kl = (tid >> 6) + 4 * j, rl = tid & 63;
    const int row = r0 + rl, k = k0 + kl;
    tile[kl][rl] = (row < n && k <= row) ? src[row + (int64_t)k * n] : 0.0;
  }
  __syncthreads();
  const int rl = tid >> 2, c8 = tid & 3;
  // two planes carry at most 22 bits: split in FP32; three planes in FP64
  typedef typename std::conditional<NS == 2, float, double>::type R;
  V out[NS];
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    R v = (R)(tile[8 * c8 + e][rl] * (H ? D_SCALE : 1.0));
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const E h = (E)(float)v;
      out[q][e] = h;
      v -= (R)(float)h;
    }
  }
  const int64_t off = (int64_t)f * fstride + ((int64_t)kt * rpad + r0 + rl) * KT + 8 * c8;
#pragma unroll
  for (int q = 0; q < NS; ++q) *reinterpret_cast<V*>(y + q * plane + off) = out[q];
}

// Rows of B^T: row n of the plane is column n of B, whose k are contiguous in
// memory. Thread: 8 consecutive k of one column (64 bytes in, 16 bytes out
// per plane).
template <int NS, bool H>
__global__ void __launch_bounds__(256)
wb_split_b_kernel(const double* __restrict__ b, int krows, int64_t ncol, int64_t ldb, int64_t npad, int nkt,
                  const double* __restrict__ scale, typename PlaneT<H>::E* __restrict__ y, int64_t plane) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  const double sc = H ? scale[0] : 1.0;
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;   // (n, kt, c8) with c8 fastest
  const int64_t total = npad * nkt * 4;
  if (e >= total) return;
  const int c8 = (int)(e & 3);
  const int kt = (int)((e >> 2) % nkt);
  const int64_t n = (e >> 2) / nkt;
  const int k0 = kt * KT + 8 * c8;
  typedef typename std::conditional<NS == 2, float, double>::type R;
  V out[NS];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int k = k0 + i;
    R v = (R)((n < ncol && k < krows) ? b[k + n * ldb] * sc : 0.0);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      const E h = (E)(float)v;
      out[q][i] = h;
      v -= (R)(float)h;
    }
  }
  const int64_t off = ((int64_t)kt * npad + n) * KT + 8 * c8;
#pragma unroll
  for (int q = 0; q < NS; ++q) *reinterpret_cast<V*>(y + q * plane + off) = out[q];
}

// ---------------------------------------------------------------- product
struct WbArgs {
  const void* a;       // D planes (bf16 or fp16)
  int64_t aplane, afstride;
  const void* b;       // B^T planes
  int64_t bplane;
  int apad, nkt, naux, tm, tn;
  int64_t npad, ncol;
  const double* bmat;  // B in FP64 for the epilogue
  int64_t ldb;
  double* out;         // out[(f * tm + bm) * ncol + n]
  const double* unscale;   // FP16X3: 1 / (D scale x B scale), else null
  int64_t tiles, nblocks;
};

// Tiles BM x BN, WM x WN waves, one 32-k slab per step. DEPTH 2: two register
// sets prefetch the next slabs while the MFMAs consume the LDS copy; DEPTH 3:
// two LDS buffers, one barrier per slab.
template <int NS, bool H, int BM, int BN, int WM, int WN, int FLUSH, int DEPTH>
__global__ void __launch_bounds__(64 * WM * WN)
wb_dot_kernel(WbArgs g) {
  typedef typename PlaneT<H>::E E;
  typedef typename PlaneT<H>::V V;
  constexpr int KP = KT;                  // LDS row: 64 bytes, chunks swizzled (lds_chunk)
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = BM / WM, TN = BN / WN;
  constexpr int FM = TM / 16, FN = TN / 16;
  constexpr int LA = BM * KT / 8 / NT, LB = BN * KT / 8 / NT;
  static_assert(BM * KT / 8 % NT == 0 && BN * KT / 8 % NT == 0, "chunks per thread");
  static_assert(DEPTH == 2 || DEPTH == 3, "prefetch depth");
  constexpr int NBUF = DEPTH == 3 ? 2 : 1;

  __shared__ __attribute__((aligned(16))) E As[NBUF][NS][BM][KP];
  __shared__ __attribute__((aligned(16))) E Bs[NBUF][NS][BN][KP];
  __shared__ double red[WM][BN];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int f = (int)(L / g.tiles);
  // groups of 8 row blocks walked column block by column block, so the
  // B^T panel of a column block is read by 8 neighbours in a row
  const int t = (int)(L % g.tiles);
  constexpr int GROUP_M = 8;
  const int per_group = GROUP_M * g.tn;
  const int group = t / per_group;
  const int first_m = group * GROUP_M;
  const int gsize = min(g.tm - first_m, GROUP_M);
  const int in_group = t - group * per_group;
  const int bm = first_m + in_group % gsize, bn = in_group / gsize;
  const int m0 = bm * BM;
  const int64_t n0 = (int64_t)bn * BN;

  const E* __restrict__ Ya = static_cast<const E*>(g.a) + (int64_t)f * g.afstride + (int64_t)m0 * KT;
  const E* __restrict__ Yb = static_cast<const E*>(g.b) + n0 * KT;
  const int64_t aslab = (int64_t)g.apad * KT, bslab = g.npad * KT;

  u4 ra0[NS][LA], rb0[NS][LB], ra1[NS][LA], rb1[NS][LB];
#define MGW_WB_LOAD(RA, RB, S_)                                                              \
  _Pragma("unroll") for (int q = 0; q < NS; ++q) {                                           \
    _Pragma("unroll") for (int i = 0; i < LA; ++i) {                                         \
      const int c = tid + i * NT;                                                            \
      RA[q][i] = *reinterpret_cast<const u4*>(Ya + q * g.aplane + (int64_t)(S_) * aslab + 8 * c); \
    }                                                                                        \
    _Pragma("unroll") for (int i = 0; i < LB; ++i) {                                         \
      const int c = tid + i * NT;                                                            \
      RB[q][i] = *reinterpret_cast<const u4*>(Yb + q * g.bplane + (int64_t)(S_) * bslab + 8 * c); \
    }                                                                                        \
  }
#define MGW_WB_STORE(RA, RB, BUF)                                                   \
  _Pragma("unroll") for (int q = 0; q < NS; ++q) {                                  \
    _Pragma("unroll") for (int i = 0; i < LA; ++i) {                                \
      const int c = tid + i * NT;                                                   \
      *reinterpret_cast<u4*>(&As[BUF][q][c >> 2][8 * lds_chunk(c >> 2, c & 3)]) = RA[q][i];  \
    }                                                                               \
    _Pragma("unroll") for (int i = 0; i < LB; ++i) {                                \
      const int c = tid + i * NT;                                                   \
      *reinterpret_cast<u4*>(&Bs[BUF][q][c >> 2][8 * lds_chunk(c >> 2, c & 3)]) = RB[q][i];  \
    }                                                                               \
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
  // D_f is lower triangular: the K range of row block bm ends at its last row
  const int nslab = min(g.nkt, (m0 + BM + KT - 1) / KT);
  auto compute = [&](int s, int buf) __attribute__((always_inline)) {
    constexpr int QA[6] = {NS == 3 ? 2 : 0, 0, 1, 1, 0, 0};
    constexpr int QB[6] = {0, NS == 3 ? 2 : 0, 1, 0, 1, 0};
    constexpr int P0 = NS == 3 ? 0 : 3;
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
    // product-major: consecutive MFMAs write different accumulators
#pragma unroll
    for (int pr = P0; pr < 6; ++pr)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          part[i][j] = mfma16x16x32<H>(b[QB[pr]][j], a[QA[pr]][i], part[i][j]);
    if constexpr (FLUSH > 0) if ((s + 1) % FLUSH == 0 || s + 1 == nslab) {
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

  if constexpr (DEPTH == 2) {
    MGW_WB_LOAD(ra0, rb0, 0)
    if (1 < nslab) { MGW_WB_LOAD(ra1, rb1, 1) }
    for (int s = 0; s < nslab; s += 2) {
      __syncthreads();
      MGW_WB_STORE(ra0, rb0, 0)
      __syncthreads();
      if (s + 2 < nslab) { MGW_WB_LOAD(ra0, rb0, s + 2) }
      compute(s, 0);
      if (s + 1 >= nslab) break;
      __syncthreads();
      MGW_WB_STORE(ra1, rb1, 0)
      __syncthreads();
      if (s + 3 < nslab) { MGW_WB_LOAD(ra1, rb1, s + 3) }
      compute(s + 1, 0);
    }
  } else {
    MGW_WB_LOAD(ra0, rb0, 0)
    MGW_WB_STORE(ra0, rb0, 0)
    if (1 < nslab) { MGW_WB_LOAD(ra0, rb0, 1) }
    __syncthreads();
    for (int s = 0; s < nslab; ++s) {
      const int cur = s & 1;
      compute(s, cur);
      if (s + 1 < nslab) {
        MGW_WB_STORE(ra0, rb0, cur ^ 1)
        if (s + 2 < nslab) { MGW_WB_LOAD(ra0, rb0, s + 2) }
      }
      __syncthreads();
    }
  }

  // epilogue: sum over the rows of the block of C (C + 2 B), B in FP64
  const double us = H ? g.unscale[0] : 1.0;
#pragma unroll
  for (int j = 0; j < FN; ++j) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int nl = wn * TN + j * 16 + 4 * fq + r;
      const int64_t n = n0 + nl;
      double sum = 0.0;
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int m = m0 + wm * TM + i * 16 + fr;
        if (m < g.naux && n < g.ncol) {
          // FLUSH 0: FP32 accumulation over the whole K (at most naux terms)
          const double c = (FLUSH > 0 ? acc[i][j][r] : (double)part[i][j][r]) * us;
          sum += c * (c + 2.0 * g.bmat[m + n * g.ldb]);
        }
      }
      sum += __shfl_xor(sum, 8, 16);
      sum += __shfl_xor(sum, 4, 16);
      sum += __shfl_xor(sum, 2, 16);
      sum += __shfl_xor(sum, 1, 16);
      if (fr == 0) red[wm][nl] = sum;
    }
  }
  __syncthreads();
  double* __restrict__ out = g.out + ((int64_t)f * g.tm + bm) * g.ncol;
  for (int nl = tid; nl < BN; nl += NT) {
    const int64_t n = n0 + nl;
    if (n < g.ncol) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < WM; ++w) s += red[w][nl];
      out[n] = s;
    }
  }
}

#undef MGW_WB_LOAD
#undef MGW_WB_STORE

struct WbLayout {
  int apad, nkt, tm, tn;
  int64_t npad;
  int64_t aplane_f;   // elements of one D plane of one frequency
  int64_t bplane;     // elements of one B^T plane
};

WbLayout wb_layout(int naux, int64_t ncol) {
  WbLayout L;
  L.apad = (naux + PAD - 1) / PAD * PAD;
  L.nkt = (naux + KT - 1) / KT;
  L.tm = (naux + 127) / 128;
  L.npad = (ncol + PAD - 1) / PAD * PAD;
  L.tn = (int)(L.npad / PAD);
  L.aplane_f = (int64_t)L.nkt * L.apad * KT;
  L.bplane = (int64_t)L.nkt * L.npad * KT;
  return L;
}

// B scale and the epilogue's unscale from the partial maxima of |B|
__global__ void wb_scale_kernel(const double* __restrict__ part, int nb, double* __restrict__ sc) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double m = 0.0;
  for (int b = tid; b < nb; b += 256) m = fmax(m, part[b]);
  red[tid] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) {
    const double bmax = red[0];
    const int e = (bmax > 0.0 && bmax <= 1.0e300) ? ilogb(bmax) : 0;
    sc[0] = ldexp(1.0, 13 - e);
    sc[1] = ldexp(1.0, e - 13) / D_SCALE;
  }
}

// bytes in front of the planes: B scale, unscale, partial maxima of |B|
constexpr size_t WB_HEAD = ((2 + ABSMAX_BLOCKS) * sizeof(double) + 255) / 256 * 256;

template <int NS, bool H>
int split_and_dot(const double* d, int naux, int nomega, const double* b, int64_t ncol, void* work,
                  size_t work_bytes, double* out, hipStream_t stream) {
  typedef typename PlaneT<H>::E E;
  constexpr int BM = 128, BN = 128;
  const WbLayout L = wb_layout(naux, ncol);
  const size_t bbytes = ((size_t)L.bplane * NS * sizeof(E) + 255) / 256 * 256;
  const size_t per_f = (size_t)L.aplane_f * NS * sizeof(E);
  MGW_REQUIRE(work_bytes >= WB_HEAD + bbytes + per_f, MGW_ERR_ARG, "W.B split: workspace too small");
  const int nf_max = (int)std::min<size_t>((size_t)nomega, (work_bytes - WB_HEAD - bbytes) / per_f);
  double* sc = static_cast<double*>(work);
  double* bpart = sc + 2;
  E* bp = reinterpret_cast<E*>(static_cast<char*>(work) + WB_HEAD);
  E* ap = reinterpret_cast<E*>(static_cast<char*>(work) + WB_HEAD + bbytes);
  if (H) {
    MGW_TRY(launch_absmax(b, (int64_t)naux * ncol, bpart, stream));
    hipLaunchKernelGGL(wb_scale_kernel, dim3(1), dim3(256), 0, stream, bpart, ABSMAX_BLOCKS, sc);
    MGW_HIP(hipGetLastError());
  }
  {
    const int64_t total = L.npad * L.nkt * 4;
    hipLaunchKernelGGL((wb_split_b_kernel<NS, H>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream, b,
                       naux, ncol, (int64_t)naux, L.npad, L.nkt, sc, bp, L.bplane);
    MGW_HIP(hipGetLastError());
  }
  for (int f0 = 0; f0 < nomega; f0 += nf_max) {
    const int nf = std::min(nf_max, nomega - f0);
    const int64_t aplane = L.aplane_f * nf;
    hipLaunchKernelGGL((wb_split_d_kernel<NS, H>), dim3(L.apad / 64, L.nkt, nf), dim3(256), 0, stream,
                       d + (int64_t)f0 * naux * naux, naux, (int64_t)naux * naux, L.apad, L.nkt, ap, aplane,
                       L.aplane_f);
    MGW_HIP(hipGetLastError());
    WbArgs g;
    g.a = ap; g.aplane = aplane; g.afstride = L.aplane_f;
    g.b = bp; g.bplane = L.bplane;
    g.apad = L.apad; g.nkt = L.nkt; g.naux = naux; g.tm = L.tm; g.tn = L.tn;
    g.npad = L.npad; g.ncol = ncol;
    g.bmat = b; g.ldb = naux;
    g.out = out + (int64_t)f0 * L.tm * ncol;
    g.unscale = H ? sc + 1 : nullptr;
    g.tiles = (int64_t)L.tm * L.tn;
    g.nblocks = g.tiles * nf;
    MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "W.B split: too many tiles");
    // two planes: FP32 accumulation over the whole K (the products of the
    // 11-bit fp16 / 8-bit bf16 planes are exact in FP32; the FP64 accumulators
    // would cost 128 VGPRs and half the occupancy)
    if (NS == 2) {
      // two LDS buffers (default) or two prefetch register sets: the same
      // products in the same order, so the tuner may pick either per shape
      auto launch = [&](int c) -> int {
        if (c == 0)
          hipLaunchKernelGGL((wb_dot_kernel<NS, H, BM, BN, 2, 2, 0, 3>), dim3((unsigned)g.nblocks), dim3(256), 0,
                             stream, g);
        else
          hipLaunchKernelGGL((wb_dot_kernel<NS, H, BM, BN, 2, 2, 0, 2>), dim3((unsigned)g.nblocks), dim3(256), 0,
                             stream, g);
        MGW_HIP(hipGetLastError());
        return MGW_OK;
      };
      MGW_TRY(autotune_select(MGW_TUNE_WB_DOT, NS, H ? 1 : 0, log2_floor(naux), log2_floor(ncol), log2_floor(nf), 2, 0,
                              launch, stream));
    } else {
      hipLaunchKernelGGL((wb_dot_kernel<NS, H, BM, BN, 2, 2, 4, 2>), dim3((unsigned)g.nblocks), dim3(256), 0, stream, g);
      MGW_HIP(hipGetLastError());
    }
  }
  return MGW_OK;
}

}  // namespace

// Workspace of the W.B split: scales, the B^T planes, and the D planes of as
// many frequencies as fit in 2 GiB (at least one).
size_t wb_split_workspace(int naux, int64_t ncol, int nomega, int precision) {
  const int ns = precision == MGW_BF16X6 ? 3 : 2;
  const WbLayout L = wb_layout(naux, ncol);
  const size_t bbytes = ((size_t)L.bplane * ns * 2 + 255) / 256 * 256;
  const size_t per_f = (size_t)L.aplane_f * ns * 2;
  const size_t budget = (size_t)2 << 30;
  const size_t nf = std::max<size_t>(1, std::min<size_t>((size_t)nomega, budget / per_f));
  return WB_HEAD + bbytes + nf * per_f;
}

int wb_split_dot(const double* d, int naux, int nomega, const double* b, int64_t ncol, int precision, void* work,
                 size_t work_bytes, double* dotpart, hipStream_t stream) {
  if (naux <= 0 || nomega <= 0 || ncol <= 0) return MGW_OK;
  if (precision == MGW_BF16X3)
    return split_and_dot<2, false>(d, naux, nomega, b, ncol, work, work_bytes, dotpart, stream);
  if (precision == MGW_BF16X6)
    return split_and_dot<3, false>(d, naux, nomega, b, ncol, work, work_bytes, dotpart, stream);
  if (precision == MGW_FP16X3)
    return split_and_dot<2, true>(d, naux, nomega, b, ncol, work, work_bytes, dotpart, stream);
  set_error("W.B split: precision must be BF16X3, BF16X6 or FP16X3");
  return MGW_ERR_ARG;
}

}  // namespace mgw
