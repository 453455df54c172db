// Batched inverse Cholesky factor of 1 - chi0(iw) for all frequencies.
//
// The reference inverts 1 - chi0 per frequency with LU (DGETRF/DGETRI) and
// forms chi = (1 - chi0)^-1 chi0 (src/m_gw_selfenergy_grid.f90:150-170);
// its E_RPA / E_GW traces come from the eigenvalues of 1 - chi0 (:171-173).
// Here 1 - chi0 = L L^T is factored for every frequency of the batch at
// once, and L is inverted in place:
//   ln det(1 - chi0) = 2 sum_i ln L_ii,   Tr (1 - chi0)^-1 = ||L^-1||_F^2,
// and the self-energy works on D = L^-1 - 1 (chi = L^-T L^-1 - 1), so it
// never needs the inverse itself.
//
// Recursive: the matrix is halved (at a multiple of 64) until 64 x 64
// leaves, which a one-wave kernel factors and inverts in registers; the
// off-diagonal work is four batched MFMA GEMMs per level (the trailing
// update has K = m/2), so the matrix is swept O(log n) times instead of once
// per 64-wide block column. The strict upper triangle of the input must be
// zero (the 1 - chi0 builds write zeros there); the recursion keeps it zero,
// so the inverted diagonal blocks are plain operands of the next products.
#include "mgw_internal.h"
#include "tile_order.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace mgw {
namespace {

constexpr int NB = CHOL_NB;

// Broadcast lane `lane` (a compile-time constant after unrolling) of a double.
__device__ inline double bcast(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// One wave per matrix, lane i holding row i of the kb x kb diagonal block at
// (k, k) in registers (rows beyond kb are identity rows): factor it, add
// sum ln L_jj to logdet, and write L_kk^-1 (zeros above the diagonal) in
// place. Fully unrolled: the column broadcasts are v_readlane, no barriers.
// A non-positive pivot records its 1-based column in info (first failure
// only) and is replaced by 1 so the batch stays finite.
__global__ void __launch_bounds__(64)
chol_diag_kernel(double* __restrict__ A, int n, int64_t strideA, int k, int kb, double* __restrict__ logdet,
                 int32_t* __restrict__ info) {
  const int f = blockIdx.x, i = threadIdx.x;
  double* __restrict__ a = A + (int64_t)f * strideA + k + (int64_t)k * n;
  double x[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j)
    x[j] = i < kb ? (j <= i ? a[i + (int64_t)j * n] : 0.0) : (i == j ? 1.0 : 0.0);
  int failed = 0;
  double ld = 0.0;
  // right-looking Cholesky: x[j] of lane i becomes L(i, j)
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double piv = bcast(x[j], j);
    if (!(piv > 0.0)) {
      if (!failed && j < kb) failed = k + j + 1;
      piv = 1.0;
    }
    const double d = sqrt(piv), rd = 1.0 / d;
    if (j < kb) ld += log(d);
    x[j] = (i == j) ? d : (i > j ? x[j] * rd : x[j]);
#pragma unroll
    for (int l = j + 1; l < NB; ++l) {
      const double llj = bcast(x[j], l);
      if (i >= l) x[l] -= x[j] * llj;
    }
  }
  // inverse by rows: T(l, 0..l) is final once lane l has divided by L(l, l);
  // every later row then subtracts L(i, l) T(l, j)
  double t[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) t[j] = (i == j) ? 1.0 : 0.0;
#pragma unroll
  for (int l = 0; l < NB; ++l) {
    const double rll = 1.0 / bcast(x[l], l);
    if (i == l) {
#pragma unroll
      for (int j = 0; j <= l; ++j) t[j] *= rll;
    }
    const double lil = x[l];
#pragma unroll
    for (int j = 0; j <= l; ++j) {
      const double tlj = bcast(t[j], l);
      if (i > l) t[j] -= lil * tlj;
    }
  }
  if (i == 0) {
    logdet[f] = (k == 0 ? 0.0 : logdet[f]) + ld;
    if (k == 0) info[f] = 0;
    if (failed && info[f] == 0) info[f] = failed;
  }
  if (i < kb) {
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < kb) a[i + (int64_t)j * n] = (j <= i) ? t[j] : 0.0;
  }
}

// FP64 reciprocal: the hardware estimate refined by two Newton steps
// (each doubles the correct bits), so the dependent chain of a leaf step is a
// few FMAs instead of the division sequence.
__device__ inline double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
  r = fma(r, fma(-d, r, 1.0), r);
  r = fma(r, fma(-d, r, 1.0), r);
  return r;
}

// The same factor and inverse with four waves per matrix, row owned, as
// L D L^T with L unit lower (no square root or division on the sequential
// path): thread t holds row i = t / 4 of the block at the columns
// l = 4 c + (t % 4), c = 0 .. 15, in registers (the four threads of a row are
// neighbouring lanes of one wave). Step j of the elimination publishes
// column j (its owners write A(i, j), i >= j, to an LDS vector), one
// barrier, d_j = A(j, j), and every thread updates its own columns j < l <= i:
//   A(i, l) -= (A(i, j) / d_j) A(l, j),   L(i, j) = A(i, j) / d_j.
// The unit lower L is inverted by rows without divisions: step l publishes
// row l of T = L^-1, one barrier, and every row i > l subtracts
// L(i, l) T(l, :), L(i, l) coming from the row's owner lane by a cross-lane
// read. Then the Cholesky factor's inverse is D^-1/2 T, row i scaled by
// d_i^-1/2, and ln det = sum ln d_j, all in parallel. The two LDS vectors
// alternate, so a step's writes never meet the previous step's reads and one
// barrier per step suffices. A non-positive pivot is recorded and replaced
// by 1, as in chol_diag_kernel.
__global__ void __launch_bounds__(256)
chol_leaf_kernel(double* __restrict__ A, int n, int64_t strideA, int k, int kb, double* __restrict__ logdet,
                 int32_t* __restrict__ info) {
  constexpr int NC = NB / 4;        // columns per thread
  __shared__ double vec[2][NB];     // published column (elimination) / row (inverse)
  __shared__ double red[4];
  const int f = blockIdx.x, t = threadIdx.x;
  const int i = t >> 2, q = t & 3, lane = t & 63;
  double* __restrict__ g = A + (int64_t)f * strideA + k + (int64_t)k * n;
  double a[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int l = 4 * c + q;
    a[c] = (i < kb && l < kb) ? (l <= i ? g[i + (int64_t)l * n] : 0.0) : (i == l ? 1.0 : 0.0);
  }
  int failed = 0;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double* __restrict__ v = vec[j & 1];
    if (q == (j & 3) && i >= j) v[i] = a[j >> 2];
    __syncthreads();
    double d = v[j];
    if (!(d > 0.0)) {
      if (!failed && j < kb) failed = k + j + 1;
      d = 1.0;
    }
    if (i > j) {
      const double lij = v[i] * rcp_nr(d);
#pragma unroll
      for (int c = j >> 2; c < NC; ++c) {
        const int l = 4 * c + q;
        if (l > j && l <= i) a[c] -= lij * v[l];
      }
      if (q == (j & 3)) a[j >> 2] = lij;
    } else if (i == j && q == (j & 3)) {
      a[j >> 2] = d;                 // D on the diagonal
    }
  }
  // d_i of this row from its owner lane; ln det(block) / 2 = sum ln d_i / 2
  double di = 0.0;
#pragma unroll
  for (int c = 0; c < NC; ++c)
    if (4 * c + q == i) di = a[c];
  di = __shfl(di, (lane & ~3) | (i & 3), 64);
  double ld = (q == 0 && i < kb) ? 0.5 * log(di) : 0.0;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) ld += __shfl_down(ld, off, 64);
  if (lane == 0) red[t >> 6] = ld;
  // T = L^-1 by rows (unit diagonal: no divisions)
  double tt[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) tt[c] = (4 * c + q == i) ? 1.0 : 0.0;
#pragma unroll
  for (int l = 0; l < NB; ++l) {
    double* __restrict__ v = vec[l & 1];
    const double lil = __shfl(a[l >> 2], (lane & ~3) | (l & 3), 64);   // L(i, l)
    if (i == l) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (4 * c + q <= l) v[4 * c + q] = tt[c];
    }
    __syncthreads();
    if (i > l) {
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (4 * c + q <= l) tt[c] -= lil * v[4 * c + q];
    }
  }
  if (i < kb) {
    const double rs = 1.0 / sqrt(di);
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int l = 4 * c + q;
      if (l < kb) g[i + (int64_t)l * n] = l <= i ? tt[c] * rs : 0.0;
    }
  }
  if (t == 0) {
    logdet[f] = (k == 0 ? 0.0 : logdet[f]) + red[0] + red[1] + red[2] + red[3];
    if (k == 0) info[f] = 0;
    if (failed && info[f] == 0) info[f] = failed;
  }
}

// Leaf kernel: MGW_CHOL_LEAF = 0 (one wave, registers) or 1 (four waves,
// LDS); read at every factorisation.
bool chol_leaf_lds() {
  const char* e = std::getenv("MGW_CHOL_LEAF");
  return e ? std::atoi(e) != 0 : true;
}

int gemm_chol(const GemmF64& g0, hipStream_t stream) {
  GemmF64 g = g0;
  g.klass = MGW_KCLASS_CHOLESKY;
  g.batch_outer = true;   // the matrices of the batch share no operand
  return gemm_f64(g, stream);
}

// Operand bounds of the FP16x3 products. With A >= I (so ||L^-1||_2 <= 1)
// and d = max_f,i A_f(i, i), in exact arithmetic at every level:
//   |A21|, |A22 - L21 L21^T| <= d        (entries of a positive definite matrix
//                                          are bounded by its largest diagonal
//                                          element, and Schur complements do
//                                          not grow the diagonal),
//   |L21|, |P = L21 (L^-1)11| <= sqrt(d)  (rows of L have norm sqrt(A_ii);
//                                          columns of L^-1 have norm <= 1),
//   |L^-1| <= 1.
// bounds = {d, sqrt(d), 1}. The FP16x3 scale puts a bound in [2^13, 2^14),
// so rounding that pushes a value past its bound stays far inside the fp16
// range (65504), and a bound that overestimates costs precision only for
// elements below 2^-17 of it.
// part[f] = max_i A_f(i, i), one workgroup per matrix
__global__ void __launch_bounds__(256)
chol_diagmax_kernel(const double* __restrict__ A, int n, int64_t strideA, double* __restrict__ part) {
  __shared__ double red[256];
  const int f = blockIdx.x;
  const double* __restrict__ a = A + (int64_t)f * strideA;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, a[i + (int64_t)i * n]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[f] = red[0];
}

__global__ void __launch_bounds__(256)
chol_bounds_kernel(const double* __restrict__ part, int nbatch, double* __restrict__ bounds) {
  __shared__ double red[256];
  double m = 0.0;
  for (int f = threadIdx.x; f < nbatch; f += 256) m = fmax(m, part[f]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double d = fmax(red[0], 1.0);
    bounds[0] = d;
    bounds[1] = sqrt(d);
    bounds[2] = 1.0;
    bounds[3] = 0.0;
  }
}

// ---------------------------------------------------------------- FP16x3 products on planes
// The large products of the recursion multiply operands split once into
// two fp16 planes (hi, lo) of 2^e x, in the k-tiled layout
// plane[f][q][kt][row][32] (q = 0 hi, 1 lo; rows padded to PL_ROWS, k to
// 32, zeros in the padding), so a row block's 32-k slab is one contiguous
// run. An operand is read in FP64 once per split instead of once per tile
// of its row or column, and the product kernel does no conversions.
constexpr int PL_K = 32;       // k per tile of the plane layout
constexpr int PL_ROWS = 128;   // rows padded to the product's tile
typedef float pf4 __attribute__((ext_vector_type(4)));
typedef _Float16 ph8 __attribute__((ext_vector_type(8)));
typedef unsigned int pu4 __attribute__((ext_vector_type(4)));

__device__ inline double plane_scale(double bound) {
  return (bound > 0.0 && bound <= 1.0e300) ? ldexp(1.0, 13 - ilogb(bound)) : 1.0;
}
// LDS rows of 64 bytes; 16-byte chunk j of row r stored at chunk j ^ ((r >> 1) & 3)
__device__ inline int pl_chunk(int r, int j) { return j ^ ((r >> 1) & 3); }

// Split op(X) (rows x k) of every matrix f into the two planes. trans = 0:
// op(X)(r, k) = X[r + k ld] (rows contiguous: staged through LDS so that a
// thread writes 8 consecutive k); trans = 1: op(X)(r, k) = X[k + r ld].
// Block: 64 rows x 32 k of one matrix. KLASS tags the caller's kernel class
// (distinct kernel symbols for the profiler).
template <int KLASS>
__global__ void __launch_bounds__(256)
plane_split_kernel(const double* __restrict__ X, int64_t ld, int64_t strideX, int trans, int rows, int k,
                  int rpad, int nkt, const double* __restrict__ bound, _Float16* __restrict__ planes) {
  __shared__ double tile[PL_K][65];
  const int tid = threadIdx.x;
  const int r0 = blockIdx.x * 64, kt = blockIdx.y, f = blockIdx.z, k0 = kt * PL_K;
  const double* __restrict__ x = X + (int64_t)f * strideX;
  const double sc = plane_scale(bound[0]);
  const int r = tid >> 2, c8 = tid & 3;
  double v[8];
  if (!trans) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int kl = (tid >> 6) + 4 * j, rl = tid & 63;
      const int R = r0 + rl, K = k0 + kl;
      tile[kl][rl] = (R < rows && K < k) ? x[R + (int64_t)K * ld] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = tile[8 * c8 + e][r];
  } else {
    const int R = r0 + r;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int K = k0 + 8 * c8 + e;
      v[e] = (R < rows && K < k) ? x[K + (int64_t)R * ld] : 0.0;
    }
  }
  ph8 hi, lo;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const float y = (float)(v[e] * sc);
    const _Float16 h = (_Float16)y;
    hi[e] = h;
    lo[e] = (_Float16)(y - (float)h);
  }
  const int64_t per_q = (int64_t)nkt * rpad * PL_K;
  _Float16* __restrict__ dst = planes + (int64_t)f * 2 * per_q + ((int64_t)kt * rpad + r0 + r) * PL_K + 8 * c8;
  *reinterpret_cast<ph8*>(dst) = hi;
  *reinterpret_cast<ph8*>(dst + per_q) = lo;
}

struct PlaneGemmArgs {
  const _Float16* a;  int64_t a_q, a_f; int a_rpad;   // A planes (M rows), plane and matrix strides
  const _Float16* b;  int64_t b_q, b_f; int b_rpad;   // B^T planes (N rows)
  int M, N, nkt;
  // K range of a tile in k tiles: A lower triangular (k <= m) stops the K
  // loop at the tile's last row; B^T(n, k) zero for k > n (kb_hi) or for
  // k < n (kb_lo) trims it at the tile's columns
  int tri_a, kb_hi, kb_lo;
  const double* amax; const double* bmax;
  double alpha, beta;
  double* C; int64_t ldc, strideC;
  // optional: the result written as the planes of a later operand instead of
  // C (op_trans = 0: plane row m, k n; 1: plane row n, k m), scaled by the
  // power of two of obound; entries of the tile outside M x N are the
  // planes' zero padding
  _Float16* op; int64_t op_q, op_f; int op_rpad, op_nkt, op_trans;
  const double* obound;
  int tiles_m, tiles_n;
  int64_t tiles, nblocks;
};

// C(m, n) = alpha sum_k A(m, k) B^T(n, k) + beta C(m, n) for every matrix,
// BM x BM tiles (128, or 64 for the middle levels, whose products would
// otherwise have too few workgroups), 4 waves, one 32-k slab per step
// (register prefetch of the next slab, one LDS buffer); hi hi + hi lo +
// lo hi on v_mfma_f32_16x16x32_f16 accumulated in FP32 over the tile's K
// range, unscaled in FP64 in the epilogue. LOWER: lower-triangle tiles
// (M == N), entries with m < n not written. Tiles of one matrix before the
// next.
template <bool LOWER, int BM, int KLASS>
__global__ void __launch_bounds__(256, 2)
plane_gemm_kernel(PlaneGemmArgs g) {
  constexpr int WM = 2, WN = 2, NT = 256;
  static_assert(BM == 64 || BM == 128, "tile");
  static_assert(PL_ROWS % BM == 0, "plane row padding covers whole tiles");
  constexpr int TM = BM / WM, TN = BM / WN, FM = TM / 16, FN = TN / 16;
  constexpr int CH = BM * PL_K / 8;        // 16-byte chunks per plane per operand slab
  constexpr int LPT = CH / NT;
  __shared__ __attribute__((aligned(16))) _Float16 As[2][BM][PL_K];
  __shared__ __attribute__((aligned(16))) _Float16 Bs[2][BM][PL_K];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int64_t L = xcd_logical(blockIdx.x, g.nblocks);
  const int f = (int)(L / g.tiles);
  const int tile = (int)(L % g.tiles);
  int bm, bn;
  if constexpr (LOWER) {
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
  const int m0 = bm * BM, n0 = bn * BM;
  int kt0 = 0, kt1 = g.nkt;
  if (g.tri_a) kt1 = min(kt1, (m0 + BM + PL_K - 1) / PL_K);
  if (g.kb_hi) kt1 = min(kt1, (n0 + BM + PL_K - 1) / PL_K);
  if (g.kb_lo) kt0 = n0 / PL_K;
  const _Float16* __restrict__ Ya = g.a + (int64_t)f * g.a_f + (int64_t)m0 * PL_K;
  const _Float16* __restrict__ Yb = g.b + (int64_t)f * g.b_f + (int64_t)n0 * PL_K;
  const int64_t sa = (int64_t)g.a_rpad * PL_K, sb = (int64_t)g.b_rpad * PL_K;
  pu4 ra[2][LPT], rb[2][LPT];
  auto load = [&](int s) __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int i = 0; i < LPT; ++i) {
        const int c = tid + i * NT;
        ra[q][i] = *reinterpret_cast<const pu4*>(Ya + q * g.a_q + (int64_t)s * sa + 8 * c);
        rb[q][i] = *reinterpret_cast<const pu4*>(Yb + q * g.b_q + (int64_t)s * sb + 8 * c);
      }
  };
  auto store = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int i = 0; i < LPT; ++i) {
        const int c = tid + i * NT;
        *reinterpret_cast<pu4*>(&As[q][c >> 2][8 * pl_chunk(c >> 2, c & 3)]) = ra[q][i];
        *reinterpret_cast<pu4*>(&Bs[q][c >> 2][8 * pl_chunk(c >> 2, c & 3)]) = rb[q][i];
      }
  };
  pf4 part[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) part[i][j] = pf4{0.f, 0.f, 0.f, 0.f};
  const int fr = lane & 15, fq = lane >> 4;
  if (kt0 < kt1) load(kt0);
  for (int s = kt0; s < kt1; ++s) {
    __syncthreads();                   // previous slab's fragments consumed
    store();
    __syncthreads();
    if (s + 1 < kt1) load(s + 1);
    ph8 a[2][FM], b[2][FN];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = wm * TM + i * 16 + fr;
        a[q][i] = *reinterpret_cast<const ph8*>(&As[q][r][8 * pl_chunk(r, fq)]);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int r = wn * TN + j * 16 + fr;
        b[q][j] = *reinterpret_cast<const ph8*>(&Bs[q][r][8 * pl_chunk(r, fq)]);
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
  // lane l, register r holds C(m = m_base + (l & 15), n = n_base + 4 (l >> 4) + r)
  const double unscale = 1.0 / (plane_scale(g.amax[0]) * plane_scale(g.bmax[0]));   // a power of two
  if (g.op) {
    // the same split as plane_split_kernel applied to the FP64 result
    // alpha part unscale (exact), so the planes are those of C
    const double so = plane_scale(g.obound[0]);
    _Float16* __restrict__ O = g.op + (int64_t)f * g.op_f;
    const int okmax = g.op_nkt * PL_K;
#pragma unroll
    for (int i = 0; i < FM; ++i) {
      const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
          const int orow = g.op_trans ? n : m, ok = g.op_trans ? m : n;
          if (orow >= g.op_rpad || ok >= okmax) continue;
          const double v = (m < g.M && n < g.N) ? g.alpha * ((double)part[i][j][r] * unscale) : 0.0;
          const float y = (float)(v * so);
          const _Float16 h = (_Float16)y;
          const int64_t idx = ((int64_t)(ok / PL_K) * g.op_rpad + orow) * PL_K + (ok % PL_K);
          O[idx] = h;
          O[idx + g.op_q] = (_Float16)(y - (float)h);
        }
    }
    return;
  }
  double* __restrict__ C = g.C + (int64_t)f * g.strideC;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int m = m0 + wm * TM + i * 16 + fr;
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int n = n0 + wn * TN + j * 16 + 4 * fq + r;
        if (m >= g.M || n >= g.N) continue;
        if (LOWER && m < n) continue;
        const int64_t idx = m + (int64_t)n * g.ldc;
        double v = g.alpha * ((double)part[i][j][r] * unscale);
        if (g.beta != 0.0) v += g.beta * C[idx];
        C[idx] = v;
      }
  }
}

// ---------------------------------------------------------------- traces
// partial[f][b] = sum over columns j = b, b + FROB_BLOCKS, ... of
// sum_{i >= j} m_f(i, j)^2.
__global__ void __launch_bounds__(256)
lower_frob2_kernel(const double* __restrict__ m, int n, double* __restrict__ partial) {
  __shared__ double red[4];
  const int b = blockIdx.x, f = blockIdx.y;
  const double* __restrict__ a = m + (int64_t)f * n * n;
  double s = 0.0;
  for (int j = b; j < n; j += FROB_BLOCKS)
    for (int i = j + threadIdx.x; i < n; i += 256) {
      const double v = a[i + (int64_t)j * n];
      s += v * v;
    }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[(int64_t)f * FROB_BLOCKS + b] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

namespace {

// Split of a diagonal block of size m > NB: m1 a multiple of the leaf size.
inline int chol_split(int m) {
  int m1 = ((m + 1) / 2 + NB - 1) / NB * NB;
  if (m1 >= m) m1 = m - NB;
  return m1;
}

// Work of the recursion on a block of size m: L21 (m2 x m1) stays alive while
// A22 is processed in the work behind it; P = L21 (L^-1)11 then reuses that
// space. A11 is processed before L21 exists.
size_t chol_need(int m) {
  if (m <= NB) return 0;
  const int m1 = chol_split(m), m2 = m - m1;
  const size_t l21 = (size_t)m2 * m1;
  return std::max(chol_need(m1), l21 + std::max(chol_need(m2), l21));
}

}  // namespace

size_t chol_work_doubles(int n) { return std::max<size_t>(chol_need(n), 1); }

// Levels of the recursion whose two halves are both at least this wide run
// their products in FP16x3 on planes when the factorisation is given
// operand bounds (128-wide tiles from 256, 64-wide below; smaller levels are
// latency-bound and stay FP64): MGW_CHOL_F16_MIN, read at every
// factorisation; 0 keeps every product in FP64.
int chol_f16_min() {
  const char* e = std::getenv("MGW_CHOL_F16_MIN");
  return e ? std::atoi(e) : 128;
}

namespace {

struct CholCtx {
  double* A;
  int n, nbatch;
  double* logdet;
  int32_t* info;
  double* work;
  int64_t sw;           // work doubles per matrix
  hipStream_t stream;
  const double* bounds; // device {d, sqrt(d), 1} for FP16x3 products, or null (FP64)
  int f16_min;          // smallest dimension of a level whose products run in FP16x3
  _Float16* pa;         // plane regions of the two operands of a product
  _Float16* pb;
  _Float16* pl21;       // stack of the levels' L21 planes (live across the A22
                        // recursion), each level's P planes above its L21
  int64_t plane_elems;  // capacity of pa, pb: elements per matrix
  int64_t l21_elems;    // capacity of the stack: elements per matrix
  bool epi;             // products write the planes of L21 and P (else: FP64, then split)
};
constexpr int BND_D = 0, BND_SQRT_D = 1, BND_ONE = 2;

// Split op(X) of every matrix into the planes at dst (layout d).
int split_planes(const CholCtx& c, const double* X, int64_t ld, int64_t strideX, int trans, int bnd,
                 const Fp16Planes& d, _Float16* dst) {
  MGW_REQUIRE(d.per_f <= c.plane_elems, MGW_ERR_ARG, "chol: plane workspace too small");
  return fp16_split_planes(X, ld, strideX, trans, d, c.nbatch, c.bounds + bnd, dst, MGW_KCLASS_CHOLESKY, c.stream);
}

int plane_gemm(const CholCtx& c, const Fp16Planes& da, const _Float16* pa, int abnd, const Fp16Planes& db,
               const _Float16* pb, int bbnd, Fp16PlaneProduct p) {
  p.nbatch = c.nbatch;
  p.a_f = da.per_f;
  p.b_f = db.per_f;
  return fp16_plane_gemm(da, pa, c.bounds + abnd, db, pb, c.bounds + bbnd, p, MGW_KCLASS_CHOLESKY, c.stream);
}

// Recursive inverse Cholesky factor of the diagonal block [k0, k0 + m):
//   [A11    ]   factor and invert A11 (recursion), L21 = A21 L11^-T,
//   [A21 A22]   A22 -= L21 L21^T, factor and invert A22 (recursion),
//               (L^-1)21 = -(L^-1)22 L21 (L^-1)11.
// The products are batched MFMA GEMMs (FP64, or FP16x3 on planes for the
// large levels when the factorisation has operand bounds); the 64 x 64
// leaves are factored and inverted by chol_leaf_kernel (or
// chol_diag_kernel). `w` is this block's work (chol_need(m) doubles per
// matrix, matrix stride c.sw); `l21_off` the L21-plane stack used by the
// enclosing levels (elements per matrix).
int chol_rec(const CholCtx& c, int k0, int m, double* w, int64_t l21_off) {
  const int n = c.n;
  const int64_t s2 = (int64_t)n * n;
  if (m <= NB) {
    if (chol_leaf_lds())
      hipLaunchKernelGGL(chol_leaf_kernel, dim3(c.nbatch), dim3(256), 0, c.stream, c.A, n, s2, k0, m, c.logdet,
                         c.info);
    else
      hipLaunchKernelGGL(chol_diag_kernel, dim3(c.nbatch), dim3(64), 0, c.stream, c.A, n, s2, k0, m, c.logdet,
                         c.info);
    MGW_HIP(hipGetLastError());
    return MGW_OK;
  }
  const int m1 = chol_split(m), m2 = m - m1;
  double* a11 = c.A + k0 + (int64_t)k0 * n;
  double* a21 = a11 + m1;
  double* a22 = a11 + m1 + (int64_t)m1 * n;
  double* l21 = w;                          // m2 x m1, leading dimension m2
  double* rest = w + (int64_t)m2 * m1;      // A22's work, then P
  const bool lowp = c.bounds && c.f16_min > 0 && std::min(m1, m2) >= c.f16_min;
  const int pbm = std::min(m1, m2) >= 256 ? 128 : 64;   // plane product tile
  MGW_TRY(chol_rec(c, k0, m1, w, l21_off));
  if (lowp && c.epi) {
    // L21 and P exist only as planes: the products that make them write
    // them, and the splits of L21 (twice) and P go away
    const Fp16Planes dl = fp16_planes(m2, m1);
    MGW_REQUIRE(l21_off + dl.per_f <= c.l21_elems, MGW_ERR_ARG, "chol: L21 plane stack too small");
    _Float16* pl = c.pl21 + l21_off * c.nbatch;
    {  // L21 = A21 (L^-1)11^T  ((L^-1)11 has exact zeros above the diagonal)
      const Fp16Planes da = fp16_planes(m2, m1), db = fp16_planes(m1, m1);
      MGW_TRY(split_planes(c, a21, n, s2, 0, BND_D, da, c.pa));
      MGW_TRY(split_planes(c, a11, n, s2, 0, BND_ONE, db, c.pb));
      Fp16PlaneProduct p;
      p.M = m2; p.N = m1; p.kb_hi = 1; p.bm = pbm;        // (L^-1)11(n, k) = 0 for k > n
      p.out_planes = pl; p.out_dims = dl; p.out_trans = 0; p.out_bound = c.bounds + BND_SQRT_D;
      MGW_TRY(plane_gemm(c, da, c.pa, BND_D, db, c.pb, BND_ONE, p));
    }
    {  // A22 -= L21 L21^T (lower triangle; the upper stays zero)
      Fp16PlaneProduct u;
      u.M = m2; u.N = m2; u.C = a22; u.ldc = n; u.strideC = s2;
      u.alpha = -1.0; u.beta = 1.0; u.lower = true; u.bm = pbm;
      MGW_TRY(plane_gemm(c, dl, pl, BND_SQRT_D, dl, pl, BND_SQRT_D, u));
    }
    MGW_TRY(chol_rec(c, k0 + m1, m2, rest, l21_off + dl.per_f));
    // the A22 recursion is done: P's planes go above this level's L21
    const Fp16Planes dp = fp16_planes(m1, m2);
    MGW_REQUIRE(l21_off + dl.per_f + dp.per_f <= c.l21_elems, MGW_ERR_ARG, "chol: plane stack too small");
    _Float16* pp = c.pl21 + (l21_off + dl.per_f) * c.nbatch;
    {  // P = L21 (L^-1)11, written as the planes of P^T (the B^T operand below)
      const Fp16Planes dt = fp16_planes(m1, m1);
      MGW_TRY(split_planes(c, a11, n, s2, 1, BND_ONE, dt, c.pb));      // B^T(n, k) = (L^-1)11(k, n)
      Fp16PlaneProduct q;
      q.M = m2; q.N = m1; q.kb_lo = 1; q.bm = pbm;        // (L^-1)11(k, n) = 0 for k < n
      q.out_planes = pp; q.out_dims = dp; q.out_trans = 1; q.out_bound = c.bounds + BND_SQRT_D;
      MGW_TRY(plane_gemm(c, dl, pl, BND_SQRT_D, dt, c.pb, BND_ONE, q));
    }
    // (L^-1)21 = -(L^-1)22 P
    const Fp16Planes da = fp16_planes(m2, m2);
    MGW_TRY(split_planes(c, a22, n, s2, 0, BND_ONE, da, c.pa));
    Fp16PlaneProduct t;
    t.M = m2; t.N = m1; t.C = a21; t.ldc = n; t.strideC = s2;
    t.alpha = -1.0; t.tri_a = 1; t.bm = pbm;
    return plane_gemm(c, da, c.pa, BND_ONE, dp, pp, BND_SQRT_D, t);
  }
  // L21 = A21 (L^-1)11^T   ((L^-1)11 has exact zeros above the diagonal)
  if (lowp) {
    const Fp16Planes da = fp16_planes(m2, m1), db = fp16_planes(m1, m1);
    MGW_TRY(split_planes(c, a21, n, s2, 0, BND_D, da, c.pa));
    MGW_TRY(split_planes(c, a11, n, s2, 0, BND_ONE, db, c.pb));
    Fp16PlaneProduct p;
    p.M = m2; p.N = m1; p.C = l21; p.ldc = m2; p.strideC = c.sw;
    p.kb_hi = 1; p.bm = pbm;                            // (L^-1)11(n, k) = 0 for k > n
    MGW_TRY(plane_gemm(c, da, c.pa, BND_D, db, c.pb, BND_ONE, p));
  } else {
    GemmF64 p;
    p.M = m2; p.N = m1; p.K = m1;
    p.A = a21; p.lda = n; p.strideA = s2;
    p.B = a11; p.ldb = n; p.strideB = s2; p.transB = true;
    p.C = l21; p.ldc = m2; p.strideC = c.sw;
    p.batch = c.nbatch;
    MGW_TRY(gemm_chol(p, c.stream));
  }
  // A22 -= L21 L21^T (lower triangle; the upper stays zero)
  if (lowp) {
    const Fp16Planes dl = fp16_planes(m2, m1);
    MGW_TRY(split_planes(c, l21, m2, c.sw, 0, BND_SQRT_D, dl, c.pa));
    Fp16PlaneProduct u;
    u.M = m2; u.N = m2; u.C = a22; u.ldc = n; u.strideC = s2;
    u.alpha = -1.0; u.beta = 1.0; u.lower = true; u.bm = pbm;
    MGW_TRY(plane_gemm(c, dl, c.pa, BND_SQRT_D, dl, c.pa, BND_SQRT_D, u));
  } else {
    GemmF64 u;
    u.M = m2; u.N = m2; u.K = m1;
    u.A = l21; u.lda = m2; u.strideA = c.sw;
    u.B = l21; u.ldb = m2; u.strideB = c.sw; u.transB = true;
    u.C = a22; u.ldc = n; u.strideC = s2;
    u.alpha = -1.0; u.beta = 1.0;
    u.lower_only = true; u.no_mirror = true;
    u.batch = c.nbatch;
    MGW_TRY(gemm_chol(u, c.stream));
  }
  MGW_TRY(chol_rec(c, k0 + m1, m2, rest, l21_off));
  // (L^-1)21 = -(L^-1)22 P,  P = L21 (L^-1)11
  if (lowp) {
    // the recursion on A22 reused the plane regions: split L21 again
    const Fp16Planes dl = fp16_planes(m2, m1), dt = fp16_planes(m1, m1);
    MGW_TRY(split_planes(c, l21, m2, c.sw, 0, BND_SQRT_D, dl, c.pa));
    MGW_TRY(split_planes(c, a11, n, s2, 1, BND_ONE, dt, c.pb));      // B^T(n, k) = (L^-1)11(k, n)
    Fp16PlaneProduct q;
    q.M = m2; q.N = m1; q.C = rest; q.ldc = m2; q.strideC = c.sw;
    q.kb_lo = 1; q.bm = pbm;                            // (L^-1)11(k, n) = 0 for k < n
    MGW_TRY(plane_gemm(c, dl, c.pa, BND_SQRT_D, dt, c.pb, BND_ONE, q));
    const Fp16Planes da = fp16_planes(m2, m2), dp = fp16_planes(m1, m2);
    MGW_TRY(split_planes(c, a22, n, s2, 0, BND_ONE, da, c.pa));
    MGW_TRY(split_planes(c, rest, m2, c.sw, 1, BND_SQRT_D, dp, c.pb));  // B^T(n, k) = P(k, n)
    Fp16PlaneProduct t;
    t.M = m2; t.N = m1; t.C = a21; t.ldc = n; t.strideC = s2;
    t.alpha = -1.0; t.tri_a = 1; t.bm = pbm;
    return plane_gemm(c, da, c.pa, BND_ONE, dp, c.pb, BND_SQRT_D, t);
  }
  GemmF64 q;
  q.M = m2; q.N = m1; q.K = m1;
  q.A = l21; q.lda = m2; q.strideA = c.sw;
  q.B = a11; q.ldb = n; q.strideB = s2;
  q.C = rest; q.ldc = m2; q.strideC = c.sw;
  q.batch = c.nbatch;
  MGW_TRY(gemm_chol(q, c.stream));
  GemmF64 t;
  t.M = m2; t.N = m1; t.K = m2;
  t.A = a22; t.lda = n; t.strideA = s2; t.tri_a = true;
  t.B = rest; t.ldb = m2; t.strideB = c.sw;
  t.C = a21; t.ldc = n; t.strideC = s2;
  t.alpha = -1.0;
  t.batch = c.nbatch;
  return gemm_chol(t, c.stream);
}

}  // namespace

Fp16Planes fp16_planes(int rows, int k) {
  Fp16Planes d;
  d.rows = rows; d.k = k;
  d.rpad = (rows + PL_ROWS - 1) / PL_ROWS * PL_ROWS;
  d.nkt = (k + PL_K - 1) / PL_K;
  d.per_q = (int64_t)d.nkt * d.rpad * PL_K;
  d.per_f = 2 * d.per_q;
  return d;
}

int fp16_split_planes(const double* X, int64_t ld, int64_t strideX, int trans, const Fp16Planes& d, int nbatch,
                      const double* bound, _Float16* dst, int klass, hipStream_t stream) {
  if (d.rows <= 0 || d.k <= 0 || nbatch <= 0) return MGW_OK;
  MGW_REQUIRE(nbatch <= 65535 && d.nkt <= 65535, MGW_ERR_ARG, "fp16 planes: grid too large");
  const dim3 grid(d.rpad / 64, d.nkt, nbatch);
  if (klass == MGW_KCLASS_GEMM_AO2MO)
    hipLaunchKernelGGL(plane_split_kernel<MGW_KCLASS_GEMM_AO2MO>, grid, dim3(256), 0, stream, X, ld, strideX, trans,
                       d.rows, d.k, d.rpad, d.nkt, bound, dst);
  else
    hipLaunchKernelGGL(plane_split_kernel<MGW_KCLASS_CHOLESKY>, grid, dim3(256), 0, stream, X, ld, strideX, trans,
                       d.rows, d.k, d.rpad, d.nkt, bound, dst);
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

int fp16_plane_gemm(const Fp16Planes& da, const _Float16* pa, const double* amax, const Fp16Planes& db,
                    const _Float16* pb, const double* bmax, const Fp16PlaneProduct& p, int klass, hipStream_t stream) {
  if (p.M <= 0 || p.N <= 0 || p.nbatch <= 0) return MGW_OK;
  MGW_REQUIRE(da.nkt == db.nkt && da.rows == p.M && db.rows == p.N && (!p.lower || p.M == p.N), MGW_ERR_ARG,
              "fp16 planes: product shapes disagree");
  MGW_REQUIRE(p.bm == 64 || p.bm == 128, MGW_ERR_ARG, "fp16 planes: product tile");
  PlaneGemmArgs g;
  g.a = pa; g.a_q = da.per_q; g.a_f = p.a_f; g.a_rpad = da.rpad;
  g.b = pb; g.b_q = db.per_q; g.b_f = p.b_f; g.b_rpad = db.rpad;
  g.M = p.M; g.N = p.N; g.nkt = da.nkt;
  g.tri_a = p.tri_a; g.kb_hi = p.kb_hi; g.kb_lo = p.kb_lo;
  g.amax = amax; g.bmax = bmax;
  g.alpha = p.alpha; g.beta = p.beta;
  g.C = p.C; g.ldc = p.ldc; g.strideC = p.strideC;
  g.op = p.out_planes; g.op_q = p.out_dims.per_q; g.op_f = p.out_dims.per_f; g.op_rpad = p.out_dims.rpad;
  g.op_nkt = p.out_dims.nkt; g.op_trans = p.out_trans; g.obound = p.out_bound;
  if (p.out_planes) {
    // the output planes: op(C) is out_dims.rows x out_dims.k, and the tiles
    // cover every entry its consumers read (the same tile size)
    const int orows = p.out_trans ? p.N : p.M, ok = p.out_trans ? p.M : p.N;
    MGW_REQUIRE(!p.lower && p.beta == 0.0 && p.out_bound && p.out_dims.rows == orows && p.out_dims.k == ok,
                MGW_ERR_ARG,
                "fp16 planes: output planes disagree with the product");
  } else {
    MGW_REQUIRE(p.C != nullptr, MGW_ERR_ARG, "fp16 planes: no output");
  }
  g.tiles_m = (p.M + p.bm - 1) / p.bm;     // tiles end inside the padded rows
  g.tiles_n = (p.N + p.bm - 1) / p.bm;
  g.tiles = p.lower ? (int64_t)g.tiles_m * (g.tiles_m + 1) / 2 : (int64_t)g.tiles_m * g.tiles_n;
  g.nblocks = g.tiles * p.nbatch;
  MGW_REQUIRE(g.nblocks < (1ll << 31), MGW_ERR_ARG, "fp16 planes: too many tiles");
  const dim3 grid((unsigned)g.nblocks);
  constexpr int KC = MGW_KCLASS_CHOLESKY, KA = MGW_KCLASS_GEMM_AO2MO;
  if (klass == KA) {
    MGW_REQUIRE(!p.lower && p.bm == 128, MGW_ERR_ARG, "fp16 planes: transform products are full 128-wide tiles");
    hipLaunchKernelGGL((plane_gemm_kernel<false, 128, KA>), grid, dim3(256), 0, stream, g);
  } else if (p.bm == 64) {
    if (p.lower) hipLaunchKernelGGL((plane_gemm_kernel<true, 64, KC>), grid, dim3(256), 0, stream, g);
    else hipLaunchKernelGGL((plane_gemm_kernel<false, 64, KC>), grid, dim3(256), 0, stream, g);
  } else {
    if (p.lower) hipLaunchKernelGGL((plane_gemm_kernel<true, 128, KC>), grid, dim3(256), 0, stream, g);
    else hipLaunchKernelGGL((plane_gemm_kernel<false, 128, KC>), grid, dim3(256), 0, stream, g);
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

// fp16 elements per matrix of each of the two plane regions: the top
// level's operands are the largest
int64_t chol_plane_elems(int n) {
  if (n <= NB) return 0;
  const int m1 = chol_split(n), mx = std::max(m1, n - m1);
  return fp16_planes(mx, mx).per_f;
}

// the regions pa, pb and the plane stack: the L21 planes of the levels on
// the A22 path (at most 4/3 of the top level's), and one level's P planes
// above its L21 once its A22 recursion is done (at most the top level's L21
// and P: two regions)
constexpr int CHOL_PLANE_REGIONS = 4;
size_t chol_plane_bytes(int n, int nbatch) {
  return (size_t)CHOL_PLANE_REGIONS * chol_plane_elems(n) * nbatch * sizeof(_Float16);
}

// The products of a level write the planes of L21 and P (default), or write
// them in FP64 and split them (MGW_CHOL_EPI_PLANES=0; the same planes bit
// for bit). Read at every factorisation.
bool chol_epi_planes() {
  const char* e = std::getenv("MGW_CHOL_EPI_PLANES");
  return e ? std::atoi(e) != 0 : true;
}

int chol_inverse_batched(double* A, int n, int nbatch, double* logdet, int32_t* info, double* work,
                         hipStream_t stream, double* bounds, void* planes) {
  if (n <= 0 || nbatch <= 0) return MGW_OK;
  MGW_REQUIRE(nbatch <= 65535, MGW_ERR_ARG, "chol_inverse_batched: batch too large");
  MGW_REQUIRE(!bounds || planes || n <= NB, MGW_ERR_ARG, "chol_inverse_batched: FP16x3 products need planes");
  const int f16_min = bounds ? chol_f16_min() : 0;
  if (bounds && f16_min > 0) {
    hipLaunchKernelGGL(chol_diagmax_kernel, dim3(nbatch), dim3(256), 0, stream, A, n, (int64_t)n * n, bounds + 4);
    hipLaunchKernelGGL(chol_bounds_kernel, dim3(1), dim3(256), 0, stream, bounds + 4, nbatch, bounds);
    MGW_HIP(hipGetLastError());
  }
  const int64_t pe = chol_plane_elems(n);
  _Float16* pa = static_cast<_Float16*>(planes);
  auto region = [&](int r) { return pa ? pa + (int64_t)r * pe * nbatch : nullptr; };
  const CholCtx c{A, n, nbatch, logdet, info, work, (int64_t)chol_work_doubles(n), stream,
                  f16_min > 0 ? bounds : nullptr, f16_min, region(0), region(1), region(2), pe,
                  (CHOL_PLANE_REGIONS - 2) * pe, chol_epi_planes()};
  return chol_rec(c, 0, n, work, 0);
}

int launch_lower_frob2(const double* m, int n, int nbatch, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(lower_frob2_kernel, dim3(FROB_BLOCKS, nbatch), dim3(256), 0, s, m, n, partial);
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}


}  // namespace mgw
