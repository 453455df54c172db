// Batched inverse Cholesky factor of 1 - chi0(iw) for all frequencies.
//
// The reference inverts 1 - chi0 per frequency with LU (DGETRF/DGETRI) and
// forms chi = (1 - chi0)^-1 chi0 (src/m_gw_selfenergy_grid.f90:150-170);
// its E_RPA / E_GW traces come from the eigenvalues of 1 - chi0 (:171-173).
// Here 1 - chi0 = L L^T is factored for every frequency of the batch at
// once, and L is inverted in place:
//   ln det(1 - chi0) = 2 sum_i ln L_ii,   Tr (1 - chi0)^-1 = ||L^-1||_F^2,
//   B^T chi B = ||L^-1 B||^2 - ||B||^2   (chi = (1 - chi0)^-1 - 1),
// so the self-energy never needs the inverse itself.
//
// Recursive: the matrix is halved (at a multiple of 64) until 64 x 64
// leaves, which a one-wave kernel factors and inverts in registers; the
// off-diagonal work is a few large batched MFMA GEMMs per level (the trailing
// update has K = m/2), so the matrix is swept O(log n) times instead of once
// per 64-wide block column.
#include "mgw_internal.h"

#include <algorithm>
#include <cmath>

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

// dst_f(rows x cols, ldd) = src_f(rows x cols, lds) for every matrix f.
__global__ void copy_block_kernel(const double* __restrict__ src, int64_t lds, int64_t strideS,
                                  double* __restrict__ dst, int64_t ldd, int64_t strideD, int rows, int cols) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int f = blockIdx.y;
  if (e >= (int64_t)rows * cols) return;
  const int i = (int)(e % rows), j = (int)(e / rows);
  dst[(int64_t)f * strideD + i + j * ldd] = src[(int64_t)f * strideS + i + j * lds];
}

int gemm_chol(const GemmF64& g0, hipStream_t stream) {
  GemmF64 g = g0;
  g.klass = MGW_KCLASS_CHOLESKY;
  g.batch_outer = true;   // the matrices of the batch share no operand
  return gemm_f64(g, stream);
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

__global__ void zero_upper_kernel(double* __restrict__ m, int n) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i < n && j < n && i < j) m[(int64_t)blockIdx.z * n * n + i + (int64_t)j * n] = 0.0;
}

}  // namespace

size_t chol_work_doubles(int n) {
  const size_t h = (size_t)(n + 1) / 2 + NB;
  return h * h;
}

namespace {

// Zero the strict upper triangle of the diagonal block [k0, k0 + m) of
// every matrix (leading dimension n, stride n^2).
__global__ void zero_upper_block_kernel(double* __restrict__ A, int n, int k0, int m) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i < m && j < m && i < j) A[(int64_t)blockIdx.z * n * n + (k0 + i) + (int64_t)(k0 + j) * n] = 0.0;
}

struct CholCtx {
  double* A;
  int n, nbatch;
  double* logdet;
  int32_t* info;
  double* work;
  int64_t sw;           // work doubles per matrix
  hipStream_t stream;
};

// Recursive inverse Cholesky factor of the diagonal block [k0, k0 + m):
//   [A11    ]   factor and invert A11 (recursion), L21 = A21 L11^-T,
//   [A21 A22]   A22 -= L21 L21^T, factor and invert A22 (recursion),
//               (L^-1)21 = -(L^-1)22 L21 (L^-1)11.
// The products are large batched MFMA GEMMs; the 64 x 64 leaves are factored
// and inverted in registers by chol_diag_kernel.
int chol_rec(const CholCtx& c, int k0, int m) {
  const int n = c.n;
  const int64_t s2 = (int64_t)n * n;
  if (m <= NB) {
    hipLaunchKernelGGL(chol_diag_kernel, dim3(c.nbatch), dim3(64), 0, c.stream, c.A, n, s2, k0, m, c.logdet,
                       c.info);
    MGW_HIP(hipGetLastError());
    return MGW_OK;
  }
  int m1 = ((m + 1) / 2 + NB - 1) / NB * NB;   // a multiple of the leaf size
  if (m1 >= m) m1 = m - NB;
  const int m2 = m - m1;
  double* a11 = c.A + k0 + (int64_t)k0 * n;
  double* a21 = a11 + m1;
  double* a22 = a11 + m1 + (int64_t)m1 * n;
  MGW_TRY(chol_rec(c, k0, m1));
  // (L^-1)11 with exact zeros above the diagonal: a plain operand below
  hipLaunchKernelGGL(zero_upper_block_kernel, dim3((m1 + 15) / 16, (m1 + 15) / 16, c.nbatch), dim3(256), 0,
                     c.stream, c.A, n, k0, m1);
  MGW_HIP(hipGetLastError());
  // L21 = A21 (L^-1)11^T   (work: m2 x m1, leading dimension m2)
  GemmF64 p;
  p.M = m2; p.N = m1; p.K = m1;
  p.A = a21; p.lda = n; p.strideA = s2;
  p.B = a11; p.ldb = n; p.strideB = s2; p.transB = true;
  p.C = c.work; p.ldc = m2; p.strideC = c.sw;
  p.batch = c.nbatch;
  MGW_TRY(gemm_chol(p, c.stream));
  // A22 -= L21 L21^T (lower triangle)
  GemmF64 u;
  u.M = m2; u.N = m2; u.K = m1;
  u.A = c.work; u.lda = m2; u.strideA = c.sw;
  u.B = c.work; u.ldb = m2; u.strideB = c.sw; u.transB = true;
  u.C = a22; u.ldc = n; u.strideC = s2;
  u.alpha = -1.0; u.beta = 1.0;
  u.lower_only = true; u.no_mirror = true;
  u.batch = c.nbatch;
  MGW_TRY(gemm_chol(u, c.stream));
  const int64_t cnt = (int64_t)m2 * m1;
  hipLaunchKernelGGL(copy_block_kernel, dim3((unsigned)((cnt + 255) / 256), c.nbatch), dim3(256), 0, c.stream,
                     c.work, (int64_t)m2, c.sw, a21, (int64_t)n, s2, m2, m1);
  MGW_HIP(hipGetLastError());
  MGW_TRY(chol_rec(c, k0 + m1, m2));
  // (L^-1)21 = -(L^-1)22 (L21 (L^-1)11)
  GemmF64 q;
  q.M = m2; q.N = m1; q.K = m1;
  q.A = a21; q.lda = n; q.strideA = s2;
  q.B = a11; q.ldb = n; q.strideB = s2;
  q.C = c.work; q.ldc = m2; q.strideC = c.sw;
  q.batch = c.nbatch;
  MGW_TRY(gemm_chol(q, c.stream));
  GemmF64 t;
  t.M = m2; t.N = m1; t.K = m2;
  t.A = a22; t.lda = n; t.strideA = s2; t.tri_a = true;
  t.B = c.work; t.ldb = m2; t.strideB = c.sw;
  t.C = a21; t.ldc = n; t.strideC = s2;
  t.alpha = -1.0;
  t.batch = c.nbatch;
  return gemm_chol(t, c.stream);
}

}  // namespace

int chol_inverse_batched(double* A, int n, int nbatch, double* logdet, int32_t* info, double* work,
                         hipStream_t stream) {
  if (n <= 0 || nbatch <= 0) return MGW_OK;
  MGW_REQUIRE(nbatch <= 65535, MGW_ERR_ARG, "chol_inverse_batched: batch too large");
  const CholCtx c{A, n, nbatch, logdet, info, work, (int64_t)chol_work_doubles(n), stream};
  return chol_rec(c, 0, n);
}

int launch_lower_frob2(const double* m, int n, int nbatch, double* partial, hipStream_t s) {
  hipLaunchKernelGGL(lower_frob2_kernel, dim3(FROB_BLOCKS, nbatch), dim3(256), 0, s, m, n, partial);
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

int launch_zero_upper(double* m, int n, int nbatch, hipStream_t s) {
  hipLaunchKernelGGL(zero_upper_kernel, dim3((n + 15) / 16, (n + 15) / 16, nbatch), dim3(256), 0, s, m, n);
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace mgw
