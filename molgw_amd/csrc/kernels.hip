// Non-GEMM kernels of the RI G0W0 path (gfx950, wave64).
//
//  * unpack_ri        -- packed eri_3center(npair, P) -> dense B(P, k, l), HBM-bound
//  * casida_*         -- transition energies, traces and eigenvalue post-processing
//  * sigma_poles      -- analytic-pole self-energy sum (FP64 VALU, lanes = frequencies)
//  * chi_*            -- imaginary-axis chi0 weights, log-det/trace reductions
//  * sigma_imag_sum   -- frequency quadrature of the imaginary-axis self-energy
//  * pade_fit         -- Thiele continued fraction per state
//  * qp_solve         -- graphical + linearised QP equation per state
#include "mgw_internal.h"

#include <algorithm>
#include <cmath>

namespace mgw {

#define MGW_LAUNCH_CHECK() MGW_HIP(hipGetLastError())

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------ fill
__global__ void fill_kernel(double* p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}

int launch_fill(double* p, int64_t n, double v, hipStream_t s) {
  if (n <= 0) return MGW_OK;
  unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(fill_kernel, dim3(g), dim3(256), 0, s, p, n, v);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ unpack
// Tile of 64 pairs x 64 auxiliary functions, transposed through LDS so that
// both the packed reads (pair fastest) and the dense writes (P fastest) are
// coalesced. Diagonal pairs hold 1/2 (kk|P): the dense diagonal gets twice
// the stored value (src/m_eri_ao_mo.f90:517-522 adds both scatter branches).
__global__ void __launch_bounds__(256)
unpack_ri_kernel(const double* __restrict__ packed, int64_t npair, int naux,
                 const int32_t* __restrict__ kl, int nbf, double* __restrict__ dense, double* __restrict__ amax) {
  __shared__ double tile[64][65];
  __shared__ double red[4];
  const int t = threadIdx.x;
  const int64_t p0 = (int64_t)blockIdx.x * 64;
  const int P0 = blockIdx.y * 64;
  const int pl = t & 63;
  for (int i = 0; i < 16; ++i) {
    const int Pl = (t >> 6) + 4 * i;
    const int64_t pair = p0 + pl;
    const int P = P0 + Pl;
    double v = 0.0;
    if (pair < npair && P < naux) v = packed[pair + npair * (int64_t)P];
    tile[Pl][pl] = v;
  }
  __syncthreads();
  const int Pl = t & 63;
  const int P = P0 + Pl;
  double m = 0.0;                            // max |value written|
  if (P < naux) {
    for (int i = 0; i < 16; ++i) {
      const int pair_l = (t >> 6) + 4 * i;
      const int64_t pair = p0 + pair_l;
      if (pair >= npair) break;
      const int k = kl[2 * pair];
      const int l = kl[2 * pair + 1];
      const double v = tile[Pl][pair_l];
      if (k == l) {
        dense[P + (int64_t)naux * (k + (int64_t)nbf * l)] = 2.0 * v;
        m = fmax(m, 2.0 * fabs(v));
      } else {
        dense[P + (int64_t)naux * (k + (int64_t)nbf * l)] = v;
        dense[P + (int64_t)naux * (l + (int64_t)nbf * k)] = v;
        m = fmax(m, fabs(v));
      }
    }
  }
  if (!amax) return;
  // block maximum, then one atomic per block: the bit patterns of
  // non-negative doubles order as unsigned integers
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((t & 63) == 0) red[t >> 6] = m;
  __syncthreads();
  if (t == 0) {
    const double b = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    atomicMax(reinterpret_cast<unsigned long long*>(amax), (unsigned long long)__double_as_longlong(b));
  }
}

int launch_unpack_ri(const double* packed, int64_t npair, int naux, const int32_t* kl, int nbf,
                     double* dense, hipStream_t s, bool zero_dense, double* amax) {
  if (zero_dense) MGW_TRY(launch_fill(dense, (int64_t)naux * nbf * nbf, 0.0, s));
  dim3 grid(nblocks(npair, 64), nblocks(naux, 64));
  hipLaunchKernelGGL(unpack_ri_kernel, grid, dim3(256), 0, s, packed, npair, naux, kl, nbf, dense, amax);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ QS self-energy
// b(s, q, i) = w(s, q, i) Re[f_i / (e_q - i eta - e_i + W_s) + (1 - f_i) / (e_q + i eta - e_i - W_s)]
// with f_i = occ_i / spin_fact (src/m_gw_selfenergy_analytic.f90:738-744).
__global__ void qs_scale_kernel(const double* __restrict__ w, double* __restrict__ b, int npole, int nse, int cnt,
                                const double* __restrict__ pole, const double* __restrict__ eq,
                                const double* __restrict__ ei, const double* __restrict__ fi, double spin_fact,
                                double eta) {
  const int64_t total = (int64_t)npole * nse * cnt;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int sp = (int)(e % npole);
    const int64_t rest = e / npole;
    const int q = (int)(rest % nse), i = (int)(rest / nse);
    const double full = fi[i] / spin_fact, empty = (spin_fact - fi[i]) / spin_fact;
    const double d = eq[q] - ei[i], om = pole[sp];
    const double a = d + om, c = d - om;
    b[e] = w[e] * (full * a / (a * a + eta * eta) + empty * c / (c * c + eta * eta));
  }
}

int launch_qs_scale(const double* w, double* b, int npole, int nse, int cnt, const double* pole, const double* eq,
                    const double* ei, const double* fi, double spin_fact, double eta, hipStream_t s) {
  const int64_t total = (int64_t)npole * nse * cnt;
  if (total <= 0) return MGW_OK;
  unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(qs_scale_kernel, dim3(g), dim3(256), 0, s, w, b, npole, nse, cnt, pole, eq, ei, fi, spin_fact,
                     eta);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// m = (m + m^T) / 2 for an n x n matrix (Kotani's hermitianization,
// src/m_selfenergy_tools.f90:821-823)
__global__ void symmetrize_kernel(double* __restrict__ m, int n) {
  const int i = blockIdx.x * 16 + (threadIdx.x & 15);
  const int j = blockIdx.y * 16 + (threadIdx.x >> 4);
  if (i >= n || j >= n || i <= j) return;
  const double v = 0.5 * (m[i + (int64_t)j * n] + m[j + (int64_t)i * n]);
  m[i + (int64_t)j * n] = v;
  m[j + (int64_t)i * n] = v;
}

int launch_symmetrize(double* m, int n, hipStream_t s) {
  dim3 grid(nblocks(n, 16), nblocks(n, 16));
  hipLaunchKernelGGL(symmetrize_kernel, grid, dim3(256), 0, s, m, n);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ max |x|
__global__ void __launch_bounds__(256)
absmax_kernel(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  // four independent loads in flight per thread (max is order-independent)
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + tid;
  for (; i + 3 * stride < n; i += 4 * stride) {
    m0 = fmax(m0, fabs(x[i]));
    m1 = fmax(m1, fabs(x[i + stride]));
    m2 = fmax(m2, fabs(x[i + 2 * stride]));
    m3 = fmax(m3, fabs(x[i + 3 * stride]));
  }
  for (; i < n; i += stride) m0 = fmax(m0, fabs(x[i]));
  red[tid] = fmax(fmax(m0, m1), fmax(m2, m3));
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) part[blockIdx.x] = red[0];
}

int launch_absmax(const double* x, int64_t n, double* part, hipStream_t s) {
  hipLaunchKernelGGL(absmax_kernel, dim3(ABSMAX_BLOCKS), dim3(256), 0, s, x, n, part);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// out[0] = max of the ABSMAX_BLOCKS partial maxima
__global__ void __launch_bounds__(256) absmax_finish_kernel(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double red[256];
  const int tid = threadIdx.x;
  double m = 0.0;
  for (int b = tid; b < ABSMAX_BLOCKS; b += 256) m = fmax(m, part[b]);
  red[tid] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmax(red[tid], red[tid + w]);
    __syncthreads();
  }
  if (tid == 0) out[0] = red[0];
}

int launch_absmax_scalar(const double* x, int64_t n, double* part, double* out, hipStream_t s) {
  MGW_TRY(launch_absmax(x, n, part, s));
  hipLaunchKernelGGL(absmax_finish_kernel, dim3(1), dim3(256), 0, s, part, out);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ diagonal
__global__ void diag_add_kernel(double* __restrict__ m, int n, int nbatch, double v) {
  const int64_t total = (int64_t)n * nbatch;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int64_t f = e / n, i = e - f * n;
    m[f * n * n + i * (n + 1)] += v;
  }
}

int launch_diag_add(double* m, int n, int nbatch, double v, hipStream_t s) {
  const int64_t total = (int64_t)n * nbatch;
  if (total <= 0) return MGW_OK;
  unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(diag_add_kernel, dim3(g), dim3(256), 0, s, m, n, nbatch, v);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ metric
// Kept columns of the eigenvectors of the 2-center metric, each divided by
// the square root of its eigenvalue (src/m_eri_calculate.f90:1206-1211).
// Eigenvalues ascend, so the dropped ones are the first `neglect`.
__global__ void invsqrt_select_kernel(const double* __restrict__ u, int n, const double* __restrict__ eig,
                                      int neglect, int kept, double* __restrict__ out) {
  const int64_t total = (int64_t)n * kept;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    const int j = (int)(e / n);
    const int64_t src = e + (int64_t)neglect * n;
    out[e] = u[src] / sqrt(eig[j + neglect]);
  }
}

int launch_invsqrt_select(const double* u, int n, const double* eig, int neglect, int kept, double* out,
                          hipStream_t s) {
  const int64_t total = (int64_t)n * kept;
  if (total <= 0) return MGW_OK;
  unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(invsqrt_select_kernel, dim3(g), dim3(256), 0, s, u, n, eig, neglect, kept, out);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ gather
// X(P, t) = eri_mo(P, a_t - aoff, i_t - ioff) for a general transition table
// (src/m_linear_response.f90:837-842).
__global__ void gather_ia_kernel(const double* __restrict__ eri, int naux, int ld_m, int64_t ld_n,
                                 const int32_t* __restrict__ ia, int npole, double* __restrict__ x) {
  const int t = blockIdx.y;
  const int i = ia[2 * t], a = ia[2 * t + 1];
  const double* src = eri + (int64_t)a * ld_m + (int64_t)i * ld_n;
  for (int P = blockIdx.x * blockDim.x + threadIdx.x; P < naux; P += gridDim.x * blockDim.x)
    x[P + (int64_t)naux * t] = src[P];
}

int launch_gather_ia(const double* eri, int naux, int nm, int nn, const int32_t* ia_dev, int npole,
                     double* x, hipStream_t s) {
  (void)nn;
  MGW_REQUIRE(npole <= 65535 * 64, MGW_ERR_ARG, "gather_ia: too many transitions");
  // one grid row per transition (gridDim.y limit 65535 handled by chunking)
  for (int t0 = 0; t0 < npole; t0 += 65535) {
    const int cnt = std::min(65535, npole - t0);
    dim3 grid(nblocks(naux, 256), cnt);
    hipLaunchKernelGGL(gather_ia_kernel, grid, dim3(256), 0, s, eri, naux, naux,
                       (int64_t)naux * nm, ia_dev + 2 * (int64_t)t0, cnt, x + (int64_t)naux * t0);
    MGW_LAUNCH_CHECK();
  }
  return MGW_OK;
}

// ------------------------------------------------------------------ reductions
// Column sums of squares: out[t] = sum_P x(P, t)^2 (one wave per column).
__global__ void colnorm2_kernel(const double* __restrict__ x, int64_t ld, int rows, int cols,
                                double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (col >= cols) return;
  const double* c = x + (int64_t)col * ld;
  double s = 0.0;
  for (int r = lane; r < rows; r += 64) s += c[r] * c[r];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) out[col] = s;
}

int launch_colnorm2(const double* x, int64_t ld, int rows, int cols, double* out, hipStream_t s) {
  if (cols <= 0) return MGW_OK;
  hipLaunchKernelGGL(colnorm2_kernel, dim3(nblocks(cols, 4)), dim3(256), 0, s, x, ld, rows, cols, out);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// Deterministic block-partial sum of a vector (optionally squared); the
// partials (one per block, <= 1024 blocks) are summed on the host.
__global__ void vsum_kernel(const double* __restrict__ x, int64_t n, int square,
                            double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double v = x[i];
    s += square ? v * v : v;
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

int launch_vsum(const double* x, int64_t n, int square, double* partial, int nblk, hipStream_t s) {
  hipLaunchKernelGGL(vsum_kernel, dim3(nblk), dim3(256), 0, s, x, n, square, partial);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ Casida
// de[t] = e_a - e_i ; sq[t] = sqrt(de) ; de2[t] = de^2.
__global__ void casida_diag_kernel(const int32_t* __restrict__ ia, int npole,
                                   const double* __restrict__ energy, double* __restrict__ de,
                                   double* __restrict__ sq, double* __restrict__ de2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npole) return;
  const double d = energy[ia[2 * t + 1]] - energy[ia[2 * t]];
  de[t] = d;
  sq[t] = sqrt(d);
  de2[t] = d * d;
}

int launch_casida_diag(const int32_t* ia, int npole, const double* energy, double* de, double* sq,
                       double* de2, hipStream_t s) {
  hipLaunchKernelGGL(casida_diag_kernel, dim3(nblocks(npole, 256)), dim3(256), 0, s, ia, npole,
                     energy, de, sq, de2);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// Omega = sqrt(eigenvalue), colscale = Omega^-1/2 (src/m_block_diago.f90:188-195).
__global__ void casida_post_kernel(const double* __restrict__ eig, int npole,
                                   double* __restrict__ pole, double* __restrict__ cs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npole) return;
  const double om = sqrt(eig[t]);
  pole[t] = om;
  cs[t] = 1.0 / sqrt(om);
}

int launch_casida_post(const double* eig, int npole, double* pole, double* cs, hipStream_t s) {
  hipLaunchKernelGGL(casida_post_kernel, dim3(nblocks(npole, 256)), dim3(256), 0, s, eig, npole, pole, cs);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ sigma poles
// sigma(w_k, p) += sum_{i in chunk} sum_s w(s,p,i)^2
//      [ f_i / (x + Omega_s - i eta) + (1 - f_i) / (x - Omega_s + i eta) ],
// x = E0_p + w_k - e_i (src/m_gw_selfenergy_analytic.f90:126-133), and
// E_GW += (1 - f_i) occ_p w^2 / (e_p - e_i - Omega_s) for filled p (:135-138).
//
// Workgroup = 4 wave64s for one state p and one block of SB poles. Lane k of
// every wave owns frequency w_k (nomega <= 63); the spare lane 63 accumulates
// the E_GW term. Wave j walks the poles s = s0 + j, s0 + j + 4, ... of the
// block, reading w^2 and Omega from LDS (one coalesced load per block per i).
// The partial sums of the block are written per (sblock, p) and reduced over
// blocks by sigma_reduce_kernel: the result does not depend on scheduling.
constexpr int SIGMA_SB = 512;

__global__ void __launch_bounds__(256)
sigma_poles_kernel(const double* __restrict__ w, int npole, int nse, int ni,
                   const double* __restrict__ pole, const double* __restrict__ omega, int nomega,
                   int with_egw, const double* __restrict__ e0p, const double* __restrict__ ei,
                   const double* __restrict__ fi, const double* __restrict__ ep,
                   const double* __restrict__ occp, double spin_fact, double eta,
                   double* __restrict__ part /* [nsblock][nse][64] complex */) {
  __shared__ double w2s[SIGMA_SB];
  __shared__ double oms[SIGMA_SB];
  __shared__ double red[4][64][2];
  const int p = blockIdx.y;
  const int sb = blockIdx.x;
  const int s0 = sb * SIGMA_SB;
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int ns = min(SIGMA_SB, npole - s0);
  const double om_k = lane < nomega ? omega[lane] : 0.0;
  const double e0 = e0p[p];
  const bool p_full = ((spin_fact - occp[p]) / spin_fact) < 1.0e-5;
  const double occ_p = occp[p];
  const double eps_p = ep[p];
  const double eta2 = eta * eta;
  double re = 0.0, im = 0.0, egw = 0.0;
  for (int s = threadIdx.x; s < ns; s += 256) oms[s] = pole[s0 + s];
  for (int ii = 0; ii < ni; ++ii) {
    __syncthreads();
    const double* wcol = w + (int64_t)npole * ((int64_t)p + (int64_t)nse * ii) + s0;
    for (int s = threadIdx.x; s < ns; s += 256) {
      const double v = wcol[s];
      w2s[s] = v * v;
    }
    __syncthreads();
    const double f_full = fi[ii] / spin_fact;
    const double f_empty = (spin_fact - fi[ii]) / spin_fact;
    const double x = e0 + om_k - ei[ii];
    if (lane < nomega) {
      if (f_full != 0.0) {
        double sre = 0.0, sim = 0.0;
        for (int s = wv; s < ns; s += 4) {
          const double a = x + oms[s];
          const double r = w2s[s] / (a * a + eta2);
          sre = fma(a, r, sre);
          sim += r;
        }
        re = fma(f_full, sre, re);
        im = fma(f_full * eta, sim, im);
      }
      if (f_empty != 0.0) {
        double sre = 0.0, sim = 0.0;
        for (int s = wv; s < ns; s += 4) {
          const double b = x - oms[s];
          const double r = w2s[s] / (b * b + eta2);
          sre = fma(b, r, sre);
          sim += r;
        }
        re = fma(f_empty, sre, re);
        im = fma(-f_empty * eta, sim, im);
      }
    } else if (lane == 63 && with_egw && p_full && f_empty != 0.0) {
      const double d0 = eps_p - ei[ii];
      double acc = 0.0;
      for (int s = wv; s < ns; s += 4) acc += w2s[s] / (d0 - oms[s]);
      egw = fma(f_empty * occ_p, acc, egw);
    }
  }
  red[wv][lane][0] = (lane == 63) ? egw : re;
  red[wv][lane][1] = (lane == 63) ? 0.0 : im;
  __syncthreads();
  if (threadIdx.x < 64) {
    const int k = threadIdx.x;
    double r0 = 0.0, r1 = 0.0;
    for (int j = 0; j < 4; ++j) { r0 += red[j][k][0]; r1 += red[j][k][1]; }
    double* out = part + 2 * (64 * ((int64_t)sb * nse + p) + k);
    out[0] = r0;
    out[1] = r1;
  }
}

// sigma[(k, p)] (+)= sum_sb part[sb][p][k]; egw partial per p in lane 63.
__global__ void sigma_reduce_kernel(const double* __restrict__ part, int nsblock, int nse, int nomega,
                                    int nomega_total, int k0, int accumulate,
                                    double* __restrict__ sigma, double* __restrict__ egw_p) {
  const int p = blockIdx.x;
  const int k = threadIdx.x;
  double r0 = 0.0, r1 = 0.0;
  for (int sb = 0; sb < nsblock; ++sb) {
    const double* q = part + 2 * (64 * ((int64_t)sb * nse + p) + k);
    r0 += q[0];
    r1 += q[1];
  }
  if (k < nomega) {
    double* o = sigma + 2 * ((int64_t)p * nomega_total + k0 + k);
    if (accumulate) { o[0] += r0; o[1] += r1; } else { o[0] = r0; o[1] = r1; }
  } else if (k == 63 && k0 == 0) {
    if (accumulate) egw_p[p] += r0; else egw_p[p] = r0;
  }
}

int launch_sigma_poles(const double* w, int npole, int nse, int ni, const double* pole,
                       const double* omega, int nomega, const double* e0p, const double* ei,
                       const double* fi, const double* ep, const double* occp, double spin_fact,
                       double eta, double* part, int accumulate, double* sigma, double* egw_p,
                       hipStream_t s) {
  MGW_REQUIRE(nomega >= 1, MGW_ERR_ARG, "sigma_poles: nomega must be positive");
  MGW_REQUIRE(nse <= 65535, MGW_ERR_ARG, "sigma_poles: too many states");
  const int nsb = (npole + SIGMA_SB - 1) / SIGMA_SB;
  // frequencies in groups of 63 lanes; lane 63 of the first group carries E_GW
  for (int k0 = 0; k0 < nomega; k0 += 63) {
    const int ng = std::min(63, nomega - k0);
    hipLaunchKernelGGL(sigma_poles_kernel, dim3(nsb, nse), dim3(256), 0, s, w, npole, nse, ni, pole,
                       omega + k0, ng, k0 == 0 ? 1 : 0, e0p, ei, fi, ep, occp, spin_fact, eta, part);
    MGW_LAUNCH_CHECK();
    hipLaunchKernelGGL(sigma_reduce_kernel, dim3(nse), dim3(64), 0, s, part, nsb, nse, ng, nomega, k0,
                       accumulate, sigma, egw_p);
    MGW_LAUNCH_CHECK();
  }
  return MGW_OK;
}

int sigma_part_doubles(int npole, int nse) {
  return ((npole + SIGMA_SB - 1) / SIGMA_SB) * nse * 64 * 2;
}

// ------------------------------------------------------------------ chi (imaginary axis)
// s(t, iw) = 2 docc_t de_t / (w^2 + de_t^2) (src/m_gw_selfenergy_grid.f90:126-137,
// applied once because both SYRK operands carry its square root).
__global__ void chi_weights_kernel(const int32_t* __restrict__ ia, int npole,
                                   const double* __restrict__ energy,
                                   const double* __restrict__ occ, const double* __restrict__ omega,
                                   int nomega, double wsign, double* __restrict__ sw) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int iw = blockIdx.y;
  if (t >= npole) return;
  const int i = ia[2 * t], a = ia[2 * t + 1];
  const double docc = occ[i] - occ[a];
  const double de = energy[a] - energy[i];
  const double w = omega[iw];
  sw[t + (int64_t)npole * iw] = 2.0 * docc * de / (wsign * w * w + de * de);
}

int launch_chi_weights(const int32_t* ia, int npole, const double* energy, const double* occ,
                       const double* omega, int nomega, double* sw, hipStream_t s, double wsign) {
  hipLaunchKernelGGL(chi_weights_kernel, dim3(nblocks(npole, 256), nomega), dim3(256), 0, s, ia,
                     npole, energy, occ, omega, nomega, wsign, sw);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ contour deformation
// Low-rank W from the eigenvalues lambda of chi0 (src/m_spectral_function.f90:892-913):
// w = lambda / (1 - lambda) where |w| > tol, else 0 (the dropped eigenpairs).
__global__ void contour_eta_kernel(const double* __restrict__ lam, int64_t n, double tol,
                                   double* __restrict__ w) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const double l = lam[k];
  const double e = l / (1.0 - l);
  w[k] = fabs(e) > tol ? e : 0.0;
}

int launch_contour_eta(const double* lam, int64_t n, double tol, double* w, hipStream_t s) {
  hipLaunchKernelGGL(contour_eta_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, lam, n, tol, w);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// sigma(k, m) of gw_selfenergy_contour (src/m_gw_selfenergy_grid.f90:429-530)
// for x = E0_m + omega_k - e_p and v = B_pm^T W B_pm:
//   imaginary axis, |x| > eta: -sum_f wt_f / (2 pi) [1/(x + i w_f) + 1/(x - i w_f)] v_f
//   occupied p (p < nlumo), de = e_p - (E0_m + omega_k) >= -eta:  -v(|de|) c
//   empty p (p > nhomo),    de = E0_m + omega_k - e_p >= -eta:    +v(|de|) c
// with c = 1/2 when |de| < eta, else 1, and v(|de|) at the real-grid point
// nearest to |de| (the first on a tie, MINLOC). v holds the imaginary
// frequencies first, then the real ones: v[(p + np m) + ncol f]. A |de|
// outside the real grid sets *err. One workgroup per (k, m).
__global__ void sigma_contour_kernel(const double* __restrict__ v, int np, int nse, int nimag, int nreal,
                                     const double* __restrict__ om_im, const double* __restrict__ wt_im,
                                     const double* __restrict__ om_re, const double* __restrict__ ep,
                                     const int32_t* __restrict__ pstate, int nhomo, int nlumo,
                                     const double* __restrict__ e0, const double* __restrict__ omega,
                                     int nomega, double eta, double* __restrict__ sigma, int32_t* __restrict__ err) {
  __shared__ double red[4];
  const int k = blockIdx.x, m = blockIdx.y;
  const int64_t ncol = (int64_t)np * nse;
  const double w = e0[m] + omega[k];
  double lo = om_re[0], hi = om_re[0];
  for (int j = 1; j < nreal; ++j) { lo = fmin(lo, om_re[j]); hi = fmax(hi, om_re[j]); }
  double acc = 0.0;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const int64_t col = p + (int64_t)np * m;
    const double x = w - ep[p];
    if (fabs(x) > eta) {
      double s = 0.0;
      for (int f = 0; f < nimag; ++f) {
        const double wf = om_im[f];
        s += wt_im[f] * (2.0 * x / (x * x + wf * wf)) * v[col + ncol * f];
      }
      acc -= s / (2.0 * 3.14159265358979323);
    }
    const int ps = pstate[p];
    for (int branch = 0; branch < 2; ++branch) {
      double de;
      if (branch == 0) {
        if (ps >= nlumo) continue;
        de = ep[p] - w;
      } else {
        if (ps <= nhomo) continue;
        de = w - ep[p];
      }
      if (de < -eta) continue;
      const double a = fabs(de);
      if (a > hi || a < lo) {
        err[0] = 1;
        continue;
      }
      int jbest = 0;
      double dbest = fabs(om_re[0] - a);
      for (int j = 1; j < nreal; ++j) {
        const double d = fabs(om_re[j] - a);
        if (d < dbest) { dbest = d; jbest = j; }
      }
      const double r = v[col + ncol * (nimag + jbest)] * (a < eta ? 0.5 : 1.0);
      acc += branch == 0 ? -r : r;
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = sigma + 2 * ((int64_t)m * nomega + k);
    o[0] = red[0] + red[1] + red[2] + red[3];
    o[1] = 0.0;
  }
}

int launch_sigma_contour(const double* v, int np, int nse, int nimag, int nreal, const double* om_im,
                         const double* wt_im, const double* om_re, const double* ep, const int32_t* pstate,
                         int nhomo, int nlumo, const double* e0, const double* omega, int nomega, double eta,
                         double* sigma, int32_t* err, hipStream_t s) {
  hipLaunchKernelGGL(sigma_contour_kernel, dim3(nomega, nse), dim3(256), 0, s, v, np, nse, nimag, nreal, om_im,
                     wt_im, om_re, ep, pstate, nhomo, nlumo, e0, omega, nomega, eta, sigma, err);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// Per frequency: out[iw] = { sum_t s_t n_t (= -Tr chi0), sum_i ln L_ii (Cholesky) }
// and later { Tr (1-chi0)^-1 }. One workgroup per frequency.
__global__ void chi_trace_kernel(const double* __restrict__ sw, const double* __restrict__ n2,
                                 int npole, const double* __restrict__ mat, int naux,
                                 int mode, double* __restrict__ out) {
  __shared__ double red[4];
  const int iw = blockIdx.x;
  double s = 0.0;
  if (mode == 0) {
    const double* w = sw + (int64_t)npole * iw;
    for (int t = threadIdx.x; t < npole; t += blockDim.x) s += w[t] * n2[t];
  } else {
    const double* m = mat + (int64_t)naux * naux * iw;
    for (int i = threadIdx.x; i < naux; i += blockDim.x) {
      const double d = m[i + (int64_t)naux * i];
      s += (mode == 1) ? log(d) : d;
    }
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[iw] = red[0] + red[1] + red[2] + red[3];
}

int launch_chi_trace(const double* sw, const double* n2, int npole, const double* mat, int naux,
                     int mode, int nomega, double* out, hipStream_t s) {
  hipLaunchKernelGGL(chi_trace_kernel, dim3(nomega), dim3(256), 0, s, sw, n2, npole, mat, naux,
                     mode, out);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ sigma imag
// v[(p, m, iw)] = sum_rowblocks dotpart[iw][rb][col] - sub[col], col = p + np * m
// (sub nullable).
__global__ void dot_reduce_kernel(const double* __restrict__ part, int nrb, int64_t ncol, int nomega,
                                  const double* __restrict__ sub, double* __restrict__ v) {
  const int64_t col = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int iw = blockIdx.y;
  if (col >= ncol) return;
  const double* q = part + (int64_t)nrb * ncol * iw;
  double s = 0.0;
  for (int rb = 0; rb < nrb; ++rb) s += q[(int64_t)rb * ncol + col];
  v[col + ncol * iw] = sub ? s - sub[col] : s;
}

int launch_dot_reduce(const double* part, int nrb, int64_t ncol, int nomega, const double* sub, double* v,
                      hipStream_t s) {
  hipLaunchKernelGGL(dot_reduce_kernel, dim3(nblocks(ncol, 256), nomega), dim3(256), 0, s, part, nrb,
                     ncol, nomega, sub, v);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// sigma_calc(kc, m) = - sum_iw weight_iw / (2 pi) sum_p
//     [1/((wc - e_p) + i w) + 1/((wc - e_p) - i w)] v(p, m, iw)
// (src/m_gw_selfenergy_grid.f90:306-317). One workgroup per (kc, m).
__global__ void sigma_imag_sum_kernel(const double* __restrict__ v, int np, int nse, int nomega,
                                      const double* __restrict__ omega, const double* __restrict__ weight,
                                      const double* __restrict__ ep, const double* __restrict__ wc,
                                      int ncalc, double* __restrict__ sigma_calc) {
  __shared__ double red[4][2];
  const int kc = blockIdx.x, m = blockIdx.y;
  const double wr = wc[2 * kc], wi = wc[2 * kc + 1];
  double sr = 0.0, si = 0.0;
  for (int iw = 0; iw < nomega; ++iw) {
    const double w = omega[iw];
    const double c = -weight[iw] / (2.0 * 3.14159265358979323);
    const double* vv = v + (int64_t)np * m + (int64_t)np * nse * iw;
    double ar = 0.0, ai = 0.0;
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const double dr = wr - ep[p];
      // 1/(dr + i(wi + w)) + 1/(dr + i(wi - w))
      const double i1 = wi + w, i2 = wi - w;
      const double d1 = dr * dr + i1 * i1, d2 = dr * dr + i2 * i2;
      const double gr = dr / d1 + dr / d2;
      const double gi = -i1 / d1 - i2 / d2;
      ar += gr * vv[p];
      ai += gi * vv[p];
    }
    sr += c * ar;
    si += c * ai;
  }
  for (int o = 32; o > 0; o >>= 1) { sr += __shfl_xor(sr, o, 64); si += __shfl_xor(si, o, 64); }
  if ((threadIdx.x & 63) == 0) { red[threadIdx.x >> 6][0] = sr; red[threadIdx.x >> 6][1] = si; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double* o = sigma_calc + 2 * ((int64_t)m * ncalc + kc);
    o[0] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    o[1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  }
}

int launch_sigma_imag_sum(const double* v, int np, int nse, int nomega, const double* omega,
                          const double* weight, const double* ep, const double* wc, int ncalc,
                          double* sigma_calc, hipStream_t s) {
  hipLaunchKernelGGL(sigma_imag_sum_kernel, dim3(ncalc, nse), dim3(256), 0, s, v, np, nse, nomega,
                     omega, weight, ep, wc, ncalc, sigma_calc);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ Pade
// Thiele continued fraction through the 2*ncalc-1 symmetrised points
// (src/m_selfenergy_tools.f90:965-999, src/m_numerical_tools.f90:86-145).
// One thread per state; the coefficient table is built in place in a
// per-thread scratch column of `work` (npts x npts complex).
struct cplx { double r, i; };
__device__ inline cplx cmul(cplx a, cplx b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
__device__ inline cplx csub(cplx a, cplx b) { return {a.r - b.r, a.i - b.i}; }
__device__ inline cplx cadd(cplx a, cplx b) { return {a.r + b.r, a.i + b.i}; }
__device__ inline cplx cdiv(cplx a, cplx b) {
  // Smith's algorithm: robust complex division
  if (fabs(b.r) >= fabs(b.i)) {
    const double r = b.i / b.r, d = b.r + b.i * r;
    return {(a.r + a.i * r) / d, (a.i - a.r * r) / d};
  } else {
    const double r = b.r / b.i, d = b.r * r + b.i;
    return {(a.r * r + a.i) / d, (a.i * r - a.r) / d};
  }
}

__global__ void pade_kernel(const double* __restrict__ wc, const double* __restrict__ sc, int ncalc,
                            int nse, const double* __restrict__ omega, int nomega, double eta,
                            double* __restrict__ work, double* __restrict__ sigma) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nse) return;
  const int np = 2 * ncalc - 1;
  cplx* z = reinterpret_cast<cplx*>(work) + (int64_t)p * (2 * np);
  cplx* g = z + np;   // holds the current row of the table; diagonal kept in aa
  cplx* aa = reinterpret_cast<cplx*>(work) + (int64_t)nse * 2 * np + (int64_t)p * np;
  const cplx* sp = reinterpret_cast<const cplx*>(sc) + (int64_t)p * ncalc;
  const cplx* w = reinterpret_cast<const cplx*>(wc);
  // symmetric points: z(nc) = wc(1); z(nc+k-1) = wc(k); z(nc-k+1) = conj(wc(k))
  z[ncalc - 1] = w[0];
  g[ncalc - 1] = sp[0];
  for (int k = 1; k < ncalc; ++k) {
    z[ncalc - 1 + k] = w[k];
    g[ncalc - 1 + k] = sp[k];
    z[ncalc - 1 - k] = {w[k].r, -w[k].i};
    g[ncalc - 1 - k] = {sp[k].r, -sp[k].i};
  }
  // g(ip, jp) = (g(ip-1, ip-1) - g(ip-1, jp)) / ((z(jp) - z(ip-1)) g(ip-1, jp))
  aa[0] = g[0];
  for (int ip = 1; ip < np; ++ip) {
    const cplx gd = g[ip - 1];
    for (int jp = ip; jp < np; ++jp) g[jp] = cdiv(csub(gd, g[jp]), cmul(csub(z[jp], z[ip - 1]), g[jp]));
    aa[ip] = g[ip];
  }
  for (int k = 0; k < nomega; ++k) {
    const double sgn = (omega[k] - w[0].r) >= 0.0 ? -1.0 : 1.0;
    const cplx zo = {omega[k], eta * sgn};
    cplx a_prev = {0.0, 0.0}, a_cur = aa[0], b_prev = {1.0, 0.0}, b_cur = {1.0, 0.0};
    for (int ip = 1; ip < np; ++ip) {
      const cplx f = cmul(csub(zo, z[ip - 1]), aa[ip]);
      const cplx a_next = cadd(a_cur, cmul(f, a_prev));
      const cplx b_next = cadd(b_cur, cmul(f, b_prev));
      a_prev = a_cur; a_cur = a_next;
      b_prev = b_cur; b_cur = b_next;
    }
    const cplx r = cdiv(a_cur, b_cur);
    sigma[2 * ((int64_t)p * nomega + k)] = r.r;
    sigma[2 * ((int64_t)p * nomega + k) + 1] = r.i;
  }
}

// The same continued fraction with one wave per state (2 ncalc - 1 <= 64):
// lane j keeps point j and its table entry in registers, the row's pivot
// g(ip-1) comes by a lane broadcast, and the lanes then evaluate the
// fraction at one real frequency each. Every element sees the operations of
// pade_kernel in the same order.
__device__ inline cplx shfl_c(cplx v, int src) { return {__shfl(v.r, src, 64), __shfl(v.i, src, 64)}; }

__global__ void __launch_bounds__(64)
pade_wave_kernel(const double* __restrict__ wc, const double* __restrict__ sc, int ncalc, int nse,
                 const double* __restrict__ omega, int nomega, double eta, double* __restrict__ sigma) {
  __shared__ cplx zs[64], aas[64];
  const int p = blockIdx.x, lane = threadIdx.x;
  const int np = 2 * ncalc - 1;
  const cplx* sp = reinterpret_cast<const cplx*>(sc) + (int64_t)p * ncalc;
  const cplx* w = reinterpret_cast<const cplx*>(wc);
  cplx zj = {0.0, 0.0}, gj = {0.0, 0.0};
  if (lane < np) {
    const int k = lane - (ncalc - 1);
    if (k >= 0) {
      zj = w[k];
      gj = sp[k];
    } else {
      zj = {w[-k].r, -w[-k].i};
      gj = {sp[-k].r, -sp[-k].i};
    }
  }
  zs[lane] = zj;
  if (lane == 0) aas[0] = gj;
  __syncthreads();
  for (int ip = 1; ip < np; ++ip) {
    const cplx gd = shfl_c(gj, ip - 1);
    if (lane >= ip && lane < np) gj = cdiv(csub(gd, gj), cmul(csub(zj, zs[ip - 1]), gj));
    if (lane == ip) aas[ip] = gj;
  }
  __syncthreads();
  for (int k = lane; k < nomega; k += 64) {
    const double sgn = (omega[k] - w[0].r) >= 0.0 ? -1.0 : 1.0;
    const cplx zo = {omega[k], eta * sgn};
    cplx a_prev = {0.0, 0.0}, a_cur = aas[0], b_prev = {1.0, 0.0}, b_cur = {1.0, 0.0};
    for (int ip = 1; ip < np; ++ip) {
      const cplx f = cmul(csub(zo, zs[ip - 1]), aas[ip]);
      const cplx a_next = cadd(a_cur, cmul(f, a_prev));
      const cplx b_next = cadd(b_cur, cmul(f, b_prev));
      a_prev = a_cur; a_cur = a_next;
      b_prev = b_cur; b_cur = b_next;
    }
    const cplx r = cdiv(a_cur, b_cur);
    sigma[2 * ((int64_t)p * nomega + k)] = r.r;
    sigma[2 * ((int64_t)p * nomega + k) + 1] = r.i;
  }
}

int pade_work_doubles(int ncalc, int nse) {
  const int np = 2 * ncalc - 1;
  return nse * (2 * np) * 2 + nse * np * 2;
}

int launch_pade(const double* wc, const double* sc, int ncalc, int nse, const double* omega,
                int nomega, double eta, double* work, double* sigma, hipStream_t s) {
  if (2 * ncalc - 1 <= 64)
    hipLaunchKernelGGL(pade_wave_kernel, dim3(nse), dim3(64), 0, s, wc, sc, ncalc, nse, omega, nomega, eta, sigma);
  else
    hipLaunchKernelGGL(pade_kernel, dim3(nblocks(nse, 64)), dim3(64), 0, s, wc, sc, ncalc, nse, omega,
                       nomega, eta, work, sigma);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

// ------------------------------------------------------------------ QP equation
// Graphical solution (src/m_selfenergy_tools.f90:253-420) and linearisation
// (:202-249) per state p of [nsemin, nsemax]. One thread per state.
__global__ void qp_kernel(const double* __restrict__ sigma, int nomega, int nse,
                          const double* __restrict__ omega, const double* __restrict__ se_e0,
                          const double* __restrict__ exc, const double* __restrict__ e0,
                          double* __restrict__ eqp, double* __restrict__ zz,
                          double* __restrict__ elin) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nse) return;
  constexpr int NFP = 4;
  double zw[NFP], ef[NFP];
  for (int j = 0; j < NFP; ++j) { zw[j] = -1.0; ef[j] = 0.0; }
  const double* sg = sigma + 2 * (int64_t)p * nomega;
  const double c = exc[p] + e0[p];
  double xprev = omega[0] + se_e0[p];
  double fprev = sg[0] + c;
  double gprev = fprev - xprev;
  for (int ix = 1; ix < nomega; ++ix) {
    const double x = omega[ix] + se_e0[p];
    const double f = sg[2 * ix] + c;
    const double gx = f - x;
    if (gprev * gx < 0.0) {
      const double z0 = 1.0 / (1.0 - (f - fprev) / (x - xprev));
      if (z0 >= 0.00001 && z0 > zw[NFP - 1]) {
        int jf = 0;
        while (!(z0 > zw[jf])) ++jf;
        for (int j = NFP - 1; j > jf; --j) { zw[j] = zw[j - 1]; ef[j] = ef[j - 1]; }
        zw[jf] = z0;
        const double gp = (gx - gprev) / (x - xprev);
        ef[jf] = xprev - gprev / gp;
      }
    }
    xprev = x; fprev = f; gprev = gx;
  }
  const int n0 = nomega / 2;
  double e, z;
  double zmax = zw[0];
  int iq = 0;
  for (int j = 1; j < NFP; ++j) if (zw[j] > zmax) { zmax = zw[j]; iq = j; }
  if (zmax > 1.0e-6) { e = ef[iq]; z = zw[iq]; } else { e = e0[p]; z = 1.0; }
  int nfixed = 0;
  for (int j = 0; j < NFP; ++j) nfixed += zw[j] > 0.0;
  if (nfixed == 0) e = e0[p] + sg[2 * n0] + exc[p];
  eqp[p] = e;
  zz[p] = z;
  // linearisation
  if (n0 > 0) {
    double zl = (sg[2 * (n0 + 1)] - sg[2 * (n0 - 1)]) / (omega[n0 + 1] - omega[n0 - 1]);
    zl = 1.0 / (1.0 - zl);
    zl = fmin(fmax(zl, 0.0), 1.0);
    elin[p] = se_e0[p] + zl * (e0[p] - se_e0[p] + sg[2 * n0] + exc[p]);
  } else {
    elin[p] = e0[p] + sg[2 * n0] + exc[p];
  }
}

int launch_qp(const double* sigma, int nomega, int nse, const double* omega, const double* se_e0,
              const double* exc, const double* e0, double* eqp, double* zz, double* elin,
              hipStream_t s) {
  hipLaunchKernelGGL(qp_kernel, dim3(nblocks(nse, 64)), dim3(64), 0, s, sigma, nomega, nse, omega,
                     se_e0, exc, e0, eqp, zz, elin);
  MGW_LAUNCH_CHECK();
  return MGW_OK;
}

}  // namespace mgw
