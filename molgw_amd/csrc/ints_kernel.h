// Device code of the two- and three-centre Coulomb integrals over contracted
// Gaussian shells on the GPU (one launcher per translation unit, ints_l*.hip), for the RI tensor of the integral front end:
//   (mu nu|P)  of calculate_integrals_eri_3center_scalapack (src/m_eri_calculate.f90:1266-1580),
//   (P|Q)      of calculate_integrals_eri_2center_scalapack (src/m_eri_calculate.f90:703-945),
// which the reference takes from libcint (cint3c2e / cint2c2e,
// src/m_libcint_tools.f90:54-260). Written from the McMurchie-Davidson
// scheme:
//   (ab|c) = 2 pi^(5/2) / (p q sqrt(p + q)) sum_tuv E^ab_tuv sum_tau,nu,phi
//            (-1)^(tau+nu+phi) E^c_(tau nu phi) R_(t+tau, u+nu, v+phi)(alpha, P - C)
// with E the Hermite expansion coefficients of the bra product and of the
// single ket Gaussian, alpha = p q / (p + q), and R the Hermite Coulomb
// integrals from the Boys function.
//
// Work decomposition. Shell pairs are grouped by (l_a, l_b), ket functions
// (shell, function) by l_c; one launch per class (l_a, l_b, l_c) with the
// three angular momenta as template parameters, so every loop below unrolls
// and indices into R, E and the accumulators are compile-time constants. A
// workgroup owns one bra shell pair and 128 ket functions of one l_c; the
// bra's primitive pairs are expanded (E coefficients, P, p) cooperatively
// into LDS in batches and read as broadcasts. Each thread keeps, per ket
// primitive, R in one in-place array (levels n = L .. 0, sums t+u+v
// processed from high to low so that level n + 1 is read before it is
// overwritten), the ket's Hermite coefficients of its final function, and
// the Cartesian bra block of its function in registers; the bra is
// transformed to the final functions once at the end.
//
// Boys function: F_L(T) by a 7-term Taylor series around the nearest point
// of a table F_n(k h), h = 0.05, T <= 36 (error below 2e-15 relative), then
// downward recursion; above T = 36, F_0 = sqrt(pi / T) / 2 (the neglected
// tail is e^-T < 3e-16) and upward recursion.
#pragma once
#include "mgw_internal.h"

#include <cmath>

namespace mgw {
namespace eri {

constexpr int BOYS_N = 17;          // orders 0 .. 16 in the table (L <= 10 plus 6 Taylor terms)
constexpr int BOYS_TAYLOR = 6;
constexpr double BOYS_H = 0.05;
constexpr double BOYS_TMAX = 36.0;
constexpr int BOYS_NT = 721;        // points 0, h, .., 36

__host__ __device__ constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
__host__ __device__ constexpr int nherm(int L) { return (L + 1) * (L + 2) * (L + 3) / 6; }
// R_{t,u,v}, t + u + v <= L: sums s in increasing order, then (u + v, v)
__host__ __device__ constexpr int hidx(int t, int u, int v) {
  const int s = t + u + v, r = u + v;
  return s * (s + 1) * (s + 2) / 6 + r * (r + 1) / 2 + v;
}
// Cartesian component c of a shell of angular momentum l: (l - i, i - k, k)
// for i = 0 .. l, k = 0 .. i (x power decreasing first), as the basis module
// orders them
__host__ __device__ constexpr int cart_i(int c) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= c) ++i;
  return i;
}
__host__ __device__ constexpr int cart_x(int l, int c) { return l - cart_i(c); }
__host__ __device__ constexpr int cart_y(int l, int c) { return cart_i(c) - (c - cart_i(c) * (cart_i(c) + 1) / 2); }
__host__ __device__ constexpr int cart_z(int l, int c) { return c - cart_i(c) * (cart_i(c) + 1) / 2; }

struct ShellsDev {
  const int32_t* l;
  const double* center;
  const int32_t* poff;
  const double* exps;
  const double* coefs;
  const int32_t* toff;
  const double* tmat;
  const int32_t* foff;
};

struct EriArgs {
  ShellsDev A, B, C;
  const int2* pairs;       // bra shell pairs (sa in A, sb in B) of the class
  const int2* kfun;        // ket functions (sc, f) of the class
  int nkfun;
  const double* boys;      // [BOYS_NT][BOYS_N]
  int metric;              // 0: packed (mu nu|P); 1: (P|Q) with B the unit shells
  const int32_t* pairmap;  // 0: pair index of (mu, nu) (nbf x nbf) or -1
  int nbf;
  int64_t npair;
  int naux;                // 1: leading dimension of (P|Q)
  double* out;
};

template <int L>
__device__ inline void boys_eval(double T, const double* __restrict__ tab, double* f) {
  if (T < BOYS_TMAX) {
    const int k = (int)(T * (1.0 / BOYS_H) + 0.5);
    const double d = k * BOYS_H - T;
    const double* __restrict__ row = tab + k * BOYS_N + L;
    // F_L(T) = sum_j F_{L+j}(t_k) (t_k - T)^j / j!
    double s = row[BOYS_TAYLOR];
#pragma unroll
    for (int j = BOYS_TAYLOR - 1; j >= 0; --j) s = row[j] + s * d / (double)(j + 1);
    f[L] = s;
    const double e = exp(-T);
#pragma unroll
    for (int n = L - 1; n >= 0; --n) f[n] = (2.0 * T * f[n + 1] + e) / (double)(2 * n + 1);
  } else {
    const double e = exp(-T);
    f[0] = 0.5 * sqrt(3.14159265358979323846 / T);
#pragma unroll
    for (int n = 0; n < L; ++n) f[n + 1] = ((double)(2 * n + 1) * f[n] - e) / (2.0 * T);
  }
}

// Hermite expansion of one dimension of a primitive pair into e[(i (LB+1) + j) (LAB+1) + t]
// (i <= LA, j <= LB, t <= i + j); e[0] carries exp(-mu X_AB^2).
template <int LA, int LB>
__device__ inline void hermite_pair(double a, double b, double xab, double* __restrict__ e) {
  constexpr int LAB = LA + LB, LD = LAB + 1;
  const double p = a + b, mu = a * b / p;
  const double xpa = -b * xab / p, xpb = a * xab / p, h = 0.5 / p;
  for (int n = 0; n < (LA + 1) * (LB + 1) * LD; ++n) e[n] = 0.0;
  e[0] = exp(-mu * xab * xab);
  for (int i = 0; i <= LA; ++i)
    for (int j = 0; j <= LB; ++j) {
      if (i == 0 && j == 0) continue;
      const double* src;
      double x;
      if (i > 0) {
        src = e + ((i - 1) * (LB + 1) + j) * LD;
        x = xpa;
      } else {
        src = e + (j - 1) * LD;
        x = xpb;
      }
      double* dst = e + (i * (LB + 1) + j) * LD;
      const int top = i + j - 1;
      for (int t = 0; t <= i + j; ++t) {
        double v = (t <= top ? x * src[t] : 0.0);
        if (t > 0 && t - 1 <= top) v += h * src[t - 1];
        if (t + 1 <= top) v += (t + 1) * src[t + 1];
        dst[t] = v;
      }
    }
}

constexpr int WG = 128;   // ket functions per workgroup
constexpr int NPB = 8;    // bra primitive pairs expanded per LDS batch

template <int LA, int LB, int LC>
__global__ void __launch_bounds__(WG) eri_kernel(EriArgs g) {
  constexpr int LAB = LA + LB, L = LAB + LC;
  constexpr int NA = ncart(LA), NB = ncart(LB), NC = ncart(LC);
  constexpr int LD = LAB + 1;
  constexpr int NE = (LA + 1) * (LB + 1) * LD;
  __shared__ double sE[NPB][3][NE];
  __shared__ double sQ[NPB][5];   // p, P (3), contraction weight

  const int2 pr = g.pairs[blockIdx.x];
  const int sa = pr.x, sb = pr.y;
  const int kf = blockIdx.y * WG + threadIdx.x;
  int sc = 0, fc = 0;
  bool active = kf < g.nkfun;
  if (active) {
    const int2 t = g.kfun[kf];
    sc = t.x;
    fc = t.y;
    // the metric is symmetric: shell blocks with sc > sa are the mirror of another thread's
    if (g.metric && sc > sa) active = false;
  }
  const double ax = g.A.center[3 * sa], ay = g.A.center[3 * sa + 1], az = g.A.center[3 * sa + 2];
  const double bx = g.B.center[3 * sb], by = g.B.center[3 * sb + 1], bz = g.B.center[3 * sb + 2];
  const int ia0 = g.A.poff[sa], nia = g.A.poff[sa + 1] - ia0;
  const int ib0 = g.B.poff[sb], nib = g.B.poff[sb + 1] - ib0;
  const int nprim = nia * nib;

  double cx = 0.0, cy = 0.0, cz = 0.0;
  int k0 = 0, k1 = 0;
  double tc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) tc[c] = 0.0;
  if (active) {
    cx = g.C.center[3 * sc];
    cy = g.C.center[3 * sc + 1];
    cz = g.C.center[3 * sc + 2];
    k0 = g.C.poff[sc];
    k1 = g.C.poff[sc + 1];
    const double* __restrict__ t = g.C.tmat + g.C.toff[sc] + fc * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) tc[c] = t[c];
  }

  double acc[NA][NB];
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = 0.0;

  for (int base = 0; base < nprim; base += NPB) {
    const int nb = min(NPB, nprim - base);
    __syncthreads();   // previous batch consumed
    if (threadIdx.x < 3 * nb) {
      const int q = threadIdx.x / 3, d = threadIdx.x % 3;
      const int ip = base + q;
      const double a = g.A.exps[ia0 + ip / nib], b = g.B.exps[ib0 + ip % nib];
      const double xab = d == 0 ? ax - bx : d == 1 ? ay - by : az - bz;
      hermite_pair<LA, LB>(a, b, xab, &sE[q][d][0]);
      if (d == 0) {
        const double p = a + b;
        sQ[q][0] = p;
        sQ[q][1] = (a * ax + b * bx) / p;
        sQ[q][2] = (a * ay + b * by) / p;
        sQ[q][3] = (a * az + b * bz) / p;
        sQ[q][4] = g.A.coefs[ia0 + ip / nib] * g.B.coefs[ib0 + ip % nib];
      }
    }
    __syncthreads();
    if (!active) continue;
    for (int q = 0; q < nb; ++q) {
      const double p = sQ[q][0], px = sQ[q][1], py = sQ[q][2], pz = sQ[q][3], wab = sQ[q][4];
      for (int k = k0; k < k1; ++k) {
        const double qe = g.C.exps[k];
        const double alpha = p * qe / (p + qe);
        const double xpc = px - cx, ypc = py - cy, zpc = pz - cz;
        double R[nherm(L)];
        {
          double f[L + 1];
          boys_eval<L>(alpha * (xpc * xpc + ypc * ypc + zpc * zpc), g.boys, f);
          // level L, then down to level 0 in place
          double pw[L + 1];
          pw[0] = 1.0;
#pragma unroll
          for (int n = 1; n <= L; ++n) pw[n] = pw[n - 1] * (-2.0 * alpha);
          R[0] = pw[L] * f[L];
#pragma unroll
          for (int n = L - 1; n >= 0; --n) {
#pragma unroll
            for (int s = L - n; s >= 1; --s)
#pragma unroll
              for (int r = s; r >= 0; --r)
#pragma unroll
                for (int v = r; v >= 0; --v) {
                  const int t = s - r, u = r - v;
                  double val;
                  if (t > 0) {
                    val = xpc * R[hidx(t - 1, u, v)];
                    if (t > 1) val += (t - 1) * R[hidx(t - 2, u, v)];
                  } else if (u > 0) {
                    val = ypc * R[hidx(t, u - 1, v)];
                    if (u > 1) val += (u - 1) * R[hidx(t, u - 2, v)];
                  } else {
                    val = zpc * R[hidx(t, u, v - 1)];
                    if (v > 1) val += (v - 1) * R[hidx(t, u, v - 2)];
                  }
                  R[hidx(t, u, v)] = val;
                }
            R[0] = pw[n] * f[n];
          }
        }
        // ket Hermite coefficients of the final function:
        // K(tau nu phi) = (-1)^(tau+nu+phi) sum_c tc[c] Ec[cx][tau] Ec[cy][nu] Ec[cz][phi]
        double K[nherm(LC)];
        {
          double ec[LC + 1][LC + 1];
#pragma unroll
          for (int i = 0; i <= LC; ++i)
#pragma unroll
            for (int t = 0; t <= LC; ++t) ec[i][t] = 0.0;
          ec[0][0] = 1.0;
          const double h = 0.5 / qe;
#pragma unroll
          for (int i = 1; i <= LC; ++i)
#pragma unroll
            for (int t = 0; t <= i; ++t) {
              double v = 0.0;
              if (t > 0) v += h * ec[i - 1][t - 1];
              if (t + 1 <= i - 1) v += (t + 1) * ec[i - 1][t + 1];
              ec[i][t] = v;
            }
#pragma unroll
          for (int n = 0; n < nherm(LC); ++n) K[n] = 0.0;
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const int ex = cart_x(LC, c), ey = cart_y(LC, c), ez = cart_z(LC, c);
#pragma unroll
            for (int t = 0; t <= ex; ++t)
#pragma unroll
              for (int u = 0; u <= ey; ++u)
#pragma unroll
                for (int v = 0; v <= ez; ++v) {
                  if (((ex - t) | (ey - u) | (ez - v)) & 1) continue;   // parity: zero coefficients
                  const double sg = ((t + u + v) & 1) ? -1.0 : 1.0;
                  K[hidx(t, u, v)] += sg * tc[c] * ec[ex][t] * ec[ey][u] * ec[ez][v];
                }
          }
        }
        // ket-contracted Hermite integrals G(tuv), t + u + v <= LAB
        double G[nherm(LAB)];
#pragma unroll
        for (int s = 0; s <= LAB; ++s)
#pragma unroll
          for (int r = 0; r <= s; ++r)
#pragma unroll
            for (int v = 0; v <= r; ++v) {
              const int t = s - r, u = r - v;
              double sum = 0.0;
#pragma unroll
              for (int s2 = 0; s2 <= LC; ++s2)
#pragma unroll
                for (int r2 = 0; r2 <= s2; ++r2)
#pragma unroll
                  for (int v2 = 0; v2 <= r2; ++v2) {
                    const int t2 = s2 - r2, u2 = r2 - v2;
                    sum += K[hidx(t2, u2, v2)] * R[hidx(t + t2, u + u2, v + v2)];
                  }
              G[hidx(t, u, v)] = sum;
            }
        const double pref = 2.0 * 17.493418327624862 / (p * qe * sqrt(p + qe)) * wab * g.C.coefs[k];   // 2 pi^(5/2)
        // bra: acc[a][b] += pref sum_tuv Ex Ey Ez G
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const int ix = cart_x(LA, a), iy = cart_y(LA, a), iz = cart_z(LA, a);
            const int jx = cart_x(LB, b), jy = cart_y(LB, b), jz = cart_z(LB, b);
            const double* __restrict__ ex = &sE[q][0][(ix * (LB + 1) + jx) * LD];
            const double* __restrict__ ey = &sE[q][1][(iy * (LB + 1) + jy) * LD];
            const double* __restrict__ ez = &sE[q][2][(iz * (LB + 1) + jz) * LD];
            double sum = 0.0;
#pragma unroll
            for (int t = 0; t <= ix + jx; ++t)
#pragma unroll
              for (int u = 0; u <= iy + jy; ++u) {
                const double etu = ex[t] * ey[u];
#pragma unroll
                for (int v = 0; v <= iz + jz; ++v) sum += etu * ez[v] * G[hidx(t, u, v)];
              }
            acc[a][b] += pref * sum;
          }
      }
    }
  }
  if (!active) return;

  // final functions of the bra: V[fa][fb] = sum_ab Ta[fa][a] Tb[fb][b] acc[a][b]
  const int nfa = g.A.foff[sa + 1] - g.A.foff[sa], nfb = g.B.foff[sb + 1] - g.B.foff[sb];
  const double* __restrict__ ta = g.A.tmat + g.A.toff[sa];
  const double* __restrict__ tb = g.B.tmat + g.B.toff[sb];
  double half[NA][NB];   // [fa][b]
#pragma unroll
  for (int fa = 0; fa < NA; ++fa)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      double s = 0.0;
      if (fa < nfa) {
#pragma unroll
        for (int a = 0; a < NA; ++a) s += ta[fa * NA + a] * acc[a][b];
      }
      half[fa][b] = s;
    }
  double V[NA][NB];
#pragma unroll
  for (int fa = 0; fa < NA; ++fa)
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
      double s = 0.0;
      if (fb < nfb) {
#pragma unroll
        for (int b = 0; b < NB; ++b) s += tb[fb * NB + b] * half[fa][b];
      }
      V[fa][fb] = s;
    }

  if (g.metric) {
    // (P|Q): the bra's second shell is the unit function; write both triangles
    const int64_t Q = g.C.foff[sc] + fc;
#pragma unroll
    for (int fa = 0; fa < NA; ++fa) {
      if (fa >= nfa) break;
      if (sc == sa && fa < fc) continue;   // one writer per element of a diagonal block
      const int64_t P = g.A.foff[sa] + fa;
      g.out[P + Q * g.naux] = V[fa][0];
      g.out[Q + P * g.naux] = V[fa][0];
    }
    return;
  }
  const int64_t P = g.C.foff[sc] + fc;
  const bool same = sa == sb && g.A.l == g.B.l;
#pragma unroll
  for (int fa = 0; fa < NA; ++fa)
#pragma unroll
    for (int fb = 0; fb < NB; ++fb) {
      if (fa >= nfa || fb >= nfb) continue;
      const int mu = g.A.foff[sa] + fa, nu = g.B.foff[sb] + fb;
      const int32_t ip = g.pairmap[(int64_t)mu * g.nbf + nu];
      if (ip < 0) continue;
      double v = V[fa][fb];
      if (same) {
        if (NA == NB) v = 0.5 * (V[fa][fb] + V[fb < NA ? fb : 0][fa < NB ? fa : 0]);
        if (mu == nu) v *= 0.5;   // diagonal pairs stored x1/2
      }
      g.out[ip + P * g.npair] = v;
    }
}

// launch one class of the translation unit that holds it
int launch_la0(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s);
int launch_la1(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s);
int launch_la2(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s);
int launch_la3(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s);
int launch_la4(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s);

}  // namespace eri
}  // namespace mgw
