// Three- and two-centre Coulomb integrals over contracted Gaussian shells by
// the McMurchie-Davidson scheme (Hermite expansion + Boys function), for the
// RI tensor of the front end: (mu nu|P) and (P|Q). Host code, OpenMP over
// shell pairs; the same algorithm as molgw_amd/frontend/integrals.py, which
// pins it.
//
// Shell description (arrays over nsh shells): l, centre (3 per shell), the
// primitive range [poff[s], poff[s+1]) into exps/coefs (coefs: contraction
// weights of the unnormalised Cartesian primitives), and the (nfunc x ncart)
// map to the final functions at toff[s] (row-major), functions of shell s
// starting at foff[s].
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int LMAX = 6;                // per shell
constexpr int LTOT = 3 * LMAX;         // l_a + l_b + l_c

struct Shells {
  int n;
  const int* l;
  const double* center;
  const int* poff;
  const double* exps;
  const double* coefs;
  const int* toff;
  const double* tmat;
  const int* foff;
};

int ncart(int l) { return (l + 1) * (l + 2) / 2; }

void cart_list(int l, int (*c)[3]) {
  int n = 0;
  for (int i = 0; i <= l; ++i)
    for (int k = 0; k <= i; ++k) {
      c[n][0] = l - i;
      c[n][1] = i - k;
      c[n][2] = k;
      ++n;
    }
}

int nfunc_of(const Shells& s, int k) { return s.foff[k + 1] - s.foff[k]; }

// F_n(t), n = 0..nmax: positive series times exp(-t) for t <= 30, then the
// downward recursion; above, F_0 = sqrt(pi/t)/2 and the upward recursion.
void boys(int nmax, double t, double* f) {
  if (t < 1e-14) {
    for (int n = 0; n <= nmax; ++n) f[n] = 1.0 / (2 * n + 1);
    return;
  }
  const double e = std::exp(-t);
  if (t <= 30.0) {
    double term = 1.0 / (2 * nmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
      term *= 2.0 * t / (2 * nmax + 2 * k + 1);
      sum += term;
      if (term < 1e-17 * sum) break;
    }
    f[nmax] = e * sum;
    for (int n = nmax - 1; n >= 0; --n) f[n] = (2.0 * t * f[n + 1] + e) / (2 * n + 1);
  } else {
    f[0] = 0.5 * std::sqrt(M_PI / t);
    for (int n = 0; n < nmax; ++n) f[n + 1] = ((2 * n + 1) * f[n] - e) / (2.0 * t);
  }
}

// E[i][j][t] (i <= la, j <= lb, t <= i + j) of one dimension.
void hermite_e(int la, int lb, double a, double b, double xab, double* e, int ld) {
  // e index: (i * (lb + 1) + j) * ld + t
  const double p = a + b, mu = a * b / p;
  const double xpa = -b * xab / p, xpb = a * xab / p, h = 0.5 / p;
  std::memset(e, 0, sizeof(double) * (la + 1) * (lb + 1) * ld);
  e[0] = std::exp(-mu * xab * xab);
  for (int i = 0; i <= la; ++i)
    for (int j = 0; j <= lb; ++j) {
      if (i == 0 && j == 0) continue;
      const double* src;
      double x;
      if (i > 0) {
        src = e + ((i - 1) * (lb + 1) + j) * ld;
        x = xpa;
      } else {
        src = e + (i * (lb + 1) + j - 1) * ld;
        x = xpb;
      }
      double* dst = e + (i * (lb + 1) + j) * ld;
      const int top = i + j - 1;   // valid t of src
      for (int t = 0; t <= i + j; ++t) {
        double v = (t <= top ? x * src[t] : 0.0);
        if (t > 0 && t - 1 <= top) v += h * src[t - 1];
        if (t + 1 <= top) v += (t + 1) * src[t + 1];
        dst[t] = v;
      }
    }
}

// R[t][u][v] (t + u + v <= L) into r (dims (L+1)^3), work of (L+1)^4.
void hermite_r(int L, double alpha, double x, double y, double z, double* r, double* work) {
  const int n1 = L + 1;
  double f[LTOT + 2];
  boys(L, alpha * (x * x + y * y + z * z), f);
  auto at = [&](double* a, int t, int u, int v) -> double& { return a[(t * n1 + u) * n1 + v]; };
  double* cur = work;
  double* nxt = work + n1 * n1 * n1;
  std::memset(cur, 0, sizeof(double) * n1 * n1 * n1);
  at(cur, 0, 0, 0) = std::pow(-2.0 * alpha, L) * f[L];
  for (int n = L - 1; n >= 0; --n) {
    const int m = L - n;
    at(nxt, 0, 0, 0) = std::pow(-2.0 * alpha, n) * f[n];
    for (int t = 0; t <= m; ++t)
      for (int u = 0; u <= m - t; ++u)
        for (int v = 0; v <= m - t - u; ++v) {
          if (t == 0 && u == 0 && v == 0) continue;
          double val;
          if (t > 0) {
            val = x * at(cur, t - 1, u, v);
            if (t > 1) val += (t - 1) * at(cur, t - 2, u, v);
          } else if (u > 0) {
            val = y * at(cur, t, u - 1, v);
            if (u > 1) val += (u - 1) * at(cur, t, u - 2, v);
          } else {
            val = z * at(cur, t, u, v - 1);
            if (v > 1) val += (v - 1) * at(cur, t, u, v - 2);
          }
          at(nxt, t, u, v) = val;
        }
    double* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  std::memcpy(r, cur, sizeof(double) * n1 * n1 * n1);
}

// (ab|c) over Cartesian components, out[(ca * ncb + cb) * ncc + cc].
void shell_eri3(const Shells& S, int sa, const Shells& B, int sb, const Shells& A, int sc, double* out,
                std::vector<double>& buf) {
  const int la = S.l[sa], lb = B.l[sb], lc = A.l[sc];
  const int L = la + lb + lc, Lab = la + lb;
  const int na = ncart(la), nb = ncart(lb), nc = ncart(lc);
  int ca[28][3], cb[28][3], cc[28][3];
  cart_list(la, ca);
  cart_list(lb, cb);
  cart_list(lc, cc);
  const double* A0 = S.center + 3 * sa;
  const double* B0 = B.center + 3 * sb;
  const double* C0 = A.center + 3 * sc;
  const int lde = Lab + 1, n1 = L + 1, m1 = lc + 1, h1 = Lab + 1;
  const size_t ne = (size_t)(la + 1) * (lb + 1) * lde;
  // E of the pair (3 ne), E of c (3 (lc+1) m1), R and the two recursion levels (3 n1^3), ket sums
  buf.resize(3 * ne + 3 * (size_t)(lc + 1) * m1 + 3 * (size_t)n1 * n1 * n1 + (size_t)nc * h1 * h1 * h1);
  double* ex = buf.data();
  double* ey = ex + ne;
  double* ez = ey + ne;
  double* gx = ez + ne;            // single-Gaussian E of c, [k][tau]
  double* gy = gx + (lc + 1) * m1;
  double* gz = gy + (lc + 1) * m1;
  double* r = gz + (lc + 1) * m1;
  double* work = r + (size_t)n1 * n1 * n1;
  double* g = work + (size_t)n1 * n1 * n1 * 2;   // ket-contracted [cc][t][u][v], t,u,v <= Lab
  std::memset(out, 0, sizeof(double) * na * nb * nc);
  for (int i = S.poff[sa]; i < S.poff[sa + 1]; ++i)
    for (int j = B.poff[sb]; j < B.poff[sb + 1]; ++j) {
      const double a = S.exps[i], b = B.exps[j], p = a + b;
      hermite_e(la, lb, a, b, A0[0] - B0[0], ex, lde);
      hermite_e(la, lb, a, b, A0[1] - B0[1], ey, lde);
      hermite_e(la, lb, a, b, A0[2] - B0[2], ez, lde);
      const double P[3] = {(a * A0[0] + b * B0[0]) / p, (a * A0[1] + b * B0[1]) / p, (a * A0[2] + b * B0[2]) / p};
      const double wab = S.coefs[i] * B.coefs[j];
      for (int k = A.poff[sc]; k < A.poff[sc + 1]; ++k) {
        const double q = A.exps[k];
        const double alpha = p * q / (p + q);
        hermite_r(L, alpha, P[0] - C0[0], P[1] - C0[1], P[2] - C0[2], r, work);
        hermite_e(lc, 0, q, 0.0, 0.0, gx, m1);
        std::memcpy(gy, gx, sizeof(double) * (lc + 1) * m1);
        std::memcpy(gz, gx, sizeof(double) * (lc + 1) * m1);
        const double pref = 2.0 * std::pow(M_PI, 2.5) / (p * q * std::sqrt(p + q)) * wab * A.coefs[k];
        // ket contraction: g[c][t][u][v] = sum sign * Gx Gy Gz * R[t+tau][u+nu][v+phi]
        for (int c = 0; c < nc; ++c) {
          double* gc = g + (size_t)c * h1 * h1 * h1;
          for (int t = 0; t <= Lab; ++t)
            for (int u = 0; u <= Lab - t; ++u)
              for (int v = 0; v <= Lab - t - u; ++v) {
                double s = 0.0;
                for (int tau = 0; tau <= cc[c][0]; ++tau) {
                  const double fx = gx[cc[c][0] * m1 + tau];
                  for (int nu = 0; nu <= cc[c][1]; ++nu) {
                    const double fxy = fx * gy[cc[c][1] * m1 + nu];
                    for (int phi = 0; phi <= cc[c][2]; ++phi) {
                      const double sg = ((tau + nu + phi) & 1) ? -1.0 : 1.0;
                      s += sg * fxy * gz[cc[c][2] * m1 + phi] * r[((t + tau) * n1 + u + nu) * n1 + v + phi];
                    }
                  }
                }
                gc[(t * h1 + u) * h1 + v] = s * pref;
              }
        }
        // bra: out[a][b][c] += sum_tuv Ex Ey Ez g[c][t][u][v]
        for (int x = 0; x < na; ++x)
          for (int y = 0; y < nb; ++y) {
            const double* fx = ex + (ca[x][0] * (lb + 1) + cb[y][0]) * lde;
            const double* fy = ey + (ca[x][1] * (lb + 1) + cb[y][1]) * lde;
            const double* fz = ez + (ca[x][2] * (lb + 1) + cb[y][2]) * lde;
            const int tx = ca[x][0] + cb[y][0], ty = ca[x][1] + cb[y][1], tz = ca[x][2] + cb[y][2];
            double* o = out + ((size_t)x * nb + y) * nc;
            for (int c = 0; c < nc; ++c) {
              const double* gc = g + (size_t)c * h1 * h1 * h1;
              double s = 0.0;
              for (int t = 0; t <= tx; ++t)
                for (int u = 0; u <= ty; ++u) {
                  const double ftu = fx[t] * fy[u];
                  for (int v = 0; v <= tz; ++v) s += ftu * fz[v] * gc[(t * h1 + u) * h1 + v];
                }
              o[c] += s;
            }
          }
      }
    }
}

// final functions: res[fa][fb][fc] = Ta Tb Tc cart
void to_final(const double* ta, int fa, int na, const double* tb, int fb, int nb, const double* tc, int fc, int nc,
              const double* cart, double* res, std::vector<double>& t1, std::vector<double>& t2) {
  t1.assign((size_t)fa * nb * nc, 0.0);
  for (int i = 0; i < fa; ++i)
    for (int x = 0; x < na; ++x) {
      const double w = ta[i * na + x];
      if (w == 0.0) continue;
      for (int yz = 0; yz < nb * nc; ++yz) t1[(size_t)i * nb * nc + yz] += w * cart[(size_t)x * nb * nc + yz];
    }
  t2.assign((size_t)fa * fb * nc, 0.0);
  for (int i = 0; i < fa; ++i)
    for (int j = 0; j < fb; ++j)
      for (int y = 0; y < nb; ++y) {
        const double w = tb[j * nb + y];
        if (w == 0.0) continue;
        for (int z = 0; z < nc; ++z) t2[((size_t)i * fb + j) * nc + z] += w * t1[((size_t)i * nb + y) * nc + z];
      }
  for (int ij = 0; ij < fa * fb; ++ij)
    for (int k = 0; k < fc; ++k) {
      double s = 0.0;
      for (int z = 0; z < nc; ++z) s += tc[k * nc + z] * t2[(size_t)ij * nc + z];
      res[(size_t)ij * fc + k] = s;
    }
}

}  // namespace

extern "C" {

// (mu nu|P) for all final functions: out[(mu * nbf + nu) * naux + P] (dense,
// both triangles). Returns 0, or -1 for an angular momentum above LMAX.
int mgwi_eri3(int nsh, const int* l, const double* center, const int* poff, const double* exps, const double* coefs,
              const int* toff, const double* tmat, const int* foff, int nash, const int* al, const double* acenter,
              const int* apoff, const double* aexps, const double* acoefs, const int* atoff, const double* atmat,
              const int* afoff, double* out) {
  const Shells S{nsh, l, center, poff, exps, coefs, toff, tmat, foff};
  const Shells A{nash, al, acenter, apoff, aexps, acoefs, atoff, atmat, afoff};
  for (int k = 0; k < nsh; ++k)
    if (l[k] > LMAX) return -1;
  for (int k = 0; k < nash; ++k)
    if (al[k] > LMAX) return -1;
  const int64_t nbf = foff[nsh], naux = afoff[nash];
  const int64_t npairs = (int64_t)nsh * (nsh + 1) / 2;
#pragma omp parallel
  {
    std::vector<double> buf, cart, fin, t1, t2;
#pragma omp for schedule(dynamic, 1)
    for (int64_t pr = 0; pr < npairs; ++pr) {
      // pair index -> (sa >= sb)
      int sa = (int)((std::sqrt(8.0 * pr + 1.0) - 1.0) / 2.0);
      while ((int64_t)sa * (sa + 1) / 2 > pr) --sa;
      while ((int64_t)(sa + 1) * (sa + 2) / 2 <= pr) ++sa;
      const int sb = (int)(pr - (int64_t)sa * (sa + 1) / 2);
      const int na = ncart(l[sa]), nb = ncart(l[sb]);
      const int fa = nfunc_of(S, sa), fb = nfunc_of(S, sb);
      for (int sc = 0; sc < nash; ++sc) {
        const int nc = ncart(al[sc]), fc = nfunc_of(A, sc);
        cart.resize((size_t)na * nb * nc);
        fin.resize((size_t)fa * fb * fc);
        shell_eri3(S, sa, S, sb, A, sc, cart.data(), buf);
        to_final(tmat + toff[sa], fa, na, tmat + toff[sb], fb, nb, atmat + atoff[sc], fc, nc, cart.data(), fin.data(),
                 t1, t2);
        for (int i = 0; i < fa; ++i)
          for (int j = 0; j < fb; ++j)
            for (int k = 0; k < fc; ++k) {
              double v = fin[((size_t)i * fb + j) * fc + k];
              if (sa == sb) v = 0.5 * (v + fin[((size_t)j * fb + i) * fc + k]);
              const int64_t mu = foff[sa] + i, nu = foff[sb] + j, P = afoff[sc] + k;
              out[(mu * nbf + nu) * naux + P] = v;
              out[(nu * nbf + mu) * naux + P] = v;
            }
      }
    }
  }
  return 0;
}

// (P|Q) over the final auxiliary functions: out[P * naux + Q], as (P 1|Q)
// with a unit s function (exponent 0) at P's centre.
int mgwi_eri2(int nash, const int* al, const double* acenter, const int* apoff, const double* aexps,
              const double* acoefs, const int* atoff, const double* atmat, const int* afoff, double* out) {
  const Shells A{nash, al, acenter, apoff, aexps, acoefs, atoff, atmat, afoff};
  for (int k = 0; k < nash; ++k)
    if (al[k] > LMAX) return -1;
  const int64_t naux = afoff[nash];
  std::vector<int> ul(nash, 0), upoff(nash + 1), utoff(nash), ufoff(nash + 1);
  std::vector<double> uexp(nash, 0.0), ucoef(nash, 1.0), utmat(nash, 1.0);
  for (int k = 0; k <= nash; ++k) {
    upoff[k] = k;
    ufoff[k] = k;
  }
  for (int k = 0; k < nash; ++k) utoff[k] = k;
  const Shells U{nash, ul.data(), acenter, upoff.data(), uexp.data(), ucoef.data(), utoff.data(), utmat.data(),
                 ufoff.data()};
#pragma omp parallel
  {
    std::vector<double> buf, cart, fin, t1, t2;
#pragma omp for schedule(dynamic, 1)
    for (int sc = 0; sc < nash; ++sc)
      for (int sd = 0; sd <= sc; ++sd) {
        const int nc = ncart(al[sc]), nd = ncart(al[sd]);
        const int fc = nfunc_of(A, sc), fd = nfunc_of(A, sd);
        cart.resize((size_t)nc * nd);
        fin.resize((size_t)fc * fd);
        shell_eri3(A, sc, U, sc, A, sd, cart.data(), buf);
        const double one = 1.0;
        to_final(atmat + atoff[sc], fc, nc, &one, 1, 1, atmat + atoff[sd], fd, nd, cart.data(), fin.data(), t1, t2);
        for (int i = 0; i < fc; ++i)
          for (int j = 0; j < fd; ++j) {
            double v = fin[(size_t)i * fd + j];
            if (sc == sd) v = 0.5 * (v + fin[(size_t)j * fd + i]);
            out[(afoff[sc] + i) * naux + afoff[sd] + j] = v;
            out[(afoff[sd] + j) * naux + afoff[sc] + i] = v;
          }
      }
  }
  return 0;
}

// Overlap, kinetic energy and nuclear attraction (sum_C -Z_C / |r - C|) over
// the final functions of one basis: s, t, v each nbf x nbf (row-major).
int mgwi_one_electron(int nsh, const int* l, const double* center, const int* poff, const double* exps,
                      const double* coefs, const int* toff, const double* tmat, const int* foff, int ncharge,
                      const double* charge, const double* ccoord, double* s_out, double* t_out, double* v_out) {
  const Shells S{nsh, l, center, poff, exps, coefs, toff, tmat, foff};
  for (int k = 0; k < nsh; ++k)
    if (l[k] > LMAX) return -1;
  const int64_t nbf = foff[nsh];
  const int64_t npairs = (int64_t)nsh * (nsh + 1) / 2;
#pragma omp parallel
  {
    std::vector<double> ex, ey, ez, r, work, cs, ct, cv, fin, t1, t2;
#pragma omp for schedule(dynamic, 4)
    for (int64_t pr = 0; pr < npairs; ++pr) {
      int sa = (int)((std::sqrt(8.0 * pr + 1.0) - 1.0) / 2.0);
      while ((int64_t)sa * (sa + 1) / 2 > pr) --sa;
      while ((int64_t)(sa + 1) * (sa + 2) / 2 <= pr) ++sa;
      const int sb = (int)(pr - (int64_t)sa * (sa + 1) / 2);
      const int la = l[sa], lb = l[sb], na = ncart(la), nb = ncart(lb), Lab = la + lb;
      int ca[28][3], cb[28][3];
      cart_list(la, ca);
      cart_list(lb, cb);
      const double* A0 = center + 3 * sa;
      const double* B0 = center + 3 * sb;
      const int lde = Lab + 3, n1 = Lab + 1;
      const size_t ne = (size_t)(la + 1) * (lb + 3) * lde;
      ex.resize(ne); ey.resize(ne); ez.resize(ne);
      r.resize((size_t)n1 * n1 * n1); work.resize(2 * (size_t)n1 * n1 * n1);
      cs.assign((size_t)na * nb, 0.0); ct.assign((size_t)na * nb, 0.0); cv.assign((size_t)na * nb, 0.0);
      for (int i = poff[sa]; i < poff[sa + 1]; ++i)
        for (int j = poff[sb]; j < poff[sb + 1]; ++j) {
          const double a = exps[i], b = exps[j], p = a + b, w = coefs[i] * coefs[j];
          // E with lb + 2 for the kinetic-energy shifts
          hermite_e(la, lb + 2, a, b, A0[0] - B0[0], ex.data(), lde);
          hermite_e(la, lb + 2, a, b, A0[1] - B0[1], ey.data(), lde);
          hermite_e(la, lb + 2, a, b, A0[2] - B0[2], ez.data(), lde);
          const double P[3] = {(a * A0[0] + b * B0[0]) / p, (a * A0[1] + b * B0[1]) / p,
                               (a * A0[2] + b * B0[2]) / p};
          const double sq = std::sqrt(M_PI / p);
          const double* E[3] = {ex.data(), ey.data(), ez.data()};
          auto s1 = [&](int d, int ia, int jb) { return E[d][(ia * (lb + 3) + jb) * lde] * sq; };
          auto k1 = [&](int d, int ia, int jb) {
            double v = -2.0 * b * b * s1(d, ia, jb + 2) + b * (2 * jb + 1) * s1(d, ia, jb);
            if (jb >= 2) v -= 0.5 * jb * (jb - 1) * s1(d, ia, jb - 2);
            return v;
          };
          for (int x = 0; x < na; ++x)
            for (int y = 0; y < nb; ++y) {
              const double sx = s1(0, ca[x][0], cb[y][0]), sy = s1(1, ca[x][1], cb[y][1]),
                           sz = s1(2, ca[x][2], cb[y][2]);
              cs[x * nb + y] += w * sx * sy * sz;
              ct[x * nb + y] += w * (k1(0, ca[x][0], cb[y][0]) * sy * sz + sx * k1(1, ca[x][1], cb[y][1]) * sz +
                                     sx * sy * k1(2, ca[x][2], cb[y][2]));
            }
          for (int c = 0; c < ncharge; ++c) {
            const double* C0 = ccoord + 3 * c;
            hermite_r(Lab, p, P[0] - C0[0], P[1] - C0[1], P[2] - C0[2], r.data(), work.data());
            const double pref = -charge[c] * 2.0 * M_PI / p * w;
            for (int x = 0; x < na; ++x)
              for (int y = 0; y < nb; ++y) {
                const double* fx = ex.data() + (ca[x][0] * (lb + 3) + cb[y][0]) * lde;
                const double* fy = ey.data() + (ca[x][1] * (lb + 3) + cb[y][1]) * lde;
                const double* fz = ez.data() + (ca[x][2] * (lb + 3) + cb[y][2]) * lde;
                double acc = 0.0;
                for (int t = 0; t <= ca[x][0] + cb[y][0]; ++t)
                  for (int u = 0; u <= ca[x][1] + cb[y][1]; ++u)
                    for (int v = 0; v <= ca[x][2] + cb[y][2]; ++v)
                      acc += fx[t] * fy[u] * fz[v] * r[(t * n1 + u) * n1 + v];
                cv[x * nb + y] += pref * acc;
              }
          }
        }
      const int fa = nfunc_of(S, sa), fb = nfunc_of(S, sb);
      const double one = 1.0;
      for (int which = 0; which < 3; ++which) {
        const std::vector<double>& cart = which == 0 ? cs : which == 1 ? ct : cv;
        double* out = which == 0 ? s_out : which == 1 ? t_out : v_out;
        fin.resize((size_t)fa * fb);
        to_final(tmat + toff[sa], fa, na, tmat + toff[sb], fb, nb, &one, 1, 1, cart.data(), fin.data(), t1, t2);
        for (int i = 0; i < fa; ++i)
          for (int j = 0; j < fb; ++j) {
            double v = fin[(size_t)i * fb + j];
            if (sa == sb) v = 0.5 * (v + fin[(size_t)j * fb + i]);
            out[(foff[sa] + i) * nbf + foff[sb] + j] = v;
            out[(foff[sb] + j) * nbf + foff[sa] + i] = v;
          }
      }
    }
  }
  return 0;
}

}  // extern "C"
