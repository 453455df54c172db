/*
 * Oracle (test infrastructure only): C restatement of the pole sum of
 * gw_selfenergy, src/m_gw_selfenergy_analytic.f90:118-139 (MOLGW 3.4), for
 * the timed CPU baseline. Same arithmetic as oracle/gw_oracle.py
 * gw_selfenergy (whose numpy form is the checked restatement; a test keeps
 * the two equal).
 *
 * For one state i (energy ei, occupation occ_i):
 *   sigma[k + nomega * p] += sum_s w2[s + npole * p] *
 *       ( f_full  / (e0[p] + omega[k] - ei + pole[s] - i eta)
 *       + f_empty / (e0[p] + omega[k] - ei - pole[s] + i eta) )
 *   egw += f_empty * occ_p * w2 / (eps[p] - ei - pole[s])   for filled p
 * OpenMP over states p; the threads are started after the BLAS products of
 * the caller have finished.
 */
#include <math.h>

/* 1/(a - i eta) = (a + i eta) / (a^2 + eta^2): real and imaginary parts are
 * accumulated separately (the same value as the complex division). */
double pole_sum_one_i(int npole, int nse, int nomega, const double *w2, const double *pole,
                      const double *omega, const double *e0, const double *eps, const double *occp,
                      double ei, double occ_i, double spin_fact, double eta,
                      double *sigma /* complex, nomega x nse interleaved */) {
  const double f_full = occ_i / spin_fact;
  const double f_empty = (spin_fact - occ_i) / spin_fact;
  const double eta2 = eta * eta;
  double egw = 0.0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : egw)
  for (int p = 0; p < nse; ++p) {
    const double *wp = w2 + (long)npole * p;
    const int p_full = ((spin_fact - occp[p]) / spin_fact) < 1.0e-5;
    for (int k = 0; k < nomega; ++k) {
      const double x = e0[p] + omega[k] - ei;
      double re = 0.0, im = 0.0;
      if (f_full != 0.0) {
        double sr = 0.0, si = 0.0;
        for (int s = 0; s < npole; ++s) {
          const double a = x + pole[s];
          const double r = wp[s] / (a * a + eta2);
          sr += a * r;
          si += r;
        }
        re += f_full * sr;
        im += f_full * eta * si;
      }
      if (f_empty != 0.0) {
        double sr = 0.0, si = 0.0;
        for (int s = 0; s < npole; ++s) {
          const double b = x - pole[s];
          const double r = wp[s] / (b * b + eta2);
          sr += b * r;
          si += r;
        }
        re += f_empty * sr;
        im -= f_empty * eta * si;
      }
      sigma[2 * ((long)p * nomega + k)] += re;
      sigma[2 * ((long)p * nomega + k) + 1] += im;
    }
    if (p_full && f_empty != 0.0) {
      double e = 0.0;
      for (int s = 0; s < npole; ++s) e += wp[s] / (eps[p] - ei - pole[s]);
      egw += f_empty * occp[p] * e;
    }
  }
  return egw;
}
