// Host side of the two- and three-centre Coulomb integrals on the GPU: the
// Boys-function table and the dispatch of one (l_a, l_b | l_c) class to the
// translation unit that instantiates its kernel (ints_l0.hip .. ints_l4.hip,
// split by l_a so that the builds run in parallel). The device code and the
// method are described in ints_kernel.h.
#include "ints_kernel.h"

#include <algorithm>
#include <vector>

namespace mgw {
namespace {

// supported classes: l_a <= 3 with l_b <= l_a, and l_a = 4 with l_b = 0 (the
// metric's bra); l_c <= 4
bool supported(int la, int lb, int lc) {
  if (lc < 0 || lc > 4 || lb < 0) return false;
  if (la <= 3) return lb <= la;
  return la == 4 && lb == 0;
}

// F_n(k h) by the series e^-T sum_j (2T)^j / ((2n+1)(2n+3)..(2n+2j+1)) in long double
std::vector<double> boys_table_host() {
  using namespace eri;
  std::vector<double> tab((size_t)BOYS_NT * BOYS_N);
  for (int k = 0; k < BOYS_NT; ++k) {
    const long double T = (long double)k * BOYS_H;
    for (int n = 0; n < BOYS_N; ++n) {
      long double term = 1.0L / (2 * n + 1), sum = term;
      for (int j = 1; j < 2000; ++j) {
        term *= 2.0L * T / (2 * n + 2 * j + 1);
        sum += term;
        if (term < 1e-22L * sum) break;
      }
      tab[(size_t)k * BOYS_N + n] = (double)(std::exp(-T) * sum);
    }
  }
  return tab;
}

}  // namespace

bool eri_class_supported(int la, int lb, int lc) { return supported(la, lb, lc); }

size_t boys_table_doubles() { return (size_t)eri::BOYS_NT * eri::BOYS_N; }

void boys_table_fill(double* host) {
  static const std::vector<double> tab = boys_table_host();
  std::copy(tab.begin(), tab.end(), host);
}

int launch_eri_class(const EriLaunch& e, hipStream_t s) {
  using namespace eri;
  MGW_REQUIRE(supported(e.la, e.lb, e.lc), MGW_ERR_ARG,
              "integrals: angular momentum class (" + std::to_string(e.la) + "," + std::to_string(e.lb) + "|" +
                  std::to_string(e.lc) + ") is not supported (l <= 3 per orbital shell, l <= 4 per auxiliary shell)");
  if (e.npairs == 0 || e.nkfun == 0) return MGW_OK;
  EriArgs g;
  auto conv = [](const EriShells& x) {
    return ShellsDev{x.l, x.center, x.poff, x.exps, x.coefs, x.toff, x.tmat, x.foff};
  };
  g.A = conv(e.A);
  g.B = conv(e.B);
  g.C = conv(e.C);
  g.pairs = reinterpret_cast<const int2*>(e.pairs);
  g.kfun = reinterpret_cast<const int2*>(e.kfun);
  g.nkfun = e.nkfun;
  g.boys = e.boys;
  g.metric = e.metric;
  g.pairmap = e.pairmap;
  g.nbf = e.nbf;
  g.npair = e.npair;
  g.naux = e.naux;
  g.out = e.out;
  MGW_REQUIRE(e.npairs < (1LL << 31) && (e.nkfun + WG - 1) / WG < 65536, MGW_ERR_ARG,
              "integrals: too many shell pairs or auxiliary functions in one class");
  const dim3 grid((unsigned)e.npairs, (unsigned)((e.nkfun + WG - 1) / WG));
  switch (e.la) {
    case 0: return launch_la0(e.lb, e.lc, grid, g, s);
    case 1: return launch_la1(e.lb, e.lc, grid, g, s);
    case 2: return launch_la2(e.lb, e.lc, grid, g, s);
    case 3: return launch_la3(e.lb, e.lc, grid, g, s);
    default: return launch_la4(e.lb, e.lc, grid, g, s);
  }
}

}  // namespace mgw
