// Coulomb-integral kernels with l_a = 3, l_b = 0, 1 (instantiations of ints_kernel.h;
// split from the other l_b so that the builds run in parallel).
#include "ints_kernel.h"

namespace mgw {
namespace eri {

int launch_la3_lb2(int lc, dim3 grid, const EriArgs& g, hipStream_t s);
int launch_la3_lb3(int lc, dim3 grid, const EriArgs& g, hipStream_t s);

int launch_la3(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s) {
  if (lb == 2) return launch_la3_lb2(lc, grid, g, s);
  if (lb == 3) return launch_la3_lb3(lc, grid, g, s);
  switch (lb * 8 + lc) {
    case 0: hipLaunchKernelGGL((eri_kernel<3, 0, 0>), grid, dim3(WG), 0, s, g); break;
    case 1: hipLaunchKernelGGL((eri_kernel<3, 0, 1>), grid, dim3(WG), 0, s, g); break;
    case 2: hipLaunchKernelGGL((eri_kernel<3, 0, 2>), grid, dim3(WG), 0, s, g); break;
    case 3: hipLaunchKernelGGL((eri_kernel<3, 0, 3>), grid, dim3(WG), 0, s, g); break;
    case 4: hipLaunchKernelGGL((eri_kernel<3, 0, 4>), grid, dim3(WG), 0, s, g); break;
    case 8: hipLaunchKernelGGL((eri_kernel<3, 1, 0>), grid, dim3(WG), 0, s, g); break;
    case 9: hipLaunchKernelGGL((eri_kernel<3, 1, 1>), grid, dim3(WG), 0, s, g); break;
    case 10: hipLaunchKernelGGL((eri_kernel<3, 1, 2>), grid, dim3(WG), 0, s, g); break;
    case 11: hipLaunchKernelGGL((eri_kernel<3, 1, 3>), grid, dim3(WG), 0, s, g); break;
    case 12: hipLaunchKernelGGL((eri_kernel<3, 1, 4>), grid, dim3(WG), 0, s, g); break;
    default:
      set_error("integrals: class not instantiated");
      return MGW_ERR_ARG;
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace eri
}  // namespace mgw
