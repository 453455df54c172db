// Coulomb-integral kernels with l_a = 2 (instantiations of ints_kernel.h).
#include "ints_kernel.h"

namespace mgw {
namespace eri {

int launch_la2(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s) {
  switch (lb * 8 + lc) {
    case 0: hipLaunchKernelGGL((eri_kernel<2, 0, 0>), grid, dim3(WG), 0, s, g); break;
    case 1: hipLaunchKernelGGL((eri_kernel<2, 0, 1>), grid, dim3(WG), 0, s, g); break;
    case 2: hipLaunchKernelGGL((eri_kernel<2, 0, 2>), grid, dim3(WG), 0, s, g); break;
    case 3: hipLaunchKernelGGL((eri_kernel<2, 0, 3>), grid, dim3(WG), 0, s, g); break;
    case 4: hipLaunchKernelGGL((eri_kernel<2, 0, 4>), grid, dim3(WG), 0, s, g); break;
    case 8: hipLaunchKernelGGL((eri_kernel<2, 1, 0>), grid, dim3(WG), 0, s, g); break;
    case 9: hipLaunchKernelGGL((eri_kernel<2, 1, 1>), grid, dim3(WG), 0, s, g); break;
    case 10: hipLaunchKernelGGL((eri_kernel<2, 1, 2>), grid, dim3(WG), 0, s, g); break;
    case 11: hipLaunchKernelGGL((eri_kernel<2, 1, 3>), grid, dim3(WG), 0, s, g); break;
    case 12: hipLaunchKernelGGL((eri_kernel<2, 1, 4>), grid, dim3(WG), 0, s, g); break;
    case 16: hipLaunchKernelGGL((eri_kernel<2, 2, 0>), grid, dim3(WG), 0, s, g); break;
    case 17: hipLaunchKernelGGL((eri_kernel<2, 2, 1>), grid, dim3(WG), 0, s, g); break;
    case 18: hipLaunchKernelGGL((eri_kernel<2, 2, 2>), grid, dim3(WG), 0, s, g); break;
    case 19: hipLaunchKernelGGL((eri_kernel<2, 2, 3>), grid, dim3(WG), 0, s, g); break;
    case 20: hipLaunchKernelGGL((eri_kernel<2, 2, 4>), grid, dim3(WG), 0, s, g); break;
    default:
      set_error("integrals: class not instantiated");
      return MGW_ERR_ARG;
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace eri
}  // namespace mgw
