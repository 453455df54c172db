// Coulomb-integral kernels with l_a = 3, l_b = 2 (instantiations of ints_kernel.h;
// split from the other l_b so that the builds run in parallel).
#include "ints_kernel.h"

namespace mgw {
namespace eri {

int launch_la3_lb2(int lc, dim3 grid, const EriArgs& g, hipStream_t s) {
  switch (lc) {
    case 0: hipLaunchKernelGGL((eri_kernel<3, 2, 0>), grid, dim3(WG), 0, s, g); break;
    case 1: hipLaunchKernelGGL((eri_kernel<3, 2, 1>), grid, dim3(WG), 0, s, g); break;
    case 2: hipLaunchKernelGGL((eri_kernel<3, 2, 2>), grid, dim3(WG), 0, s, g); break;
    case 3: hipLaunchKernelGGL((eri_kernel<3, 2, 3>), grid, dim3(WG), 0, s, g); break;
    case 4: hipLaunchKernelGGL((eri_kernel<3, 2, 4>), grid, dim3(WG), 0, s, g); break;
    default:
      set_error("integrals: class not instantiated");
      return MGW_ERR_ARG;
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace eri
}  // namespace mgw
