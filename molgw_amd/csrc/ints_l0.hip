// Coulomb-integral kernels with l_a = 0 (instantiations of ints_kernel.h).
#include "ints_kernel.h"

namespace mgw {
namespace eri {

int launch_la0(int lb, int lc, dim3 grid, const EriArgs& g, hipStream_t s) {
  switch (lb * 8 + lc) {
    case 0: hipLaunchKernelGGL((eri_kernel<0, 0, 0>), grid, dim3(WG), 0, s, g); break;
    case 1: hipLaunchKernelGGL((eri_kernel<0, 0, 1>), grid, dim3(WG), 0, s, g); break;
    case 2: hipLaunchKernelGGL((eri_kernel<0, 0, 2>), grid, dim3(WG), 0, s, g); break;
    case 3: hipLaunchKernelGGL((eri_kernel<0, 0, 3>), grid, dim3(WG), 0, s, g); break;
    case 4: hipLaunchKernelGGL((eri_kernel<0, 0, 4>), grid, dim3(WG), 0, s, g); break;
    default:
      set_error("integrals: class not instantiated");
      return MGW_ERR_ARG;
  }
  MGW_HIP(hipGetLastError());
  return MGW_OK;
}

}  // namespace eri
}  // namespace mgw
