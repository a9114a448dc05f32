// hyb_k_za.hip -- widening of .a records staged on the device (hyb_za.cpp):
// 32-bit words in file order (idm x jdm) to FP64 in the halo layout.
#include "hyb_launch.h"

namespace hyb {

__global__ void __launch_bounds__(kBlock) k_za_widen(const float* src, double* out, int idm,
                                                      int jdm, int nbdy) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  if (i > idm) return;
  const int64_t NI = (int64_t)idm + 2 * nbdy;
  out[(i - 1 + nbdy) + NI * (j - 1 + nbdy)] = (double)src[(i - 1) + (int64_t)(j - 1) * idm];
}

hipError_t launch_za_widen(const float* src, double* out, int idm, int jdm, int nbdy,
                           hipStream_t s) {
  dim3 g((idm + kBlock - 1) / kBlock, jdm);
  hipLaunchKernelGGL(k_za_widen, g, dim3(kBlock), 0, s, src, out, idm, jdm, nbdy);
  return hipGetLastError();
}

}  // namespace hyb
