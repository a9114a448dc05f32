// hyb_k_stokes.hip -- Stokes drift onto the model layers (hyb_stokes.h):
// the PLM column walk, blending the two snapshots as it reads them.  Lanes
// along i on 128-byte-aligned rows, blockIdx.y = j.
#include "hyb_launch.h"
#include "hyb_stokes.h"

namespace hyb {

HYB_HD int stokes_i0(const KStokes& Q, int j) {
  return 1 - (int)(((int64_t)Q.nb + Q.NI * (int64_t)(j - 1 + Q.nb)) & 15);
}

__global__ void __launch_bounds__(kWave) k_stokes_walk(KStokes Q) {
  const int j = blockIdx.y + 1;
  const int i = stokes_i0(Q, j) + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > Q.ii) return;
  const int64_t c0 = stokes_ij(Q, i, j);
  if (Q.ip[c0] == 0) return;
  stokes_column<2>(Q, c0);
}

hipError_t launch_stokes(const KStokes& Q, hipStream_t s) {
  dim3 gw((Q.ii + kAlignPad + kWave - 1) / kWave, Q.jj);
  hipLaunchKernelGGL(k_stokes_walk, gw, dim3(kWave), 0, s, Q);
  return hipGetLastError();
}

}  // namespace hyb
