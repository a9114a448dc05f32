// hyb_k_trc.hip -- the tracer column remaps of trcupd.F90 (hyb_trc.h): one
// column per lane on 128-byte-aligned rows, blockIdx.y = j.
#include <limits.h>

#include "hyb_launch.h"
#include "hyb_trc.h"

namespace hyb {

// plmtrc at 3 waves per SIMD: its chain of 4 fields needs about 200 VGPRs,
// and at 168 the few that spill cost less than a wave fewer per SIMD (13.3
// against 16.7 ms at GLBb); pcmtrc fits 4 waves
template <int METHOD, int CHUNK, int WPE>
__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(WPE))) k_trc(KTrc Q) {
  const int j = blockIdx.y + 1;
  const int i0 = 1 - (int)(((int64_t)Q.nb + Q.NI * (int64_t)(j - 1 + Q.nb)) & 15);
  const int i = i0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > Q.ii) return;
  const int64_t c0 = trc_ij(Q, i, j);
  if (Q.mask[c0] == 0) return;
  if (!trc_column_m<METHOD, CHUNK>(Q, c0))
    atomicMin(Q.bad, (int32_t)(c0 < INT_MAX ? c0 : INT_MAX - 1));
}

// the first-bad-column word starts at INT_MAX (a kernel, not a memset, so
// that a caller may capture the call in a graph)
__global__ void k_trc_reset(int32_t* bad) { *bad = INT_MAX; }

hipError_t launch_trc(const KTrc& Q, hipStream_t s) {
  dim3 g((Q.ii + kAlignPad + kWave - 1) / kWave, Q.jj);
  hipLaunchKernelGGL(k_trc_reset, dim3(1), dim3(1), 0, s, Q.bad);
  if (Q.method == 0)
    hipLaunchKernelGGL((k_trc<0, kTrcChunk, 4>), g, dim3(kWave), 0, s, Q);
  else
    hipLaunchKernelGGL((k_trc<1, kTrcChunk, 3>), g, dim3(kWave), 0, s, Q);
  return hipGetLastError();
}

}  // namespace hyb
