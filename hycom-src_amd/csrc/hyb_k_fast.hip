// hyb_k_fast.hip -- register-resident kernels for kdm in {22,32,41}
// (hyb_fast.h): every k-loop unrolled for a fixed K.
#include "hyb_fast.h"
#include "hyb_launch.h"

namespace hyb {

template <int K>
__global__ void __launch_bounds__(kBlock) k_regrid(KParams P, View V, Scratch W, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  regrid_fast<K>(P, V, W, i, j, n);
}

template <int K, int METHOD>
__global__ void __launch_bounds__(kWave) k_remap(KParams P, View V, Scratch W, int n) {
  extern __shared__ double lds[];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  LaneRow row{lds + threadIdx.x, kWave};
  remap_fast<K, METHOD>(P, V, W, i, j, n, row);
}

template <int K>
__global__ void __launch_bounds__(kWave) k_uv(KParams P, View V, int nl) {
  extern __shared__ double lds[];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  bool isv = blockIdx.z != 0;
  if (i > P.ii) return;
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[ij_of(V, i, j)] == 0) return;
  LaneRow row{lds + threadIdx.x, kWave};
  uv_fast<K>(P, V, i, j, nl, isv, row);
}

template <int K>
static hipError_t launch_fast(const KParams& P, const View& V, const Scratch& W, int m,
                              int n, int raflag, hipStream_t s, PhaseTimer* T) {
  if (T)
    for (int q = 0; q < kPhases; q++) T->used[q] = false;
  dim3 blk(kBlock);
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  dim3 w1(kWave);
  dim3 gw2((P.ii + kWave - 1) / kWave, P.jj);
  dim3 gw3((P.ii + kWave - 1) / kWave, P.jj, 2);
  const size_t lds = sizeof(double) * (K + 1) * kWave;
  launch_diag_pre(P, V, n, s);
  {
    PhaseMark pm(T, s, 0);
    hipLaunchKernelGGL(k_regrid<K>, g2, blk, 0, s, P, V, W, n);
    switch (P.hybmap) {
      case 0: hipLaunchKernelGGL((k_remap<K, 0>), gw2, w1, lds, s, P, V, W, n); break;
      case 1: hipLaunchKernelGGL((k_remap<K, 1>), gw2, w1, lds, s, P, V, W, n); break;
      case 2: hipLaunchKernelGGL((k_remap<K, 2>), gw2, w1, lds, s, P, V, W, n); break;
      default: hipLaunchKernelGGL((k_remap<K, 3>), gw2, w1, lds, s, P, V, W, n); break;
    }
  }
  launch_diag_post(P, V, n, s);
  hipError_t e;
  {
    PhaseMark pm(T, s, 1);
    e = launch_halo(P, V, s);
  }
  if (e != hipSuccess) return e;
  {
    PhaseMark pm(T, s, 2);
    hipLaunchKernelGGL(k_uv<K>, gw3, w1, lds, s, P, V, n);
  }
  if (raflag) {
    {
      PhaseMark pm(T, s, 3);
      launch_ra_lane(P, V, m, n, s);
    }
    {
      PhaseMark pm(T, s, 4);
      e = launch_halo(P, V, s);
    }
    if (e != hipSuccess) return e;
    PhaseMark pm(T, s, 5);
    hipLaunchKernelGGL(k_uv<K>, gw3, w1, lds, s, P, V, m);
  }
  return hipGetLastError();
}

bool fast_path_ok(const KParams& P) {
  if (P.ldenfix || P.nhybrd != P.kk) return false;
  return P.kk == 22 || P.kk == 32 || P.kk == 41;
}

hipError_t launch_hybgen_fast(const KParams& P, const View& V, const Scratch& W, int m,
                              int n, int raflag, hipStream_t s, PhaseTimer* T) {
  switch (P.kk) {
    case 22: return launch_fast<22>(P, V, W, m, n, raflag, s, T);
    case 32: return launch_fast<32>(P, V, W, m, n, raflag, s, T);
    case 41: return launch_fast<41>(P, V, W, m, n, raflag, s, T);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace hyb
