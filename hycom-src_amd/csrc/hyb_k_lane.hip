// hyb_k_lane.hip -- the per-lane-array kernels (hyb_column.h): every kdm,
// every option (HYBGEN_DENFIX, nhybrd < kdm); also the copy-first kernels
// for the points the streaming kernels leave to them.
#include "hyb_column.h"
#include "hyb_launch.h"
#include "hyb_stream.h"

namespace hyb {

// initial lcm of a row when isopcm is off (hybgen.F90:386-394)
HYB_HD uint64_t lcm_init(const KParams& P) {
  uint64_t m = 0;
  if (!P.isopcm)
    for (int k = P.nhybrd + 1; k <= P.kk; k++) m |= 1ull << (k - 1);
  return m;
}

template <int KC>
__global__ void __launch_bounds__(kBlock) k_phase_a(KParams P, View V, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  uint64_t lcm = lcm_init(P);
  column_a<KC>(P, V, i, j, n, lcm);
}

template <int KC>
__global__ void __launch_bounds__(kBlock) k_phase_ra(KParams P, View V, int m, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  column_ra<KC>(P, V, i, j, m, n, lcm_init(P));
}

template <int KC>
__global__ void __launch_bounds__(kBlock) k_phase_c(KParams P, View V, int nl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  bool isv = blockIdx.z != 0;
  if (i > P.ii) return;
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[ij_of(V, i, j)] == 0) return;
  point_c<KC>(P, V, i, j, nl, isv);
}

__global__ void __launch_bounds__(kWave) k_remap_fb(KParams P, View V, Scratch W, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  if (V.ip[c0] == 0) return;
  if (!(W.kp[c0] & kUnsafe)) return;
  remap_fallback(P, V, W, i, j, n);
}

template <int KC>
__global__ void __launch_bounds__(kWave) k_uv_fb(KParams P, View V, Scratch W, int nl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  bool isv = blockIdx.z != 0;
  if (i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[c0] == 0) return;
  if (W.uvf[c0 + (isv ? V.NIJ : 0)] == 0) return;
  point_c<KC>(P, V, i, j, nl, isv);
}

template <int KC>
__global__ void __launch_bounds__(kWave) k_ra_fb(KParams P, View V, Scratch W, int m, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  if (V.ip[c0] == 0) return;
  if (!(W.kp[c0] & kUnsafe)) return;
  column_ra<KC>(P, V, i, j, m, n, lcm_init(P));
}

#define HYB_KC_DISPATCH(KK, CALL) \
  do {                            \
    if ((KK) <= 24) {             \
      constexpr int KC = 24;      \
      CALL;                       \
    } else if ((KK) <= 32) {      \
      constexpr int KC = 32;      \
      CALL;                       \
    } else if ((KK) <= 48) {      \
      constexpr int KC = 48;      \
      CALL;                       \
    } else {                      \
      constexpr int KC = 64;      \
      CALL;                       \
    }                             \
  } while (0)

void launch_uv_lane(const KParams& P, const View& V, int nl, hipStream_t s) {
  dim3 g3((P.ii + kBlock - 1) / kBlock, P.jj, 2);
  HYB_KC_DISPATCH(P.kk, hipLaunchKernelGGL(k_phase_c<KC>, g3, dim3(kBlock), 0, s, P, V, nl));
}

void launch_ra_lane(const KParams& P, const View& V, int m, int n, hipStream_t s) {
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  HYB_KC_DISPATCH(P.kk, hipLaunchKernelGGL(k_phase_ra<KC>, g2, dim3(kBlock), 0, s, P, V, m, n));
}

void launch_remap_fallback(const KParams& P, const View& V, const Scratch& W, int n,
                           hipStream_t s) {
  dim3 gw2((P.ii + kWave - 1) / kWave, P.jj);
  hipLaunchKernelGGL(k_remap_fb, gw2, dim3(kWave), 0, s, P, V, W, n);
}

void launch_ra_fallback(const KParams& P, const View& V, const Scratch& W, int m, int n,
                        hipStream_t s) {
  dim3 gw2((P.ii + kWave - 1) / kWave, P.jj);
  HYB_KC_DISPATCH(P.kk, hipLaunchKernelGGL(k_ra_fb<KC>, gw2, dim3(kWave), 0, s, P, V, W, m, n));
}

void launch_uv_fallback(const KParams& P, const View& V, const Scratch& W, int nl,
                        hipStream_t s) {
  dim3 gw3((P.ii + kWave - 1) / kWave, P.jj, 2);
  HYB_KC_DISPATCH(P.kk, hipLaunchKernelGGL(k_uv_fb<KC>, gw3, dim3(kWave), 0, s, P, V, W, nl));
}

template <int KC>
static hipError_t launch_all(const KParams& P, const View& V, int m, int n,
                             int raflag, hipStream_t s, PhaseTimer* T) {
  if (T)
    for (int q = 0; q < kPhases; q++) T->used[q] = false;
  dim3 blk(kBlock);
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  launch_diag_pre(P, V, n, s);
  {
    PhaseMark pm(T, s, 0);
    hipLaunchKernelGGL(k_phase_a<KC>, g2, blk, 0, s, P, V, n);
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
    launch_uv_lane(P, V, n, s);
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
    launch_uv_lane(P, V, m, s);
  }
  return hipGetLastError();
}

hipError_t launch_hybgen_fp64(const KParams& P, const View& V, int m, int n,
                              int raflag, hipStream_t s, PhaseTimer* T) {
  HYB_KC_DISPATCH(P.kk, return launch_all<KC>(P, V, m, n, raflag, s, T));
}

}  // namespace hyb
