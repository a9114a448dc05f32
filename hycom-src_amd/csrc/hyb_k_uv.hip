// hyb_k_uv.hip -- the streaming u/v kernels (uv_stream, hyb_stream.h):
// dpudpv, the old and new interfaces and the remap of u and v at every u-
// and v-point (hybgen.F90:171-195, 1758-1907), any kdm <= 64.  A file of
// its own so that it can take its own scheduling options.
#include "hyb_launch.h"
#include "hyb_stream.h"

namespace hyb {

template <int METHOD, int PD, bool MIXED = false, bool RECOMP = false>
__global__ void __launch_bounds__(kWave) k_uv_s(KParams P, View V, Scratch W, int nl) {
  extern __shared__ double lds[];
  int j = blockIdx.y + 1;
  int i = i_aligned0(V, j) + blockIdx.x * blockDim.x + threadIdx.x;
  bool isv = blockIdx.z != 0;
  if (i < 1 || i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  const int32_t* msk = isv ? V.iv : V.iu;
  if (msk[c0] == 0) return;
  LaneRow dcol{lds + threadIdx.x, kWave};
  const bool ok = uv_stream<METHOD, PD, MIXED, RECOMP>(P, V, i, j, nl, isv, dcol);
  W.uvf[c0 + (isv ? V.NIJ : 0)] = ok ? 0 : 1;
}

template <int M, bool RECOMP>
static void uv_stream_m(const KParams& P, const View& V, const Scratch& W, int nl, int pd,
                        hipStream_t s) {
  dim3 w1(kWave);
  dim3 gw3((P.ii + kAlignPad + kWave - 1) / kWave, P.jj, 2);
  const size_t lds = RECOMP ? 0 : sizeof(double) * P.kk * kWave;
  if (P.precision == 1)
    hipLaunchKernelGGL((k_uv_s<M, 4, true, RECOMP>), gw3, w1, lds, s, P, V, W, nl);
  else if (pd == 4)
    hipLaunchKernelGGL((k_uv_s<M, 4, false, RECOMP>), gw3, w1, lds, s, P, V, W, nl);
  else if (pd > 0)
    hipLaunchKernelGGL((k_uv_s<M, 2, false, RECOMP>), gw3, w1, lds, s, P, V, W, nl);
  else
    hipLaunchKernelGGL((k_uv_s<M, 0, false, RECOMP>), gw3, w1, lds, s, P, V, W, nl);
}

void launch_uv_stream(const KParams& P, const View& V, const Scratch& W, int nl,
                      const StreamTuning& tune, hipStream_t s) {
  const int m = uv_method(P);
  const int pd = tune.uv_prefetch;
  if (tune.uv_recompute) {
    if (m == 0) uv_stream_m<0, true>(P, V, W, nl, pd, s);
    else if (m == 1) uv_stream_m<1, true>(P, V, W, nl, pd, s);
    else uv_stream_m<3, true>(P, V, W, nl, pd, s);
  } else {
    if (m == 0) uv_stream_m<0, false>(P, V, W, nl, pd, s);
    else if (m == 1) uv_stream_m<1, false>(P, V, W, nl, pd, s);
    else uv_stream_m<3, false>(P, V, W, nl, pd, s);
  }
  launch_uv_fallback(P, V, W, nl, s);
}

}  // namespace hyb
