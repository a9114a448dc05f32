// hyb_k_tile.hip -- the tiled p-column remap (hyb_tile.h): one workgroup of
// 256 threads per tile of 16 columns of a row.
#include "hyb_launch.h"
#include "hyb_tile.h"

namespace hyb {

// the workgroup's threads, each with its registers, separated by barriers
template <class ST>
struct TileTeamDev {
  ST st;
  int t;
  template <class F>
  __device__ __attribute__((always_inline)) void each(F&& f) {
    f(st, t);
    __syncthreads();
  }
  template <class F>
  __device__ __attribute__((always_inline)) bool each_or(F&& f) {
    const bool p = f(st, t);
    return __syncthreads_or(p ? 1 : 0) != 0;
  }
};

template <int Q>
__global__ void __launch_bounds__(kTileT) k_remap_t(KParams P, View V, Scratch W, int n) {
  extern __shared__ double lds[];
  TileTeamDev<TileThread<Q>> team;
  team.t = threadIdx.x;
  const int j = blockIdx.y + 1;
  const int i0 = i_aligned0(V, j) + blockIdx.x * kTileW;
  tile_remap_ppm<Q>(P, V, W, n, j, i0, lds, team);
}

bool tile_remap_ok(const KParams& P) {
  return P.hybmap == 2 && P.precision == 0 && P.kk >= 4 && P.kk + 1 <= 4 * kTileR;
}

void launch_remap_tiled(const KParams& P, const View& V, const Scratch& W, int n,
                        hipStream_t s) {
  dim3 blk(kTileT);
  dim3 grid((P.ii + kAlignPad + kTileW - 1) / kTileW, P.jj);
  const size_t lds = TileLayout{P.kk}.bytes();
  if (P.kk + 1 <= 2 * kTileR)
    hipLaunchKernelGGL((k_remap_t<2>), grid, blk, lds, s, P, V, W, n);
  else if (P.kk + 1 <= 3 * kTileR)
    hipLaunchKernelGGL((k_remap_t<3>), grid, blk, lds, s, P, V, W, n);
  else
    hipLaunchKernelGGL((k_remap_t<4>), grid, blk, lds, s, P, V, W, n);
}

}  // namespace hyb
