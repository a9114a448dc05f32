// hyb_k_misc.hip -- p halo, diagnostic tracers and the synthetic-input
// generator.
#include "hyb_launch.h"
#include "hyb_synth.h"

namespace hyb {

// p halo, levels 2..kk+1: first the j rows (mod_xc_sm.h:1374-1392), then
// the i columns including the corners (1394-1424).
__global__ void k_halo_j(KParams P, View V) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int k = blockIdx.y + 2;
  if (i > P.ii) return;
  p_halo_jrow(P, V, i, k);
}
__global__ void k_halo_i(KParams P, View V) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;  // 0..jj+1
  int k = blockIdx.y + 2;
  if (j > P.jj + 1) return;
  p_halo_icol(P, V, j, k);
}

// diagnostic tracers itracr(801..803): copy before (hybgen.F90:44-87),
// rate after (97-164).
__global__ void k_diag_pre(KParams P, View V, int n, int t, const double* src) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  int64_t c0 = ij_of(V, i, j);
  for (int k = 1; k <= P.kk; k++)
    V.tracer[c0 + levt(V, k, n, t)] = src[c0 + lev4(V, k, n)];
}
__global__ void k_diag_post(KParams P, View V, int n, int t, const double* src, double q) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  int64_t c0 = ij_of(V, i, j);
  if (V.ip[c0] == 0) return;
  for (int k = 1; k <= P.kk; k++) {
    int64_t a = c0 + levt(V, k, n, t);
    V.tracer[a] = q * (src[c0 + lev4(V, k, n)] - V.tracer[a]);
  }
}

hipError_t launch_halo(const KParams& P, const View& V, hipStream_t s) {
  dim3 gj((P.ii + kBlock - 1) / kBlock, P.kk);
  hipLaunchKernelGGL(k_halo_j, gj, dim3(kBlock), 0, s, P, V);
  dim3 gi((P.jj + 2 + kBlock - 1) / kBlock, P.kk);
  hipLaunchKernelGGL(k_halo_i, gi, dim3(kBlock), 0, s, P, V);
  return hipGetLastError();
}

void launch_diag_pre(const KParams& P, const View& V, int n, hipStream_t s) {
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  const double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) hipLaunchKernelGGL(k_diag_pre, g2, dim3(kBlock), 0, s, P, V, n, itr[d], src[d]);
}

void launch_diag_post(const KParams& P, const View& V, int n, hipStream_t s) {
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  const double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      hipLaunchKernelGGL(k_diag_post, g2, dim3(kBlock), 0, s, P, V, n, itr[d], src[d],
                         d == 0 ? P.diag_q801 : P.diag_q80x);
}

// synthetic inputs on the device (same arithmetic as the host version)
__global__ void k_synth_pass1(SynthDomain D, View V, int ii, int jj, int m, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;  // 0..ii+1
  int j = blockIdx.y;                             // 0..jj+1
  if (i > ii + 1) return;
  synth_point_pass1(D, V, ii, jj, i, j, m, n);
}
__global__ void k_synth_pass2(SynthDomain D, View V, int ii, int jj, int m, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > ii) return;
  synth_point_pass2(D, V, ii, jj, i, j, m, n);
}

hipError_t launch_synth(const SynthDomain& D, const View& V, int ii, int jj,
                        int m, int n, hipStream_t s) {
  dim3 blk(64);
  dim3 g1((ii + 2 + 63) / 64, jj + 2);
  hipLaunchKernelGGL(k_synth_pass1, g1, blk, 0, s, D, V, ii, jj, m, n);
  dim3 g2((ii + 63) / 64, jj);
  hipLaunchKernelGGL(k_synth_pass2, g2, blk, 0, s, D, V, ii, jj, m, n);
  return hipGetLastError();
}

}  // namespace hyb
