// hyb_k_misc.hip -- p halo, diagnostic tracers and the synthetic-input
// generator.
#include <stdint.h>

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

// Scan of the inputs of one call for NaN/Inf (hyb_params.check_finite):
// sea p-points (dp, T, S, th3d, theta, tracers at level n; depths, topiso),
// u-points (u, dpu at n; depthu, ubavg) and v-points alike.  Each block
// writes the smallest key i*128 + field of its bad columns (INT32_MAX if
// none) to out[block]; the host reads the first bad block in row order.
__global__ void __launch_bounds__(kBlock) k_scan_finite(KParams P, View V, int n,
                                                         int32_t* out) {
  __shared__ int32_t red[kBlock];
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y + 1;
  int32_t key = INT32_MAX;
  if (i <= P.ii) {
    const int64_t c0 = ij_of(V, i, j);
    int field = -1;
    auto check = [&](const double* a, int id) {
      if (field < 0 && !__builtin_isfinite(*a)) field = id;
    };
    if (V.ip[c0]) {
      for (int k = 1; k <= P.kk; k++) {
        const int64_t a = c0 + lev4(V, k, n);
        check(V.dp + a, 0);
        check(V.temp + a, 1);
        check(V.saln + a, 2);
        check(V.th3d + a, 3);
        check(V.theta + c0 + lev(V, k), 4);
        for (int t = 1; t <= P.ntracr; t++) check(V.tracer + c0 + levt(V, k, n, t), 16 + t);
      }
      check(V.depths + c0, 5);
      check(V.topiso + c0, 6);
    }
    if (V.iu[c0]) {
      for (int k = 1; k <= P.kk; k++) {
        check(V.u + c0 + lev4(V, k, n), 7);
        check(V.dpu + c0 + lev4(V, k, n), 8);
      }
      check(V.depthu + c0, 9);
      check(V.ubavg + c0 + lev(V, n), 10);
    }
    if (V.iv[c0]) {
      for (int k = 1; k <= P.kk; k++) {
        check(V.v + c0 + lev4(V, k, n), 11);
        check(V.dpv + c0 + lev4(V, k, n), 12);
      }
      check(V.depthv + c0, 13);
      check(V.vbavg + c0 + lev(V, n), 14);
    }
    if (field >= 0) key = i * 128 + field;
  }
  red[threadIdx.x] = key;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w && red[threadIdx.x + w] < red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

int scan_blocks(const KParams& P) { return ((P.ii + kBlock - 1) / kBlock) * P.jj; }

hipError_t launch_scan_finite(const KParams& P, const View& V, int n, int32_t* out,
                              hipStream_t s) {
  dim3 g((P.ii + kBlock - 1) / kBlock, P.jj);
  hipLaunchKernelGGL(k_scan_finite, g, dim3(kBlock), 0, s, P, V, n, out);
  return hipGetLastError();
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

// The quotient helpers beside the compiler's division, element by element
// (testing only: the hot kernels inline the same helpers).
__global__ void k_selftest_quot(int which, const double* a, const double* b, double* fast,
                                double* ieee, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const double x = a[q], y = b[q];
  double f, e;
  if (which == 4) {
    f = div3(x);
    e = x / 3.0;
  } else if (which == 2) {
    f = qrcp(rcp64(y));
    e = 1.0 / y;
  } else {
    const Rcp64 r = rcp64(y);
    f = which == 0 ? qdiv(x, r) : which == 1 ? qdiv_pos(x, r) : qdiv_any(x, r);
    e = x / y;
  }
  fast[q] = f;
  ieee[q] = e;
}

hipError_t launch_selftest_quot(int which, const double* a, const double* b, double* fast,
                                double* ieee, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int64_t nb = (n + 255) / 256;
  hipLaunchKernelGGL(k_selftest_quot, dim3((unsigned)nb), dim3(256), 0, s, which, a, b, fast,
                     ieee, n);
  return hipGetLastError();
}

}  // namespace hyb
