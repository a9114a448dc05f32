// hyb_kernels.hip -- gfx950 kernels of the FP64 hybgen path and their
// launchers.  Grid mapping: blockIdx.y = j row, threads along i, so every
// k-level access of a field is a coalesced row of the Fortran layout.
#include <hip/hip_runtime.h>

#include "hyb_column.h"
#include "hyb_fast.h"
#include "hyb_kernels.h"
#include "hyb_stream.h"
#include "hyb_synth.h"

namespace hyb {

constexpr int kBlock = 256;

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

// ---------------------------------------------------------------------
// register-resident kernels for fixed K (hyb_fast.h)
// ---------------------------------------------------------------------
constexpr int kWave = 64;

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

// ---------------------------------------------------------------------
// streaming kernels (hyb_stream.h), any kdm
// ---------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_regrid_s(KParams P, View V, Scratch W, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  regrid_stream(P, V, W, i, j, n);
}

// one wave per block; each lane keeps its column's new interfaces in its
// own LDS column (stride 64 doubles: conflict-free for any per-lane k)
template <int METHOD>
__global__ void __launch_bounds__(kWave) k_remap_s(KParams P, View V, Scratch W, int n) {
  extern __shared__ double lds[];
  int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  int j = blockIdx.y + 1;
  if (i > P.ii) return;
  const int64_t c0 = ij_of(V, i, j);
  if (V.ip[c0] == 0) return;
  if (W.kp[c0] & kUnsafe) return;
  double* row = lds + threadIdx.x;
  const double* pg = V.p + c0;
  for (int k = 1; k <= P.kk + 1; k++) row[(k - 1) * kWave] = pg[(int64_t)(k - 1) * V.NIJ];
  LaneRow po{row, kWave};
  remap_stream<METHOD>(P, V, W, i, j, n, po);
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

static hipError_t launch_halo(const KParams& P, const View& V, hipStream_t s) {
  dim3 gj((P.ii + kBlock - 1) / kBlock, P.kk);
  hipLaunchKernelGGL(k_halo_j, gj, dim3(kBlock), 0, s, P, V);
  dim3 gi((P.jj + 2 + kBlock - 1) / kBlock, P.kk);
  hipLaunchKernelGGL(k_halo_i, gi, dim3(kBlock), 0, s, P, V);
  return hipGetLastError();
}

struct PhaseMark {
  PhaseTimer* T;
  hipStream_t s;
  int ph;
  PhaseMark(PhaseTimer* t, hipStream_t st, int p) : T(t), s(st), ph(p) {
    if (T) { (void)hipEventRecord(T->start[ph], s); T->used[ph] = true; }
  }
  ~PhaseMark() {
    if (T) (void)hipEventRecord(T->stop[ph], s);
  }
};

template <int KC>
static hipError_t launch_all(const KParams& P, const View& V, int m, int n,
                             int raflag, hipStream_t s, PhaseTimer* T) {
  if (T)
    for (int q = 0; q < kPhases; q++) T->used[q] = false;
  dim3 blk(kBlock);
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  dim3 g3((P.ii + kBlock - 1) / kBlock, P.jj, 2);
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  const double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) hipLaunchKernelGGL(k_diag_pre, g2, blk, 0, s, P, V, n, itr[d], src[d]);
  {
    PhaseMark pm(T, s, 0);
    hipLaunchKernelGGL(k_phase_a<KC>, g2, blk, 0, s, P, V, n);
  }
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      hipLaunchKernelGGL(k_diag_post, g2, blk, 0, s, P, V, n, itr[d], src[d],
                         d == 0 ? P.diag_q801 : P.diag_q80x);
  hipError_t e;
  {
    PhaseMark pm(T, s, 1);
    e = launch_halo(P, V, s);
  }
  if (e != hipSuccess) return e;
  {
    PhaseMark pm(T, s, 2);
    hipLaunchKernelGGL(k_phase_c<KC>, g3, blk, 0, s, P, V, n);
  }
  if (raflag) {
    {
      PhaseMark pm(T, s, 3);
      hipLaunchKernelGGL(k_phase_ra<KC>, g2, blk, 0, s, P, V, m, n);
    }
    {
      PhaseMark pm(T, s, 4);
      e = launch_halo(P, V, s);
    }
    if (e != hipSuccess) return e;
    PhaseMark pm(T, s, 5);
    hipLaunchKernelGGL(k_phase_c<KC>, g3, blk, 0, s, P, V, m);
  }
  return hipGetLastError();
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
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  const double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) hipLaunchKernelGGL(k_diag_pre, g2, blk, 0, s, P, V, n, itr[d], src[d]);
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
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      hipLaunchKernelGGL(k_diag_post, g2, blk, 0, s, P, V, n, itr[d], src[d],
                         d == 0 ? P.diag_q801 : P.diag_q80x);
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
      hipLaunchKernelGGL(k_phase_ra<K>, g2, blk, 0, s, P, V, m, n);
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

static void launch_uv(const KParams& P, const View& V, int nl, hipStream_t s) {
  dim3 blk(kBlock);
  dim3 g3((P.ii + kBlock - 1) / kBlock, P.jj, 2);
  dim3 w1(kWave);
  dim3 gw3((P.ii + kWave - 1) / kWave, P.jj, 2);
  const size_t lds = sizeof(double) * (P.kk + 1) * kWave;
  switch (P.kk) {
    case 22: hipLaunchKernelGGL(k_uv<22>, gw3, w1, lds, s, P, V, nl); return;
    case 32: hipLaunchKernelGGL(k_uv<32>, gw3, w1, lds, s, P, V, nl); return;
    case 41: hipLaunchKernelGGL(k_uv<41>, gw3, w1, lds, s, P, V, nl); return;
    default: break;
  }
  if (P.kk <= 24) hipLaunchKernelGGL(k_phase_c<24>, g3, blk, 0, s, P, V, nl);
  else if (P.kk <= 32) hipLaunchKernelGGL(k_phase_c<32>, g3, blk, 0, s, P, V, nl);
  else if (P.kk <= 48) hipLaunchKernelGGL(k_phase_c<48>, g3, blk, 0, s, P, V, nl);
  else hipLaunchKernelGGL(k_phase_c<64>, g3, blk, 0, s, P, V, nl);
}

static void launch_ra(const KParams& P, const View& V, int m, int n, hipStream_t s) {
  dim3 blk(kBlock);
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  if (P.kk <= 24) hipLaunchKernelGGL(k_phase_ra<24>, g2, blk, 0, s, P, V, m, n);
  else if (P.kk <= 32) hipLaunchKernelGGL(k_phase_ra<32>, g2, blk, 0, s, P, V, m, n);
  else if (P.kk <= 48) hipLaunchKernelGGL(k_phase_ra<48>, g2, blk, 0, s, P, V, m, n);
  else hipLaunchKernelGGL(k_phase_ra<64>, g2, blk, 0, s, P, V, m, n);
}

bool stream_path_ok(const KParams& P) { return !P.ldenfix && P.nhybrd == P.kk; }

hipError_t launch_hybgen_stream(const KParams& P, const View& V, const Scratch& W, int m,
                                int n, int raflag, hipStream_t s, PhaseTimer* T) {
  if (T)
    for (int q = 0; q < kPhases; q++) T->used[q] = false;
  dim3 blk(kBlock);
  dim3 g2((P.ii + kBlock - 1) / kBlock, P.jj);
  dim3 w1(kWave);
  dim3 gw2((P.ii + kWave - 1) / kWave, P.jj);
  const size_t lds = sizeof(double) * (P.kk + 1) * kWave;
  const int itr[3] = {P.itr801, P.itr802, P.itr803};
  const double* src[3] = {V.dp, V.temp, V.saln};
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0) hipLaunchKernelGGL(k_diag_pre, g2, blk, 0, s, P, V, n, itr[d], src[d]);
  {
    PhaseMark pm(T, s, 0);
    hipLaunchKernelGGL(k_regrid_s, g2, blk, 0, s, P, V, W, n);
    switch (P.hybmap) {
      case 0: hipLaunchKernelGGL(k_remap_s<0>, gw2, w1, lds, s, P, V, W, n); break;
      case 1: hipLaunchKernelGGL(k_remap_s<1>, gw2, w1, lds, s, P, V, W, n); break;
      case 2: hipLaunchKernelGGL(k_remap_s<2>, gw2, w1, lds, s, P, V, W, n); break;
      default: hipLaunchKernelGGL(k_remap_s<3>, gw2, w1, lds, s, P, V, W, n); break;
    }
    hipLaunchKernelGGL(k_remap_fb, gw2, w1, 0, s, P, V, W, n);
  }
  for (int d = 0; d < 3; d++)
    if (itr[d] > 0)
      hipLaunchKernelGGL(k_diag_post, g2, blk, 0, s, P, V, n, itr[d], src[d],
                         d == 0 ? P.diag_q801 : P.diag_q80x);
  hipError_t e;
  {
    PhaseMark pm(T, s, 1);
    e = launch_halo(P, V, s);
  }
  if (e != hipSuccess) return e;
  {
    PhaseMark pm(T, s, 2);
    launch_uv(P, V, n, s);
  }
  if (raflag) {
    {
      PhaseMark pm(T, s, 3);
      launch_ra(P, V, m, n, s);
    }
    {
      PhaseMark pm(T, s, 4);
      e = launch_halo(P, V, s);
    }
    if (e != hipSuccess) return e;
    PhaseMark pm(T, s, 5);
    launch_uv(P, V, m, s);
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

hipError_t launch_hybgen_fp64(const KParams& P, const View& V, int m, int n,
                              int raflag, hipStream_t s, PhaseTimer* T) {
  if (P.kk <= 24) return launch_all<24>(P, V, m, n, raflag, s, T);
  if (P.kk <= 32) return launch_all<32>(P, V, m, n, raflag, s, T);
  if (P.kk <= 48) return launch_all<48>(P, V, m, n, raflag, s, T);
  return launch_all<64>(P, V, m, n, raflag, s, T);
}

// ---------------------------------------------------------------------
// synthetic inputs on the device (same arithmetic as the host version)
// ---------------------------------------------------------------------
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
