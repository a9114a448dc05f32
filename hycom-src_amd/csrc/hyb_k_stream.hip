// hyb_k_stream.hip -- the streaming family (hyb_stream.h): regrid kernels
// and the launcher of the whole call, any kdm <= 64 (the u/v kernels are in
// hyb_k_uv.hip).
#include <mutex>

#include "hyb_launch.h"
#include "hyb_stream.h"

namespace hyb {

// A second stream per device for the u/v phase, which needs only the
// regrid's interfaces and runs beside the p-column remap: the two kernels
// touch disjoint outputs and are both bound by latency at low occupancy.
// The fork/join events order it after the regrid and before whatever the
// caller enqueues next; the mutex keeps one call's record/wait pairs
// together when several host threads call at once.
namespace {
struct Fork {
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex g_fork_mu;
Fork g_fork[64];
hipError_t fork_of(Fork** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
  Fork& F = g_fork[dev];
  if (!F.s2) {
    if ((e = hipStreamCreateWithFlags(&F.s2, hipStreamNonBlocking)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&F.fork, hipEventDisableTiming)) != hipSuccess) return e;
    if ((e = hipEventCreateWithFlags(&F.join, hipEventDisableTiming)) != hipSuccess) return e;
  }
  *out = &F;
  return hipSuccess;
}
}  // namespace

__global__ void __launch_bounds__(kBlock) k_regrid_s(KParams P, View V, Scratch W, int n) {
  int j = blockIdx.y + 1;
  int i = i_aligned0(V, j) + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  regrid_stream(P, V, W, i, j, n);
}

// the regrid with its window loaded one iteration ahead, held to 3 waves
// per SIMD
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3)))
k_regrid_a(KParams P, View V, Scratch W, int n) {
  int j = blockIdx.y + 1;
  int i = i_aligned0(V, j) + blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 1 || i > P.ii) return;
  if (V.ip[ij_of(V, i, j)] == 0) return;
  regrid_stream<1>(P, V, W, i, j, n);
}

bool stream_path_ok(const KParams& P) { return P.nhybrd == P.kk; }

int uv_method_of(const KParams& P) { return uv_method(P); }

hipError_t launch_hybgen_stream(const KParams& P, const View& V, const Scratch& W, int m,
                                int n, int raflag, const StreamTuning& tune, hipStream_t s,
                                PhaseTimer* T) {
  if (T)
    for (int q = 0; q < kPhases; q++) T->used[q] = false;
  dim3 blk(kBlock);
  dim3 g2((P.ii + kAlignPad + kBlock - 1) / kBlock, P.jj);
  launch_diag_pre(P, V, n, s);
  std::unique_lock<std::mutex> fork_lock(g_fork_mu, std::defer_lock);
  Fork* F = nullptr;
  hipStream_t su = s;  // stream of the u/v phase
  {
    PhaseMark pm(T, s, 0);
    {
      PhaseMark pk(T, s, 6);
      if (tune.regrid_ahead)
        hipLaunchKernelGGL(k_regrid_a, g2, blk, 0, s, P, V, W, n);
      else
        hipLaunchKernelGGL(k_regrid_s, g2, blk, 0, s, P, V, W, n);
    }
    if (tune.concurrent_uv) {
      hipError_t e = fork_of(&F);
      if (e != hipSuccess) return e;
      fork_lock.lock();
      if ((e = hipEventRecord(F->fork, s)) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(F->s2, F->fork, 0)) != hipSuccess) return e;
      su = F->s2;
      {
        PhaseMark pu(T, su, 1);
        if ((e = launch_halo(P, V, su)) != hipSuccess) return e;
      }
      PhaseMark pu(T, su, 2);
      launch_uv_stream(P, V, W, n, tune, su);
    }
    {
      PhaseMark pk(T, s, 7);
      const int rv = tune.remap_variant;
      switch (P.hybmap) {
        case 0: launch_remap_stream_m0(P, V, W, n, rv, s); break;
        case 1: launch_remap_stream_m1(P, V, W, n, rv, s); break;
        case 2: launch_remap_stream_m2(P, V, W, n, rv, s); break;
        default: launch_remap_stream_m3(P, V, W, n, rv, s); break;
      }
    }
    launch_remap_fallback(P, V, W, n, s);
  }
  launch_diag_post(P, V, n, s);
  hipError_t e;
  if (F) {  // join: the caller's stream waits for the u/v phase
    if ((e = hipEventRecord(F->join, su)) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(s, F->join, 0)) != hipSuccess) return e;
    fork_lock.unlock();
  } else {
    {
      PhaseMark pm(T, s, 1);
      e = launch_halo(P, V, s);
    }
    if (e != hipSuccess) return e;
    PhaseMark pm(T, s, 2);
    launch_uv_stream(P, V, W, n, tune, s);
  }
  if (raflag) {
    {
      PhaseMark pm(T, s, 3);
      switch (P.hybmap) {
        case 0: launch_ra_stream_m0(P, V, W, m, n, s); break;
        case 1: launch_ra_stream_m1(P, V, W, m, n, s); break;
        case 2: launch_ra_stream_m2(P, V, W, m, n, s); break;
        default: launch_ra_stream_m3(P, V, W, m, n, s); break;
      }
      launch_ra_fallback(P, V, W, m, n, s);
    }
    {
      PhaseMark pm(T, s, 4);
      e = launch_halo(P, V, s);
    }
    if (e != hipSuccess) return e;
    PhaseMark pm(T, s, 5);
    launch_uv_stream(P, V, W, m, tune, s);
  }
  return hipGetLastError();
}

}  // namespace hyb
