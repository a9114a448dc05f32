// hyb_launch.h -- launch helpers shared by the kernel translation units
// (hyb_k_*.hip).  Grid mapping everywhere: blockIdx.y = j row, threads
// along i, so every k-level access of a field is a coalesced row of the
// Fortran layout.
#pragma once

#include <hip/hip_runtime.h>

#include "hyb_kernels.h"

namespace hyb {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// HIP events around one phase (PhaseTimer), recorded on the call's stream
struct PhaseMark {
  PhaseTimer* T;
  hipStream_t s;
  int ph;
  PhaseMark(PhaseTimer* t, hipStream_t st, int p) : T(t), s(st), ph(p) {
    if (T) {
      (void)hipEventRecord(T->start[ph], s);
      T->used[ph] = true;
    }
  }
  ~PhaseMark() {
    if (T) (void)hipEventRecord(T->stop[ph], s);
  }
};

// hyb_k_uv.hip: the streaming u/v kernels of time level nl, then the
// fallback points they flagged
void launch_uv_stream(const KParams& P, const View& V, const Scratch& W, int nl,
                      const StreamTuning& tune, hipStream_t s);
// hyb_k_misc.hip
hipError_t launch_halo(const KParams& P, const View& V, hipStream_t s);
void launch_diag_pre(const KParams& P, const View& V, int n, hipStream_t s);
void launch_diag_post(const KParams& P, const View& V, int n, hipStream_t s);
// hyb_k_lane.hip: per-lane-array kernels used by the other families too
void launch_uv_lane(const KParams& P, const View& V, int nl, hipStream_t s);
void launch_ra_lane(const KParams& P, const View& V, int m, int n, hipStream_t s);
void launch_remap_fallback(const KParams& P, const View& V, const Scratch& W, int n,
                           hipStream_t s);
void launch_uv_fallback(const KParams& P, const View& V, const Scratch& W, int nl,
                        hipStream_t s);
// column_ra for the columns the streaming RA kernel flagged in W.kp
void launch_ra_fallback(const KParams& P, const View& V, const Scratch& W, int m, int n,
                        hipStream_t s);
// hyb_k_remap_m*.hip: streaming remap of the p-columns for one remapper;
// variant selects fields per walk and the prefetch depth
void launch_remap_stream_m0(const KParams& P, const View& V, const Scratch& W, int n,
                            int variant, hipStream_t s);
void launch_remap_stream_m1(const KParams& P, const View& V, const Scratch& W, int n,
                            int variant, hipStream_t s);
void launch_remap_stream_m2(const KParams& P, const View& V, const Scratch& W, int n,
                            int variant, hipStream_t s);
void launch_remap_stream_m3(const KParams& P, const View& V, const Scratch& W, int n,
                            int variant, hipStream_t s);
// the Robert-Asselin branch (hybgencj) on the streaming design, per remapper
void launch_ra_stream_m0(const KParams& P, const View& V, const Scratch& W, int m, int n,
                         hipStream_t s);
void launch_ra_stream_m1(const KParams& P, const View& V, const Scratch& W, int m, int n,
                         hipStream_t s);
void launch_ra_stream_m2(const KParams& P, const View& V, const Scratch& W, int m, int n,
                         hipStream_t s);
void launch_ra_stream_m3(const KParams& P, const View& V, const Scratch& W, int m, int n,
                         hipStream_t s);

}  // namespace hyb
