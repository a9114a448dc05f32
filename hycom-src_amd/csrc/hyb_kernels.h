// hyb_kernels.h -- launchers exported by hyb_kernels.hip to the C-ABI layer.
#pragma once

#include <hip/hip_runtime_api.h>

#include "hyb_column.h"
#include "hyb_synth.h"

namespace hyb {

// Enqueues the whole FP64 hybgen(m,n,raflag) on stream s.
hipError_t launch_hybgen_fp64(const KParams& P, const View& V, int m, int n,
                              int raflag, hipStream_t s);

// Enqueues the synthetic-input generator for window (1..ii, 1..jj).
hipError_t launch_synth(const SynthDomain& D, const View& V, int ii, int jj,
                        int m, int n, hipStream_t s);

}  // namespace hyb
