// hyb_kernels.h -- launchers exported by hyb_kernels.hip to the C-ABI layer.
#pragma once

#include <hip/hip_runtime_api.h>

#include "hyb_column.h"
#include "hyb_synth.h"

namespace hyb {

// Optional HIP events around each phase of one call, for measurement:
// 0 column kernel (hybgenaj), 1 p halo, 2 u/v kernel (dpudpv+hybgenbj),
// 3 Robert-Asselin kernel (hybgencj), 4 second halo, 5 second u/v kernel.
constexpr int kPhases = 8;
struct PhaseTimer {
  hipEvent_t start[kPhases], stop[kPhases];
  bool used[kPhases];
};

// Enqueues the whole FP64 hybgen(m,n,raflag) on stream s.
hipError_t launch_hybgen_fp64(const KParams& P, const View& V, int m, int n,
                              int raflag, hipStream_t s, PhaseTimer* T = nullptr);

// True when the streaming kernels cover this parameter pack (any kdm <= 64
// with nhybrd==kdm; HYBGEN_DENFIX and the RA branch included).
bool stream_path_ok(const KParams& P);
// remapper of the u/v phase: 0 PCM, 1 PLM, 3 WENO (hybgen.F90:1811-1824)
int uv_method_of(const KParams& P);
// Kernel variants of the streaming family (measurement and tuning):
// remap_variant picks the p-column remap kernel (hyb_k_remap.inc),
// uv_prefetch the prefetch depth of the u/v walk (0, 2 or 4).
struct StreamTuning {
  int remap_variant = 3;  // 3 fields per walk, 2 levels loaded ahead (fastest measured)
  int uv_prefetch = 4;
  int uv_recompute = 1;  // u/v new thicknesses formed again from p, no LDS (fastest measured)
  int regrid_ahead = 1;  // the regrid's window loaded one iteration ahead (fastest measured)
  int concurrent_uv = 0; // the u/v phase on a second stream beside the p-column remap
};
// Enqueues hybgen with the streaming kernels; W must hold NIJ*(kk+1)
// doubles (pres), NIJ int32 (kp), NIJ uint64 (lcm) and 2*NIJ int32 (uvf).
hipError_t launch_hybgen_stream(const KParams& P, const View& V, const Scratch& W, int m,
                                int n, int raflag, const StreamTuning& tune, hipStream_t s,
                                PhaseTimer* T = nullptr);

// ---------------------------------------------------------------------
// Batched ragged domains (hyb_k_batch.hip).  Device arrays of the batch's
// domains, and the block layout of each launch: a group lists domains
// (dom[0..ndom)) and their block prefix (prefix[0..ndom], prefix[0] = 0,
// prefix[ndom] = nblocks); key selects the kernel template of the group.
// ---------------------------------------------------------------------
struct BatchDevice {
  const KParams* P;
  const View* V;
  const Scratch* W;
};
struct BatchGroup {
  const int32_t* dom;
  const int32_t* prefix;
  int32_t ndom, nblocks, key;
};
constexpr int kMaxGroups = 8;
struct BatchLaunch {
  int32_t kmax;        // largest kdm of the batch (LDS size)
  int32_t remap_cols;  // columns per remap workgroup: 64, 128 or 256
  int32_t regrid_cols; // columns per regrid workgroup: 64, 128 or 256
  int32_t remap_stage; // interface levels a remap lane stages in LDS (0: all)
  BatchGroup regrid;                 // row blocks of regrid_cols, all domains
  BatchGroup remap[kMaxGroups];      // row blocks of remap_cols, by hybmap + 4*precision
  int32_t nremap;
  BatchGroup rows64;                 // 64-wide row blocks, all domains
  BatchGroup halo_j, halo_i;         // kk x 256-wide slices
  BatchGroup uv[kMaxGroups];         // u+v 64-wide blocks, by uv method + 4*precision
  int32_t nuv;
  BatchGroup uvall;                  // u+v 64-wide blocks, all domains
};
// Enqueues one hybgen(m,n,.false.) step of every domain of the batch.
hipError_t launch_batch(const BatchDevice& B, const BatchLaunch& L, int n, hipStream_t s);
// the interface levels per column a remap workgroup of L stages in LDS
int batch_stage_levels(const BatchLaunch& L);

// NaN/Inf scan of a call's inputs: scan_blocks(P) int32 results in out
// (hyb_k_misc.hip; the key of a bad column is i*128 + field).
int scan_blocks(const KParams& P);
hipError_t launch_scan_finite(const KParams& P, const View& V, int n, int32_t* out,
                              hipStream_t s);

// Self-test of the FP64 quotient helpers of hyb_math.h on the device:
// fast[q] from the helper `which` (0 qdiv, 1 qdiv_pos, 2 qrcp, 3 qdiv_any,
// 4 div3), ieee[q] from the compiler's a/b (1/b, a/3) on the same operands.
hipError_t launch_selftest_quot(int which, const double* a, const double* b, double* fast,
                                double* ieee, int64_t n, hipStream_t s);

// Enqueues the synthetic-input generator for window (1..ii, 1..jj).
hipError_t launch_synth(const SynthDomain& D, const View& V, int ii, int jj,
                        int m, int n, hipStream_t s);

}  // namespace hyb
