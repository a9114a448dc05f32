// hyb_host.h -- host-side helpers shared by the C-ABI translation units
// (hyb_capi.cpp, hyb_server.cpp): error reporting, device and argument
// checks, and the block layout of batched launches.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/hycom_hybgen.h"
#include "hyb_kernels.h"

namespace hyb {

// Sets the calling thread's last-error text (hycom_hybgen_last_error) and
// returns code.
int fail(int code, const char* fmt, ...);
// Text of the calling thread's last error.
const char* last_error();
// 0 once a gfx950 device is current, HYB_ENODEVICE otherwise.
int device_ready();
// Every field the call reads or writes is present.
int check_fields(const hyb_params* P, const hyb_fields* F, int raflag);
// Domain d of a batch can be stepped by the batched kernels.
int batch_domain_ok(const hyb_params* P, const hyb_fields* F, int d);

// Block layout of every launch of launch_batch for nd domains with
// parameters hP and the widths in L.remap_cols / L.regrid_cols.  The
// groups' dom/prefix arrays are appended to I; their pointers in L hold
// offsets into I until batch_bind() turns them into device addresses.
void batch_layout(const KParams* hP, int nd, BatchLaunch& L, std::vector<int32_t>& I);
void batch_bind(BatchLaunch& L, const int32_t* dI);

}  // namespace hyb

#define HIPCHK(expr)                                                           \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return ::hyb::fail(e_ == hipErrorOutOfMemory ? HYB_ENOMEM : HYB_EDEVICE, \
                         "%s: %s", #expr, hipGetErrorString(e_));              \
  } while (0)
