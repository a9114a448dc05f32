// hyb_k_remap_m0.hip -- streaming p-column remap kernels, hybmap = 0
#include "hyb_launch.h"
#include "hyb_stream.h"
#define HYB_M 0
#define HYB_LAUNCH launch_remap_stream_m0
#define HYB_LAUNCH_RA launch_ra_stream_m0
#include "hyb_k_remap.inc"
