// hyb_k_remap_m1.hip -- streaming p-column remap kernels, hybmap = 1
#include "hyb_launch.h"
#include "hyb_stream.h"
#define HYB_M 1
#define HYB_LAUNCH launch_remap_stream_m1
#define HYB_LAUNCH_RA launch_ra_stream_m1
#include "hyb_k_remap.inc"
