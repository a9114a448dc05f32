// hyb_k_remap_m2.hip -- streaming p-column remap kernels, hybmap = 2
#include "hyb_launch.h"
#include "hyb_stream.h"
#define HYB_M 2
#define HYB_LAUNCH launch_remap_stream_m2
#define HYB_LAUNCH_RA launch_ra_stream_m2
#include "hyb_k_remap.inc"
