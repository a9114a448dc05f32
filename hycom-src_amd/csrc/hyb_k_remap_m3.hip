// hyb_k_remap_m3.hip -- streaming p-column remap kernels, hybmap = 3
#include "hyb_launch.h"
#include "hyb_stream.h"
#define HYB_M 3
#define HYB_LAUNCH launch_remap_stream_m3
#define HYB_LAUNCH_RA launch_ra_stream_m3
#include "hyb_k_remap.inc"
