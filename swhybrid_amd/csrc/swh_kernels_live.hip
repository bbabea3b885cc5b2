// swh_kernels_live.hip -- the live-queue instantiations of the scan kernel
// (k_scan<..., LIVE = true>, fed by the serving loop while it runs), built
// from swh_kernels.hip as a second object so the two halves compile in
// parallel.
#define SWH_LIVE_ONLY
#include "swh_kernels.hip"
