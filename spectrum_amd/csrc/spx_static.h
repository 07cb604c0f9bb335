// spx_static.h -- the run that allocates nothing: device state in one static array (spx_static.hip).
#pragma once

#include "spx_kernels.h"

namespace spx {

struct StaticResult {
   RunCtl ctl;          // the control block at the end of the run
   double wall_s = 0.0;
   int launches = 0;
   int64_t lanes = 0;
};

// Integrate trajectories first .. first+n_traj-1 of `plan` (counter streams keyed by `seed`) on the
// SPX_STATIC_LANES slots of the static array of the current device, and copy the histograms into the
// caller's host arrays (counts at hist_off, sums at acc_off, one event count per distribution).
// hipErrorInvalidValue when the plan's histograms exceed SPX_STATIC_BINS / SPX_STATIC_COMPONENTS.
hipError_t static_run(const spx_plan_desc& plan, int precision, int steps_per_launch, uint64_t first, uint64_t n_traj, uint64_t seed,
                      int64_t* counts, double* sums, int64_t* n_events, StaticResult& res);

}  // namespace spx
