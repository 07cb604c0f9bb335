// spx_kernels_f64.hip -- double instantiations of the transport kernels.
//
// The plain FP64 kernels (the drop-in): built with contraction off so that every floating-point
// operation rounds where the reference's does.
#include "spx_parker.h"
#include "spx_gc.h"
#include "spx_lorentz.h"
#include "spx_focused.h"

namespace spx {

hipError_t launch_f64(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st,
                      hipEvent_t mid)
{
   if (hp.traj_type == SPX_TRAJ_PARKER || hp.traj_type == SPX_TRAJ_PARKER_SOURCE || hp.traj_type == SPX_TRAJ_FIELDLINE) return dispatch_parker<double>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type >= SPX_TRAJ_GUIDING_DIFF_SCATT && hp.traj_type <= SPX_TRAJ_GUIDING_DIFF) return dispatch_gc<double>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_LORENTZ) return dispatch_lorentz<double>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_FOCUSED) return dispatch_focused<double>(hp, n_lanes, d_launch, st, mid);
   return hipErrorInvalidValue;
}

}  // namespace spx
