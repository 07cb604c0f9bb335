// spx_kernels_f64.hip -- double instantiations of the transport kernels.
//
// The plain FP64 kernels (the drop-in): built with contraction off so that every floating-point
// operation rounds where the reference's does.
#include "spx_parker.h"
#include "spx_gc.h"

namespace spx {

hipError_t launch_f64(const spx_plan_desc& hp, const spx_plan_desc* d_plan, const Launch& L, const Launch* d_launch, hipStream_t st,
                      hipEvent_t mid)
{
   if (hp.traj_type == SPX_TRAJ_PARKER || hp.traj_type == SPX_TRAJ_PARKER_SOURCE) return dispatch_parker<double>(hp, d_plan, L, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_GUIDING_DIFF_SCATT) return dispatch_gc<double>(hp, d_plan, L, d_launch, st, mid);
   return hipErrorInvalidValue;
}

}  // namespace spx
