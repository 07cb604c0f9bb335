// spx_kernels_f32.hip -- float instantiations of the transport kernels.
//
// The FP32-step / FP64-accumulate kernels: built with reciprocal and approximate-function math
// and multiply-add contraction (the step arithmetic is FP32; positions, time, momentum and the
// weights accumulate in FP64).
//
// The gridded-background step kernel fits 4 resident waves per SIMD in 128 registers (a few spilled):
// measured 39% faster than at its natural 3 on the reference's Cartesian benchmark.  The FP64 one
// stays at its natural 2: forcing 3 spills ~100 registers and runs slower.  The guiding-centre step
// kernel likewise (137 registers natural, 5 dwords spilled at 128): 10% faster on the solar wind.
#define SPX_GRID_WAVES 4
#define SPX_GC_WAVES 4
#include "spx_parker.h"
#include "spx_gc.h"
#include "spx_lorentz.h"
#include "spx_focused.h"

namespace spx {

hipError_t launch_f32(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st,
                      hipEvent_t mid)
{
   if (hp.traj_type == SPX_TRAJ_PARKER || hp.traj_type == SPX_TRAJ_PARKER_SOURCE || hp.traj_type == SPX_TRAJ_FIELDLINE) return dispatch_parker<float>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type >= SPX_TRAJ_GUIDING_DIFF_SCATT && hp.traj_type <= SPX_TRAJ_GUIDING_DIFF) return dispatch_gc<float>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_LORENTZ) return dispatch_lorentz<float>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_FOCUSED) return dispatch_focused<float>(hp, n_lanes, d_launch, st, mid);
   return hipErrorInvalidValue;
}

}  // namespace spx
