// spx_kernels_f64.hip -- double instantiations of the transport kernels.
//
// The plain FP64 kernels (the drop-in): built with contraction off so that every floating-point
// operation rounds where the reference's does.
#include "spx_parker.h"
#include "spx_gc.h"
#include "spx_lorentz.h"
#include "spx_focused.h"

namespace spx {

__global__ __launch_bounds__(256) void grid_cells_kernel(const double* __restrict__ zones, int32_t n0, int32_t n1, int32_t n2,
                                                         float* __restrict__ c32)
{
   const int64_t total = (int64_t)n0 * n1 * n2 * 8;
   for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
      const int64_t cell = t >> 3;
      const int c = (int)(t & 7);
      const int64_t gx = cell % n0, gy = (cell / n0) % n1, gz = cell / ((int64_t)n0 * n1);
      const int64_t x = gx + (c & 1) < n0 ? gx + (c & 1) : n0 - 1;
      const int64_t y = gy + ((c >> 1) & 1) < n1 ? gy + ((c >> 1) & 1) : n1 - 1;
      const int64_t z = gz + (c >> 2) < n2 ? gz + (c >> 2) : n2 - 1;
      const double* src = zones + (x + n0 * (y + (int64_t)n1 * z)) * SPX_GRID_REC;
#pragma unroll
      for (int m = 0; m < SPX_GRID_REC; m++) c32[t * SPX_GRID_REC + m] = (float)src[m];
   }
}

hipError_t grid_expand_cells(const double* zones, const int32_t nz[3], float* c32, hipStream_t st)
{
   const int64_t total = (int64_t)nz[0] * nz[1] * nz[2] * 8;
   const int64_t blocks = (total + 255) / 256;
   hipLaunchKernelGGL(grid_cells_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, st, zones, nz[0], nz[1], nz[2],
                      c32);
   return hipGetLastError();
}

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
