// spx_kernels.h -- launch-side types shared by the kernels and the host runtime.
#pragma once

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "spx_plan.h"
#include "spx_acc.h"
#include "../../include/spectrum_amd.h"

namespace spx {

// Device-side control block of one run or served query.
struct RunCtl {
   unsigned long long next;             // next trajectory id handed out
   unsigned long long done;             // trajectories finished (accepted or failed)
   unsigned long long failed;           // trajectories that exhausted their attempts
   unsigned long long steps_done;       // accepted steps of finished trajectories
   unsigned long long steps_executed;   // every executed step (throughput counter)
   unsigned long long discards[8];      // discarded attempts by reason
   unsigned long long tmin_bits, tmax_bits;
};

// One run, or one served query: what its trajectories share.  Lives in device memory and is read
// with scalar loads; each slot pool holds a table of them.
struct QueryDesc {
   const spx_plan_desc* plan;
   RunCtl* ctl;
   spx_rng_desc rng;                    // replay pointers are device pointers here
   uint64_t first, end;                 // trajectory id range
   int32_t n_dist, pad_;
   int64_t hist_off[SPX_MAX_DIST];      // offsets of each distribution in this query's counts (bins)
   int64_t acc_off[SPX_MAX_DIST];       // and in its exact sums (bin components)
   unsigned long long* hist_counts;
   unsigned long long* hist_acc;        // SPX_ACC_LIMBS limbs per bin component (spx_acc.h)
   unsigned long long* hist_events;     // per distribution
   spx_traj_record* records;            // optional per-trajectory records
   uint64_t rec_first, rec_n;
};

// A pool of trajectory slots (kernel argument block, in device memory).  The slots of one wave always
// work for the same query; a wave moves on to the oldest admitted query with trajectories left when
// all of its slots are idle and its own query has none.
struct Launch {
   const QueryDesc* queries;            // query table
   const int32_t* queue;                // admitted query indices, oldest first
   int32_t n_queue;
   int32_t steps_per_launch;
   uint64_t n_lanes;                    // a multiple of the block size: waves are always full
   int32_t* slot_q;                     // query index of each slot (-1: none)
   // lane state (structure of arrays, n_lanes each; boundary arrays SPX_MAX_BND * n_lanes)
   int32_t* active;
   double *t, *x, *y, *z, *m0, *m1, *m2, *amp, *wgt, *dt_prev, *src_old, *ds_prev, *dmax, *dt_adapt;
   double* pf;                          // momentum magnitude of the slot's last field evaluation (guiding centre)
   float* kc;                           // SPX_KC * n_lanes: compensation terms of the FP32-step kernels' FP64 sums
   double *bdelta, *bdelta_old;
   int32_t *bcross, *bstat;
   uint64_t* traj;
   uint32_t *nn, *nu;
   int32_t* attempt;
   int64_t* steps;
   int32_t* n_events;
   // start state (read when an event is binned)
   double *t0, *x0, *y0, *z0, *m00, *m01, *m02;
   // last event per distribution (SPX_MAX_DIST * n_lanes)
   int32_t* ev_bin;
   double* ev_w;
   // events queued by the step kernel for the service kernel: 3 classes x 9 fields x n_lanes
   double* evs;
   int32_t* ev_bnd;                     // 3 * n_lanes
   int32_t* ev_pending;                 // bit per class
   int32_t* ex_code;                    // discard reason of the last step kernel
};

// Compensated FP64 accumulation of the FP32-step kernels: the sums t, x, y, z and p (Parker) or
// t, x, y, z (guiding centre) each carry a Kahan compensation term, kept per slot across launches.
#define SPX_KC 5

// Host-side dispatch.  One engine launch is the service kernel followed by the step kernel on
// `stream`; when `mid` is not null it is recorded between the two, so the caller can time each
// kernel with events.  launch_f64 / launch_f32 pick the instantiation for (transport, background,
// flow, diffusion) of each precision; they live in separately compiled units.
hipError_t launch_f64(const spx_plan_desc& host_plan, uint64_t n_lanes, const Launch* dev_launch,
                      hipStream_t stream, hipEvent_t mid);
hipError_t launch_f32(const spx_plan_desc& host_plan, uint64_t n_lanes, const Launch* dev_launch,
                      hipStream_t stream, hipEvent_t mid);

// FP32 cell records of a Cartesian grid (spx_grid_desc::c32) from its device zone records (nz[0] x nz[1] x
// nz[2], x fastest, SPX_GRID_REC doubles each): for every zone, the 8 zones of the stencil whose lowest
// corner it is, in corner order i + 2j + 4k (corners beyond the last zone repeat it; no stencil reads them).
hipError_t grid_expand_cells(const double* zones, const int32_t nz[3], float* c32, hipStream_t st);

inline hipError_t launch_transport(const spx_plan_desc& host_plan, uint64_t n_lanes, const Launch* dev_launch,
                                   int precision, hipStream_t stream, hipEvent_t mid)
{
   return precision == SPX_PREC_MIXED ? launch_f32(host_plan, n_lanes, dev_launch, stream, mid)
                                      : launch_f64(host_plan, n_lanes, dev_launch, stream, mid);
}

}  // namespace spx
