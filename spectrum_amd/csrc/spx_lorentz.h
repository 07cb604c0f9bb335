// spx_lorentz.h -- full-orbit (Newton-Lorentz) transport: TrajectoryLorentz::Advance = RKAdvance with the
// Euler table (src/trajectory_lorentz.cc:44-83, src/trajectory_base.cc:545-579), forward or backward in
// time, on the engine of spx_engine.h.  The momentum is the Cartesian vector (px, py, pz).
//
// After SetStart the trajectory evaluates its fields with the mask E | B only (src/trajectory_lorentz.cc:36-37).
// BackgroundBase keeps its SpatialData between calls (src/background_base.cc:425-456) and the solar wind
// forms E = -U x B / c from whatever flow that structure holds when U is not in the mask
// (src/background_solarwind.cc:252): the flow of the last evaluation that included U, SetStart's at the
// starting point.  The kernel keeps that flow per slot and forms E from it, as the reference does.
#pragma once

#include <hip/hip_runtime.h>

#include "spx_engine.h"

namespace spx {

template <typename R>
struct LzFields {
   v3<R> E, B;
   R Bmag, dmax;
};

// GetFields with the mask E | B at a position: B, |B| (an ExFieldError below sp_tiny), E from the kept flow
// U0 (SolarWind) or the uniform field's E, and dmax.
template <typename R, int BG>
__device__ __forceinline__ bool lz_fields(const SPX_CP spx_bg_t<R>& bg, double t, double x, double y, double z, v3<R> U0, LzFields<R>& F)
{
   F.dmax = bg_dmax<R, BG>(bg, x, y, z, R(0));
   if constexpr (BG == SPX_BG_GRID) {
      // the server interpolates U and B and forms E from them; |B| is interpolated, dmax limited by the zone
      // (BackgroundServer, src/background_server.cc:80-110; ServerCartesianFront, src/server_cartesian.cc)
      v3<R> U, gU[3], gB[3], gm;
      double zl_min;
      if (!grid_eval_full<R, false>(bg.grid, x, y, z, U, F.B, F.Bmag, F.E, gU, gB, gm, zl_min)) return false;
      F.dmax = R(fmin((double)F.dmax, zl_min));
      return F.Bmag >= R(sp_tiny);
   }
   v3<R> U, B;
   R ds = R(0);
   bg_eval<R, BG>(bg, t, x, y, z, 2, U, B, ds);
   F.B = B;
   F.Bmag = norm(B);
   if (F.Bmag < R(sp_tiny)) return false;
   if (BG == SPX_BG_UNIFORM) F.E = v3of(bg.Eu);                    // src/background_uniform.cc:58-77
   else if (BG == SPX_BG_SHOCK)                                     // src/background_shock.cc:91-111: E0 / E1 by side
      F.E = (dot(rel(bg, x, y, z), v3of(bg.n_shock)) - bg.v_shock * R(t) > R(0)) ? v3of(bg.E0) : v3of(bg.E1);
   else F.E = (-cross(U0, B)) / R(c_code);                          // src/background_solarwind.cc:252
   return true;
}

// Initial momentum of the full orbit (src/initial_momentum.cc:300-325, :381-413 with TRAJ_LORENTZ): the
// magnitude of the configured class in a direction drawn isotropically in the global frame, mu = 2u - 1
// then phi = 2 pi u'; or the fixed vector.
__device__ __forceinline__ void lz_mom_sample(PlanPtr P, const spx_rng_desc& rng, uint64_t first, uint64_t traj,
                                              uint32_t& nu, bool& ex, double m[3])
{
   double p;
   if (P->im_kind == SPX_IM_FIXED) { m[0] = P->im_mom[0]; m[1] = P->im_mom[1]; m[2] = P->im_mom[2]; return; }
   if (P->im_kind == SPX_IM_THICKSHELL) {
      const double u = next_uniform(rng, first, traj, nu, ex);
      p = P->im_log ? pow(10.0, P->im_p1 + (P->im_p2 - P->im_p1) * u) : P->im_p1 + (P->im_p2 - P->im_p1) * u;
   }
   else p = P->im_p0;
   const double mu = -1.0 + 2.0 * next_uniform(rng, first, traj, nu, ex);
   const double st = sqrt(1.0 - mu * mu);
   const double phi = (2.0 * M_PI) * next_uniform(rng, first, traj, nu, ex);
   m[0] = p * st * cos(phi);
   m[1] = p * st * sin(phi);
   m[2] = p * mu;
}

// TrajectoryBase::SetStart (src/trajectory_base.cc:867-929) for the full orbit: two momentum samples
// around the first field evaluation (all fields there, so the flow U0 at the start is kept).
template <typename R, int BG>
__device__ int lz_set_start(PlanPtr P, QueryPtr Q, LaunchPtr L, const spx_rng_desc& rng, uint64_t slot,
                            uint64_t traj, uint32_t& nu)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const uint64_t first = Q->first;
   const double t = P->it_t;
   bool ex = false;
   double x, y, z;
   pos_sample(P, rng, first, traj, nu, ex, x, y, z);
   double m1[3], m[3];
   lz_mom_sample(P, rng, first, traj, nu, ex, m1);
   v3<R> U0 = mk<R>(R(0), R(0), R(0)), B0;
   R ds = R(0);
   if (BG != SPX_BG_GRID) bg_eval<R, BG>(bg, t, x, y, z, 3, U0, B0, ds);   // the server's E needs no kept flow
   LzFields<R> F;
   if (!lz_fields<R, BG>(bg, t, x, y, z, U0, F)) return 1;
   lz_mom_sample(P, rng, first, traj, nu, ex, m);
   if (ex) return 5;
   const uint64_t n = L->n_lanes;
   L->t[slot] = t; L->x[slot] = x; L->y[slot] = y; L->z[slot] = z; L->m0[slot] = m[0]; L->m1[slot] = m[1]; L->m2[slot] = m[2];
   L->amp[slot] = 0.0; L->wgt[slot] = 0.0;
   // the kept flow U0 (three otherwise unused per-slot doubles of this transport)
   L->dt_prev[slot] = (double)U0.x; L->src_old[slot] = (double)U0.y; L->ds_prev[slot] = (double)U0.z;
   L->dmax[slot] = (double)F.dmax;
   L->pf[slot] = 0.0;
   L->steps[slot] = 0; L->n_events[slot] = 0;   // the draw counters continue across attempts
   L->dt_adapt[slot] = sp_large * (double)F.dmax / c_code;
   L->t0[slot] = t; L->x0[slot] = x; L->y0[slot] = y; L->z0[slot] = z; L->m00[slot] = m[0]; L->m01[slot] = m[1]; L->m02[slot] = m[2];
   for (int d = 0; d < SPX_MAX_DIST; d++) L->ev_bin[(uint64_t)d * n + slot] = -1;
   const double dmax = (double)F.dmax;
   const double pmag = norm(mk<double>(m[0], m[1], m[2]));
   for (int b = 0; b < P->n_bnd; b++) {
      const SPX_CP spx_bnd_desc& Bd = P->bnd[b];
      const double scale = (Bd.cls == SPX_BC_TIME) ? fabs(dmax / c_code) : (Bd.cls == SPX_BC_SPACE ? fabs(dmax) : fabs(pmag));
      const int left = Bd.max_crossings;
      const double d = bnd_eval(Bd, left, true, t, x, y, z, m[0], m[1], m[2], nullptr);
      L->bcross[b * n + slot] = left;
      L->bdelta[b * n + slot] = d;
      L->bdelta_old[b * n + slot] = (d == 0.0) ? -sp_tiny * scale : d;
      L->bstat[b * n + slot] = (Bd.status0 & ~(SPX_BND_TERMINAL | SPX_BND_CROSSED)) | (left == 1 ? SPX_BND_TERMINAL : 0);
   }
   return 0;
}

template <typename R, int BG>
__global__ __launch_bounds__(BLOCK) void lz_service(const Launch* __restrict__ Lg)
{
   const LaunchPtr L = (LaunchPtr)Lg;
   service_slot(L, [&](PlanPtr P, QueryPtr Q, const spx_rng_desc& rng, uint64_t slot, uint64_t traj, uint32_t& nu) {
      return lz_set_start<R, BG>(P, Q, L, rng, slot, traj, nu);
   });
}

// ---- step kernel -----------------------------------------------------------------------------
template <typename R>
struct LzHot {
   double t, x, y, z, px, py, pz;
   LzFields<R> F;
   v3<R> U0;
   double bd[SPX_MAX_BND], bdo[SPX_MAX_BND];
   int bst[SPX_MAX_BND];
   int32_t steps;
   typedef R real;
   __device__ __forceinline__ double m0() const { return px; }
   __device__ __forceinline__ double m1() const { return py; }
   __device__ __forceinline__ double m2() const { return pz; }
   // TrajectoryBase::ReverseMomentum (src/trajectory_base.cc:139-142)
   __device__ __forceinline__ void reflect_mom() { px = -px; py = -py; pz = -pz; }
};

// One step: Slopes (src/trajectory_lorentz.cc:66-70) with Particle::Vel in Cartesian coordinates
// (common/physics.hh:330-336) and LorentzForce (:409-412), PhysicalStep (:52-57: 100 steps per gyration,
// at most 0.1 dmax per step), TimeBoundaryProximityCheck, the Euler RKStep, HandleBoundaries and the
// end-of-step fields; no random numbers.
template <typename R, int BG, int FWD>
__device__ __forceinline__ int lz_step(PlanPtr P, LaunchPtr L, uint64_t slot, LzHot<R>& H, ColdLDS& C, int tid, int& bflags)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const v3<R> mom = mk<R>(R(H.px), R(H.py), R(H.pz));
   const R mm = norm(mom);
   const R a = VelP(mm) / mm;
   const v3<R> vel = mk<R>(mom.x * a, mom.y * a, mom.z * a);
   const R vmag = norm(vel);
   const R Omega = fmax(CyclotronFrequency(vmag, H.F.Bmag), R(sp_tiny));
   R dtp = R(2.0 * M_PI) / Omega / R(100);
   dtp = fmin(dtp, R(0.1) * H.F.dmax / vmag);
   double dt = fmin((double)dtp, C.dt_adapt[tid]);
   dt = time_proximity<FWD>(P, H, dt, (double)H.F.dmax);
   const v3<R> force = R(sp_charge) * (H.F.E + cross(vel, H.F.B) / R(c_code));
   const R dtv = R(dt) * R(1);
   if (FWD) {
      H.t += dt;
      H.x += (double)(dtv * vel.x); H.y += (double)(dtv * vel.y); H.z += (double)(dtv * vel.z);
      H.px += (double)(dtv * force.x); H.py += (double)(dtv * force.y); H.pz += (double)(dtv * force.z);
   }
   else {
      H.t -= dt;
      H.x -= (double)(dtv * vel.x); H.y -= (double)(dtv * vel.y); H.z -= (double)(dtv * vel.z);
      H.px -= (double)(dtv * force.x); H.py -= (double)(dtv * force.y); H.pz -= (double)(dtv * force.z);
   }
   bflags = handle_boundaries<true>(P, L, slot, H, C, tid);
   if (!(bflags & 1)) {
      if (!lz_fields<R, BG>(bg, H.t, H.x, H.y, H.z, H.U0, H.F)) return 1;
   }
   H.steps++;
   if (H.steps > P->max_steps) return 4;
   if (dt < sp_tiny * (double)H.F.dmax / c_code) return 2;
   if (!__builtin_isnormal(dt)) return 3;
   return 0;
}

template <typename R, int BG, int FWD>
__global__ __launch_bounds__(BLOCK) void lz_step_kernel(const Launch* __restrict__ Lg)
{
   const LaunchPtr Lp = (LaunchPtr)Lg;
   __shared__ ColdLDS C;
   const int tid = threadIdx.x;
   const uint64_t slot = (uint64_t)blockIdx.x * BLOCK + tid;
   const int q = __builtin_amdgcn_readfirstlane(Lp->slot_q[slot]);
   const bool run = q >= 0 && Lp->active[slot] == SLOT_RUN;
   if (__ballot(run) == 0) return;
   const QueryPtr Q = query_of(Lp, q);
   const PlanPtr P = plan_of(Q);
   LzHot<R> H;
   if (run) {
      const uint64_t n = Lp->n_lanes;
      H.t = Lp->t[slot]; H.x = Lp->x[slot]; H.y = Lp->y[slot]; H.z = Lp->z[slot];
      H.px = Lp->m0[slot]; H.py = Lp->m1[slot]; H.pz = Lp->m2[slot];
      H.U0 = mk<R>(R(Lp->dt_prev[slot]), R(Lp->src_old[slot]), R(Lp->ds_prev[slot]));
      H.steps = (int32_t)Lp->steps[slot];
      C.amp[tid] = Lp->amp[slot]; C.wgt[tid] = Lp->wgt[slot]; C.dt_adapt[tid] = Lp->dt_adapt[slot];
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         H.bd[b] = Lp->bdelta[b * n + slot]; H.bdo[b] = Lp->bdelta_old[b * n + slot];
         C.cross[b][tid] = Lp->bcross[b * n + slot]; H.bst[b] = Lp->bstat[b * n + slot];
      }
      // the fields are a function of (t, x) and the kept flow: recomputed bit for bit
      (void)lz_fields<R, BG>(PlanView<R>::bg(P), H.t, H.x, H.y, H.z, H.U0, H.F);
   }
   int st = run ? SLOT_RUN : SLOT_IDLE;
   int exc = 0;
   unsigned long long steps_here = 0;
   for (int it = 0; it < Lp->steps_per_launch; it++) {
      if (__ballot(st == SLOT_RUN) == 0) break;
      if (st != SLOT_RUN) continue;
      PlanPtr Pi = P;
      asm volatile("" : "+s"(Pi));
      LaunchPtr Li = Lp;
      asm volatile("" : "+s"(Li));
      int bflags = 0;
      const int e = lz_step<R, BG, FWD>(Pi, Li, slot, H, C, tid, bflags);
      steps_here++;
      if (e != 0) { exc = e; st = SLOT_DISCARD; }
      else if (bflags & 1) st = SLOT_DONE;
      else if (bflags & 2) st = SLOT_EVENT;
   }
   LaunchPtr Ls = Lp;
   asm volatile("" : "+s"(Ls));
   unsigned long long s = steps_here;
   for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
   if (lane_id() == 0 && s) atomicAdd(&Q->ctl->steps_executed, s);
   if (!run) return;
   const uint64_t n = Ls->n_lanes;
   Ls->active[slot] = st;
   Ls->ex_code[slot] = exc;
   Ls->t[slot] = H.t; Ls->x[slot] = H.x; Ls->y[slot] = H.y; Ls->z[slot] = H.z;
   Ls->m0[slot] = H.px; Ls->m1[slot] = H.py; Ls->m2[slot] = H.pz;
   Ls->steps[slot] = H.steps;
   Ls->dmax[slot] = (double)H.F.dmax;
   Ls->amp[slot] = C.amp[tid]; Ls->wgt[slot] = C.wgt[tid];
#pragma unroll
   for (int b = 0; b < SPX_MAX_BND; b++) {
      Ls->bdelta[b * n + slot] = H.bd[b]; Ls->bdelta_old[b * n + slot] = H.bdo[b];
      Ls->bcross[b * n + slot] = C.cross[b][tid]; Ls->bstat[b * n + slot] = H.bst[b];
   }
}

template <typename R, int BG, int FWD>
static hipError_t launch_lz_bg(uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const unsigned grid = (unsigned)((n_lanes + BLOCK - 1) / BLOCK);
   hipLaunchKernelGGL((lz_service<R, BG>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   if (mid) {
      hipError_t e = hipEventRecord(mid, st);
      if (e != hipSuccess) return e;
   }
   hipLaunchKernelGGL((lz_step_kernel<R, BG, FWD>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   return hipGetLastError();
}

template <typename R>
hipError_t dispatch_lorentz(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const bool fwd = hp.time_flow == 0;
   switch (hp.bg.kind) {
   case SPX_BG_SOLARWIND:
      return fwd ? launch_lz_bg<R, SPX_BG_SOLARWIND, 1>(n_lanes, d_launch, st, mid) : launch_lz_bg<R, SPX_BG_SOLARWIND, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_UNIFORM:
      return fwd ? launch_lz_bg<R, SPX_BG_UNIFORM, 1>(n_lanes, d_launch, st, mid) : launch_lz_bg<R, SPX_BG_UNIFORM, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_SHOCK:
      return fwd ? launch_lz_bg<R, SPX_BG_SHOCK, 1>(n_lanes, d_launch, st, mid) : launch_lz_bg<R, SPX_BG_SHOCK, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_GRID:
      return fwd ? launch_lz_bg<R, SPX_BG_GRID, 1>(n_lanes, d_launch, st, mid) : launch_lz_bg<R, SPX_BG_GRID, 0>(n_lanes, d_launch, st, mid);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace spx
