// spx_focused.h -- focused transport: TrajectoryFocused::Advance = RKAdvance with the Euler table
// (src/trajectory_focused.cc:31-158, src/trajectory_base.cc:545-579) with the magnetic drifts
// (TRAJ_FOCUSED_USE_B_DRIFTS) and the pitch-cosine slope (PPERP_METHOD 1) the reference builds with
// (src/trajectory_focused.hh:17-20), forward or backward in time, on the engine of spx_engine.h.  The
// momentum is (p, mu, 0); the velocity (p (v(p) / p), mu, 0) (Particle::Vel with the SphericalRMP metric,
// common/physics.hh:330-336, common/coordinates.hh:48,379-382).  No random numbers after the start.
//
// The fields carry the mask ALL | gradB | gradU | dUdt (src/trajectory_focused.cc:36-37).  The solar wind
// differentiates numerically (BackgroundBase::NumericalDerivatives, src/background_base.cc:228-312) in the
// field-aligned frame at r_g = min(r_L(p), dmax), p the momentum magnitude of the evaluation; it is static,
// so its time derivative is exactly zero, as is every derivative of the uniform background and of the
// discontinuous shock (BackgroundShock::EvaluateBackgroundDerivatives, src/background_shock.cc:116-128:
// U and B jump at the shock and are constant on either side).  The smooth
// shock's derivatives are analytic (BackgroundSmoothShock::EvaluateBackgroundDerivatives,
// src/background_smooth_shock.cc:159-187), dU/dt included when the shock moves; its dmax reads the shock
// coordinate of the previous evaluation (:192-196), which each slot carries.
#pragma once

#include <hip/hip_runtime.h>

#include "spx_engine.h"

namespace spx {

// What the slopes read of SpatialData at one point, contracted as far as the step allows.
template <typename R>
struct FcFields {
   v3<R> U, bhat, gradBmag, BgB;   // flow, b, grad |B|, (B . grad) B = B^T gradB
   R Bmag, dmax;
   R BcB;                          // B . curl B
   R divU, bbgU;                   // div U, bb : grad U
   R bcdU2;                        // 2 b . (dU/dt + (U . grad) U)
   R divbhat;                      // (div B - grad|B| . b) / |B| (common/spatial_data.hh:304-307)
};

// The contractions the slopes read, from grad U (rows gU), grad B (rows gB), dU/dt and F.gradBmag.
template <typename R>
__device__ __forceinline__ void fc_contract(FcFields<R>& F, v3<R> U, v3<R> B, const v3<R> (&gU)[3], const v3<R> (&gB)[3], v3<R> dUdt)
{
   const v3<R> b = F.bhat;
   F.divU = (gU[0].x + gU[1].y) + gU[2].z;
   F.bbgU = ((R(0) + dot(b.x * b, gU[0])) + dot(b.y * b, gU[1])) + dot(b.z * b, gU[2]);
   const v3<R> UgU = (U.x * gU[0] + U.y * gU[1]) + U.z * gU[2];
   F.bcdU2 = dot(R(2) * b, dUdt + UgU);
   const v3<R> curlB = mk<R>(gB[1].z - gB[2].y, gB[2].x - gB[0].z, gB[0].y - gB[1].x);
   F.BcB = dot(B, curlB);
   F.BgB = (B.x * gB[0] + B.y * gB[1]) + B.z * gB[2];
   const R divB = (gB[0].x + gB[1].y) + gB[2].z;
   F.divbhat = (divB - dot(F.gradBmag, b)) / F.Bmag;
}

// GetFields with the focused mask at a position for momentum magnitude pm; ds_prev is the smooth shock's
// hidden coordinate (read for dmax, then set to this evaluation's).  Returns false on
// ExFieldError (|B| below sp_tiny).
template <typename R, int BG>
__device__ __forceinline__ bool fc_fields(const SPX_CP spx_bg_t<R>& bg, double t, double x, double y, double z, R pm, R& ds_prev, FcFields<R>& F)
{
   F.dmax = bg_dmax<R, BG>(bg, x, y, z, ds_prev);
   if constexpr (BG == SPX_BG_GRID) {
      // the server's interpolated U, B, |B| and its stencil gradients; a static grid has dU/dt = 0
      v3<R> U, B, E, gU[3], gB[3];
      double zl_min;
      if (!grid_eval_full<R, true>(bg.grid, x, y, z, U, B, F.Bmag, E, gU, gB, F.gradBmag, zl_min)) return false;
      F.dmax = R(fmin((double)F.dmax, zl_min));
      if (F.Bmag < R(sp_tiny)) return false;
      F.U = U;
      F.bhat = B / F.Bmag;
      fc_contract(F, U, B, gU, gB, mk<R>(R(0), R(0), R(0)));
      return true;
   }
   v3<R> U, B;
   R ds = ds_prev;
   bg_eval<R, BG>(bg, t, x, y, z, 3, U, B, ds);
   if (BG == SPX_BG_SMOOTHSHOCK) ds_prev = ds;
   F.U = U;
   F.Bmag = norm(B);
   if (F.Bmag < R(sp_tiny)) return false;
   F.bhat = B / F.Bmag;
   const R z0 = R(0);
   if (BG == SPX_BG_SMOOTHSHOCK) {
      // grad X = n (X_up - X_down) f and dX/dt = -v_sh (X_up - X_down) f with f = tanh-transition' / width
      const R f = shock_transition_deriv(ds) / bg.width_shock;
      const v3<R> n = v3of(bg.n_shock);
      const v3<R> du = v3of(bg.u0) - v3of(bg.u1), db = v3of(bg.B0) - v3of(bg.B1);
      v3<R> gU[3], gB[3];
      gU[0] = f * (n.x * du); gU[1] = f * (n.y * du); gU[2] = f * (n.z * du);
      gB[0] = f * (n.x * db); gB[1] = f * (n.y * db); gB[2] = f * (n.z * db);
      const v3<R> b = F.bhat;
      F.gradBmag = mk<R>(dot(gB[0], b), dot(gB[1], b), dot(gB[2], b));
      const v3<R> dUdt = (shock_transition_deriv(ds) * bg.v_shock / bg.width_shock) * (v3of(bg.u1) - v3of(bg.u0));
      fc_contract(F, U, B, gU, gB, dUdt);
      return true;
   }
   if (BG != SPX_BG_SOLARWIND) {
      F.gradBmag = mk<R>(z0, z0, z0); F.BgB = F.gradBmag;
      F.BcB = z0; F.divU = z0; F.bbgU = z0; F.bcdU2 = z0; F.divbhat = z0;
      return true;
   }
   const R r_g = fmin(LarmorRadius(pm, F.Bmag), F.dmax);
   v3<R> fa[3];
   fa[2] = F.bhat;
   fa[0] = second_unit(F.bhat);
   fa[1] = cross(fa[2], fa[0]);
   // grad X = transpose(fa) * G with G's row k the central difference along fa[k], accumulated over k
   v3<R> gU[3], gB[3], gbm;
   v3<R> dc = mk<R>(z0, z0, z0);
   if constexpr (is_f32<R>()) dc = rel(bg, x, y, z);   // FP32 stencil points relative to the origin
#pragma unroll 1
   for (int k = 0; k < 3; k++) {
      const v3<R> e = k == 0 ? fa[0] : (k == 1 ? fa[1] : fa[2]);
      R dr = r_g;
      v3<R> Uf, Bf, Ub, Bb;
      R dsd;
      if constexpr (is_f32<R>()) {
         const v3<R> pf = dc + dr * e;
         bg_eval_rel<R, BG>(bg, t, pf, 3, Uf, Bf, dsd);
         dr *= R(2);
         bg_eval_rel<R, BG>(bg, t, pf - dr * e, 3, Ub, Bb, dsd);
      }
      else {
         const double xf = x + (double)(dr * e.x), yf = y + (double)(dr * e.y), zf = z + (double)(dr * e.z);
         bg_eval<R, BG>(bg, t, xf, yf, zf, 3, Uf, Bf, dsd);
         dr *= R(2);
         bg_eval<R, BG>(bg, t, xf - (double)(dr * e.x), yf - (double)(dr * e.y), zf - (double)(dr * e.z), 3, Ub, Bb, dsd);
      }
      const v3<R> GU = (Uf - Ub) / dr;
      const v3<R> GB = (Bf - Bb) / dr;
      const R gm = (norm(Bf) - norm(Bb)) / dr;
      const v3<R> ek = mk<R>(e.x, e.y, e.z);
      if (k == 0) {
         gU[0] = ek.x * GU; gU[1] = ek.y * GU; gU[2] = ek.z * GU;
         gB[0] = ek.x * GB; gB[1] = ek.y * GB; gB[2] = ek.z * GB;
         gbm = gm * ek;
      }
      else {
         gU[0] = gU[0] + ek.x * GU; gU[1] = gU[1] + ek.y * GU; gU[2] = gU[2] + ek.z * GU;
         gB[0] = gB[0] + ek.x * GB; gB[1] = gB[1] + ek.y * GB; gB[2] = gB[2] + ek.z * GB;
         gbm = gbm + gm * ek;
      }
   }
   F.gradBmag = gbm;
   // the static wind's dU/dt is exactly zero
   fc_contract(F, U, B, gU, gB, mk<R>(z0, z0, z0));
   return true;
}

// Initial momentum of the focused transport (src/initial_momentum.cc:203-220, :300-315, :390-402 with
// TRAJ_FOCUSED): (p, mu, 0) with mu = 2u - 1 drawn after the magnitude; the ring keeps its mu0 and draws
// nothing; a fixed vector is taken as it is.
__device__ __forceinline__ void fc_mom_sample(PlanPtr P, const spx_rng_desc& rng, uint64_t first, uint64_t traj,
                                              uint32_t& nu, bool& ex, double& p, double& mu)
{
   switch (P->im_kind) {
   case SPX_IM_FIXED: p = P->im_mom[0]; mu = P->im_mom[1]; return;
   case SPX_IM_RING: p = P->im_p0; mu = P->im_mu0; return;
   case SPX_IM_SHELL: p = P->im_p0; break;
   case SPX_IM_THICKSHELL: {
      const double u = next_uniform(rng, first, traj, nu, ex);
      p = P->im_log ? pow(10.0, P->im_p1 + (P->im_p2 - P->im_p1) * u) : P->im_p1 + (P->im_p2 - P->im_p1) * u;
      break;
   }
   default: {
      const int k = (int)(traj % (uint64_t)P->im_n);
      const double lT1 = log10(P->im_T1), lT2 = log10(P->im_T2);
      const double Tk = (P->im_n > 1) ? pow(10.0, lT1 + (lT2 - lT1) * k / (double)(P->im_n - 1)) : P->im_T1;
      p = MomT(Tk);
   }
   }
   mu = -1.0 + 2.0 * next_uniform(rng, first, traj, nu, ex);
}

// TrajectoryBase::SetStart (src/trajectory_base.cc:867-929) then TrajectoryFocused::SetStart (:31-41): two
// momentum samples around the first field evaluation, whose derivatives use the first sample's p.
template <typename R, int BG>
__device__ int fc_set_start(PlanPtr P, QueryPtr Q, LaunchPtr L, const spx_rng_desc& rng, uint64_t slot,
                            uint64_t traj, uint32_t& nu)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const uint64_t first = Q->first;
   const double t = P->it_t;
   bool ex = false;
   double x, y, z;
   pos_sample(P, rng, first, traj, nu, ex, x, y, z);
   double p1, mu1, p, mu;
   fc_mom_sample(P, rng, first, traj, nu, ex, p1, mu1);
   FcFields<R> F;
   // the smooth shock's hidden coordinate starts at the starting point's
   R ds = R(0);
   if (BG == SPX_BG_SMOOTHSHOCK) ds = (dot(rel(bg, x, y, z), v3of(bg.n_shock)) - bg.v_shock * R(t)) / bg.width_shock;
   const R ds_dm = ds;
   if (!fc_fields<R, BG>(bg, t, x, y, z, R(p1), ds, F)) return 1;
   fc_mom_sample(P, rng, first, traj, nu, ex, p, mu);
   if (ex) return 5;
   const uint64_t n = L->n_lanes;
   L->t[slot] = t; L->x[slot] = x; L->y[slot] = y; L->z[slot] = z; L->m0[slot] = p; L->m1[slot] = mu; L->m2[slot] = 0.0;
   L->amp[slot] = 0.0; L->wgt[slot] = 0.0; L->dt_prev[slot] = 0.0; L->src_old[slot] = 0.0;
   L->ds_prev[slot] = (double)ds_dm;   // the coordinate the current fields' dmax read
   L->dmax[slot] = (double)F.dmax;
   L->pf[slot] = p1;
   L->steps[slot] = 0; L->n_events[slot] = 0;   // the draw counters continue across attempts
   L->dt_adapt[slot] = sp_large * (double)F.dmax / c_code;
   L->t0[slot] = t; L->x0[slot] = x; L->y0[slot] = y; L->z0[slot] = z; L->m00[slot] = p; L->m01[slot] = mu; L->m02[slot] = 0.0;
   for (int d = 0; d < SPX_MAX_DIST; d++) L->ev_bin[(uint64_t)d * n + slot] = -1;
   const double dmax = (double)F.dmax;
   const double pmag = norm(mk<double>(p, mu, 0.0));   // TrajectoryBase::ResetAllBoundaries scales by _mom.Norm()
   for (int b = 0; b < P->n_bnd; b++) {
      const SPX_CP spx_bnd_desc& Bd = P->bnd[b];
      const double scale = (Bd.cls == SPX_BC_TIME) ? fabs(dmax / c_code) : (Bd.cls == SPX_BC_SPACE ? fabs(dmax) : fabs(pmag));
      const int left = Bd.max_crossings;
      const double d = bnd_eval(Bd, left, false, t, x, y, z, p, mu, 0.0, nullptr);
      L->bcross[b * n + slot] = left;
      L->bdelta[b * n + slot] = d;
      L->bdelta_old[b * n + slot] = (d == 0.0) ? -sp_tiny * scale : d;
      L->bstat[b * n + slot] = (Bd.status0 & ~(SPX_BND_TERMINAL | SPX_BND_CROSSED)) | (left == 1 ? SPX_BND_TERMINAL : 0);
   }
   return 0;
}

template <typename R, int BG>
__global__ __launch_bounds__(BLOCK) void fc_service(const Launch* __restrict__ Lg)
{
   const LaunchPtr L = (LaunchPtr)Lg;
   service_slot(L, [&](PlanPtr P, QueryPtr Q, const spx_rng_desc& rng, uint64_t slot, uint64_t traj, uint32_t& nu) {
      return fc_set_start<R, BG>(P, Q, L, rng, slot, traj, nu);
   });
}

// ---- step kernel -----------------------------------------------------------------------------
template <typename R>
struct FcHot {
   double t, x, y, z, p, mu;
   double pf;                                   // p of the last field evaluation
   R ds, ds_dm;                                 // smooth shock: coordinate of the last evaluation, and the one its dmax read
   FcFields<R> F;
   double bd[SPX_MAX_BND], bdo[SPX_MAX_BND];
   int bst[SPX_MAX_BND];
   int32_t steps;
   typedef R real;
   __device__ __forceinline__ double m0() const { return p; }
   __device__ __forceinline__ double m1() const { return mu; }
   __device__ __forceinline__ double m2() const { return 0.0; }
   // TrajectoryFocused::ReverseMomentum (src/trajectory_focused.hh:126-129)
   __device__ __forceinline__ void reflect_mom() { mu = -mu; }
};

// One step: DriftCoeff (src/trajectory_focused.cc:55-79), Slopes (:98-125), PhysicalStep (:87-91),
// TimeBoundaryProximityCheck, the Euler RKStep, HandleBoundaries, then the end-of-step fields and
// MomentumCorrection's clamp of mu to [-1, 1] (:142-152).
template <typename R, int BG, int FWD>
__device__ __forceinline__ int fc_step(PlanPtr P, LaunchPtr L, uint64_t slot, FcHot<R>& H, ColdLDS& C, int tid, int& bflags)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const FcFields<R>& F = H.F;
   const R p = R(H.p), mu = R(H.mu);
   const R v = p * (VelP(p) / p);
   const R ct2 = mu * mu, st2 = R(1) - ct2;
   v3<R> drift = ((R(0.5) * st2) * cross(F.bhat, F.gradBmag)) / F.Bmag;
   drift = drift + (F.BcB * ((R(0.5) * st2) * F.bhat) + ct2 * cross(F.bhat, F.BgB)) / (F.Bmag * F.Bmag);
   drift = (LarmorRadius(p, F.Bmag) * v) * drift;
   drift = drift + (F.U + (v * mu) * F.bhat);
   const R bcdU = F.bcdU2 / v;
   const R slope_p = R(0.5) * p * ((R(3) * st2 - R(2)) * F.bbgU - st2 * F.divU - mu * bcdU);
   const R slope_mu = R(0.5) * st2 * (v * F.divbhat - bcdU + mu * (F.divU - R(3) * F.bbgU));
   const R dtp = R(0.5) * F.dmax / (norm(drift) + R(0.5) * v);
   double dt = fmin((double)dtp, C.dt_adapt[tid]);
   dt = time_proximity<FWD>(P, H, dt, (double)F.dmax);
   const R dtv = R(dt) * R(1);
   if (FWD) {
      H.t += dt;
      H.x += (double)(dtv * drift.x); H.y += (double)(dtv * drift.y); H.z += (double)(dtv * drift.z);
      H.p += (double)(dtv * slope_p); H.mu += (double)(dtv * slope_mu);
   }
   else {
      H.t -= dt;
      H.x -= (double)(dtv * drift.x); H.y -= (double)(dtv * drift.y); H.z -= (double)(dtv * drift.z);
      H.p -= (double)(dtv * slope_p); H.mu -= (double)(dtv * slope_mu);
   }
   bflags = handle_boundaries<false>(P, L, slot, H, C, tid);
   if (!(bflags & 1)) {
      H.ds_dm = H.ds;
      if (!fc_fields<R, BG>(bg, H.t, H.x, H.y, H.z, R(H.p), H.ds, H.F)) return 1;
      H.pf = H.p;
      if (H.mu > 1.0) H.mu = 1.0;
      else if (H.mu < -1.0) H.mu = -1.0;
   }
   H.steps++;
   if (H.steps > P->max_steps) return 4;
   if (dt < sp_tiny * (double)H.F.dmax / c_code) return 2;
   if (!__builtin_isnormal(dt)) return 3;
   return 0;
}

template <typename R, int BG, int FWD>
__global__ __launch_bounds__(BLOCK) void fc_step_kernel(const Launch* __restrict__ Lg)
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
   FcHot<R> H;
   if (run) {
      const uint64_t n = Lp->n_lanes;
      H.t = Lp->t[slot]; H.x = Lp->x[slot]; H.y = Lp->y[slot]; H.z = Lp->z[slot];
      H.p = Lp->m0[slot]; H.mu = Lp->m1[slot];
      H.pf = Lp->pf[slot];
      H.steps = (int32_t)Lp->steps[slot];
      C.amp[tid] = Lp->amp[slot]; C.wgt[tid] = Lp->wgt[slot]; C.dt_adapt[tid] = Lp->dt_adapt[slot];
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         H.bd[b] = Lp->bdelta[b * n + slot]; H.bdo[b] = Lp->bdelta_old[b * n + slot];
         C.cross[b][tid] = Lp->bcross[b * n + slot]; H.bst[b] = Lp->bstat[b * n + slot];
      }
      // the fields are a function of (t, x, p_f, the dmax coordinate): recomputed bit for bit
      H.ds_dm = R(Lp->ds_prev[slot]);
      H.ds = H.ds_dm;
      (void)fc_fields<R, BG>(PlanView<R>::bg(P), H.t, H.x, H.y, H.z, R(H.pf), H.ds, H.F);
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
      const int e = fc_step<R, BG, FWD>(Pi, Li, slot, H, C, tid, bflags);
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
   Ls->m0[slot] = H.p; Ls->m1[slot] = H.mu; Ls->m2[slot] = 0.0;
   Ls->pf[slot] = H.pf;
   Ls->ds_prev[slot] = (double)H.ds_dm;
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
static hipError_t launch_fc_bg(uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const unsigned grid = (unsigned)((n_lanes + BLOCK - 1) / BLOCK);
   hipLaunchKernelGGL((fc_service<R, BG>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   if (mid) {
      hipError_t e = hipEventRecord(mid, st);
      if (e != hipSuccess) return e;
   }
   hipLaunchKernelGGL((fc_step_kernel<R, BG, FWD>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   return hipGetLastError();
}

template <typename R>
hipError_t dispatch_focused(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const bool fwd = hp.time_flow == 0;
   switch (hp.bg.kind) {
   case SPX_BG_SOLARWIND:
      return fwd ? launch_fc_bg<R, SPX_BG_SOLARWIND, 1>(n_lanes, d_launch, st, mid) : launch_fc_bg<R, SPX_BG_SOLARWIND, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_UNIFORM:
      return fwd ? launch_fc_bg<R, SPX_BG_UNIFORM, 1>(n_lanes, d_launch, st, mid) : launch_fc_bg<R, SPX_BG_UNIFORM, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_SMOOTHSHOCK:
      return fwd ? launch_fc_bg<R, SPX_BG_SMOOTHSHOCK, 1>(n_lanes, d_launch, st, mid) : launch_fc_bg<R, SPX_BG_SMOOTHSHOCK, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_SHOCK:
      return fwd ? launch_fc_bg<R, SPX_BG_SHOCK, 1>(n_lanes, d_launch, st, mid) : launch_fc_bg<R, SPX_BG_SHOCK, 0>(n_lanes, d_launch, st, mid);
   case SPX_BG_GRID:
      return fwd ? launch_fc_bg<R, SPX_BG_GRID, 1>(n_lanes, d_launch, st, mid) : launch_fc_bg<R, SPX_BG_GRID, 0>(n_lanes, d_launch, st, mid);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace spx
