// spx_parker.h -- Parker and Parker-with-source transport kernels for gfx950 (MI355X).
//
// One lane per trajectory slot; the slots of a run are resident in HBM as structure-of-arrays state.
// Each engine launch is a pair of kernels, both captured together in a HIP graph for steady loops:
//   * the service kernel refills idle slots from the device work counter (one atomic per wave),
//     runs SetStart, bins pending distribution events, retries discarded attempts and writes the
//     per-trajectory records;
//   * the step kernel advances every running slot by up to `steps_per_launch` steps with the hot
//     state in registers; a slot that produces a distribution event or ends stops there and waits
//     for the next service pass.
// The step reproduces TrajectoryParker::Advance (src/trajectory_parker.cc:255-306) with the Euler
// table and TrajectoryBase::Integrate's checks (src/trajectory_base.cc:946-995).
#pragma once

#include <hip/hip_runtime.h>

#include "spx_engine.h"

namespace spx {

// ---------------------------------------------------------------------------------------------
// Parker transport
// ---------------------------------------------------------------------------------------------
template <typename R>
struct PkFields {
   v3<R> U, bhat;
   R Bmag, divU, dmax;
   v3<R> fa0, fa1;   // SolarWind: the field-aligned frame of bhat (second unit vector, bhat x it), reused by the step
};

// CommonFields with mask U|B|gradU|gradB for the Parker transport (BackgroundBase::GetFields,
// src/background_base.cc:425-456, with NumericalDerivatives :228-312 for SolarWind and the analytic
// derivatives of src/background_smooth_shock.cc:159-187).  The transport consumes only div(U) of the
// gradients, so only U is evaluated at the derivative points.  Returns false on ExFieldError.
template <typename R, int BG>
__device__ __forceinline__ bool parker_fields(const SPX_CP spx_bg_t<R>& bg, double t, double x, double y, double z, R p, R& ds_prev, PkFields<R>& F)
{
   F.dmax = bg_dmax<R, BG>(bg, x, y, z, ds_prev);
   if constexpr (BG == SPX_BG_GRID) {
      // BackgroundServer: interpolated U, B and |B|, the stencil's grad U, dmax limited by the zone size
      GridFields<R> G;
      if (!grid_eval<R, true>(bg.grid, x, y, z, G)) return false;
      F.dmax = fmin(F.dmax, R(G.zl_min));
      F.U = G.U;
      F.Bmag = G.Bmag;
      if (F.Bmag < R(sp_tiny)) return false;
      if constexpr (is_f32<R>()) F.bhat = (R(1) / F.Bmag) * G.B;
      else F.bhat = G.B / F.Bmag;
      F.divU = G.divU;
      return true;
   }
   v3<R> B;
   R ds = ds_prev;
   bg_eval<R, BG>(bg, t, x, y, z, 3, F.U, B, ds);
   if (BG == SPX_BG_SMOOTHSHOCK) ds_prev = ds;
   if constexpr (is_f32<R>()) {
      const R b2 = norm2(B), ib = frsq(b2);
      F.Bmag = b2 * ib;
      if (!(F.Bmag >= R(sp_tiny))) return false;   // a zero field gives 0 * inf = NaN here: invalid, as |B| < tiny
      F.bhat = ib * B;
   }
   else {
      F.Bmag = norm(B);
      if (F.Bmag < R(sp_tiny)) return false;
      F.bhat = B / F.Bmag;
   }
   if (BG == SPX_BG_SOLARWIND) {
      const R r_g = fmin(LarmorRadius(p, F.Bmag), F.dmax);
      v3<R> fa[3];
      fa[2] = F.bhat;
      fa[0] = second_unit(F.bhat);
      fa[1] = cross(fa[2], fa[0]);
      F.fa0 = fa[0];
      F.fa1 = fa[1];
      // trace of transpose(fa) * G (common/matrix.hh:509-519, :186-189), summed per diagonal element
      R d0 = R(0), d1 = R(0), d2 = R(0);
      v3<R> dc = mk<R>(R(0), R(0), R(0));
      if constexpr (is_f32<R>()) {
         dc = rel(bg, x, y, z);   // FP32 stencil points relative to the origin; the radial wind in any basis
         F.divU = (R(0.5) / r_g) * sw_divu_stencil<R>(bg, dc, r_g, fa);
         return true;
      }
      // the frame vectors rotate through e (moves instead of a select of each component per point)
      v3<R> e = fa[0], e_next = fa[1];
#pragma unroll 1
      for (int k = 0; k < 3; k++) {
         R dr = r_g;
         v3<R> Uf, Ub, Bd;
         R dsd;
         if constexpr (is_f32<R>()) {
            const v3<R> pf = dc + dr * e;
            bg_eval_rel<R, BG>(bg, t, pf, 1, Uf, Bd, dsd);
            dr *= R(2);
            bg_eval_rel<R, BG>(bg, t, pf - dr * e, 1, Ub, Bd, dsd);
         }
         else {
            const double xf = x + (double)(dr * e.x), yf = y + (double)(dr * e.y), zf = z + (double)(dr * e.z);
            bg_eval<R, BG>(bg, t, xf, yf, zf, 1, Uf, Bd, dsd);
            dr *= R(2);
            bg_eval<R, BG>(bg, t, xf - (double)(dr * e.x), yf - (double)(dr * e.y), zf - (double)(dr * e.z), 1, Ub, Bd, dsd);
         }
         const v3<R> G = is_f32<R>() ? (R(0.5) / r_g) * (Uf - Ub) : (Uf - Ub) / dr;   // FP32: one reciprocal per point
         if (k == 0) { d0 = e.x * G.x; d1 = e.y * G.y; d2 = e.z * G.z; }
         else { d0 += e.x * G.x; d1 += e.y * G.y; d2 += e.z * G.z; }
         e = e_next;
         e_next = fa[2];
      }
      F.divU = d0 + d1 + d2;
   }
   else if (BG == SPX_BG_SMOOTHSHOCK) {
      const R f = shock_transition_deriv(ds) / bg.width_shock;
      const v3<R> du = v3of(bg.u0) - v3of(bg.u1);
      F.divU = (bg.n_shock[0] * du.x) * f + (bg.n_shock[1] * du.y) * f + (bg.n_shock[2] * du.z) * f;
   }
   else F.divU = R(0);
   return true;
}

// LISM_ind of the Strauss / Potgieter models at x (src/diffusion_other.cc:891-892, :973-974): SolarWind sets
// region[0] = +1 inside the heliopause and -1 beyond it, region[1] = -1 (src/background_solarwind.cc:144-146);
// the other analytic backgrounds set positive regions only (src/background_uniform.cc:68,
// src/background_shock.cc:99-107, src/background_smooth_shock.cc:151).  r is the norm in the rotated frame;
// the diagonal +-1 bases leave it unchanged.
template <typename R, int BG>
__device__ __forceinline__ bool lism_outer(const SPX_CP spx_bg_t<R>& bg, const SPX_CP spx_df_t<R>& df, double x, double y, double z)
{
   if (BG != SPX_BG_SOLARWIND || df.region_idx < 0) return false;
   if (df.region_idx == 1) return true;
   const v3<R> d = rel(bg, x, y, z);
   R r;
   if (bg.ep_diag != 0) r = norm(d);
   else {
      const v3<R> ep[3] = {v3of(bg.eprime[0]), v3of(bg.eprime[1]), v3of(bg.eprime[2])};
      r = norm(to_basis(d, ep));
   }
   return !(r < R(hp_rad_sw));
}

// DiffusionBase::GetComponent for components 0 and 1 at a position.  DF is the diffusion class when it
// is known at compile time (0 = any, dispatched at run time); km2 is the LISM factor of the region models,
// or with km3 the two perpendicular factors of the empirical SOQLT model (lam_perp, lam_perp kap_rat_red).
template <typename R, int BG, int DF>
__device__ __forceinline__ void parker_kappa(const SPX_CP spx_bg_t<R>& bg, const SPX_CP spx_df_t<R>& df, R km, R km2, R km3, R Bmag,
                                             double x, double y, double z, v3<R> U, R& K0, R& K1)
{
   if (DF == SPX_DF_RIGBMAG) {   // src/diffusion_other.cc:804-809; pow(x, 0) == 1 exactly
      K1 = km;
      K1 = K1 * ((df.pow_law_B != R(0)) ? rpow(Bmag / df.B0, df.pow_law_B) : R(1));
      K0 = df.kap_rat * K1;
      return;
   }
   if (DF == SPX_DF_FULL) { K0 = df.Dperp; K1 = df.Dpara; return; }   // :538-543
   if (df.kind == SPX_DF_STRAUSS || df.kind == SPX_DF_POTGIETER) {       // :886-906, :968-988
      const bool outer = lism_outer<R, BG>(bg, df, x, y, z);
      K1 = (outer ? km2 : km) * ((outer ? Bmag : df.B0) / Bmag);
      K0 = (outer ? df.kap_rat_out : df.kap_rat_in) * K1;
      return;
   }
   if (df.kind == SPX_DF_SOQLT) {                                         // :1053-1072
      // the reduced perpendicular path inside r_lim_red where region[Bmix_idx] < 0 (lism_outer reads it)
      const bool reduced = lism_outer<R, BG>(bg, df, x, y, z) && norm(mk<double>(x, y, z)) < (double)df.r_lim_red;
      const R bf = df.B0 / Bmag;
      K1 = km * bf;
      K1 = K1 / df.sc_den;
      K0 = (reduced ? km3 : km2) * bf;
      K0 = K0 / df.sc_den;
      return;
   }
   R r = R(0), Umag = R(0);
   if (df.kind == SPX_DF_KINRAD) r = R(norm(mk<double>(x, y, z)));
   if (df.kind == SPX_DF_FLOWMOM || (df.kind == SPX_DF_KINRAD && df.stream_dep_idx != 0)) Umag = norm(U);
   kappa01(df, km, Bmag, r, Umag, K0, K1);
}

// Momentum sample of the initial condition (src/initial_momentum.cc:283-410), Parker layout (p, 0, 0).
__device__ __forceinline__ double parker_mom_sample(PlanPtr P, const spx_rng_desc& rng, uint64_t first,
                                                    uint64_t traj, uint32_t& nu, bool& ex)
{
   switch (P->im_kind) {
   case SPX_IM_FIXED: return P->traj_type == SPX_TRAJ_FIELDLINE ? P->im_mom[2] : P->im_mom[0];   // (0, 0, p_par) on a field line
   case SPX_IM_SHELL: case SPX_IM_RING: return P->im_p0;
   case SPX_IM_THICKSHELL: {
      const double u = next_uniform(rng, first, traj, nu, ex);
      return P->im_log ? pow(10.0, P->im_p1 + (P->im_p2 - P->im_p1) * u) : P->im_p1 + (P->im_p2 - P->im_p1) * u;
   }
   default: {   // shell grid: log-spaced kinetic energies, grid point traj mod n
      const int k = (int)(traj % (uint64_t)P->im_n);
      const double lT1 = log10(P->im_T1), lT2 = log10(P->im_T2);
      const double Tk = (P->im_n > 1) ? pow(10.0, lT1 + (lT2 - lT1) * k / (double)(P->im_n - 1)) : P->im_T1;
      return MomT(Tk);
   }
   }
}

// ---- service kernel: refill, SetStart, events, discards, records ------------------------------
// TrajectoryBase::SetStart (src/trajectory_base.cc:867-929) + TrajectoryParker::SetStart (:58-66)
// + TrajectoryParkerSource::SetStart (src/trajectory_parker_source.cc:43-52).  Writes the slot's full
// state to HBM.  Returns an exception code.
template <typename R, int BG, int TT>
__device__ int parker_set_start(PlanPtr P, QueryPtr Q, LaunchPtr L, const spx_rng_desc& rng, uint64_t slot,
                                uint64_t traj, uint32_t& nu)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const uint64_t first = Q->first;
   const double t = P->it_t;
   bool ex = false;
   double x, y, z;
   pos_sample(P, rng, first, traj, nu, ex, x, y, z);
   // first sample along an arbitrary axis: only its magnitude enters the first field evaluation
   const double p1 = parker_mom_sample(P, rng, first, traj, nu, ex);
   R ds = R(0);
   if (BG == SPX_BG_SMOOTHSHOCK) {
      const v3<R> d = rel(bg, x, y, z);
      ds = (dot(d, v3of(bg.n_shock)) - bg.v_shock * R(t)) / bg.width_shock;
   }
   PkFields<R> F;
   if (!parker_fields<R, BG>(bg, t, x, y, z, R(p1), ds, F)) return 1;
   const double p = parker_mom_sample(P, rng, first, traj, nu, ex);
   if (ex) return 5;
   const uint64_t n = L->n_lanes;
   // Parker keeps (p, 0, 0), the field line (0, 0, p_par) (src/initial_momentum.cc:289-292)
   const double pa = (TT == 2) ? 0.0 : p, pc = (TT == 2) ? p : 0.0;
   L->t[slot] = t; L->x[slot] = x; L->y[slot] = y; L->z[slot] = z; L->m0[slot] = pa; L->m1[slot] = 0.0; L->m2[slot] = pc;
   L->amp[slot] = 0.0; L->wgt[slot] = 0.0; L->dt_prev[slot] = 0.0;
   L->ds_prev[slot] = (double)ds; L->dmax[slot] = (double)F.dmax;
   L->pf[slot] = p1;   // the first step uses these fields, evaluated with the first sample's magnitude
   L->steps[slot] = 0; L->n_events[slot] = 0;   // the draw counters continue across attempts
   // the adaptive step starts large so that dt is the physical step (src/trajectory_base.cc:894-895)
   L->dt_adapt[slot] = sp_large * (double)F.dmax / c_code;
   L->t0[slot] = t; L->x0[slot] = x; L->y0[slot] = y; L->z0[slot] = z; L->m00[slot] = pa; L->m01[slot] = 0.0; L->m02[slot] = pc;
   for (int d = 0; d < SPX_MAX_DIST; d++) L->ev_bin[(uint64_t)d * n + slot] = -1;
   // ResetAllBoundaries (src/trajectory_base.cc:67-81, src/boundary_base.cc:138-162)
   const double dmax = (double)F.dmax;
   for (int b = 0; b < P->n_bnd; b++) {
      const SPX_CP spx_bnd_desc& B = P->bnd[b];
      const double scale = (B.cls == SPX_BC_TIME) ? fabs(dmax / c_code) : (B.cls == SPX_BC_SPACE ? fabs(dmax) : fabs(p));
      const int left = B.max_crossings;
      const double d = bnd_eval(B, left, false, t, x, y, z, pa, 0.0, pc, nullptr);
      L->bcross[b * n + slot] = left;
      L->bdelta[b * n + slot] = d;
      L->bdelta_old[b * n + slot] = (d == 0.0) ? -sp_tiny * scale : d;
      L->bstat[b * n + slot] = (B.status0 & ~(SPX_BND_TERMINAL | SPX_BND_CROSSED)) | (left == 1 ? SPX_BND_TERMINAL : 0);
   }
   // SourceMomentumInjection primed by GetSource(..., 1.0)
   L->src_old[slot] = (TT == 1) ? p - P->src_p_inj : 0.0;
   return 0;
}

template <typename R, int BG, int TT>
__global__ __launch_bounds__(BLOCK) void parker_service(const Launch* __restrict__ Lg)
{
   const LaunchPtr L = (LaunchPtr)Lg;
   service_slot(L, [&](PlanPtr P, QueryPtr Q, const spx_rng_desc& rng, uint64_t slot, uint64_t traj, uint32_t& nu) {
      return parker_set_start<R, BG, TT>(P, Q, L, rng, slot, traj, nu);
   });
}

// ---- step kernel -----------------------------------------------------------------------------
template <typename R>
struct PkHot {
   double t, x, y, z, p;
   PkFields<R> F;
   R ds_prev;                                   // BackgroundSmoothShock's hidden coordinate (smooth shock only)
   double bd[SPX_MAX_BND], bdo[SPX_MAX_BND];    // boundary _delta and _delta_old
   int bst[SPX_MAX_BND];                        // boundary status bits
   NormalStream<R> rs;
   int32_t steps;
   bool fieldline;                              // TrajectoryFieldline: p is the signed p_par of (0, 0, p_par)
   typedef R real;
   __device__ __forceinline__ double m0() const { return fieldline ? 0.0 : p; }
   __device__ __forceinline__ double m1() const { return 0.0; }
   __device__ __forceinline__ double m2() const { return fieldline ? p : 0.0; }
   // TrajectoryParker::ReverseMomentum does nothing (src/trajectory_parker.hh:129-131); the field line
   // keeps TrajectoryBase's, which reverses the momentum (src/trajectory_base.cc:139-142)
   __device__ __forceinline__ void reflect_mom() { if (fieldline) p = -p; }
};

// One step of TrajectoryFieldline::Advance = RKAdvance with the Euler table (src/trajectory_fieldline.cc:41-77,
// src/trajectory_base.cc:545-579): the centre follows the magnetic field, slope = v_par bhat with v_par the
// third component of Particle::Vel(0, 0, p_par) (common/physics.hh:330-336), PhysicalStep
// dt = 0.5 dmax / |v_par|; the momentum is constant and no random numbers are drawn.
template <typename R, int BG, int FWD>
__device__ __forceinline__ int fieldline_step(PlanPtr P, LaunchPtr L, uint64_t slot, PkHot<R>& H, ColdLDS& C, int tid, int& bflags)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const R p = R(H.p), pm = fabs(p);
   const R v = p * (VelP(pm) / pm);
   const R dtp = R(0.5) * H.F.dmax / fabs(v);
   double dt = fmin((double)dtp, C.dt_adapt[tid]);
   dt = time_proximity<FWD>(P, H, dt, (double)H.F.dmax);
   const v3<R> slope_pos = v * H.F.bhat;
   const R dtv = R(dt) * R(1);
   if constexpr (is_f32<R>()) {   // compensated FP64 sums of the FP32 increments
      const R sg = FWD ? R(1) : R(-1);
      kahan_add(H.t, C.kc[0][tid], FWD ? dt : -dt);
      kahan_add(H.x, C.kc[1][tid], (double)(sg * (dtv * slope_pos.x)));
      kahan_add(H.y, C.kc[2][tid], (double)(sg * (dtv * slope_pos.y)));
      kahan_add(H.z, C.kc[3][tid], (double)(sg * (dtv * slope_pos.z)));
   }
   else if (FWD) {
      H.t += dt;
      H.x += (double)(dtv * slope_pos.x); H.y += (double)(dtv * slope_pos.y); H.z += (double)(dtv * slope_pos.z);
   }
   else {
      H.t -= dt;
      H.x -= (double)(dtv * slope_pos.x); H.y -= (double)(dtv * slope_pos.y); H.z -= (double)(dtv * slope_pos.z);
   }
   bflags = handle_boundaries<false>(P, L, slot, H, C, tid);
   if (!(bflags & 1)) {
      if (!parker_fields<R, BG>(bg, H.t, H.x, H.y, H.z, R(fabs(H.p)), H.ds_prev, H.F)) return 1;
   }
   H.steps++;
   if (H.steps > P->max_steps) return 4;
   if (dt < divc_d<R>(sp_tiny * (double)H.F.dmax, c_code)) return 2;
   if (!__builtin_isnormal(dt)) return 3;
   return 0;
}

// One Parker step: TrajectoryParker::Advance (src/trajectory_parker.cc:255-306) with the Euler table
// (common/rk_config.hh:72-80) and TrajectoryBase::Integrate's checks (src/trajectory_base.cc:964-992).
// Returns an exception code; `bflags` gets handle_boundaries' result.
template <typename R, int BG, int FWD, int TT, int DF>
__device__ __forceinline__ int parker_step(PlanPtr P, LaunchPtr L, const spx_rng_desc& rng,
                                           uint64_t first, uint64_t slot, PkHot<R>& H, ColdLDS& C, int tid, int& bflags)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const SPX_CP spx_df_t<R>& df = PlanView<R>::df(P);
   const R p = R(H.p);
   // DriftCoeff + DiffusionCoeff (src/trajectory_parker.cc:189-206, 83-153)
   R K0 = R(0), K1 = R(0);
   const R km = (DF == SPX_DF_FULL) ? R(0) : kappa_mom(df, p);
   // the LISM parameters of the Strauss / Potgieter models, when a region indicator can select them
   R km2 = R(0), km3 = R(0);
   if (DF == 0) {
      if ((df.kind == SPX_DF_STRAUSS || df.kind == SPX_DF_POTGIETER) && df.region_idx >= 0) km2 = kappa_mom_region(df, p, true);
      else if (df.kind == SPX_DF_SOQLT) { km2 = kappa_mom_soqlt(df, p, 0, false); km3 = kappa_mom_soqlt(df, p, 0, true); }
   }
   parker_kappa<R, BG, DF>(bg, df, km, km2, km3, H.F.Bmag, H.x, H.y, H.z, H.F.U, K0, K1);
   const R delta = fmin(LarmorRadius(p, H.F.Bmag), H.F.dmax);
   const R hdi = is_f32<R>() ? R(0.5) / delta : R(0);
   const double ydel = is_f32<R>() ? 0.0 : 1.0 / (double)delta;   // FP64: the stencil's nine quotients by delta share it
   v3<R> divK = mk<R>(R(0), R(0), R(0));
   const bool pos_dep = (DF == SPX_DF_FULL) ? false : (df.pos_dep != 0);
   if (BG == SPX_BG_GRID && (pos_dep || K0 != K1)) {
      // the same stencil on the gridded background: interpolated B and |B| at the derivative points
#pragma unroll 1
      for (int j = 0; j < 3; j++) {
         const double ex = (j == 0) ? 1.0 : 0.0, ey = (j == 1) ? 1.0 : 0.0, ez = (j == 2) ? 1.0 : 0.0;
         double xf = H.x + (double)delta * ex, yf = H.y + (double)delta * ey, zf = H.z + (double)delta * ez;
         // with fd_fields == 0 the fields at these points cannot change the result (the host has shown it):
         // only the positions are checked, and bhat = 0 meets a zero (kappa_par - kappa_perp)
         const bool loads = bg.grid.fd_fields != 0;
         GridFields<R> Gf, Gb;
         Gf.U = Gb.U = Gf.B = Gb.B = mk<R>(R(0), R(0), R(0));
         Gf.Bmag = Gb.Bmag = H.F.Bmag;
         if (loads ? !grid_eval<R, false>(bg.grid, xf, yf, zf, Gf) : !grid_inside<R>(bg.grid, xf, yf, zf)) return 1;
         if (loads && Gf.Bmag < R(sp_tiny)) return 1;
         const v3<R> bf = !loads ? Gf.B : (is_f32<R>() ? (R(1) / Gf.Bmag) * Gf.B : Gf.B / Gf.Bmag);
         R K0f = K0, K1f = K1;
         if (pos_dep) parker_kappa<R, BG, DF>(bg, df, km, km2, km3, Gf.Bmag, xf, yf, zf, Gf.U, K0f, K1f);
         xf -= 2.0 * (double)delta * ex;
         yf -= 2.0 * (double)delta * ey;
         zf -= 2.0 * (double)delta * ez;
         if (loads ? !grid_eval<R, false>(bg.grid, xf, yf, zf, Gb) : !grid_inside<R>(bg.grid, xf, yf, zf)) return 1;
         if (loads && Gb.Bmag < R(sp_tiny)) return 1;
         const v3<R> bb = !loads ? Gb.B : (is_f32<R>() ? (R(1) / Gb.Bmag) * Gb.B : Gb.B / Gb.Bmag);
         R K0b = K0, K1b = K1;
         if (pos_dep) parker_kappa<R, BG, DF>(bg, df, km, km2, km3, Gb.Bmag, xf, yf, zf, Gb.U, K0b, K1b);
         const R bfj = j == 0 ? bf.x : (j == 1 ? bf.y : bf.z);
         const R bbj = j == 0 ? bb.x : (j == 1 ? bb.y : bb.z);
         const R cf = (K1f - K0f) * bfj, cb = (K1b - K0b) * bbj;
         if constexpr (is_f32<R>()) {   // one reciprocal of delta for the whole stencil
            divK.x += hdi * ((K0f * (j == 0 ? R(1) : R(0)) + cf * bf.x) - (K0b * (j == 0 ? R(1) : R(0)) + cb * bb.x));
            divK.y += hdi * ((K0f * (j == 1 ? R(1) : R(0)) + cf * bf.y) - (K0b * (j == 1 ? R(1) : R(0)) + cb * bb.y));
            divK.z += hdi * ((K0f * (j == 2 ? R(1) : R(0)) + cf * bf.z) - (K0b * (j == 2 ? R(1) : R(0)) + cb * bb.z));
         }
         else {
            divK.x += div_rn(R(0.5) * ((K0f * (j == 0 ? R(1) : R(0)) + cf * bf.x) - (K0b * (j == 0 ? R(1) : R(0)) + cb * bb.x)), delta, ydel);
            divK.y += div_rn(R(0.5) * ((K0f * (j == 1 ? R(1) : R(0)) + cf * bf.y) - (K0b * (j == 1 ? R(1) : R(0)) + cb * bb.y)), delta, ydel);
            divK.z += div_rn(R(0.5) * ((K0f * (j == 2 ? R(1) : R(0)) + cf * bf.z) - (K0b * (j == 2 ? R(1) : R(0)) + cb * bb.z)), delta, ydel);
         }
      }
   }
   else if (BG == SPX_BG_GRID) {
      // divK vanishes identically, but the reference still evaluates the six derivative points, and one
      // outside the grid discards the attempt
#pragma unroll 1
      for (int j = 0; j < 3; j++) {
         const double ex = (j == 0) ? 1.0 : 0.0, ey = (j == 1) ? 1.0 : 0.0, ez = (j == 2) ? 1.0 : 0.0;
         double xf = H.x + (double)delta * ex, yf = H.y + (double)delta * ey, zf = H.z + (double)delta * ez;
         const bool loads = bg.grid.fd_fields != 0;
         GridFields<R> G;
         if (loads ? (!grid_eval<R, false>(bg.grid, xf, yf, zf, G) || G.Bmag < R(sp_tiny)) : !grid_inside<R>(bg.grid, xf, yf, zf)) return 1;
         xf -= 2.0 * (double)delta * ex;
         yf -= 2.0 * (double)delta * ey;
         zf -= 2.0 * (double)delta * ez;
         if (loads ? (!grid_eval<R, false>(bg.grid, xf, yf, zf, G) || G.Bmag < R(sp_tiny)) : !grid_inside<R>(bg.grid, xf, yf, zf)) return 1;
      }
   }
   else if (is_f32<R>() && BG == SPX_BG_SOLARWIND && !pos_dep && K0 != K1 && bg.ep_diag != 0 &&
            sw_stencil_clear<R>(bg, rel(bg, H.x, H.y, H.z), delta)) {
      // kappa is the same at every point: the K0 terms cancel, and (K1 - K0) b b is left
      if constexpr (is_f32<R>()) divK = ((K1 - K0) * hdi) * sw_bb_stencil<R>(bg, rel(bg, H.x, H.y, H.z), delta);
   }
   else if (pos_dep || K0 != K1) {
      const int want = (df.kind == SPX_DF_FLOWMOM || df.kind == SPX_DF_KINRAD) ? 3 : 2;
      // FP32 steps with a position-independent kappa evaluate the stencil relative to the origin
      v3<R> dc = mk<R>(R(0), R(0), R(0));
      if constexpr (is_f32<R>()) dc = rel(bg, H.x, H.y, H.z);
#pragma unroll 1
      for (int j = 0; j < 3; j++) {
         const double ex = (j == 0) ? 1.0 : 0.0, ey = (j == 1) ? 1.0 : 0.0, ez = (j == 2) ? 1.0 : 0.0;
         const v3<R> ej = mk<R>(R(ex), R(ey), R(ez));
         double xf = H.x + (double)delta * ex, yf = H.y + (double)delta * ey, zf = H.z + (double)delta * ez;
         v3<R> U2, B2;
         R dsd = H.ds_prev;
         if (is_f32<R>() && !pos_dep) bg_eval_rel<R, BG>(bg, H.t, dc + delta * ej, want, U2, B2, dsd);
         else bg_eval<R, BG>(bg, H.t, xf, yf, zf, want, U2, B2, dsd);
         R Bmf;
         v3<R> bf;
         if constexpr (is_f32<R>()) { const R b2 = norm2(B2), ib = frsq(b2); Bmf = b2 * ib; bf = ib * B2; }
         else { Bmf = norm(B2); bf = B2 / Bmf; }
         if (Bmf < R(sp_tiny)) return 1;
         R K0f = K0, K1f = K1;
         if (pos_dep) parker_kappa<R, BG, DF>(bg, df, km, km2, km3, Bmf, xf, yf, zf, U2, K0f, K1f);
         xf -= 2.0 * (double)delta * ex;
         yf -= 2.0 * (double)delta * ey;
         zf -= 2.0 * (double)delta * ez;
         if (is_f32<R>() && !pos_dep) bg_eval_rel<R, BG>(bg, H.t, dc + delta * ej - (R(2) * delta) * ej, want, U2, B2, dsd);
         else bg_eval<R, BG>(bg, H.t, xf, yf, zf, want, U2, B2, dsd);
         if (BG == SPX_BG_SMOOTHSHOCK) H.ds_prev = dsd;
         R Bmb;
         v3<R> bb;
         if constexpr (is_f32<R>()) { const R b2 = norm2(B2), ib = frsq(b2); Bmb = b2 * ib; bb = ib * B2; }
         else { Bmb = norm(B2); bb = B2 / Bmb; }
         if (Bmb < R(sp_tiny)) return 1;
         R K0b = K0, K1b = K1;
         if (pos_dep) parker_kappa<R, BG, DF>(bg, df, km, km2, km3, Bmb, xf, yf, zf, U2, K0b, K1b);
         const R bfj = j == 0 ? bf.x : (j == 1 ? bf.y : bf.z);
         const R bbj = j == 0 ? bb.x : (j == 1 ? bb.y : bb.z);
         const R cf = (K1f - K0f) * bfj, cb = (K1b - K0b) * bbj;
         if constexpr (is_f32<R>()) {   // one reciprocal of delta for the whole stencil
            divK.x += hdi * ((K0f * (j == 0 ? R(1) : R(0)) + cf * bf.x) - (K0b * (j == 0 ? R(1) : R(0)) + cb * bb.x));
            divK.y += hdi * ((K0f * (j == 1 ? R(1) : R(0)) + cf * bf.y) - (K0b * (j == 1 ? R(1) : R(0)) + cb * bb.y));
            divK.z += hdi * ((K0f * (j == 2 ? R(1) : R(0)) + cf * bf.z) - (K0b * (j == 2 ? R(1) : R(0)) + cb * bb.z));
         }
         else {
            divK.x += div_rn(R(0.5) * ((K0f * (j == 0 ? R(1) : R(0)) + cf * bf.x) - (K0b * (j == 0 ? R(1) : R(0)) + cb * bb.x)), delta, ydel);
            divK.y += div_rn(R(0.5) * ((K0f * (j == 1 ? R(1) : R(0)) + cf * bf.y) - (K0b * (j == 1 ? R(1) : R(0)) + cb * bb.y)), delta, ydel);
            divK.z += div_rn(R(0.5) * ((K0f * (j == 2 ? R(1) : R(0)) + cf * bf.z) - (K0b * (j == 2 ? R(1) : R(0)) + cb * bb.z)), delta, ydel);
         }
      }
   }
   else if (BG == SPX_BG_SMOOTHSHOCK) {
      // divK vanishes identically; the last derivative point still sets the shock's hidden coordinate
      const double zb = (H.z + (double)delta * 1.0) - 2.0 * (double)delta * 1.0;
      const v3<R> d = rel(bg, H.x + (double)delta * 0.0, H.y + (double)delta * 0.0, zb);
      H.ds_prev = (dot(d, v3of(bg.n_shock)) - bg.v_shock * R(H.t)) / bg.width_shock;
   }
   const v3<R> slope_pos = FWD ? H.F.U + divK : H.F.U - divK;
   const R slope_mom = divc<R>(-p * H.F.divU, R(3));
   double slope_amp = 0.0;
   if (TT == 1) {   // SourceMomentumInjection::EvaluateSource (src/source_other.cc:113-119) with the previous dt
      const double del = H.p - P->src_p_inj;
      slope_amp = (del * C.src_old[tid] < 0.0) ? P->src_rate / C.dt_prev[tid] : 0.0;
      C.src_old[tid] = del;
   }
   // PhysicalStep (src/trajectory_parker.cc:212-221)
   R dtp = R(0.5) * H.F.dmax / norm(slope_pos);
   dtp = fmin(dtp, R(0.5) * sqr(H.F.dmax) / fmax(K0, K1));
   dtp = fmin(dtp, R(0.5) * R(3) * R(0.01) / fabs(H.F.divU));
   double dt = fmin(P->dt_scale == 1.0 ? (double)dtp : (double)dtp * P->dt_scale, C.dt_adapt[tid]);   // 1: the reference's step
   dt = time_proximity<FWD>(P, H, dt, (double)H.F.dmax);
   // EulerDiffSlopes (src/trajectory_parker.cc:159-180); the solar wind's fields carry the frame of bhat
   v3<R> fa[3];
   fa[2] = H.F.bhat;
   if (BG == SPX_BG_SOLARWIND) { fa[0] = H.F.fa0; fa[1] = H.F.fa1; }
   else {
      fa[0] = second_unit(H.F.bhat);
      fa[1] = cross(fa[2], fa[0]);
   }
   const R sdt = rsqrt_(R(dt));
   R nrm[3];
   next_normals<R, 3>(rng, first, H.rs, nrm);
   const R dWx = sdt * nrm[0], dWy = sdt * nrm[1], dWz = sdt * nrm[2];
   if (H.rs.exhausted) return 5;
   const v3<R> drp = from_basis(mk<R>(rsqrt_(R(2) * K0) * dWx, rsqrt_(R(2) * K0) * dWy, rsqrt_(R(2) * K1) * dWz), fa);
   // RKStep, Euler (src/trajectory_base.cc:362-389), then the stochastic displacement
   const R dtv = R(dt) * R(1);
   if constexpr (is_f32<R>()) {
      // compensated FP64 sums; the drift and the stochastic displacement of a coordinate are added in FP32
      // first (their sum carries FP32 rounding either way)
      const R sg = FWD ? R(1) : R(-1);
      kahan_add(H.t, C.kc[0][tid], FWD ? dt : -dt);
      kahan_add(H.x, C.kc[1][tid], (double)(sg * (dtv * slope_pos.x) + drp.x));
      kahan_add(H.y, C.kc[2][tid], (double)(sg * (dtv * slope_pos.y) + drp.y));
      kahan_add(H.z, C.kc[3][tid], (double)(sg * (dtv * slope_pos.z) + drp.z));
      kahan_add(H.p, C.kc[4][tid], (double)(sg * (dtv * slope_mom)));
   }
   else {
      if (FWD) {
         H.t += dt;
         H.x += (double)(dtv * slope_pos.x); H.y += (double)(dtv * slope_pos.y); H.z += (double)(dtv * slope_pos.z);
         H.p += (double)(dtv * slope_mom);
      }
      else {
         H.t -= dt;
         H.x -= (double)(dtv * slope_pos.x); H.y -= (double)(dtv * slope_pos.y); H.z -= (double)(dtv * slope_pos.z);
         H.p -= (double)(dtv * slope_mom);
      }
      H.x += (double)drp.x; H.y += (double)drp.y; H.z += (double)drp.z;
   }
   if (TT == 1) {
      if (slope_amp != 0.0) C.amp[tid] += dt * 1.0 * slope_amp * exp(C.wgt[tid]);
      C.dt_prev[tid] = dt;
   }
   // HandleBoundaries and the end-of-step fields
   bflags = handle_boundaries<false>(P, L, slot, H, C, tid);
   if (!(bflags & 1)) {
      if (!parker_fields<R, BG>(bg, H.t, H.x, H.y, H.z, R(H.p), H.ds_prev, H.F)) return 1;
   }
   H.steps++;
   // Integrate's safety checks
   if (H.steps > P->max_steps) return 4;
   if (dt < divc_d<R>(sp_tiny * (double)H.F.dmax, c_code)) return 2;
   if (!__builtin_isnormal(dt)) return 3;
   return 0;
}

template <typename R, int BG, int FWD, int TT, int DF>
__device__ __forceinline__ void parker_step_body(const Launch* __restrict__ Lg)
{
   const LaunchPtr Lp = (LaunchPtr)Lg;
   __shared__ ColdLDS C;
   const int tid = threadIdx.x;
   const uint64_t slot = (uint64_t)blockIdx.x * BLOCK + tid;
   // the slots of a wave work for one query (the pool size is a multiple of the block size)
   const int q = __builtin_amdgcn_readfirstlane(Lp->slot_q[slot]);
   const bool run = q >= 0 && Lp->active[slot] == SLOT_RUN;
   if (__ballot(run) == 0) return;
   const QueryPtr Q = query_of(Lp, q);
   const PlanPtr P = plan_of(Q);
   const spx_rng_desc rng = load_rng(Q);
   const uint64_t first = Q->first;
   PkHot<R> H;
   if (run) {
      const uint64_t n = Lp->n_lanes;
      H.fieldline = (TT == 2);
      H.t = Lp->t[slot]; H.x = Lp->x[slot]; H.y = Lp->y[slot]; H.z = Lp->z[slot]; H.p = (TT == 2) ? Lp->m2[slot] : Lp->m0[slot];
      if constexpr (is_f32<R>()) {
#pragma unroll
         for (int k = 0; k < SPX_KC; k++) C.kc[k][tid] = Lp->kc[k * n + slot];
      }
      H.steps = (int32_t)Lp->steps[slot];
      H.ds_prev = R(Lp->ds_prev[slot]);
      H.rs.traj = Lp->traj[slot]; H.rs.nn = Lp->nn[slot]; H.rs.exhausted = false;
      prime_normals<R>(rng, H.rs);
      C.amp[tid] = Lp->amp[slot]; C.wgt[tid] = Lp->wgt[slot]; C.dt_adapt[tid] = Lp->dt_adapt[slot];
      C.dt_prev[tid] = Lp->dt_prev[slot]; C.src_old[tid] = Lp->src_old[slot];
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         H.bd[b] = Lp->bdelta[b * n + slot]; H.bdo[b] = Lp->bdelta_old[b * n + slot];
         C.cross[b][tid] = Lp->bcross[b * n + slot]; H.bst[b] = Lp->bstat[b * n + slot];
      }
      // The fields at the current point are a function of (t, x, p of their evaluation) and are recomputed
      // bit for bit: p is the trajectory's momentum after a step, the first momentum sample before the
      // first (src/trajectory_base.cc:874-879); dmax (which may depend on the previous evaluation) is restored.
      R ds_tmp = H.ds_prev;
      (void)parker_fields<R, BG>(PlanView<R>::bg(P), H.t, H.x, H.y, H.z, R(Lp->pf[slot]), ds_tmp, H.F);
      H.F.dmax = R(Lp->dmax[slot]);
   }
   int st = run ? SLOT_RUN : SLOT_IDLE;
   int exc = 0;
   unsigned long long steps_here = 0;
   for (int it = 0; it < Lp->steps_per_launch; it++) {
      if (__ballot(st == SLOT_RUN) == 0) break;
      if (st != SLOT_RUN) continue;
      // Re-read the plan and launch block each step instead of pinning their fields in registers: they are
      // uniform scalar loads, cheap to repeat, while keeping them live would exhaust the register file.
      PlanPtr Pi = P;
      asm volatile("" : "+s"(Pi));
      LaunchPtr Li = Lp;
      asm volatile("" : "+s"(Li));
      int bflags = 0;
      const int e = (TT == 2) ? fieldline_step<R, BG, FWD>(Pi, Li, slot, H, C, tid, bflags)
                              : parker_step<R, BG, FWD, TT, DF>(Pi, Li, rng, first, slot, H, C, tid, bflags);
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
   if constexpr (is_f32<R>()) {
#pragma unroll
      for (int k = 0; k < SPX_KC; k++) Ls->kc[k * n + slot] = C.kc[k][tid];
   }
   if (TT == 2) Ls->m2[slot] = H.p;
   else Ls->m0[slot] = H.p;
   Ls->steps[slot] = H.steps; Ls->nn[slot] = H.rs.nn;
   if (steps_here) Ls->pf[slot] = (TT == 2) ? fabs(H.p) : H.p;   // every step ends with fields at the new momentum
   Ls->ds_prev[slot] = (double)H.ds_prev; Ls->dmax[slot] = (double)H.F.dmax;
   Ls->amp[slot] = C.amp[tid]; Ls->wgt[slot] = C.wgt[tid];
   Ls->dt_prev[slot] = C.dt_prev[tid]; Ls->src_old[slot] = C.src_old[tid];
#pragma unroll
   for (int b = 0; b < SPX_MAX_BND; b++) {
      Ls->bdelta[b * n + slot] = H.bd[b]; Ls->bdelta_old[b * n + slot] = H.bdo[b];
      Ls->bcross[b * n + slot] = C.cross[b][tid]; Ls->bstat[b * n + slot] = H.bst[b];
   }
}

// Resident waves per SIMD asked of the compiler for a step kernel (1: its own register budget).  The FP64
// smooth-shock source kernel (diffusive shock acceleration) runs 7.7% faster at 3 waves (168 VGPRs, 40 B
// spilled) than at its natural 2 (177 VGPRs); the FP64 modulation kernel loses 3% at 3 (196 B spilled) and
// keeps its natural 2, as do the smooth-shock source kernels of the other diffusion classes (192-216 B would
// spill).  SPX_PARKER_WAVES, when defined, asks it of every instantiation.
template <typename R, int BG, int TT, int DF>
constexpr int pk_waves()
{
#ifdef SPX_PARKER_WAVES
   return SPX_PARKER_WAVES;
#else
   return (!is_f32<R>() && BG == SPX_BG_SMOOTHSHOCK && TT == 1 && DF == SPX_DF_FULL) ? 3 : 1;
#endif
}
template <typename R, int BG, int FWD, int TT, int DF>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(pk_waves<R, BG, TT, DF>())))
void parker_step_kernel(const Launch* __restrict__ Lg)
{
   parker_step_body<R, BG, FWD, TT, DF>(Lg);
}

// The gridded background's step kernel waits on gathers: it asks the compiler for room for
// SPX_GRID_WAVES resident waves per SIMD (0: the compiler's own register budget).
#ifndef SPX_GRID_WAVES
#define SPX_GRID_WAVES 0
#endif
#if SPX_GRID_WAVES > 0
#define SPX_GRID_ATTR __attribute__((amdgpu_waves_per_eu(SPX_GRID_WAVES)))
#else
#define SPX_GRID_ATTR
#endif
template <typename R, int FWD, int TT>
__global__ __launch_bounds__(BLOCK) SPX_GRID_ATTR void parker_grid_step_kernel(const Launch* __restrict__ Lg)
{
   parker_step_body<R, SPX_BG_GRID, FWD, TT, 0>(Lg);
}

// ---------------------------------------------------------------------------------------------
// Dispatch: one engine launch = service kernel + step kernel on the same stream
// ---------------------------------------------------------------------------------------------
template <typename R, int BG, int FWD, int TT, int DF>
static hipError_t launch_parker(uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const unsigned grid = (unsigned)((n_lanes + BLOCK - 1) / BLOCK);
   hipLaunchKernelGGL((parker_service<R, BG, TT>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   if (mid) {
      hipError_t e = hipEventRecord(mid, st);
      if (e != hipSuccess) return e;
   }
   if constexpr (BG == SPX_BG_GRID) hipLaunchKernelGGL((parker_grid_step_kernel<R, FWD, TT>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   else hipLaunchKernelGGL((parker_step_kernel<R, BG, FWD, TT, DF>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   return hipGetLastError();
}

template <typename R, int BG, int FWD, int TT>
static hipError_t dispatch_df(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   switch (hp.df.kind) {
   case SPX_DF_RIGBMAG: return launch_parker<R, BG, FWD, TT, SPX_DF_RIGBMAG>(n_lanes, d_launch, st, mid);
   case SPX_DF_FULL: return launch_parker<R, BG, FWD, TT, SPX_DF_FULL>(n_lanes, d_launch, st, mid);
   default: return launch_parker<R, BG, FWD, TT, 0>(n_lanes, d_launch, st, mid);
   }
}

template <typename R, int BG>
static hipError_t dispatch_flow(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const bool fwd = hp.time_flow == 0;
   if (hp.traj_type == 1) return fwd ? dispatch_df<R, BG, 1, 1>(hp, n_lanes, d_launch, st, mid) : dispatch_df<R, BG, 0, 1>(hp, n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_FIELDLINE)
      return fwd ? launch_parker<R, BG, 1, 2, 0>(n_lanes, d_launch, st, mid) : launch_parker<R, BG, 0, 2, 0>(n_lanes, d_launch, st, mid);
   return fwd ? dispatch_df<R, BG, 1, 0>(hp, n_lanes, d_launch, st, mid) : dispatch_df<R, BG, 0, 0>(hp, n_lanes, d_launch, st, mid);
}

// The gridded background runs the generic diffusion dispatch only (every class through kappa01).
template <typename R>
static hipError_t dispatch_grid(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const bool fwd = hp.time_flow == 0;
   constexpr int G = SPX_BG_GRID;
   if (hp.traj_type == 1) return fwd ? launch_parker<R, G, 1, 1, 0>(n_lanes, d_launch, st, mid) : launch_parker<R, G, 0, 1, 0>(n_lanes, d_launch, st, mid);
   if (hp.traj_type == SPX_TRAJ_FIELDLINE)
      return fwd ? launch_parker<R, G, 1, 2, 0>(n_lanes, d_launch, st, mid) : launch_parker<R, G, 0, 2, 0>(n_lanes, d_launch, st, mid);
   return fwd ? launch_parker<R, G, 1, 0, 0>(n_lanes, d_launch, st, mid) : launch_parker<R, G, 0, 0, 0>(n_lanes, d_launch, st, mid);
}

template <typename R>
hipError_t dispatch_parker(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   switch (hp.bg.kind) {
   case SPX_BG_GRID: return dispatch_grid<R>(hp, n_lanes, d_launch, st, mid);
   case SPX_BG_SOLARWIND: return dispatch_flow<R, SPX_BG_SOLARWIND>(hp, n_lanes, d_launch, st, mid);
   case SPX_BG_SMOOTHSHOCK: return dispatch_flow<R, SPX_BG_SMOOTHSHOCK>(hp, n_lanes, d_launch, st, mid);
   case SPX_BG_SHOCK: return dispatch_flow<R, SPX_BG_SHOCK>(hp, n_lanes, d_launch, st, mid);
   case SPX_BG_UNIFORM: return dispatch_flow<R, SPX_BG_UNIFORM>(hp, n_lanes, d_launch, st, mid);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace spx
