// spx_gc.h -- guiding-centre transport with perpendicular diffusion and pitch-angle scattering.
//
// TrajectoryGuidingDiffScatt::Advance (src/trajectory_guiding_diff_scatt.cc:76-174), forward in
// time, on the engine of spx_engine.h: one lane per trajectory slot, service kernel + step kernel
// per engine launch.  The momentum is kept in the reference's guiding-centre layout
// (p_perp, 0, p_par).  Static backgrounds only (SolarWind, Uniform, the Cartesian grid): their time derivatives
// vanish exactly, so the dB/dt terms of the modified fields and of dp_perp/dt are left out,
// which changes no result bit beyond the sign of a zero.
#pragma once

#include <hip/hip_runtime.h>

#include "spx_engine.h"

namespace spx {

// The SpatialData members the guiding-centre step reads, at one point.
template <typename R>
struct GcFields {
   v3<R> E, B, bhat, gradBmag, curlbhat;
   R Bmag, dmax;
   R rg;               // the solar wind's derivative step r_g (half of SpatialData::_dr)
   v3<R> dbhatdt;      // time-dependent backgrounds (the moving smooth shock): d bhat / dt
   R dBmagdt;          // and d|B| / dt = bhat . dB/dt
   v3<R> fa0, fa1;     // FP32 step on the solar wind: the field-aligned frame of bhat, reused by the step
};

// CommonFields with mask ALL|gradB|dBdt (src/trajectory_guiding.cc:51): BackgroundBase::GetFields
// (src/background_base.cc:425-456) with NumericalDerivatives (:228-312) for SolarWind in the
// field-aligned frame at r_g = min(r_L(pm), dmax).  gradB enters only through curl(bhat) and
// grad|B| (common/spatial_data.hh:304-341), which are accumulated here row by row in the order
// of transpose(fa) * G.  The smooth shock's derivatives are analytic (src/background_smooth_shock.cc:159-187)
// and its dmax reads the shock coordinate of the previous evaluation, ds_prev (set to this one's); the
// discontinuous shock has none (src/background_shock.cc:116-128).  Returns false on ExFieldError.
template <typename R, int BG>
__device__ __forceinline__ bool gc_fields(const SPX_CP spx_bg_t<R>& bg, double t, double x, double y, double z, R pm, R& ds_prev,
                                          GcFields<R>& F)
{
   F.dmax = bg_dmax<R, BG>(bg, x, y, z, ds_prev);
   F.dbhatdt = mk<R>(R(0), R(0), R(0));
   F.dBmagdt = R(0);
   if constexpr (BG == SPX_BG_GRID) {
      // BackgroundServer: interpolated fields and the stencil's derivatives, no finite differences
      v3<R> U;
      R gB[6];
      double zl_min;
      if (!grid_eval_gc<R>(bg.grid, x, y, z, U, F.B, F.Bmag, F.E, F.gradBmag, gB, zl_min)) return false;
      F.dmax = fmin(F.dmax, R(zl_min));
      F.rg = R(0);
      if (F.Bmag < R(sp_tiny)) return false;
      F.bhat = F.B / F.Bmag;
      const v3<R> curlB = mk<R>(gB[3] - gB[5], gB[4] - gB[1], gB[0] - gB[2]);
      F.curlbhat = (curlB - cross(F.gradBmag, F.bhat)) / F.Bmag;
      return true;
   }
   v3<R> U, B;
   R ds = ds_prev;
   bg_eval<R, BG>(bg, t, x, y, z, 3, U, B, ds);
   if (BG == SPX_BG_SMOOTHSHOCK) ds_prev = ds;
   F.B = B;
   F.Bmag = norm(B);
   if (F.Bmag < R(sp_tiny)) return false;
   F.bhat = B / F.Bmag;
   if (BG == SPX_BG_UNIFORM) F.E = v3of(bg.Eu);                    // src/background_uniform.cc:58-77
   else if (BG == SPX_BG_SHOCK)                                     // src/background_shock.cc:91-111: E0 / E1 by side
      F.E = (dot(rel(bg, x, y, z), v3of(bg.n_shock)) - bg.v_shock * R(t) > R(0)) ? v3of(bg.E0) : v3of(bg.E1);
   else F.E = (-cross(U, B)) / R(c_code);                          // src/background_solarwind.cc:250-253, smooth_shock.cc:151
   if constexpr (BG == SPX_BG_SMOOTHSHOCK) {
      // grad B = n (B0 - B1) f, grad|B| = grad B . bhat, dB/dt = f v_sh (B1 - B0) with f = transition' / width;
      // d|B|/dt = bhat . dB/dt (the class sets no d|B|/dt; this is its value), d bhat/dt = (dB/dt - d|B|/dt bhat) / |B|
      const R f = shock_transition_deriv(ds) / bg.width_shock;
      const v3<R> n = v3of(bg.n_shock), db = v3of(bg.B0) - v3of(bg.B1);
      const v3<R> g0 = f * (n.x * db), g1 = f * (n.y * db), g2 = f * (n.z * db);
      F.gradBmag = mk<R>(dot(g0, F.bhat), dot(g1, F.bhat), dot(g2, F.bhat));
      const v3<R> curlB = mk<R>(g1.z - g2.y, g2.x - g0.z, g0.y - g1.x);
      F.curlbhat = (curlB - cross(F.gradBmag, F.bhat)) / F.Bmag;
      const v3<R> dBdt = (shock_transition_deriv(ds) * bg.v_shock / bg.width_shock) * (v3of(bg.B1) - v3of(bg.B0));
      F.dBmagdt = dot(dBdt, F.bhat);
      F.dbhatdt = (dBdt - F.dBmagdt * F.bhat) / F.Bmag;
      F.rg = R(0);
      return true;
   }
   if (BG == SPX_BG_SOLARWIND) {
      const R r_g = fmin(LarmorRadius(pm, F.Bmag), F.dmax);
      F.rg = r_g;
      v3<R> fa[3];
      fa[2] = F.bhat;
      fa[0] = second_unit(F.bhat);
      fa[1] = cross(fa[2], fa[0]);
      if constexpr (is_f32<R>()) { F.fa0 = fa[0]; F.fa1 = fa[1]; }
      R g01 = R(0), g02 = R(0), g10 = R(0), g12 = R(0), g20 = R(0), g21 = R(0);
      v3<R> gbm = mk<R>(R(0), R(0), R(0));
      v3<R> dc = mk<R>(R(0), R(0), R(0));
      if constexpr (is_f32<R>()) dc = rel(bg, x, y, z);   // FP32 stencil points relative to the origin
      // the frame vectors rotate through e (moves instead of a select of each component per point)
      v3<R> e = fa[0], e_next = fa[1];
#pragma unroll 1
      for (int k = 0; k < 3; k++) {
         R dr = r_g;
         v3<R> Ud, Bf, Bb;
         R dsd;
         if constexpr (is_f32<R>()) {
            const v3<R> pf = dc + dr * e;
            bg_eval_rel<R, BG>(bg, t, pf, 2, Ud, Bf, dsd);
            dr *= R(2);
            bg_eval_rel<R, BG>(bg, t, pf - dr * e, 2, Ud, Bb, dsd);
         }
         else {
            const double xf = x + (double)(dr * e.x), yf = y + (double)(dr * e.y), zf = z + (double)(dr * e.z);
            bg_eval<R, BG>(bg, t, xf, yf, zf, 2, Ud, Bf, dsd);
            dr *= R(2);
            bg_eval<R, BG>(bg, t, xf - (double)(dr * e.x), yf - (double)(dr * e.y), zf - (double)(dr * e.z), 2, Ud, Bb, dsd);
         }
         const R idr = is_f32<R>() ? R(0.5) / r_g : R(0);   // FP32: one reciprocal per point
         const v3<R> gB = is_f32<R>() ? idr * (Bf - Bb) : (Bf - Bb) / dr;
         const R gm = is_f32<R>() ? (norm(Bf) - norm(Bb)) * idr : (norm(Bf) - norm(Bb)) / dr;
         if (k == 0) {
            g01 = e.x * gB.y; g02 = e.x * gB.z; g10 = e.y * gB.x; g12 = e.y * gB.z; g20 = e.z * gB.x; g21 = e.z * gB.y;
            gbm = mk<R>(e.x * gm, e.y * gm, e.z * gm);
         }
         else {
            g01 += e.x * gB.y; g02 += e.x * gB.z; g10 += e.y * gB.x; g12 += e.y * gB.z; g20 += e.z * gB.x; g21 += e.z * gB.y;
            gbm.x += e.x * gm; gbm.y += e.y * gm; gbm.z += e.z * gm;
         }
         e = e_next;
         e_next = fa[2];
      }
      F.gradBmag = gbm;
      const v3<R> curlB = mk<R>(g12 - g21, g20 - g02, g01 - g10);
      F.curlbhat = (curlB - cross(gbm, F.bhat)) / F.Bmag;
   }
   else {
      F.gradBmag = mk<R>(R(0), R(0), R(0));
      F.curlbhat = mk<R>(R(0), R(0), R(0));
      F.rg = R(0);
   }
   return true;
}

// Gauss's 2F1(1, 1; c; z), 0 <= z < 1: the series (Abramowitz & Stegun 15.1.1) below z = 1/2, above it the
// transformation to 1 - z (A&S 15.3.6 with a = b = 1):
//   2F1(1, 1; c; z) = (c-1)/(c-2) 2F1(1, 1; 3-c; 1-z) + Gamma(c) Gamma(2-c) (1-z)^(c-2) z^(1-c),
// with hyp_g = Gamma(c) Gamma(2-c) and hyp_r = (c-1)/(c-2) from the host.
template <typename R>
__device__ __forceinline__ R hyp_series_11(R c, R z)
{
   const R tol = is_f32<R>() ? R(1e-8) : R(1e-17);
   R sum = R(1), term = R(1);
#pragma unroll 1
   for (int n = 0; n < 400; n++) {
      term *= (R(n) + R(1)) * z / (c + R(n));
      sum += term;
      if (fabs(term) < tol * fabs(sum)) break;
   }
   return sum;
}

template <typename R>
__device__ __forceinline__ R hyp2f1_11(const SPX_CP spx_df_t<R>& df, R z)
{
   if (z < R(0.5)) return hyp_series_11(df.hyp_c, z);
   const R w = R(1) - z;
   return df.hyp_r * hyp_series_11(R(3) - df.hyp_c, w) + df.hyp_g * rpow(w, df.hyp_c - R(2)) * rpow(z, R(1) - df.hyp_c);
}

// DiffusionWNLTConstant::EvaluateDiffusion (src/diffusion_other.cc:215-245) over DiffusionQLTConstant's
// component 2 (:145-149), with DiffusionWNLTRampVLISM's k_min and amplitudes at the position (:301-328):
// kappa_perp into K0 and, when comp2, D_mumu into K2.
template <typename R>
__device__ __noinline__ void wnlt_k(const SPX_CP spx_df_t<R>& df, bool comp2, double x, double y, double z, R pmag, R mu, R Bmag, R& K0, R& K2)
{
   R k_min = df.k_min, A2A = df.A2A, A2T = df.A2T, A2L = df.A2L;
   if (df.kind == SPX_DF_WNLT_RAMP) {
      const double r0 = norm(mk<double>(x, y, z));
      const R z0 = R(sqrt(0.5 * r0 * (r0 + z)));
      if (z0 > df.z_sheath) { k_min = df.k_min_ref; A2A = df.A2A_ref; A2T = df.A2T_ref; A2L = df.A2L_ref; }
      else {
         k_min = R(2.0 * M_PI) / (df.l_max_HP + df.dl_max * (z0 - df.z_nose) / df.dz);
         const R k_ratio = rpow(df.k_min_ref / k_min, df.ps_minus);
         A2A = df.A2A_ref * k_ratio; A2T = df.A2T_ref * k_ratio; A2L = df.A2L_ref * k_ratio;
      }
   }
   const R vmag = VelP(pmag);
   const R st2 = R(1) - sqr(mu);
   const R Omega = CyclotronFrequency(vmag, Bmag);
   const R CT = R(0.5) * df.ps_minus / df.ps_plus * sqr(mu / k_min) * A2T;
   const R CL = R(0.125) * sqr(vmag * st2 / Omega) * A2L;
   K0 = vmag * rsqrt_(CT + CL);
   if (!comp2) return;
   K2 = R(0.25) * R(M_PI) * df.ps_minus * R(fabs(Omega)) * st2 * rpow(R(vmag * k_min * R(fabs(mu / Omega))), df.ps_minus) * A2A;
   if (A2T > R(sp_tiny)) {
      const R xi1 = vmag * k_min * rsqrt_(st2) / R(M_SQRT2) / R(fabs(Omega));
      R F21 = hyp2f1_11(df, R(1) / (R(1) + sqr(xi1)));
      const R DT1 = R(1) / (R(1) + sqr(xi1)) * F21;
      const R xi2 = K0 * k_min * k_min / R(fabs(Omega));
      F21 = hyp2f1_11(df, R(1) / (R(1) + sqr(xi2)));
      const R DT2 = R(1) / (R(1) + sqr(xi2)) * F21;
      K2 += (df.ps_minus / df.ps_plus) * Omega * st2 * A2T * DT1 * DT2 / rsqrt_(sqr(DT1) + sqr(DT2));
   }
}

// DiffusionBase::GetComponent(0) (src/diffusion_base.cc:105-121) of the models usable with the
// guiding centre; |p| = pmag, mu the pitch cosine, (x, y, z) the position, outer the LISM indicator the
// Strauss / Potgieter models read (src/diffusion_other.cc:886-906, :968-988).
// XD compiles in the models that depend on mu and position (WNLT) or on the region (Strauss / Potgieter,
// SOQLT), kept out of the kernels of the other models.
template <typename R, bool XD>
__device__ __forceinline__ R gc_kperp(const SPX_CP spx_df_t<R>& df, R pmag, R mu, R Bmag, double x, double y, double z, bool outer)
{
   if constexpr (XD) {
      if (df.kind == SPX_DF_STRAUSS || df.kind == SPX_DF_POTGIETER) {
         const R B0_eff = outer ? Bmag : df.B0;
         return (outer ? df.kap_rat_out : df.kap_rat_in) * (kappa_mom_region(df, pmag, outer) * (B0_eff / Bmag));
      }
      if (df.kind == SPX_DF_WNLT || df.kind == SPX_DF_WNLT_RAMP) {
         R K0, K2;
         wnlt_k(df, false, x, y, z, pmag, mu, Bmag, K0, K2);
         return K0;
      }
      if (df.kind == SPX_DF_SOQLT) {   // :1053-1072; `outer` is region[Bmix_idx] < 0, the radius is this point's
         const bool reduced = outer && norm(mk<double>(x, y, z)) < (double)df.r_lim_red;
         R K0 = kappa_mom_soqlt(df, pmag, 0, reduced) * (df.B0 / Bmag);
         K0 = K0 / df.sc_den;
         return K0;
      }
   }
   switch (df.kind) {
   case SPX_DF_PERP: return df.D0;                                  // src/diffusion_other.cc:459-463
   case SPX_DF_FULL: return df.Dperp;                               // :538-543
   case SPX_DF_PERPISO: return df.Dperp;
   case SPX_DF_RIGBMAG: {                                           // :804-809
      R K1 = (df.lam0 * VelP(pmag) / R(3)) * rpow(Rigidity(pmag) / df.R0, df.pow_law_R);
      K1 = K1 * ((df.pow_law_B != R(0)) ? rpow(Bmag / df.B0, df.pow_law_B) : R(1));
      return df.kap_rat * K1;
   }
   default: return R(0);                                            // IsotropicConstant, ParaConstant
   }
}

// TrajectoryGuidingScatt::DiffusionCoeff (src/trajectory_guiding_scatt.cc:61-67): D_mumu = GetComponent(2)
// and its mu derivative, analytic for the isotropic model (src/diffusion_other.cc:64-79) and the forward
// difference of DiffusionBase::GetMuDerivative (src/diffusion_base.cc:226-255) for DiffusionQLTConstant
// (src/diffusion_other.cc:145-149); the other models have no component 2.
template <typename R, bool XD>
__device__ __forceinline__ void gc_dmumu(const SPX_CP spx_df_t<R>& df, R pmag, R Bmag, R mu, double x, double y, double z, R& Dmumu, R& Vmu)
{
   if (XD && (df.kind == SPX_DF_WNLT || df.kind == SPX_DF_WNLT_RAMP)) {
      R K0, K2;
      wnlt_k(df, true, x, y, z, pmag, mu, Bmag, K0, Dmumu);
      const R dmu = R(sp_small) * ((mu + R(sp_small) < R(1)) ? R(1) : R(-1));
      wnlt_k(df, true, x, y, z, pmag, mu + dmu, Bmag, K0, K2);
      Vmu = (K2 - Dmumu) / dmu;
      return;
   }
   if (df.kind == SPX_DF_ISO || df.kind == SPX_DF_PERPISO) {
      Dmumu = df.D0 * (R(1) - sqr(mu));
      Vmu = R(-2) * df.D0 * mu;
   }
   else if (df.kind == SPX_DF_QLT) {
      const R vmag = VelP(pmag);
      const R Omega = CyclotronFrequency(vmag, Bmag);
      auto k2 = [&](R m) {
         const R st2 = R(1) - sqr(m);
         return R(0.25) * R(M_PI) * df.ps_minus * R(fabs(Omega)) * st2 * rpow(R(vmag * df.k_min * R(fabs(m / Omega))), R(df.ps_minus)) * df.A2A;
      };
      Dmumu = k2(mu);
      const R dmu = R(sp_small) * ((mu + R(sp_small) < R(1)) ? R(1) : R(-1));
      Vmu = (k2(mu + dmu) - Dmumu) / dmu;
   }
}

// TrajectoryGuidingDiff::DiffusionCoeff (src/trajectory_guiding_diff.cc:72-98): D_perp and the
// drift V_perp = grad_perp D_perp in the field-aligned frame.  Only the rigidity power law has a
// spatial derivative (src/diffusion_other.cc:818-823), and only through |B|.
template <typename R, bool XD>
__device__ __forceinline__ void gc_perp(const SPX_CP spx_df_t<R>& df, const GcFields<R>& F, const v3<R>* fa, R pmag, R mu,
                                        double x, double y, double z, bool outer, R& Dperp, v3<R>& Vperp)
{
   Dperp = gc_kperp<R, XD>(df, pmag, mu, F.Bmag, x, y, z, outer);
   if (XD && (df.kind == SPX_DF_WNLT || df.kind == SPX_DF_WNLT_RAMP || df.kind == SPX_DF_STRAUSS || df.kind == SPX_DF_POTGIETER ||
              df.kind == SPX_DF_SOQLT)) {
      // DiffusionBase::GetDirectionalDerivative (src/diffusion_base.cc:130-170): along each Cartesian axis,
      // kappa_perp at the position moved by r_g with |B| carried by its gradient, and back by 2 r_g; the
      // region indicator stays the point's (the background is not evaluated again)
      R g[3];
#pragma unroll 1
      for (int k = 0; k < 3; k++) {
         R dr = F.rg;
         const R gk = k == 0 ? F.gradBmag.x : (k == 1 ? F.gradBmag.y : F.gradBmag.z);
         double xs[3] = {x, y, z};
         xs[k] += (double)dr;
         R Bm = F.Bmag + gk * dr;
         const R Kf = gc_kperp<R, XD>(df, pmag, mu, Bm, xs[0], xs[1], xs[2], outer);
         dr *= R(2);
         xs[k] -= (double)dr;
         Bm -= gk * dr;
         const R Kb = gc_kperp<R, XD>(df, pmag, mu, Bm, xs[0], xs[1], xs[2], outer);
         g[k] = (Kf - Kb) / dr;
      }
      const v3<R> gv = mk<R>(g[0], g[1], g[2]);
      Vperp = from_basis(mk<R>(dot(gv, fa[0]), dot(gv, fa[1]), R(0)), fa);
      return;
   }
   if (df.kind == SPX_DF_RIGBMAG && df.pow_law_B != R(0)) {
      const R c = Dperp * df.pow_law_B;
      const v3<R> g = mk<R>(c * F.gradBmag.x / F.Bmag, c * F.gradBmag.y / F.Bmag, c * F.gradBmag.z / F.Bmag);
      Vperp = from_basis(mk<R>(dot(g, fa[0]), dot(g, fa[1]), R(0)), fa);
   }
   else Vperp = mk<R>(R(0), R(0), R(0));
}

// TrajectoryGuiding::DriftCoeff / ModifiedFields (src/trajectory_guiding.cc:65-146): E*, B* and the
// guiding-centre velocity for velocity components (v_perp, 0, v_par).
template <typename R, int BG>
__device__ __forceinline__ void gc_drift(const GcFields<R>& F, v3<R> vel, R pp, R pl, v3<R>& drift, v3<R>& Es, v3<R>& Bs)
{
   const R rL = LarmorRadius(pp, F.Bmag), rR = LarmorRadius(pl, F.Bmag);
   if (BG == SPX_BG_SMOOTHSHOCK)   // the time-dependent term -r_R |B| (d bhat/dt) / c of E*
      Es = (F.E - ((rR * F.Bmag) * F.dbhatdt) / R(c_code)) - ((rL * vel.x) * F.gradBmag) / (R(2) * R(c_code));
   else Es = F.E - ((rL * vel.x) * F.gradBmag) / (R(2) * R(c_code));
   Bs = F.B + (rR * F.Bmag) * F.curlbhat;
   drift = (vel.z * Bs + R(c_code) * cross(Es, F.bhat)) / dot(Bs, F.bhat);
}

// Particle velocity for a guiding-centre momentum (p_perp, 0, p_par) (common/physics.hh:290-297).
template <typename R>
__device__ __forceinline__ v3<R> gc_vel(R pp, R pl)
{
   const R mm = norm(mk<R>(pp, R(0), pl));
   const R a = VelP(mm) / mm;
   return mk<R>(pp * a, R(0), pl * a);
}

// Initial momentum (src/initial_momentum.cc:283-410) in the guiding-centre layout: the magnitude
// of the configured class and mu = 2u - 1 from one uniform draw.
__device__ __forceinline__ void gc_mom_sample(PlanPtr P, const spx_rng_desc& rng, uint64_t first, uint64_t traj,
                                              uint32_t& nu, bool& ex, double& pp, double& pl)
{
   double p;
   switch (P->im_kind) {
   case SPX_IM_FIXED: pp = P->im_mom[0]; pl = P->im_mom[2]; return;
   case SPX_IM_RING: pp = P->im_p0 * P->im_st0; pl = P->im_p0 * P->im_mu0; return;   // no draw
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
   const double mu = -1.0 + 2.0 * next_uniform(rng, first, traj, nu, ex);
   const double st = sqrt(1.0 - mu * mu);
   pp = p * st;
   pl = p * mu;
}

// TrajectoryBase::SetStart (src/trajectory_base.cc:867-929) for the guiding centre: the momentum is
// sampled twice (before and after the first field evaluation, :875 and :891), each time drawing mu.
template <typename R, int BG>
__device__ int gc_set_start(PlanPtr P, QueryPtr Q, LaunchPtr L, const spx_rng_desc& rng, uint64_t slot,
                            uint64_t traj, uint32_t& nu)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const uint64_t first = Q->first;
   const double t = P->it_t;
   bool ex = false;
   double x, y, z;
   pos_sample(P, rng, first, traj, nu, ex, x, y, z);
   double pp1, pl1, pp, pl;
   gc_mom_sample(P, rng, first, traj, nu, ex, pp1, pl1);
   const double pm1 = sqrt(pp1 * pp1 + pl1 * pl1);
   R ds = R(0);   // the smooth shock's coordinate at the start stands for the "previous" evaluation's
   if (BG == SPX_BG_SMOOTHSHOCK) {
      const v3<R> d = rel(bg, x, y, z);
      ds = (dot(d, v3of(bg.n_shock)) - bg.v_shock * R(t)) / bg.width_shock;
   }
   GcFields<R> F;
   if (!gc_fields<R, BG>(bg, t, x, y, z, R(pm1), ds, F)) return 1;
   gc_mom_sample(P, rng, first, traj, nu, ex, pp, pl);
   if (ex) return 5;
   const uint64_t n = L->n_lanes;
   L->t[slot] = t; L->x[slot] = x; L->y[slot] = y; L->z[slot] = z; L->m0[slot] = pp; L->m1[slot] = 0.0; L->m2[slot] = pl;
   L->amp[slot] = 0.0; L->wgt[slot] = 0.0; L->dt_prev[slot] = 0.0; L->src_old[slot] = 0.0; L->ds_prev[slot] = (double)ds;
   L->dmax[slot] = (double)F.dmax;
   L->pf[slot] = pm1;
   L->steps[slot] = 0; L->n_events[slot] = 0;   // the draw counters continue across attempts
   L->dt_adapt[slot] = sp_large * (double)F.dmax / c_code;
   L->t0[slot] = t; L->x0[slot] = x; L->y0[slot] = y; L->z0[slot] = z; L->m00[slot] = pp; L->m01[slot] = 0.0; L->m02[slot] = pl;
   for (int d = 0; d < SPX_MAX_DIST; d++) L->ev_bin[(uint64_t)d * n + slot] = -1;
   const double dmax = (double)F.dmax;
   const double pmag = norm(mk<double>(pp, 0.0, pl));
   for (int b = 0; b < P->n_bnd; b++) {
      const SPX_CP spx_bnd_desc& B = P->bnd[b];
      const double scale = (B.cls == SPX_BC_TIME) ? fabs(dmax / c_code) : (B.cls == SPX_BC_SPACE ? fabs(dmax) : fabs(pmag));
      const int left = B.max_crossings;
      const double d = bnd_eval(B, left, true, t, x, y, z, pp, 0.0, pl, nullptr);
      L->bcross[b * n + slot] = left;
      L->bdelta[b * n + slot] = d;
      L->bdelta_old[b * n + slot] = (d == 0.0) ? -sp_tiny * scale : d;
      L->bstat[b * n + slot] = (B.status0 & ~(SPX_BND_TERMINAL | SPX_BND_CROSSED)) | (left == 1 ? SPX_BND_TERMINAL : 0);
   }
   return 0;
}

template <typename R, int BG>
__global__ __launch_bounds__(BLOCK) void gc_service(const Launch* __restrict__ Lg)
{
   const LaunchPtr L = (LaunchPtr)Lg;
   service_slot(L, [&](PlanPtr P, QueryPtr Q, const spx_rng_desc& rng, uint64_t slot, uint64_t traj, uint32_t& nu) {
      return gc_set_start<R, BG>(P, Q, L, rng, slot, traj, nu);
   });
}

// ---- step kernel -----------------------------------------------------------------------------
template <typename R>
struct GcHot {
   double t, x, y, z, pp, pl;                   // time, position, (p_perp, p_par)
   double pf;                                   // |p| of the last field evaluation
   GcFields<R> F;
   R ds_prev;                                   // BackgroundSmoothShock's hidden coordinate (smooth shock only)
   double bd[SPX_MAX_BND], bdo[SPX_MAX_BND];
   int bst[SPX_MAX_BND];
   NormalStream<R> rs;
   int32_t steps;
   typedef R real;
   __device__ __forceinline__ double m0() const { return pp; }
   __device__ __forceinline__ double m2() const { return pl; }
   __device__ __forceinline__ double m1() const { return 0.0; }
   __device__ __forceinline__ void reflect_mom() { pl = -pl; }
};

// One step of TrajectoryGuidingDiffScatt::Advance (src/trajectory_guiding_diff_scatt.cc:76-174):
// DiffusionCoeff, DriftCoeff, PhysicalStep (src/trajectory_guiding_diff.cc:324-332 then
// src/trajectory_guiding_scatt.cc:255-262), EulerPitchAngleScatt (:81-106), Slopes, EulerPerpDiffSlopes
// (src/trajectory_guiding_diff.cc:123-140), the Euler RKStep, boundaries and the end-of-step fields.
template <typename R, int BG, bool XD>
__device__ __forceinline__ int gc_step(PlanPtr P, LaunchPtr L, const spx_rng_desc& rng,
                                       uint64_t first, uint64_t slot, GcHot<R>& H, ColdLDS& C, int tid, int& bflags)
{
   const SPX_CP spx_bg_t<R>& bg = PlanView<R>::bg(P);
   const SPX_CP spx_df_t<R>& df = PlanView<R>::df(P);
   // the four guiding-centre transports differ in two terms: pitch-angle scattering (GuidingScatt,
   // GuidingDiffScatt) and perpendicular diffusion (GuidingDiff, GuidingDiffScatt); without them the
   // step draws no numbers for them and its coefficients are zero
   const int tt = P->traj_type;
   const bool pa_step = (tt == SPX_TRAJ_GUIDING_DIFF_SCATT || tt == SPX_TRAJ_GUIDING_SCATT);
   const bool perp_step = (tt == SPX_TRAJ_GUIDING_DIFF_SCATT || tt == SPX_TRAJ_GUIDING_DIFF);
   R pp = R(H.pp), pl = R(H.pl);
   const double pmag_d = sqrt_d<R>(H.pp * H.pp + H.pl * H.pl);
   v3<R> vel = gc_vel(pp, pl);
   v3<R> fa[3];
   fa[2] = H.F.bhat;
   if (is_f32<R>() && BG == SPX_BG_SOLARWIND) { fa[0] = H.F.fa0; fa[1] = H.F.fa1; }   // built by gc_fields from this bhat
   else {
      fa[0] = second_unit(H.F.bhat);
      fa[1] = cross(fa[2], fa[0]);
   }
   // DiffusionCoeff: D_perp, V_perp, D_mumu and its mu derivative (src/diffusion_other.cc:64-79)
   R pmag = R(pmag_d);
   const R mu = pl / norm(mk<R>(pp, R(0), pl));
   R Dperp = R(0);
   v3<R> Vperp = mk<R>(R(0), R(0), R(0));
   // the LISM indicator of the Strauss / Potgieter models (the mixing indicator of the SOQLT model) at the
   // point of the last field evaluation
   const bool outer = XD && (df.kind == SPX_DF_STRAUSS || df.kind == SPX_DF_POTGIETER || df.kind == SPX_DF_SOQLT) &&
                      lism_outer<R, BG>(bg, df, H.x, H.y, H.z);
   if (perp_step) gc_perp<R, XD>(df, H.F, fa, pmag, mu, H.x, H.y, H.z, outer, Dperp, Vperp);
   R Dmumu = R(0), Vmu = R(0);
   if (pa_step) gc_dmumu<R, XD>(df, pmag, H.F.Bmag, mu, H.x, H.y, H.z, Dmumu, Vmu);
   v3<R> drift, Es, Bs;
   gc_drift<R, BG>(H.F, vel, pp, pl, drift, Es, Bs);
   // PhysicalStep: TrajectoryGuiding's (src/trajectory_guiding.cc:96-100), GuidingDiff's with V_perp and
   // D_perp (src/trajectory_guiding_diff.cc:306-314), GuidingScatt's pitch-angle limits (:255-262)
   R dtp = R(0.5) * H.F.dmax / ((perp_step ? norm(drift + Vperp) : norm(drift)) + R(0.5) * norm(vel));
   if (perp_step) dtp = fmin(dtp, R(0.5) * sqr(H.F.dmax) / Dperp);
   if (pa_step) {
      const R dthetamax = R(2.0 * M_PI / 180.0);
      const R amu = fabs(pl / norm(mk<R>(pp, R(0), pl)));
      const R dmumax = rsqrt_(R(1) - amu) * dthetamax + R(0.5) * amu * sqr(dthetamax);
      dtp = fmin(dtp, R(0.5) * sqr(dmumax) / Dmumu);
      dtp = fmin(dtp, R(0.5) * dmumax / fabs(Vmu));
   }
   double dt = fmin(P->dt_scale == 1.0 ? (double)dtp : (double)dtp * P->dt_scale, C.dt_adapt[tid]);   // 1: the reference's step
   dt = time_proximity<1>(P, H, dt, (double)H.F.dmax);
   const R dtr = R(dt);
   // The step's normal draws in stream order, taken together: the pitch-angle draw, then the two of the
   // perpendicular displacement (a pitch-angle redraw moves them up and draws one more)
   R q0 = R(0), q1 = R(0), q2 = R(0);
   if (pa_step && perp_step) {
      R nrm[3];
      next_normals<R, 3>(rng, first, H.rs, nrm);
      q0 = nrm[0]; q1 = nrm[1]; q2 = nrm[2];
   }
   else if (pa_step) {
      R nrm[1];
      next_normals<R, 1>(rng, first, H.rs, nrm);
      q0 = nrm[0];
   }
   else if (perp_step) {
      R nrm[2];
      next_normals<R, 2>(rng, first, H.rs, nrm);
      q0 = nrm[0]; q1 = nrm[1];
   }
   if (H.rs.exhausted) return 5;
   // EulerPitchAngleScatt: redraw until |mu'| <= 1
   double ppd = H.pp, pld = H.pl;
   if (pa_step) {
      const R sd = rsqrt_(R(2) * Dmumu * dtr);
      R mu_new;
      int guard = 0;
      while (true) {
         mu_new = mu + Vmu * dtr + sd * q0;
         if (!(fabs(mu_new) > R(1))) break;
         if (++guard >= 100000) return 3;   // the first 100000 draws rejected
         R nrm[1];
         next_normals<R, 1>(rng, first, H.rs, nrm);
         if (H.rs.exhausted) return 5;
         if (perp_step) { q0 = q1; q1 = q2; q2 = nrm[0]; }
         else q0 = nrm[0];
      }
      // the new momentum components are formed in double from |p| (an FP32 step keeps |p| exact here)
      ppd = pmag_d * sqrt_d<R>(1.0 - (double)mu_new * (double)mu_new);
      pld = pmag_d * (double)mu_new;
      pp = R(ppd);
      pl = R(pld);
      vel = gc_vel(pp, pl);
   }
   // Slopes with the scattered momentum
   gc_drift<R, BG>(H.F, vel, pp, pl, drift, Es, Bs);
   const R slope_pp = (BG == SPX_BG_SMOOTHSHOCK) ? R(0.5) * pp / H.F.Bmag * (H.F.dBmagdt + dot(drift, H.F.gradBmag))
                                                 : R(0.5) * pp / H.F.Bmag * dot(drift, H.F.gradBmag);
   const R slope_pl = R(sp_charge) * dot(Es, Bs) / dot(Bs, H.F.bhat);
   pmag = rsqrt_(sqr(pp) + sqr(pl));
   v3<R> slope_pos = drift;
   v3<R> dr = mk<R>(R(0), R(0), R(0));
   if (perp_step) {
      const R mu2 = pl / norm(mk<R>(pp, R(0), pl));
      gc_perp<R, XD>(df, H.F, fa, pmag, mu2, H.x, H.y, H.z, outer, Dperp, Vperp);
      slope_pos = drift + Vperp;
      // EulerPerpDiffSlopes
      const R sdt = rsqrt_(dtr);
      const R dWx = sdt * (pa_step ? q1 : q0), dWy = sdt * (pa_step ? q2 : q1);
      Dperp = gc_kperp<R, XD>(df, pmag, mu2, H.F.Bmag, H.x, H.y, H.z, outer);
      dr = from_basis(mk<R>(rsqrt_(R(2) * Dperp) * dWx, rsqrt_(R(2) * Dperp) * dWy, R(0)), fa);
   }
   // RKStep (Euler), then the perpendicular displacement
   const R dtv = dtr * R(1);
   if constexpr (is_f32<R>()) {   // compensated FP64 sums; drift and perpendicular displacement added in FP32 first
      const v3<R> d = perp_step ? dtv * slope_pos + dr : dtv * slope_pos;
      kahan_add(H.t, C.kc[0][tid], dt);
      kahan_add(H.x, C.kc[1][tid], (double)d.x); kahan_add(H.y, C.kc[2][tid], (double)d.y); kahan_add(H.z, C.kc[3][tid], (double)d.z);
      H.pp = ppd + (double)(dtv * slope_pp);
      H.pl = pld + (double)(dtv * slope_pl);
   }
   else {
      H.t += dt;
      H.x += (double)(dtv * slope_pos.x); H.y += (double)(dtv * slope_pos.y); H.z += (double)(dtv * slope_pos.z);
      H.pp = ppd + (double)(dtv * slope_pp);
      H.pl = pld + (double)(dtv * slope_pl);
      if (perp_step) { H.x += (double)dr.x; H.y += (double)dr.y; H.z += (double)dr.z; }
   }
   bflags = handle_boundaries<true>(P, L, slot, H, C, tid);
   if (!(bflags & 1)) {
      H.pf = sqrt_d<R>(H.pp * H.pp + H.pl * H.pl);
      if (!gc_fields<R, BG>(bg, H.t, H.x, H.y, H.z, R(H.pf), H.ds_prev, H.F)) return 1;
   }
   H.steps++;
   if (H.steps > P->max_steps) return 4;
   if (dt < sp_tiny * (double)H.F.dmax / c_code) return 2;
   if (!__builtin_isnormal(dt)) return 3;
   return 0;
}

// SPX_GC_WAVES resident waves per SIMD for the step kernel (0: the compiler's own register budget).
#ifndef SPX_GC_WAVES
#define SPX_GC_WAVES 0
#endif
#if SPX_GC_WAVES > 0
#define SPX_GC_ATTR __attribute__((amdgpu_waves_per_eu(SPX_GC_WAVES)))
#else
#define SPX_GC_ATTR
#endif

template <typename R, int BG, bool XD>
__global__ __launch_bounds__(BLOCK) SPX_GC_ATTR void gc_step_kernel(const Launch* __restrict__ Lg)
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
   GcHot<R> H;
   if (run) {
      const uint64_t n = Lp->n_lanes;
      H.t = Lp->t[slot]; H.x = Lp->x[slot]; H.y = Lp->y[slot]; H.z = Lp->z[slot]; H.pp = Lp->m0[slot]; H.pl = Lp->m2[slot];
      if constexpr (is_f32<R>()) {
#pragma unroll
         for (int k = 0; k < 4; k++) C.kc[k][tid] = Lp->kc[k * n + slot];
      }
      H.pf = Lp->pf[slot];
      H.steps = (int32_t)Lp->steps[slot];
      H.rs.traj = Lp->traj[slot]; H.rs.nn = Lp->nn[slot]; H.rs.exhausted = false;
      prime_normals<R>(rng, H.rs);
      C.amp[tid] = Lp->amp[slot]; C.wgt[tid] = Lp->wgt[slot]; C.dt_adapt[tid] = Lp->dt_adapt[slot];
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         H.bd[b] = Lp->bdelta[b * n + slot]; H.bdo[b] = Lp->bdelta_old[b * n + slot];
         C.cross[b][tid] = Lp->bcross[b * n + slot]; H.bst[b] = Lp->bstat[b * n + slot];
      }
      // the fields are a function of (t, x, |p| of their evaluation) and are recomputed bit for bit; dmax
      // (which may depend on the previous evaluation) is restored
      H.ds_prev = (BG == SPX_BG_SMOOTHSHOCK) ? R(Lp->ds_prev[slot]) : R(0);
      R ds_tmp = H.ds_prev;
      (void)gc_fields<R, BG>(PlanView<R>::bg(P), H.t, H.x, H.y, H.z, R(H.pf), ds_tmp, H.F);
      if (BG == SPX_BG_SMOOTHSHOCK) H.F.dmax = R(Lp->dmax[slot]);
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
      const int e = gc_step<R, BG, XD>(Pi, Li, rng, first, slot, H, C, tid, bflags);
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
   Ls->t[slot] = H.t; Ls->x[slot] = H.x; Ls->y[slot] = H.y; Ls->z[slot] = H.z; Ls->m0[slot] = H.pp; Ls->m2[slot] = H.pl;
   if constexpr (is_f32<R>()) {
#pragma unroll
      for (int k = 0; k < 4; k++) Ls->kc[k * n + slot] = C.kc[k][tid];
   }
   Ls->pf[slot] = H.pf;
   Ls->steps[slot] = H.steps; Ls->nn[slot] = H.rs.nn;
   Ls->dmax[slot] = (double)H.F.dmax;
   if (BG == SPX_BG_SMOOTHSHOCK) Ls->ds_prev[slot] = (double)H.ds_prev;
   Ls->amp[slot] = C.amp[tid]; Ls->wgt[slot] = C.wgt[tid];
#pragma unroll
   for (int b = 0; b < SPX_MAX_BND; b++) {
      Ls->bdelta[b * n + slot] = H.bd[b]; Ls->bdelta_old[b * n + slot] = H.bdo[b];
      Ls->bcross[b * n + slot] = C.cross[b][tid]; Ls->bstat[b * n + slot] = H.bst[b];
   }
}

template <typename R, int BG, bool XD>
static hipError_t launch_gc_bg(uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   const unsigned grid = (unsigned)((n_lanes + BLOCK - 1) / BLOCK);
   hipLaunchKernelGGL((gc_service<R, BG>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   if (mid) {
      hipError_t e = hipEventRecord(mid, st);
      if (e != hipSuccess) return e;
   }
   hipLaunchKernelGGL((gc_step_kernel<R, BG, XD>), dim3(grid), dim3(BLOCK), 0, st, d_launch);
   return hipGetLastError();
}

template <typename R>
hipError_t dispatch_gc(const spx_plan_desc& hp, uint64_t n_lanes, const Launch* d_launch, hipStream_t st, hipEvent_t mid)
{
   if (hp.time_flow != 0) return hipErrorInvalidValue;
   // the WNLT, Strauss / Potgieter and SOQLT models run in kernels of their own (SolarWind, Uniform)
   const int k = hp.df.kind;
   const bool xd = (k == SPX_DF_WNLT || k == SPX_DF_WNLT_RAMP || k == SPX_DF_STRAUSS || k == SPX_DF_POTGIETER || k == SPX_DF_SOQLT);
   switch (hp.bg.kind) {
   case SPX_BG_SOLARWIND:
      return xd ? launch_gc_bg<R, SPX_BG_SOLARWIND, true>(n_lanes, d_launch, st, mid) : launch_gc_bg<R, SPX_BG_SOLARWIND, false>(n_lanes, d_launch, st, mid);
   case SPX_BG_UNIFORM:
      return xd ? launch_gc_bg<R, SPX_BG_UNIFORM, true>(n_lanes, d_launch, st, mid) : launch_gc_bg<R, SPX_BG_UNIFORM, false>(n_lanes, d_launch, st, mid);
   case SPX_BG_GRID: return xd ? hipErrorInvalidValue : launch_gc_bg<R, SPX_BG_GRID, false>(n_lanes, d_launch, st, mid);
   case SPX_BG_SHOCK: return xd ? hipErrorInvalidValue : launch_gc_bg<R, SPX_BG_SHOCK, false>(n_lanes, d_launch, st, mid);
   case SPX_BG_SMOOTHSHOCK: return xd ? hipErrorInvalidValue : launch_gc_bg<R, SPX_BG_SMOOTHSHOCK, false>(n_lanes, d_launch, st, mid);
   default: return hipErrorInvalidValue;
   }
}

}  // namespace spx
