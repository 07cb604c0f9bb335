// spx_physics.h -- device-side physics of the trajectory hot path.
//
// Each routine states the reference function it reproduces (file:line in vflorins/SPECTRUM).
// The step arithmetic type R is double for the drop-in kernel (bit-level agreement with the
// reference method) and float for the mixed-precision kernel; time, position, momentum,
// amplitude and weight are always accumulated in double.
#pragma once

#include <type_traits>

#include "spx_plan.h"
#include "spx_vec.h"

namespace spx {

// ---- code units (common/specie.hh:30-85, common/physical_constants.hh:20-81) --------------------
constexpr double C_CGS = 2.99792458e+10;
constexpr double MP_CGS = 1.67262192596E-24;
constexpr double E_CGS = 4.8032044E-10;
constexpr double AU_CGS = 1.495978707E+13;
constexpr double unit_length = 1.0E10;
constexpr double unit_velocity = 0.1 * C_CGS;
constexpr double c_code = C_CGS / unit_velocity;
constexpr double sp_mass = MP_CGS / MP_CGS;
constexpr double sp_charge = E_CGS / E_CGS;
constexpr double hp_rad_sw = 117.0 * AU_CGS / unit_length;   // src/background_solarwind.hh:33

// common/definitions.hh:54-66
constexpr double sp_large = 1.0E+20;
constexpr double sp_little = 1.0E-5;
constexpr double sp_small = 1.0E-8;
constexpr double sp_tiny = 1.0E-15;

// ---- arithmetic helpers: IEEE in double, hardware approximations in float -------------------------
SPX_HD double rsqrt_(double x) { return sqrt_rn(x); }
SPX_HD float rsqrt_(float x) { return __builtin_sqrtf(x); }
SPX_HD double rdiv(double a, double b) { return a / b; }
SPX_HD float rdiv(float a, float b) { return a / b; }
// pow with the identities every pow meets exactly (x^1 = x, x^0 = 1 as C99 F.9.4.4 fixes), and x^2 and
// x^(1/2) correctly rounded: x * x and the correctly rounded square root (+ 0.0 gives pow's +0 at -0),
// for x >= -0 or +inf (the ratios of positive quantities these exponents apply to).  No libm pow is
// correctly rounded: glibc's pow(x, 0.5) differs from the square root in about 1 of 1200 arguments by one
// ulp, the device library's in about 1 of 4, so the square root follows the host's pow far more closely.
// The exponents are plan constants, so these branches are uniform.
SPX_HD double rpow(double a, double b)
{
   if (b == 1.0) return a;
   if (b == 0.0) return 1.0;
   if (b == 2.0) return a * a;
   if (b == 0.5) return sqrt_rn(a) + 0.0;
   return pow(a, b);
}
SPX_HD float rpow(float a, float b)
{
   if (b == 1.0f) return a;
   if (b == 0.0f) return 1.0f;
   if (b == 2.0f) return a * a;
   if (b == 0.5f) return __builtin_sqrtf(a);
   return __powf(a, b);
}
SPX_HD double rtanh(double a) { return tanh(a); }
SPX_HD float rtanh(float a) { return tanhf(a); }

// The FP32 step may restructure its arithmetic (one reciprocal square root for a length and its
// inverse, stencil points relative to the background origin); the FP64 step keeps the reference's.
template <typename R> constexpr bool is_f32() { return std::is_same<R, float>::value; }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }

// a / c for a constant c: the FP64 kernels round it as the division does, through div_rn with the
// reciprocal the compiler folds (correctly rounded); the FP32 unit, built with reciprocal math, keeps its
// product by the reciprocal.  divc for the step type R, divc_d for an FP64 quantity in a kernel of step type R.
template <typename R> SPX_HD R divc(R a, R c)
{
   if constexpr (is_f32<R>()) return a / c;
   else return div_rn(a, c, 1.0 / c);
}
template <typename R> SPX_HD double divc_d(double a, double c)
{
   if constexpr (is_f32<R>()) return a / c;
   else return div_rn(a, c, 1.0 / c);
}
// sqrt of an FP64 quantity in a kernel of step type R: correctly rounded (sqrt_rn) in the FP64 kernels,
// the unit's own square root in the FP32-step kernels.
template <typename R> SPX_HD double sqrt_d(double x)
{
   if constexpr (is_f32<R>()) return sqrt(x);
   else return sqrt_rn(x);
}

// Compensated (Kahan) addition of an increment to an FP64 sum, for the FP32-step kernels: c carries what
// the rounding of s dropped (less than an ulp of s, so FP32 holds it to 24 bits) into the next addition.
// No reassociation is enabled for the FP32 unit, so (t - s) - y is evaluated as written.
__device__ __forceinline__ void kahan_add(double& s, float& c, double inc)
{
   const double y = inc - (double)c;
   const double t = s + y;
   c = (float)((t - s) - y);
   s = t;
}

// ---- relativistic kinematics (common/physics.hh:246-456) ---------------------------------------
template <typename R> SPX_HD R EnrTot(R p) { return rsqrt_(sqr(p) + sqr(R(sp_mass * c_code))) * R(c_code); }
template <typename R> SPX_HD R EnrKin(R p) { return EnrTot(p) - R(sp_mass * (c_code * c_code)); }
template <typename R> SPX_HD R VelP(R p) { return p * R(c_code * c_code) / EnrTot(p); }
template <typename R> SPX_HD R Rigidity(R p) { return p * R(c_code) / R(fabs(sp_charge)); }
template <typename R> SPX_HD R LarmorRadius(R p, R B) { return p * R(c_code) / (R(fabs(sp_charge)) * B); }
template <typename R> SPX_HD R RelFactor(R v) { return R(1) / rsqrt_(R(1) - sqr(v / R(c_code))); }
template <typename R> SPX_HD R CyclotronFrequency(R v, R B) { return R(sp_charge) * B / (RelFactor(v) * R(sp_mass) * R(c_code)); }
// Particle::Mom (common/physics.hh:262-266)
SPX_HD double MomT(double T) { return sqrt(T * (T + 2.0 * sp_mass * (c_code * c_code))) / c_code; }

// ---- fields at one point: the SpatialData members the path reads --------------------------------
template <typename R>
struct Fields {
   v3<R> U, B, E, bhat;
   R Bmag, dmax, divU;
   m3<R> gradB;        // guiding centre only
   v3<R> gradBmag;     // guiding centre only
   v3<R> dBdt;         // guiding centre only
   R dBmagdt;
};

// Plan access.  The kernels read the plan through constant-address-space pointers: the plan never
// changes while a kernel runs, so every read is a scalar load through the scalar cache instead of a
// per-lane vector load.
#define SPX_CP __attribute__((address_space(4)))
typedef const SPX_CP spx_plan_desc* PlanPtr;

// The descriptors of the step type R.
template <typename R> struct PlanView;
template <> struct PlanView<double> {
   static __device__ const SPX_CP spx_bg_t<double>& bg(PlanPtr P) { return P->bg; }
   static __device__ const SPX_CP spx_df_t<double>& df(PlanPtr P) { return P->df; }
};
template <> struct PlanView<float> {
   static __device__ const SPX_CP spx_bg_t<float>& bg(PlanPtr P) { return P->bgf; }
   static __device__ const SPX_CP spx_df_t<float>& df(PlanPtr P) { return P->dff; }
};

template <typename R> __device__ v3<R> v3of(const SPX_CP R* a) { return mk<R>(a[0], a[1], a[2]); }

// Relative position x - r0 (double subtraction, then the step type).
template <typename R>
SPX_HD v3<R> rel(const SPX_CP spx_bg_t<R>& b, double x, double y, double z) { return mk<R>(R(x - b.r0[0]), R(y - b.r0[1]), R(z - b.r0[2])); }

// BackgroundSolarWind::EvaluateBackground (src/background_solarwind.cc:111-255), every switch at 0, in
// the frame of the rotation axis: pp is the position in that frame, u and Bl the fields in it.
template <typename R>
SPX_HD void solarwind_local(const SPX_CP spx_bg_t<R>& b, v3<R> pp, int want, v3<R>& u, v3<R>& Bl)
{
   if constexpr (is_f32<R>()) {
      // The spiral in closed form: sin(theta) cos(phi) = x / r, sin(theta) sin(phi) = y / r, cos(theta) = z / r,
      // so B = B0 r_ref^2 / r^3 (x + q y, y - q x, z) with q = (r - r_ref) w0 / ur0 (one reciprocal square root)
      const R r2 = norm2(pp), ir = frsq(r2), r = r2 * ir;
      if (want & 1) u = (b.ur0 * ir) * pp;
      if (want & 2) {
         const R Bri = b.sw_c1 * (ir * ir) * ir;
         const R q = (r - b.r_ref) * b.sw_k;
         Bl = mk<R>(Bri * (pp.x + q * pp.y), Bri * (pp.y - q * pp.x), Bri * pp.z);
      }
      return;
   }
   else {
   // the quotients by r (those of unit(pp), cos(theta), r_ref / r) and by s share one correctly rounded
   // reciprocal each; div_rn rounds every quotient as the division does
   R r = norm(pp);
   const R yr = R(1) / r;
   if (want & 1) u = b.ur0 * mk<R>(div_rn(pp.x, r, yr), div_rn(pp.y, r, yr), div_rn(pp.z, r, yr));
   if (want & 2) {
      R r_mns = r - b.r_ref;
      R costheta = div_rn(pp.z, r, yr);
      R sintheta = rsqrt_(fmax(R(1) - sqr(costheta), R(0)));
      R s = rsqrt_(sqr(pp.x) + sqr(pp.y));
      R cosphi, sinphi;
      if (s < r * R(sp_tiny)) { cosphi = R(0); sinphi = R(0); }
      else { const R ys = R(1) / s; cosphi = div_rn(pp.x, s, ys); sinphi = div_rn(pp.y, s, ys); }
      R Br = b.B0[0] * sqr(div_rn(b.r_ref, r, yr));
      R Bp = div_rn(-Br * sintheta * r_mns * b.w0, b.ur0, b.ur0_rcp);   // / ur0, rounded as the division
      Bl = mk<R>(Br * sintheta * cosphi - Bp * sinphi, Br * sintheta * sinphi + Bp * cosphi, Br * costheta);
   }
   }
}

// Flags select what is evaluated: U (bit 0), B (bit 1).  With a diagonal +-1 basis, to_basis /
// from_basis reduce to sign flips: every product with an off-diagonal zero adds a signed zero, so the
// results agree bit for bit up to the sign of a zero.  The identity basis (rotation axis along +z,
// a uniform branch) needs no products at all: x * 1 == x exactly.
template <typename R>
SPX_HD void solarwind_eval(const SPX_CP spx_bg_t<R>& b, v3<R> d, int want, v3<R>& U, v3<R>& B)
{
   if (is_f32<R>() && b.ep_diag != 0) {
      // the closed form in the global frame: a diagonal basis S (det 1) turns S w(S d) into
      // (d0 + s q d1, d1 - s q d0, d2) with s = eprime[0][0] eprime[1][1], folded into sw_kg
      const R r2 = norm2(d), ir = frsq(r2), r = r2 * ir;
      if (want & 1) U = (b.ur0 * ir) * d;
      if (want & 2) {
         const R Bri = b.sw_c1 * (ir * ir) * ir;
         const R q = (r - b.r_ref) * b.sw_kg;
         B = mk<R>(Bri * (d.x + q * d.y), Bri * (d.y - q * d.x), Bri * d.z);
      }
      return;
   }
   if (b.ep_diag == 2) {
      solarwind_local<R>(b, d, want, U, B);
      return;
   }
   v3<R> ep[3] = {v3of(b.eprime[0]), v3of(b.eprime[1]), v3of(b.eprime[2])};
   const bool dg = b.ep_diag != 0;
   auto diag = [&](v3<R> a) { return mk<R>(a.x * b.eprime[0][0], a.y * b.eprime[1][1], a.z * b.eprime[2][2]); };
   v3<R> u, Bl;
   solarwind_local<R>(b, dg ? diag(d) : to_basis(d, ep), want, u, Bl);
   if (want & 1) U = dg ? diag(u) : from_basis(u, ep);
   if (want & 2) B = dg ? diag(Bl) : from_basis(Bl, ep);
}

// The Parker step's two SolarWind stencils in the FP32 step, restated from the closed forms above in the
// global frame.  At the points c +- h e (c = x - r0, |e| = 1), |c +- h e|^2 = |c|^2 + h^2 +- 2h e.c,
// so each point costs one reciprocal square root and no vector products.  U = ur0 p / |p| does not depend
// on the basis; the b b stencil needs a diagonal one (the rotation axis along +-z), whose winding sense
// in the global frame is the sign of eprime[0][0] eprime[1][1] (sw_kg).
//
// NumericalDerivatives' trace for div U (src/background_base.cc:228-312 with the basis e_k):
// sum_k e_k . [U(c + h e_k) - U(c - h e_k)], with e . U(p) = ur0 (e . p) / |p|.
template <typename R>
SPX_HD R sw_divu_stencil(const SPX_CP spx_bg_t<R>& b, v3<R> c, R h, const v3<R> (&e)[3])
{
   const R c2h = norm2(c) + h * h, h2 = R(2) * h;
   R acc = R(0);
#pragma unroll
   for (int k = 0; k < 3; k++) {
      const R s = dot(e[k], c);
      acc += frsq(c2h + h2 * s) * (s + h) - frsq(c2h - h2 * s) * (s - h);
   }
   return b.ur0 * acc;
}

// The b b part of div(kappa)'s stencil (TrajectoryParker::DiffusionCoeff, src/trajectory_parker.cc:120-151,
// with kappa position independent): sum_j [b_j b](c + h e_j) - [b_j b](c - h e_j) over the Cartesian axes.
// b is the direction of w = (x + q y, y - q x, z), q = (r - r_ref) sw_kg: the radial factor of B cancels
// in b_j b = w_j w / |w|^2.
// True when no point c +- h e_j can meet GetFields' |B| < sp_tiny (src/background_base.cc:444-448): every
// |c +- h e_j|^2 is at most |c|^2 + h^2 + 2h (|c_x| + |c_y| + |c_z|).
template <typename R>
SPX_HD bool sw_stencil_clear(const SPX_CP spx_bg_t<R>& b, v3<R> c, R h)
{
   return norm2(c) + h * h + (R(2) * h) * (fabs(c.x) + fabs(c.y) + fabs(c.z)) <= b.sw_r2_ok;
}

template <typename R>
SPX_HD v3<R> sw_bb_stencil(const SPX_CP spx_bg_t<R>& b, v3<R> c, R h)
{
   const R c2h = norm2(c) + h * h, h2 = R(2) * h;
   v3<R> acc = mk<R>(R(0), R(0), R(0));
#pragma unroll
   for (int j = 0; j < 3; j++) {
#pragma unroll
      for (int s = 0; s < 2; s++) {
         const R sh = s == 0 ? h : -h;
         const R cj = j == 0 ? c.x : (j == 1 ? c.y : c.z);
         const v3<R> p = mk<R>(j == 0 ? c.x + sh : c.x, j == 1 ? c.y + sh : c.y, j == 2 ? c.z + sh : c.z);
         const R r2 = c2h + (s == 0 ? h2 : -h2) * cj;
         const R ir = frsq(r2), r = r2 * ir;
         const R q = (r - b.r_ref) * b.sw_kg;
         const v3<R> w = mk<R>(p.x + q * p.y, p.y - q * p.x, p.z);
         const R wj = j == 0 ? w.x : (j == 1 ? w.y : w.z);
         const R t = wj / norm2(w);
         acc = acc + (s == 0 ? t : -t) * w;
      }
   }
   return acc;
}

// src/background_smooth_shock.cc:46-75 (SMOOTH_SHOCK_ORDER 4)
template <typename R> SPX_HD R shock_transition(R x) { return R(0.5) * (R(1) + rtanh(R(8) * x)); }
template <typename R> SPX_HD R shock_transition_deriv(R x) { return R(0.5) * R(8) * (R(1) - sqr(rtanh(R(8) * x))); }

// Generic EvaluateBackground for U and B (want bits) at d = x - r0, used at the finite-difference
// points.  `ds` receives BackgroundSmoothShock's transition coordinate of this evaluation.
template <typename R, int BG>
SPX_HD void bg_eval_rel(const SPX_CP spx_bg_t<R>& b, double t, v3<R> d, int want, v3<R>& U, v3<R>& B, R& ds)
{
   if (BG == SPX_BG_SOLARWIND) {
      solarwind_eval<R>(b, d, want, U, B);
   }
   else if (BG == SPX_BG_UNIFORM) {   // src/background_uniform.cc:70-77
      U = v3of(b.u0); B = v3of(b.B0);
   }
   else if (BG == SPX_BG_SHOCK) {     // src/background_shock.cc:91-111
      if (dot(d, v3of(b.n_shock)) - b.v_shock * R(t) > R(0)) { U = v3of(b.u0); B = v3of(b.B0); }
      else { U = v3of(b.u1); B = v3of(b.B1); }
   }
   else if (BG == SPX_BG_SMOOTHSHOCK) {   // src/background_smooth_shock.cc:139-154
      ds = (dot(d, v3of(b.n_shock)) - b.v_shock * R(t)) / b.width_shock;
      R a1 = shock_transition(ds), a2 = R(1) - a1;
      U = a1 * v3of(b.u0) + a2 * v3of(b.u1);
      B = a1 * v3of(b.B0) + a2 * v3of(b.B1);
   }
}

template <typename R, int BG>
SPX_HD void bg_eval(const SPX_CP spx_bg_t<R>& b, double t, double x, double y, double z, int want, v3<R>& U, v3<R>& B, R& ds)
{
   bg_eval_rel<R, BG>(b, t, rel(b, x, y, z), want, U, B, ds);
}

// EvaluateDmax per class (src/background_solarwind.cc:310-314, src/background_smooth_shock.cc:192-196,
// src/background_base.cc:378-382).  `ds_prev` is the smooth shock coordinate of the previous evaluation.
template <typename R, int BG>
SPX_HD R bg_dmax(const SPX_CP spx_bg_t<R>& b, double x, double y, double z, R ds_prev)
{
   if (BG == SPX_BG_SOLARWIND) return fmin(b.dmax_fraction * norm(rel(b, x, y, z)), b.dmax0);
   if (BG == SPX_BG_SMOOTHSHOCK) return fmin(b.dmax_fraction * b.width_shock * fmax(R(1), fabs(ds_prev)), b.dmax0);
   return b.dmax0;
}

// ---- gridded background: BackgroundServerCartesian, SERVER_INTERP_ORDER 1, no ghost cells -----------
// The block of a position follows ReadCartesianGetNode (src/reader_cartesian.cc:255-270) on the
// position in fluid units; its geometry ReadCartesianGetBlockCorners (src/reader_cartesian.cc:277-290),
// BlockBase::LoadDimensions and ConfigureProperties (src/block_base.cc:110-146); the offsets
// BlockBase::GetZoneOffset (src/block_base.cc:164-176); the weights and derivative weights
// ServerCartesianFront::InteriorInterpolationStencil (src/server_cartesian.cc:135-183), which the
// reference keeps when the stencil reaches into neighbour blocks (:215-246), so a corner's zone is
// 4 * block + local zone in the global zone array; the sums GetVariablesInterp1 (:454-560) and
// GetGradientsInterp1 (:630-705) in corner order; the unit conversions GetVariables / GetGradients
// (:570-610, :715-740).  A position outside the domain (ExServerError, src/server_cartesian.cc:1058)
// or whose stencil leaves the grid (the reference then reads a block that does not exist) returns
// false: a field error that discards the attempt.
template <typename R>
struct GridFields {
   v3<R> U, B;
   R Bmag;
   R divU;            // trace of grad U (GRAD only)
   double zl_min;     // smallest zone length of the position's block (ServerCartesianFront dmax limit)
};

typedef double grid_d2 __attribute__((ext_vector_type(2)));
typedef float grid_f4 __attribute__((ext_vector_type(4)));

// Corner c = i + 2j + 4k of the stencil whose lowest zone is `cell`: that zone record (U, B, |B|, fluid
// units).  FP64 reads it from the zone array, FP32 from the cell's contiguous copy of its 8 corners.
template <typename R> __device__ __forceinline__ void grid_corner(const SPX_CP spx_grid_desc& g, int64_t cell, int c, R v[7]);
template <> __device__ __forceinline__ void grid_corner<double>(const SPX_CP spx_grid_desc& g, int64_t cell, int c, double v[7])
{
   const int64_t zone = cell + (c & 1) + ((c >> 1) & 1) * (int64_t)g.nz[0] + (c >> 2) * ((int64_t)g.nz[0] * g.nz[1]);
   const grid_d2* r = reinterpret_cast<const grid_d2*>(g.z64) + zone * (SPX_GRID_REC / 2);
   const grid_d2 a = r[0], b = r[1], d = r[2], e = r[3];
   v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = d.x; v[5] = d.y; v[6] = e.x;
}
template <> __device__ __forceinline__ void grid_corner<float>(const SPX_CP spx_grid_desc& g, int64_t cell, int c, float v[7])
{
   const grid_f4* r = reinterpret_cast<const grid_f4*>(g.c32) + (cell * 8 + c) * (SPX_GRID_REC / 4);
   const grid_f4 a = r[0], b = r[1];
   v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z;
}

// The position's block, its stencil's first zone and the offsets within it.  False where the
// reference fails (outside the domain, or a stencil beyond the grid).
template <typename R>
__device__ __forceinline__ bool grid_locate(const SPX_CP spx_grid_desc& g, double x, double y, double z, R lo[3], R hi[3], R dl[3],
                                            int64_t& zone0, double& zl_min)
{
   const double pos[3] = {x, y, z};
   zone0 = 0;
   zl_min = 0.0;
#pragma unroll
   for (int a = 0; a < 3; a++) {
      const double pf = pos[a] / AU_CGS * unit_length;          // server side: position in fluid units
      if (!(pf >= g.dom_min[a] && pf <= g.dom_max[a])) return false;
      const int ib = (int)((pf - g.dom_min[a]) / g.block_length[a]);
      if (ib >= g.nb[a]) return false;                          // on the upper face of the domain
      const double f0 = g.dom_min[a] + g.block_length[a] * (double)ib;
      const double face_lo = f0 * g.f_len, face_hi = (f0 + g.block_length[a]) * g.f_len;
      const double zl = (face_hi - face_lo) / (double)SPX_GRID_BLOCK;
      const double cmin = face_lo + 0.5 * zl;
      double off = (pos[a] - cmin) / zl;
      int zi = (int)off;
      if (off < 0.0) zi--;
      off = off - (double)zi;
      const int gz = SPX_GRID_BLOCK * ib + zi;
      if (gz < 0 || gz + 1 >= g.nz[a]) return false;
      lo[a] = R(off);
      hi[a] = R(1.0 - off);
      dl[a] = R(zl);
      zl_min = (a == 0) ? zl : fmin(zl_min, zl);
      zone0 += (a == 0) ? (int64_t)gz : (a == 1 ? (int64_t)gz * g.nz[0] : (int64_t)gz * g.nz[0] * g.nz[1]);
   }
   return true;
}

template <typename R>
__device__ __forceinline__ bool grid_inside(const SPX_CP spx_grid_desc& g, double x, double y, double z)
{
   // well inside the zone-centre box the answer is yes without the exact block arithmetic
   if (x > g.inner_lo[0] && x < g.inner_hi[0] && y > g.inner_lo[1] && y < g.inner_hi[1] && z > g.inner_lo[2] && z < g.inner_hi[2])
      return true;
   R lo[3], hi[3], dl[3];
   int64_t zone0;
   double zl_min;
   return grid_locate<R>(g, x, y, z, lo, hi, dl, zone0, zl_min);
}

template <typename R, bool GRAD>
__device__ __forceinline__ bool grid_eval(const SPX_CP spx_grid_desc& g, double x, double y, double z, GridFields<R>& o)
{
   R lo[3], hi[3], dl[3];
   int64_t zone0;
   double zl_min;
   if (!grid_locate<R>(g, x, y, z, lo, hi, dl, zone0, zl_min)) return false;
   // corners in the reference's order c = i + 2j + 4k, one row (j, k) of two x-neighbours at a time: the
   // sums start from zero as the reference's do, and a rolled loop keeps two records in flight instead
   // of eight, which leaves registers for more resident waves to hide the gather latency
   R acc[7], dUx = R(0), dUy = R(0), dUz = R(0);
#pragma unroll
   for (int m = 0; m < 7; m++) acc[m] = R(0);
#pragma unroll 1
   for (int r = 0; r < 4; r++) {
      const int j = r & 1, k = r >> 1;
      R v0[7], v1[7];
      grid_corner<R>(g, zone0, 2 * r, v0);
      grid_corner<R>(g, zone0, 2 * r + 1, v1);
      const R fy = j ? lo[1] : hi[1], fz = k ? lo[2] : hi[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
         const R* v = i ? v1 : v0;
         const R fx = i ? lo[0] : hi[0];
         const R w = fx * fy * fz;
#pragma unroll
         for (int m = 0; m < 7; m++) acc[m] = acc[m] + w * v[m];
         if (GRAD) {   // the diagonal of grad U: d(Ux)/dx, d(Uy)/dy, d(Uz)/dz
            const R sx = i ? R(1) : R(-1), sy_ = j ? R(1) : R(-1), sz_ = k ? R(1) : R(-1);
            const R ddx = sx * (fy * fz) / dl[0], ddy = sy_ * (fx * fz) / dl[1], ddz = sz_ * (fx * fy) / dl[2];
            dUx = dUx + ddx * v[0];
            dUy = dUy + ddy * v[1];
            dUz = dUz + ddz * v[2];
         }
      }
   }
   const R fv = R(g.f_vel), fm = R(g.f_mag);
   o.U = mk<R>(acc[0] * fv, acc[1] * fv, acc[2] * fv);
   o.B = mk<R>(acc[3] * fm, acc[4] * fm, acc[5] * fm);
   o.Bmag = acc[6] * fm;
   if (GRAD) o.divU = dUx * fv + dUy * fv + dUz * fv;
   o.zl_min = zl_min;
   return true;
}

// The guiding centre's fields from the grid: U, B, the interpolated |B|, E = -(U x B) / c formed in fluid
// units before the unit conversion (GetVariablesInterp1, src/server_cartesian.cc:538-560, :590-609), and
// from the stencil's derivative weights (GetGradientsInterp1, :630-705) grad|B| and the off-diagonal
// d B_j / d x_i that curl B needs.  gB holds (01, 02, 10, 12, 20, 21) of gradB[i][j] = d B_j / d x_i.
template <typename R>
__device__ __forceinline__ bool grid_eval_gc(const SPX_CP spx_grid_desc& g, double x, double y, double z, v3<R>& U, v3<R>& B, R& Bmag,
                                             v3<R>& E, v3<R>& gradBmag, R gB[6], double& zl_min)
{
   R lo[3], hi[3], dl[3];
   int64_t zone0;
   if (!grid_locate<R>(g, x, y, z, lo, hi, dl, zone0, zl_min)) return false;
   R acc[7], gm[3], g6[6];
#pragma unroll
   for (int m = 0; m < 7; m++) acc[m] = R(0);
#pragma unroll
   for (int m = 0; m < 3; m++) gm[m] = R(0);
#pragma unroll
   for (int m = 0; m < 6; m++) g6[m] = R(0);
#pragma unroll 1
   for (int r = 0; r < 4; r++) {
      const int j = r & 1, k = r >> 1;
      R v0[7], v1[7];
      grid_corner<R>(g, zone0, 2 * r, v0);
      grid_corner<R>(g, zone0, 2 * r + 1, v1);
      const R fy = j ? lo[1] : hi[1], fz = k ? lo[2] : hi[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
         const R* v = i ? v1 : v0;
         const R fx = i ? lo[0] : hi[0];
         const R w = fx * fy * fz;
#pragma unroll
         for (int m = 0; m < 7; m++) acc[m] = acc[m] + w * v[m];
         const R sx = i ? R(1) : R(-1), sy_ = j ? R(1) : R(-1), sz_ = k ? R(1) : R(-1);
         const R dx = sx * (fy * fz) / dl[0], dy = sy_ * (fx * fz) / dl[1], dz = sz_ * (fx * fy) / dl[2];
         gm[0] = gm[0] + dx * v[6]; gm[1] = gm[1] + dy * v[6]; gm[2] = gm[2] + dz * v[6];
         g6[0] = g6[0] + dx * v[4];   // d By / dx
         g6[1] = g6[1] + dx * v[5];   // d Bz / dx
         g6[2] = g6[2] + dy * v[3];   // d Bx / dy
         g6[3] = g6[3] + dy * v[5];   // d Bz / dy
         g6[4] = g6[4] + dz * v[3];   // d Bx / dz
         g6[5] = g6[5] + dz * v[4];   // d By / dz
      }
   }
   const R fv = R(g.f_vel), fm = R(g.f_mag), cf = R(C_CGS / 1.0E7);
   const v3<R> Uf = mk<R>(acc[0], acc[1], acc[2]), Bf = mk<R>(acc[3], acc[4], acc[5]);
   const v3<R> Ef = (-cross(Uf, Bf)) / cf;
   U = mk<R>(Uf.x * fv, Uf.y * fv, Uf.z * fv);
   B = mk<R>(Bf.x * fm, Bf.y * fm, Bf.z * fm);
   E = mk<R>(Ef.x * fm, Ef.y * fm, Ef.z * fm);
   Bmag = acc[6] * fm;
   gradBmag = mk<R>(gm[0] * fm, gm[1] * fm, gm[2] * fm);
#pragma unroll
   for (int m = 0; m < 6; m++) gB[m] = g6[m] * fm;
   return true;
}

// All the fields the server gives for the full orbit (GRADS = false: U, B, the interpolated |B|, E) and the
// focused transport (GRADS = true: also grad U and grad B row by row, gU[i] = d U / d x_i, and grad |B|),
// from the stencil's weights and derivative weights (GetVariablesInterp1 / GetGradientsInterp1,
// src/server_cartesian.cc:538-609, :630-705); E = -(U x B) / c in fluid units before the unit conversion.
template <typename R, bool GRADS>
__device__ __forceinline__ bool grid_eval_full(const SPX_CP spx_grid_desc& g, double x, double y, double z, v3<R>& U, v3<R>& B, R& Bmag,
                                               v3<R>& E, v3<R> (&gU)[3], v3<R> (&gB)[3], v3<R>& gradBmag, double& zl_min)
{
   R lo[3], hi[3], dl[3];
   int64_t zone0;
   if (!grid_locate<R>(g, x, y, z, lo, hi, dl, zone0, zl_min)) return false;
   R acc[7], gu[9], gb[9], gm[3];
#pragma unroll
   for (int m = 0; m < 7; m++) acc[m] = R(0);
#pragma unroll
   for (int m = 0; m < 9; m++) { gu[m] = R(0); gb[m] = R(0); }
#pragma unroll
   for (int m = 0; m < 3; m++) gm[m] = R(0);
#pragma unroll 1
   for (int r = 0; r < 4; r++) {
      const int j = r & 1, k = r >> 1;
      R v0[7], v1[7];
      grid_corner<R>(g, zone0, 2 * r, v0);
      grid_corner<R>(g, zone0, 2 * r + 1, v1);
      const R fy = j ? lo[1] : hi[1], fz = k ? lo[2] : hi[2];
#pragma unroll
      for (int i = 0; i < 2; i++) {
         const R* v = i ? v1 : v0;
         const R fx = i ? lo[0] : hi[0];
         const R w = fx * fy * fz;
#pragma unroll
         for (int m = 0; m < 7; m++) acc[m] = acc[m] + w * v[m];
         if (GRADS) {
            const R sx = i ? R(1) : R(-1), sy_ = j ? R(1) : R(-1), sz_ = k ? R(1) : R(-1);
            const R d[3] = {sx * (fy * fz) / dl[0], sy_ * (fx * fz) / dl[1], sz_ * (fx * fy) / dl[2]};
#pragma unroll
            for (int a = 0; a < 3; a++) {
#pragma unroll
               for (int c = 0; c < 3; c++) {
                  gu[3 * a + c] = gu[3 * a + c] + d[a] * v[c];
                  gb[3 * a + c] = gb[3 * a + c] + d[a] * v[3 + c];
               }
               gm[a] = gm[a] + d[a] * v[6];
            }
         }
      }
   }
   const R fv = R(g.f_vel), fm = R(g.f_mag), cf = R(C_CGS / 1.0E7);
   const v3<R> Uf = mk<R>(acc[0], acc[1], acc[2]), Bf = mk<R>(acc[3], acc[4], acc[5]);
   const v3<R> Ef = (-cross(Uf, Bf)) / cf;
   U = mk<R>(Uf.x * fv, Uf.y * fv, Uf.z * fv);
   B = mk<R>(Bf.x * fm, Bf.y * fm, Bf.z * fm);
   E = mk<R>(Ef.x * fm, Ef.y * fm, Ef.z * fm);
   Bmag = acc[6] * fm;
   if (GRADS) {
#pragma unroll
      for (int a = 0; a < 3; a++) {
         gU[a] = mk<R>(gu[3 * a] * fv, gu[3 * a + 1] * fv, gu[3 * a + 2] * fv);
         gB[a] = mk<R>(gb[3 * a] * fm, gb[3 * a + 1] * fm, gb[3 * a + 2] * fm);
      }
      gradBmag = mk<R>(gm[0] * fm, gm[1] * fm, gm[2] * fm);
   }
   return true;
}

// ---- diffusion: DiffusionBase::GetComponent (src/diffusion_base.cc:105-121) + models ----------------
SPX_HD double rcbrt(double x) { return cbrt(x); }
SPX_HD float rcbrt(float x) { return cbrtf(x); }

// DiffusionStraussEtAl2013 (src/diffusion_other.cc:886-906) and DiffusionPotgieterEtAl2015 (:968-988):
// kappa_par = (lam v / 3) f(R) (B0_eff / |B|), resp. kappa (v / c) g(R) (B0_eff / |B|); this is the factor
// before (B0_eff / |B|) for the inner (heliosphere) or outer (LISM) parameters, in the reference's order.
template <typename R>
SPX_HD R kappa_mom_region(const SPX_CP spx_df_t<R>& d, R p, bool outer)
{
   const R lam = outer ? d.lam_out : d.lam_in;
   const R rig = Rigidity(p), vmag = VelP(p);
   if (d.kind == SPX_DF_STRAUSS) return (lam * vmag / R(3)) * (rig < d.R0 ? rcbrt(rig / d.R0) : rig / d.R0);
   const R x = rig / d.R0;
   return lam * (vmag / R(c_code)) * (rig < d.R0 ? R(1) : rsqrt_(x * x * x));
}

// DiffusionEmpiricalSOQLTandUNLT (src/diffusion_other.cc:1053-1072): kappa = (lam v / 3) f(R) (B0 / |B|) / S
// with the bent power laws f = cbrt(x (1 + x^4)) (parallel), cbrt(x (1 + x^2)) (perpendicular), x = R / R0.
// This is the factor before (B0 / |B|), for the parallel component or for the perpendicular one with the
// mean free path lam_perp (reduced = false) or lam_perp kap_rat_red (reduced = true), in the reference's order.
template <typename R>
SPX_HD R kappa_mom_soqlt(const SPX_CP spx_df_t<R>& d, R p, int comp, bool reduced)
{
   const R rig = Rigidity(p), vmag = VelP(p);
   const R x = rig / d.R0;
   if (comp == 1) return (d.lam_para * vmag / R(3)) * rcbrt(x * (R(1) + sqr(sqr(x))));
   // lam_perp (Bmix_ind + (1 - Bmix_ind) kap_rat_red) with Bmix_ind 1 or 0 is lam_perp or lam_perp kap_rat_red exactly
   const R lam = reduced ? d.lam_perp * d.kap_rat_red : d.lam_perp;
   return (lam * vmag / R(3)) * rcbrt(x * (R(1) + sqr(x)));
}

// The momentum-only factor of kappa_par: the same at every point of a step's derivative stencil, so
// it is formed once per step (the same inputs give the same value).
template <typename R>
SPX_HD R kappa_mom(const SPX_CP spx_df_t<R>& d, R p)
{
   switch (d.kind) {
   case SPX_DF_FLOWMOM: return rpow(p / d.p0, d.pow_law_p);                                   // :622-627
   case SPX_DF_KINRAD: return d.kap0 * rpow(EnrKin(p) / d.T0, d.pow_law_T);                   // :713-723
   case SPX_DF_RIGBMAG: return divc<R>(d.lam0 * VelP(p), R(3)) * rpow(Rigidity(p) / d.R0, d.pow_law_R);   // :804-809
   case SPX_DF_STRAUSS: case SPX_DF_POTGIETER: return kappa_mom_region(d, p, false);
   case SPX_DF_SOQLT: return kappa_mom_soqlt(d, p, 1, false);
   default: return R(0);
   }
}

// Kappa components 0 (perp) and 1 (para) at one point, in the reference's operation order; km is
// kappa_mom, r is |x| (KINRAD), Umag |U|.
template <typename R>
SPX_HD void kappa01(const SPX_CP spx_df_t<R>& d, R km, R Bmag, R r, R Umag, R& K0, R& K1)
{
   switch (d.kind) {
   case SPX_DF_PARA: K1 = d.D0; break;                                      // :381-385
   case SPX_DF_PERP: K0 = d.D0; break;                                      // :459-463
   case SPX_DF_FULL: K0 = d.Dperp; K1 = d.Dpara; break;                     // :538-543
   case SPX_DF_FLOWMOM:                                                      // :622-627
      K1 = d.kap0 * rpow(Umag / d.U0, d.pow_law_U) * km;
      K0 = d.kap_rat * K1;
      break;
   case SPX_DF_KINRAD:                                                       // :713-723
      if (d.stream_dep_idx == 0 || r < d.r0) K1 = km * rpow(r / d.r0, d.pow_law_r);
      else if (r < d.r0 + d.w_sh) K1 = km * sqr(Umag / d.u_up);
      else K1 = km * d.dn_up_rat;
      K0 = d.kap_rat * K1;
      break;
   case SPX_DF_RIGBMAG:                                                      // :804-809
      K1 = km * rpow(Bmag / d.B0, d.pow_law_B);
      K0 = d.kap_rat * K1;
      break;
   case SPX_DF_PERPISO: K0 = d.Dperp; break;
   default: break;
   }
}

// ---- counter-based RNG: Philox4x32-10 (Salmon et al., SC'11) ------------------------------------
SPX_HD void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4])
{
#pragma unroll
   for (int r = 0; r < 10; r++) {
      uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
      uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
      uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
      uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
      c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
      k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
   }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

SPX_HD void philox_block(uint64_t seed, uint64_t traj, uint64_t draw, uint32_t kind, uint32_t o[4])
{
   philox4x32_10((uint32_t)draw, ((uint32_t)(draw >> 32) & 0x0FFFFFFFu) | (kind << 28), (uint32_t)traj, (uint32_t)(traj >> 32),
                 (uint32_t)seed, (uint32_t)(seed >> 32), o);
}

// uniform in (0,1]
SPX_HD double u53(uint32_t lo, uint32_t hi)
{
   uint64_t x = ((uint64_t)hi << 32) | lo;
   return (double)((x >> 11) + 1) * 0x1.0p-53;
}
SPX_HD float u24(uint32_t x) { return (float)((x >> 8) + 1) * 0x1.0p-24f; }

SPX_HD double counter_uniform(uint64_t seed, uint64_t traj, uint64_t draw)
{
   uint32_t o[4];
   philox_block(seed, traj, draw, 1u, o);
   return u53(o[0], o[1]);
}

// FP64 step: Box-Muller pairs, normal draw k comes from Philox block k/2 of the normal stream; even k
// takes the cosine branch, odd k the sine branch of the same radius and angle (53-bit uniforms).
__device__ __forceinline__ void normal_pair_f64(uint64_t seed, uint64_t traj, uint64_t block, double& c, double& s)
{
   uint32_t o[4];
   philox_block(seed, traj, block, 2u, o);
   const double u1 = u53(o[0], o[1]), u2 = u53(o[2], o[3]);
   const double r = sqrt_rn(-2.0 * log(u1));   // -2 ln(u1) is 0 or above 2^-53
   double sn, cs;
   sincospi(2.0 * u2, &sn, &cs);
   c = r * cs;
   s = r * sn;
}
// FP32 step: one Philox block gives four normals, two Box-Muller pairs (o[0], o[1]) and (o[2], o[3]).
// The radius uniform is the whole 32-bit word, (o + 1/2) 2^-32 in (0, 1]: its smallest value 2^-33 cuts
// the normal tails at sqrt(66 ln 2) = 6.76 sigma (probability 1.4e-11 per draw); the angle uses 24 bits.
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& c, float& s)
{
   const float u1 = __builtin_fmaf((float)a, 0x1.0p-32f, 0x1.0p-33f);
   const float u2 = u24(b);
   const float r = __builtin_sqrtf(-2.0f * __logf(u1));
   float sn, cs;
   __sincosf(6.2831853071795864f * u2, &sn, &cs);
   c = r * cs;
   s = r * sn;
}
__device__ __forceinline__ void normal_quad_f32(uint64_t seed, uint64_t traj, uint64_t block, float& n0, float& n1, float& n2, float& n3)
{
   uint32_t o[4];
   philox_block(seed, traj, block, 2u, o);
   box_muller_f32(o[0], o[1], n0, n1);
   box_muller_f32(o[2], o[3], n2, n3);
}

}  // namespace spx
