// spx_api.cpp -- host runtime behind the C ABI (include/spectrum_amd.h).
//
// Replaces the reference's MPI master/worker task farm (src/simulation.cc:253-305, 743-835) with a
// single-GPU scheduler: trajectories are queued on a device counter, resident lane slots are refilled
// by the persistent kernels, and distributions accumulate in device histograms.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "spx_kernels.h"

using namespace spx;

namespace {

constexpr double C_CGS_H = 2.99792458e+10;
constexpr double unit_velocity_h = 0.1 * C_CGS_H;
constexpr double c_code_h = C_CGS_H / unit_velocity_h;
constexpr double sp_tiny_h = 1.0E-15;

// ---- DataContainer reader (common/data_container.cc:214-238): [u8 len][len bytes] records -------------
struct Reader {
   const uint8_t* p;
   const uint8_t* end;
   bool err = false;
   void raw(void* dst, size_t want)
   {
      if (err || p >= end) { err = true; return; }
      size_t len = *p;
      if (p + 1 + len > end || len != want) { err = true; return; }
      std::memcpy(dst, p + 1, len);
      p += len + 1;
   }
   double d() { double x = 0.0; raw(&x, 8); return x; }
   int32_t i() { int32_t x = 0; raw(&x, 4); return x; }
   bool b() { uint8_t x = 0; raw(&x, 1); return x != 0; }
   void v(double* out) { out[0] = out[1] = out[2] = 0.0; raw(out, 24); }
   void mi(int32_t* out) { out[0] = out[1] = out[2] = 0; raw(out, 12); }
   int vi(int32_t* out, int maxn)
   {
      if (err || p >= end) { err = true; return 0; }
      size_t len = *p;
      if (p + 1 + len > end || (len % 4) || (int)(len / 4) > maxn) { err = true; return 0; }
      std::memcpy(out, p + 1, len);
      p += len + 1;
      return (int)(len / 4);
   }
   void skip()
   {
      if (err || p >= end) { err = true; return; }
      size_t len = *p;
      if (p + 1 + len > end) { err = true; return; }
      p += len + 1;
   }
};

inline void vcross(const double* a, const double* b, double* o)
{
   double x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
   o[0] = x; o[1] = y; o[2] = z;
}
inline double vdot(const double* a, const double* b) { double s = a[0] * b[0]; s += a[1] * b[1]; s += a[2] * b[2]; return s; }
inline double vnorm(const double* a) { return std::sqrt(vdot(a, a)); }
inline void vunit(double* a) { double n = vnorm(a); a[0] /= n; a[1] /= n; a[2] /= n; }
inline void second_unit(const double* f, double* o)
{
   int k;
   if (std::fabs(f[0]) < std::fabs(f[1])) k = (std::fabs(f[0]) < std::fabs(f[2])) ? 0 : 2;
   else k = (std::fabs(f[1]) < std::fabs(f[2])) ? 1 : 2;
   double e[3] = {k == 0 ? 1.0 : 0.0, k == 1 ? 1.0 : 0.0, k == 2 ? 1.0 : 0.0};
   vcross(f, e, o);
   vunit(o);
}
inline void induced_e(const double* u, const double* B, double* E)
{
   double c[3];
   vcross(u, B, c);
   E[0] = -c[0] / c_code_h; E[1] = -c[1] / c_code_h; E[2] = -c[2] / c_code_h;
}

template <typename T>
struct DevBuf {
   T* p = nullptr;
   size_t n = 0;
   hipError_t alloc(size_t count)
   {
      if (count <= n && p) return hipSuccess;
      release();
      n = count;
      return hipMalloc((void**)&p, sizeof(T) * (count ? count : 1));
   }
   void release()
   {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
   }
};

}  // namespace

struct spx_sim {
   spx_plan_desc plan;
   bool have_bg = false, have_df = false, have_it = false, have_is = false, have_im = false;
   int precision = SPX_PREC_FP64;
   int block = 256;
   int64_t lanes_opt = 0;
   int steps_per_launch = 2000;
   int graph_launches = 0;
   std::string err;
   // device resources
   bool dev_ready = false;
   hipStream_t stream = nullptr;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   std::vector<hipEvent_t> tev;        // per-launch events: [2i] before launch i, [2i+1] between its two kernels
   double step_ms = 0.0, service_ms = 0.0;
   spx_plan_desc* d_plan = nullptr;
   RunCtl* d_ctl = nullptr;
   Launch* d_launch = nullptr;
   uint64_t n_lanes = 0;
   DevBuf<int32_t> active, bcross, bstat, attempt, n_events, ev_bin, ev_bnd, ev_pending, ex_code;
   DevBuf<double> t, x, y, z, m0, m1, m2, amp, wgt, dt_prev, src_old, ds_prev, dmax, dt_adapt, pf, bdelta, bdelta_old;
   DevBuf<double> t0, x0, y0, z0, m00, m01, m02, ev_w, evs;
   DevBuf<uint64_t> traj;
   DevBuf<uint32_t> nn, nu;
   DevBuf<int64_t> steps;
   DevBuf<unsigned long long> hist_counts, hist_events;
   DevBuf<double> hist_sums;
   DevBuf<double> rep_normals, rep_uniforms;
   DevBuf<spx_traj_record> records;
   int64_t hist_bins = 0;
   // run state
   bool running = false;
   uint64_t run_first = 0, run_n = 0;
   spx_rng_desc run_rng{};
   bool keep_records = false;
   int64_t launches_total = 0;
   hipGraphExec_t graph_exec = nullptr;
   hipGraph_t graph = nullptr;
};

static int fail(spx_sim* s, int code, const std::string& msg)
{
   if (s) s->err = msg;
   return code;
}

#define HIPCHK(sim, expr)                                                                   \
   do {                                                                                     \
      hipError_t e_ = (expr);                                                               \
      if (e_ != hipSuccess) return fail((sim), SPX_ERR_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
   } while (0)

// ---------------------------------------------------------------------------------------------------
// Plugin lowering: the reference's Setup* methods read the container in this order.
// ---------------------------------------------------------------------------------------------------
static int add_background(spx_sim* s, const std::string& cls, Reader& r)
{
   spx_bg_desc& g = s->plan.bg;
   std::memset(&g, 0, sizeof(g));
   // BackgroundBase::SetupBackground (src/background_base.cc:334-352)
   g.t0 = r.d(); r.v(g.r0); r.v(g.u0); r.v(g.B0); g.dmax0 = r.d();
   if (cls == "BackgroundUniform") {   // src/background_uniform.cc:58-64
      g.kind = SPX_BG_UNIFORM;
      induced_e(g.u0, g.B0, g.Eu);
   }
   else if (cls == "BackgroundSolarWind") {   // src/background_solarwind.cc:58-84
      g.kind = SPX_BG_SOLARWIND;
      double Omega[3];
      r.v(Omega); g.r_ref = r.d(); g.dmax_fraction = r.d();
      g.w0 = vnorm(Omega);
      double e2[3] = {0.0, 0.0, 1.0};
      if (!(g.w0 < sp_tiny_h)) { e2[0] = Omega[0]; e2[1] = Omega[1]; e2[2] = Omega[2]; vunit(e2); }
      double e0[3], e1[3];
      second_unit(e2, e0);
      vcross(e2, e0, e1);
      for (int k = 0; k < 3; k++) { g.eprime[0][k] = e0[k]; g.eprime[1][k] = e1[k]; g.eprime[2][k] = e2[k]; }
      // a diagonal +-1 basis turns the rotations into sign flips with the same results
      g.ep_diag = 1;
      for (int k = 0; k < 3; k++)
         for (int j = 0; j < 3; j++)
            if ((j != k && g.eprime[k][j] != 0.0) || (j == k && std::fabs(g.eprime[k][j]) != 1.0)) g.ep_diag = 0;
      g.ur0 = std::fabs(g.u0[0]);
      g.hp_rad = 117.0 * 1.495978707E+13 / 1.0E10;
   }
   else if (cls == "BackgroundShock" || cls == "BackgroundSmoothShock") {   // src/background_shock.cc:58-86
      g.kind = (cls == "BackgroundShock") ? SPX_BG_SHOCK : SPX_BG_SMOOTHSHOCK;
      r.v(g.n_shock); g.v_shock = r.d(); g.compression = r.d();
      vunit(g.n_shock);
      const double* n = g.n_shock;
      double un = vdot(g.u0, n), bn = vdot(g.B0, n);
      double u0n[3], u0t[3], B0n[3], B0t[3];
      for (int k = 0; k < 3; k++) {
         u0n[k] = (un - g.v_shock) * n[k];
         u0t[k] = g.u0[k] - un * n[k];
         B0n[k] = bn * n[k];
      }
      for (int k = 0; k < 3; k++) B0t[k] = g.B0[k] - B0n[k];
      for (int k = 0; k < 3; k++) {
         g.u1[k] = (u0n[k] / g.compression + g.v_shock * n[k]) + u0t[k];
         g.B1[k] = B0n[k] + g.compression * B0t[k];
      }
      induced_e(g.u0, g.B0, g.E0);
      induced_e(g.u1, g.B1, g.E1);
      if (g.kind == SPX_BG_SMOOTHSHOCK) { g.width_shock = r.d(); g.dmax_fraction = r.d(); }   // src/background_smooth_shock.cc:139-146
   }
   else return fail(s, SPX_ERR_CLASS, "unknown background class " + cls);
   s->have_bg = true;
   return SPX_OK;
}

static int add_diffusion(spx_sim* s, const std::string& cls, Reader& r)
{
   spx_df_desc& d = s->plan.df;
   std::memset(&d, 0, sizeof(d));
   if (cls == "DiffusionIsotropicConstant") { d.kind = SPX_DF_ISO; d.D0 = r.d(); }
   else if (cls == "DiffusionParaConstant") { d.kind = SPX_DF_PARA; d.D0 = r.d(); }
   else if (cls == "DiffusionPerpConstant") { d.kind = SPX_DF_PERP; d.D0 = r.d(); }
   else if (cls == "DiffusionFullConstant") { d.kind = SPX_DF_FULL; d.Dperp = r.d(); d.Dpara = r.d(); }
   else if (cls == "DiffusionFlowMomentumPowerLaw") {
      d.kind = SPX_DF_FLOWMOM;
      d.kap0 = r.d(); d.U0 = r.d(); d.pow_law_U = r.d(); d.p0 = r.d(); d.pow_law_p = r.d(); d.kap_rat = r.d();
   }
   else if (cls == "DiffusionKineticEnergyRadialDistancePowerLaw") {
      d.kind = SPX_DF_KINRAD;
      d.kap0 = r.d(); d.T0 = r.d(); d.r0 = r.d(); d.pow_law_T = r.d(); d.pow_law_r = r.d(); d.kap_rat = r.d();
      d.stream_dep_idx = r.i(); d.u_up = r.d(); d.w_sh = r.d(); d.s_sh = r.d();
      d.dn_up_rat = (1.0 / d.s_sh) * (1.0 / d.s_sh);
   }
   else if (cls == "DiffusionRigidityMagneticFieldPowerLaw") {
      d.kind = SPX_DF_RIGBMAG;
      d.lam0 = r.d(); d.R0 = r.d(); d.B0 = r.d(); d.pow_law_R = r.d(); d.pow_law_B = r.d(); d.kap_rat = r.d();
   }
   else if (cls == "DiffusionPerpConstantPitchIsotropic") { d.kind = SPX_DF_PERPISO; d.Dperp = r.d(); d.D0 = r.d(); }
   else return fail(s, SPX_ERR_CLASS, "unknown diffusion class " + cls);
   s->have_df = true;
   return SPX_OK;
}

static int add_boundary(spx_sim* s, const std::string& cls, Reader& r)
{
   spx_bnd_desc b;
   std::memset(&b, 0, sizeof(b));
   // BoundaryBase::SetupBoundary (src/boundary_base.cc:73-83)
   b.max_crossings = r.i();
   b.n_act = r.vi(b.actions, SPX_MAX_ACT);
   if (cls == "BoundaryTimeExpire") { b.cls = SPX_BC_TIME; b.kind = SPX_BK_TIME; b.timemark = r.d(); b.max_crossings = 1; }
   else if (cls == "BoundaryTimePass") { b.cls = SPX_BC_TIME; b.kind = SPX_BK_TIME; b.timemark = r.d(); b.max_crossings = -1; }
   else if (cls == "BoundaryTimeRecurrent") { b.cls = SPX_BC_TIME; b.kind = SPX_BK_TIME; b.timemark = r.d(); b.status0 = SPX_BND_RECURRENT; }
   else if (cls == "BoundarySphereAbsorb") { b.cls = SPX_BC_SPACE; b.kind = SPX_BK_SPHERE; r.v(b.origin); b.radius = r.d(); b.max_crossings = 1; }
   else if (cls == "BoundarySphereReflect") { b.cls = SPX_BC_SPACE; b.kind = SPX_BK_SPHERE; r.v(b.origin); b.radius = r.d(); b.status0 = SPX_BND_REFLECT; }
   else if (cls == "BoundaryPlaneAbsorb" || cls == "BoundaryPlaneReflect" || cls == "BoundaryPlanePass") {
      b.cls = SPX_BC_SPACE; b.kind = SPX_BK_PLANE; r.v(b.origin); r.v(b.norm); vunit(b.norm);
      if (cls == "BoundaryPlaneAbsorb") b.max_crossings = 1;
      if (cls == "BoundaryPlaneReflect") b.status0 = SPX_BND_REFLECT;
   }
   else if (cls == "BoundaryMomentumInject") { b.cls = SPX_BC_MOMENTUM; b.kind = SPX_BK_MOMENTUM; b.momentum = r.d(); b.max_crossings = 1; }
   else if (cls == "BoundaryMomentumPass") { b.cls = SPX_BC_MOMENTUM; b.kind = SPX_BK_MOMENTUM; b.momentum = r.d(); }
   else return fail(s, SPX_ERR_CLASS, "unknown boundary class " + cls);
   if (s->plan.n_bnd >= SPX_MAX_BND) return fail(s, SPX_ERR_LIMIT, "more than 4 boundaries");
   // keep slots ordered time, space, momentum (the order of ComputeAllBoundaries, src/trajectory_base.cc:87-93)
   int pos = s->plan.n_bnd;
   while (pos > 0 && s->plan.bnd[pos - 1].cls > b.cls) { s->plan.bnd[pos] = s->plan.bnd[pos - 1]; pos--; }
   s->plan.bnd[pos] = b;
   s->plan.n_bnd++;
   if (b.cls == SPX_BC_TIME) s->plan.n_bt++;
   else if (b.cls == SPX_BC_SPACE) s->plan.n_bs++;
   else s->plan.n_bm++;
   return SPX_OK;
}

static int add_initial(spx_sim* s, const std::string& cls, Reader& r)
{
   spx_plan_desc& P = s->plan;
   if (cls == "InitialTimeFixed") { P.it_t = r.d(); s->have_it = true; }                        // src/initial_time.cc:48-58
   else if (cls == "InitialSpaceFixed") { r.v(P.is_pos); s->have_is = true; }                    // src/initial_space.cc:49-59
   else if (cls == "InitialMomentumFixed") { P.im_kind = SPX_IM_FIXED; r.v(P.im_mom); s->have_im = true; }
   else if (cls == "InitialMomentumShell") { P.im_kind = SPX_IM_SHELL; P.im_p0 = r.d(); s->have_im = true; }   // src/initial_momentum.cc:283-295
   else if (cls == "InitialMomentumRing") {                                                       // src/initial_momentum.cc:203-220
      P.im_kind = SPX_IM_RING; P.im_p0 = r.d();
      const double th = r.d();
      P.im_mu0 = std::cos(th); P.im_st0 = std::sin(th);
      s->have_im = true;
   }
   else if (cls == "InitialMomentumThickShell") {                                                 // src/initial_momentum.cc:363-377
      P.im_kind = SPX_IM_THICKSHELL; P.im_p1 = r.d(); P.im_p2 = r.d(); P.im_log = r.b() ? 1 : 0;
      if (P.im_log) { P.im_p1 = std::log10(P.im_p1); P.im_p2 = std::log10(P.im_p2); }
      s->have_im = true;
   }
   else if (cls == "InitialMomentumShellGrid") {
      P.im_kind = SPX_IM_SHELLGRID; P.im_T1 = r.d(); P.im_T2 = r.d(); P.im_n = r.i();
      if (P.im_n < 1) return fail(s, SPX_ERR_CONTAINER, "InitialMomentumShellGrid needs n >= 1");
      s->have_im = true;
   }
   else return fail(s, SPX_ERR_CLASS, "unknown initial class " + cls);
   return SPX_OK;
}

static int add_distribution(spx_sim* s, const std::string& cls, Reader& r)
{
   spx_plan_desc& P = s->plan;
   if (P.n_dist >= SPX_MAX_DIST) return fail(s, SPX_ERR_LIMIT, "more than 4 distributions");
   spx_ds_desc& D = P.dist[P.n_dist];
   std::memset(&D, 0, sizeof(D));
   int width = 1;
   if (cls == "DistributionPositionCumulativeOrder1") width = 3;
   else if (cls == "DistributionPositionCumulativeOrder2") width = 9;
   // DistributionTemplated::SetupDistribution (src/distribution_templated.cc:92-139)
   double minval[3], maxval[3], unit_val[3];
   r.mi(D.n_bins); r.v(minval); r.v(maxval); r.mi(D.log_bins); r.mi(D.bin_outside);
   r.skip();   // unit_distro
   r.v(unit_val); (void)r.b();
   D.dims = 0;
   for (int k = 0; k < 3; k++) {
      if (D.n_bins[k] > 0) D.dims |= (1 << k);
      if (D.n_bins[k] < 1) D.n_bins[k] = 1;
   }
   for (int k = 0; k < 3; k++) {
      double l0 = D.log_bins[k] ? std::log10(minval[k]) : minval[k];
      double l1 = D.log_bins[k] ? std::log10(maxval[k]) : maxval[k];
      D.limits0[k] = l0;
      D.bin_size[k] = (l1 - l0) / D.n_bins[k];
   }
   int need = 1;
   if (cls == "DistributionSpectrumKineticEnergyPowerLaw") { D.kind = SPX_DS_SPECTRUM; D.J0 = r.d(); D.T0 = r.d(); D.pow_law = r.d(); D.val_cold = r.d(); }
   else if (cls == "DistributionTimeUniform") { D.kind = SPX_DS_TIME; D.val_hot = r.d(); D.val_cold = r.d(); D.val_time = r.i(); }
   else if (cls == "DistributionMomentumUniform") { D.kind = SPX_DS_MOMENTUM; D.val_hot = r.d(); D.val_cold = r.d(); D.val_time = r.i(); D.val_coord = r.i(); need = 7; }
   else if (cls == "DistributionPositionMomentumUniform") {
      D.kind = SPX_DS_POSMOM; D.val_hot = r.d(); D.val_cold = r.d(); D.val_time = r.i(); D.pos_idx = r.i(); D.mom_idx = r.i(); need = 5;
      if (D.pos_idx < 0 || D.pos_idx > 3 || D.mom_idx < 0 || D.mom_idx > 2) return fail(s, SPX_ERR_CONTAINER, "bad coordinate index");
   }
   else if (cls == "DistributionPositionPitchUniform") { D.kind = SPX_DS_POSPITCH; D.val_hot = r.d(); D.val_cold = r.d(); D.val_time = r.i(); need = 3; }
   else if (cls == "DistributionPositionCumulativeOrder1") D.kind = SPX_DS_CUMO1;
   else if (cls == "DistributionPositionCumulativeOrder2") D.kind = SPX_DS_CUMO2;
   else return fail(s, SPX_ERR_CLASS, "unknown distribution class " + cls);
   if (D.dims != need) return fail(s, SPX_ERR_CONTAINER, cls + ": wrong set of active dimensions");
   D.width = width;
   D.nb = D.n_bins[0] * D.n_bins[1] * D.n_bins[2];
   if (D.nb <= 0 || D.nb > (1 << 24)) return fail(s, SPX_ERR_LIMIT, "distribution size out of range");
   P.hist_off[P.n_dist] = s->hist_bins;
   s->hist_bins += D.nb;
   P.n_dist++;
   return SPX_OK;
}

// ---------------------------------------------------------------------------------------------------
// Device resources
// ---------------------------------------------------------------------------------------------------
static int ensure_device(spx_sim* s)
{
   if (s->dev_ready) return SPX_OK;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return fail(s, SPX_ERR_DEVICE, "no HIP device");
   HIPCHK(s, hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
   HIPCHK(s, hipEventCreate(&s->ev0));
   HIPCHK(s, hipEventCreate(&s->ev1));
   HIPCHK(s, hipMalloc((void**)&s->d_plan, sizeof(spx_plan_desc)));
   HIPCHK(s, hipMalloc((void**)&s->d_ctl, sizeof(RunCtl)));
   HIPCHK(s, hipMalloc((void**)&s->d_launch, sizeof(Launch)));
   s->dev_ready = true;
   return SPX_OK;
}

static uint64_t auto_lanes(spx_sim* s)
{
   hipDeviceProp_t prop;
   if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 65536;
   // a few waves per SIMD keep the VALU busy; the state is register resident
   uint64_t per_cu = 4 * 4 * 64;   // 4 SIMDs x 4 waves x 64 lanes
   return (uint64_t)prop.multiProcessorCount * per_cu;
}

static int alloc_lanes(spx_sim* s, uint64_t n)
{
   if (s->n_lanes == n && s->active.p) return SPX_OK;
   s->n_lanes = n;
   HIPCHK(s, s->active.alloc(n)); HIPCHK(s, s->attempt.alloc(n)); HIPCHK(s, s->n_events.alloc(n));
   HIPCHK(s, s->bcross.alloc(n * SPX_MAX_BND)); HIPCHK(s, s->bstat.alloc(n * SPX_MAX_BND));
   HIPCHK(s, s->ev_bin.alloc(n * SPX_MAX_DIST)); HIPCHK(s, s->ev_w.alloc(n * SPX_MAX_DIST));
   HIPCHK(s, s->evs.alloc(n * 3 * 9)); HIPCHK(s, s->ev_bnd.alloc(n * 3)); HIPCHK(s, s->ev_pending.alloc(n)); HIPCHK(s, s->ex_code.alloc(n));
   DevBuf<double>* dv[] = {&s->t, &s->x, &s->y, &s->z, &s->m0, &s->m1, &s->m2, &s->amp, &s->wgt, &s->dt_prev, &s->src_old,
                           &s->ds_prev, &s->dmax, &s->dt_adapt, &s->pf, &s->t0, &s->x0, &s->y0, &s->z0, &s->m00, &s->m01, &s->m02};
   for (auto* b : dv) HIPCHK(s, b->alloc(n));
   HIPCHK(s, s->bdelta.alloc(n * SPX_MAX_BND)); HIPCHK(s, s->bdelta_old.alloc(n * SPX_MAX_BND));
   HIPCHK(s, s->traj.alloc(n)); HIPCHK(s, s->nn.alloc(n)); HIPCHK(s, s->nu.alloc(n)); HIPCHK(s, s->steps.alloc(n));
   return SPX_OK;
}

static Launch make_launch(spx_sim* s)
{
   Launch L;
   std::memset(&L, 0, sizeof(L));
   L.ctl = s->d_ctl;
   L.rng = s->run_rng;
   L.first = s->run_first;
   L.end = s->run_first + s->run_n;
   L.n_lanes = s->n_lanes;
   L.steps_per_launch = s->steps_per_launch;
   L.n_dist = s->plan.n_dist;
   for (int d = 0; d < SPX_MAX_DIST; d++) L.plan_host_hist_off[d] = s->plan.hist_off[d];
   L.active = s->active.p;
   L.t = s->t.p; L.x = s->x.p; L.y = s->y.p; L.z = s->z.p; L.m0 = s->m0.p; L.m1 = s->m1.p; L.m2 = s->m2.p;
   L.amp = s->amp.p; L.wgt = s->wgt.p; L.dt_prev = s->dt_prev.p; L.src_old = s->src_old.p; L.ds_prev = s->ds_prev.p;
   L.dmax = s->dmax.p; L.dt_adapt = s->dt_adapt.p; L.pf = s->pf.p;
   L.bdelta = s->bdelta.p; L.bdelta_old = s->bdelta_old.p; L.bcross = s->bcross.p; L.bstat = s->bstat.p;
   L.traj = s->traj.p; L.nn = s->nn.p; L.nu = s->nu.p; L.attempt = s->attempt.p; L.steps = s->steps.p; L.n_events = s->n_events.p;
   L.t0 = s->t0.p; L.x0 = s->x0.p; L.y0 = s->y0.p; L.z0 = s->z0.p; L.m00 = s->m00.p; L.m01 = s->m01.p; L.m02 = s->m02.p;
   L.ev_bin = s->ev_bin.p; L.ev_w = s->ev_w.p;
   L.evs = s->evs.p; L.ev_bnd = s->ev_bnd.p; L.ev_pending = s->ev_pending.p; L.ex_code = s->ex_code.p;
   L.hist_counts = s->hist_counts.p; L.hist_sums = s->hist_sums.p; L.hist_events = s->hist_events.p;
   L.records = s->keep_records ? s->records.p : nullptr;
   L.rec_first = s->run_first;
   L.rec_n = s->keep_records ? s->run_n : 0;
   return L;
}

// Float copies of the background / diffusion parameters for the mixed-precision kernel, and the
// position-dependence flag of kappa (pow(x, 0) == 1 exactly, so a zero exponent removes it bit for bit).
static void finalize_plan(spx_sim* s)
{
   spx_plan_desc& P = s->plan;
   const spx_bg_t<double>& g = P.bg;
   spx_bg_t<float>& f = P.bgf;
   std::memset(&f, 0, sizeof(f));
   f.kind = g.kind; f.ep_diag = g.ep_diag; f.t0 = g.t0;
   for (int k = 0; k < 3; k++) {
      f.r0[k] = g.r0[k]; f.u0[k] = (float)g.u0[k]; f.B0[k] = (float)g.B0[k]; f.n_shock[k] = (float)g.n_shock[k];
      f.u1[k] = (float)g.u1[k]; f.B1[k] = (float)g.B1[k]; f.E0[k] = (float)g.E0[k]; f.E1[k] = (float)g.E1[k]; f.Eu[k] = (float)g.Eu[k];
      for (int j = 0; j < 3; j++) f.eprime[k][j] = (float)g.eprime[k][j];
   }
   f.dmax0 = (float)g.dmax0; f.r_ref = (float)g.r_ref; f.dmax_fraction = (float)g.dmax_fraction; f.ur0 = (float)g.ur0;
   f.w0 = (float)g.w0; f.hp_rad = (float)g.hp_rad; f.v_shock = (float)g.v_shock; f.compression = (float)g.compression;
   f.width_shock = (float)g.width_shock;
   spx_df_t<double>& d = P.df;
   d.pos_dep = 0;
   if (d.kind == SPX_DF_FLOWMOM) d.pos_dep = (d.pow_law_U != 0.0);
   if (d.kind == SPX_DF_KINRAD) d.pos_dep = 1;
   if (d.kind == SPX_DF_RIGBMAG) d.pos_dep = (d.pow_law_B != 0.0);
   spx_df_t<float>& e = P.dff;
   e.kind = d.kind; e.stream_dep_idx = d.stream_dep_idx; e.pos_dep = d.pos_dep;
   e.D0 = (float)d.D0; e.Dperp = (float)d.Dperp; e.Dpara = (float)d.Dpara; e.kap0 = (float)d.kap0; e.U0 = (float)d.U0;
   e.pow_law_U = (float)d.pow_law_U; e.p0 = (float)d.p0; e.pow_law_p = (float)d.pow_law_p; e.kap_rat = (float)d.kap_rat;
   e.T0 = (float)d.T0; e.r0 = (float)d.r0; e.pow_law_T = (float)d.pow_law_T; e.pow_law_r = (float)d.pow_law_r;
   e.u_up = (float)d.u_up; e.w_sh = (float)d.w_sh; e.s_sh = (float)d.s_sh; e.dn_up_rat = (float)d.dn_up_rat;
   e.lam0 = (float)d.lam0; e.R0 = (float)d.R0; e.B0 = (float)d.B0; e.pow_law_R = (float)d.pow_law_R; e.pow_law_B = (float)d.pow_law_B;
}

static int check_ready(spx_sim* s)
{
   // TrajectoryBase::IsSimmulationReady (src/trajectory_base.cc:608-632), TrajectoryParker (src/trajectory_parker.cc:44-52)
   if (!s->have_bg) return fail(s, SPX_ERR_NOT_READY, "no background");
   if (!s->have_it || !s->have_is || !s->have_im) return fail(s, SPX_ERR_NOT_READY, "initial condition missing");
   if (s->plan.n_bt < 1) return fail(s, SPX_ERR_NOT_READY, "at least one time boundary is required");
   if (!s->have_df) return fail(s, SPX_ERR_NOT_READY, "no diffusion model");
   if (s->plan.traj_type == SPX_TRAJ_PARKER_SOURCE && s->plan.src_kind == SPX_SRC_NONE) return fail(s, SPX_ERR_NOT_READY, "no source");
   return SPX_OK;
}

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

int spx_create(int traj_type, int time_flow, spx_sim** out)
{
   if (!out || traj_type < 0 || traj_type > 2 || (time_flow != 0 && time_flow != 1)) return SPX_ERR_ARG;
   spx_sim* s = new spx_sim();
   std::memset(&s->plan, 0, sizeof(s->plan));
   s->plan.traj_type = traj_type;
   s->plan.time_flow = time_flow;
   s->plan.max_steps = 10000000;
   s->plan.max_attempts = 16;
   *out = s;
   return SPX_OK;
}

void spx_destroy(spx_sim* s)
{
   if (!s) return;
   if (s->graph_exec) (void)hipGraphExecDestroy(s->graph_exec);
   if (s->graph) (void)hipGraphDestroy(s->graph);
   DevBuf<int32_t>* di[] = {&s->active, &s->bcross, &s->bstat, &s->attempt, &s->n_events, &s->ev_bin, &s->ev_bnd, &s->ev_pending, &s->ex_code};
   for (auto* b : di) b->release();
   DevBuf<double>* dv[] = {&s->t, &s->x, &s->y, &s->z, &s->m0, &s->m1, &s->m2, &s->amp, &s->wgt, &s->dt_prev, &s->src_old, &s->ds_prev,
                           &s->dmax, &s->dt_adapt, &s->pf, &s->bdelta, &s->bdelta_old, &s->t0, &s->x0, &s->y0, &s->z0, &s->m00, &s->m01,
                           &s->m02, &s->ev_w, &s->evs, &s->hist_sums, &s->rep_normals, &s->rep_uniforms};
   for (auto* b : dv) b->release();
   s->traj.release(); s->nn.release(); s->nu.release(); s->steps.release();
   s->hist_counts.release(); s->hist_events.release(); s->records.release();
   if (s->d_plan) (void)hipFree(s->d_plan);
   if (s->d_ctl) (void)hipFree(s->d_ctl);
   if (s->d_launch) (void)hipFree(s->d_launch);
   for (hipEvent_t e : s->tev) (void)hipEventDestroy(e);
   if (s->ev0) (void)hipEventDestroy(s->ev0);
   if (s->ev1) (void)hipEventDestroy(s->ev1);
   if (s->stream) (void)hipStreamDestroy(s->stream);
   delete s;
}

int spx_add_plugin(spx_sim* s, int role, const char* class_name, const uint8_t* container, size_t len)
{
   if (!s || !class_name || (!container && len)) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "plugins cannot change during a run");
   Reader r{container, container + len};
   std::string cls(class_name);
   int rc;
   switch (role) {
   case SPX_BACKGROUND: rc = add_background(s, cls, r); break;
   case SPX_DIFFUSION: rc = add_diffusion(s, cls, r); break;
   case SPX_SOURCE:
      if (cls == "SourceMomentumInjection") { s->plan.src_kind = SPX_SRC_MOMINJ; s->plan.src_p_inj = r.d(); s->plan.src_rate = r.d(); rc = SPX_OK; }   // src/source_other.cc:98-108
      else rc = fail(s, SPX_ERR_CLASS, "unknown source class " + cls);
      break;
   case SPX_BOUNDARY: rc = add_boundary(s, cls, r); break;
   case SPX_INITIAL: rc = add_initial(s, cls, r); break;
   case SPX_DISTRIBUTION: rc = add_distribution(s, cls, r); break;
   default: return fail(s, SPX_ERR_ARG, "unknown role");
   }
   if (rc == SPX_OK && r.err) return fail(s, SPX_ERR_CONTAINER, cls + ": container does not match the class parameters");
   return rc;
}

int spx_set_option(spx_sim* s, const char* name, double v)
{
   if (!s || !name) return SPX_ERR_ARG;
   std::string n(name);
   if (n == "DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE") { if (v < 0 || v > 2) return SPX_ERR_ARG; s->plan.spectrum_type = (int)v; }
   else if (n == "max_trajectory_steps") { if (v < 1) return SPX_ERR_ARG; s->plan.max_steps = (int64_t)v; }
   else if (n == "max_attempts") { if (v < 1) return SPX_ERR_ARG; s->plan.max_attempts = (int)v; }
   else if (n == "precision") { if (v != 0 && v != 1) return SPX_ERR_ARG; s->precision = (int)v; }
   else if (n == "block") { if (v < 64 || v > 1024 || ((int)v % 64)) return SPX_ERR_ARG; s->block = (int)v; }
   else if (n == "lanes") { if (v < 0) return SPX_ERR_ARG; s->lanes_opt = (int64_t)v; }
   else if (n == "steps_per_launch") { if (v < 1 || v > 1e7) return SPX_ERR_ARG; s->steps_per_launch = (int)v; }
   else if (n == "graph_launches") { if (v < 0 || v > 256) return SPX_ERR_ARG; s->graph_launches = (int)v; }
   else return fail(s, SPX_ERR_ARG, "unknown option " + n);
   return SPX_OK;
}

int spx_begin(spx_sim* s, uint64_t first, uint64_t n_traj, const spx_rng* rng, int keep_records)
{
   if (!s || !rng) return SPX_ERR_ARG;
   int rc = check_ready(s);
   if (rc) return rc;
   finalize_plan(s);
   rc = ensure_device(s);
   if (rc) return rc;
   if (s->plan.traj_type == SPX_TRAJ_GUIDING_DIFF_SCATT) {
      if (s->plan.time_flow != SPX_FLOW_FORWARD) return fail(s, SPX_ERR_ARG, "guiding-centre transport runs forward in time");
      if (s->plan.bg.kind != SPX_BG_SOLARWIND && s->plan.bg.kind != SPX_BG_UNIFORM)
         return fail(s, SPX_ERR_ARG, "guiding-centre transport supports the static SolarWind and Uniform backgrounds");
      if (s->plan.df.kind == SPX_DF_FLOWMOM || s->plan.df.kind == SPX_DF_KINRAD)
         return fail(s, SPX_ERR_ARG, "guiding-centre transport: diffusion model needs grad U, which its field mask does not compute");
   }
   uint64_t lanes = s->lanes_opt > 0 ? (uint64_t)s->lanes_opt : auto_lanes(s);
   lanes = (lanes + s->block - 1) / s->block * s->block;
   rc = alloc_lanes(s, lanes);
   if (rc) return rc;
   if (s->hist_counts.n < (size_t)std::max<int64_t>(s->hist_bins, 1)) {
      HIPCHK(s, s->hist_counts.alloc(std::max<int64_t>(s->hist_bins, 1)));
      HIPCHK(s, s->hist_sums.alloc(std::max<int64_t>(s->hist_bins, 1) * 9));
      HIPCHK(s, s->hist_events.alloc(SPX_MAX_DIST));
      HIPCHK(s, hipMemsetAsync(s->hist_counts.p, 0, sizeof(unsigned long long) * s->hist_counts.n, s->stream));
      HIPCHK(s, hipMemsetAsync(s->hist_sums.p, 0, sizeof(double) * s->hist_sums.n, s->stream));
      HIPCHK(s, hipMemsetAsync(s->hist_events.p, 0, sizeof(unsigned long long) * SPX_MAX_DIST, s->stream));
   }
   // run description
   s->run_first = first;
   s->run_n = n_traj;
   std::memset(&s->run_rng, 0, sizeof(s->run_rng));
   s->run_rng.mode = rng->mode;
   s->run_rng.seed = rng->seed;
   if (rng->mode == 1) {
      if (!rng->normals || rng->normal_stride < 1 || rng->uniform_stride < 0) return fail(s, SPX_ERR_ARG, "replay needs draws");
      size_t nn = (size_t)n_traj * (size_t)rng->normal_stride, nu = (size_t)n_traj * (size_t)std::max<int64_t>(rng->uniform_stride, 1);
      HIPCHK(s, s->rep_normals.alloc(nn));
      HIPCHK(s, s->rep_uniforms.alloc(nu));
      HIPCHK(s, hipMemcpyAsync(s->rep_normals.p, rng->normals, sizeof(double) * nn, hipMemcpyHostToDevice, s->stream));
      if (rng->uniforms && rng->uniform_stride > 0)
         HIPCHK(s, hipMemcpyAsync(s->rep_uniforms.p, rng->uniforms, sizeof(double) * n_traj * rng->uniform_stride, hipMemcpyHostToDevice, s->stream));
      s->run_rng.normals = s->rep_normals.p;
      s->run_rng.uniforms = s->rep_uniforms.p;
      s->run_rng.normal_stride = rng->normal_stride;
      s->run_rng.uniform_stride = (rng->uniforms && rng->uniform_stride > 0) ? rng->uniform_stride : 0;
   }
   else if (rng->mode != 0) return fail(s, SPX_ERR_ARG, "rng mode");
   s->keep_records = keep_records != 0;
   if (s->keep_records) {
      HIPCHK(s, s->records.alloc(n_traj ? n_traj : 1));
      HIPCHK(s, hipMemsetAsync(s->records.p, 0, sizeof(spx_traj_record) * (n_traj ? n_traj : 1), s->stream));
   }
   RunCtl ctl;
   std::memset(&ctl, 0, sizeof(ctl));
   ctl.next = first;
   ctl.tmin_bits = 0x7FF0000000000000ull;
   HIPCHK(s, hipMemcpyAsync(s->d_plan, &s->plan, sizeof(spx_plan_desc), hipMemcpyHostToDevice, s->stream));
   HIPCHK(s, hipMemcpyAsync(s->d_ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice, s->stream));
   HIPCHK(s, hipMemsetAsync(s->active.p, 0, sizeof(int32_t) * s->n_lanes, s->stream));
   HIPCHK(s, hipMemsetAsync(s->ev_pending.p, 0, sizeof(int32_t) * s->n_lanes, s->stream));
   s->running = true;
   Launch Lh = make_launch(s);
   HIPCHK(s, hipMemcpyAsync(s->d_launch, &Lh, sizeof(Launch), hipMemcpyHostToDevice, s->stream));
   HIPCHK(s, hipStreamSynchronize(s->stream));
   if (s->graph_exec) { (void)hipGraphExecDestroy(s->graph_exec); s->graph_exec = nullptr; }
   if (s->graph) { (void)hipGraphDestroy(s->graph); s->graph = nullptr; }
   s->running = true;
   s->launches_total = 0;
   return SPX_OK;
}

static int read_stats(spx_sim* s, spx_run_stats* st, double wall, double kms, int launches)
{
   RunCtl ctl;
   HIPCHK(s, hipMemcpyAsync(&ctl, s->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, s->stream));
   HIPCHK(s, hipStreamSynchronize(s->stream));
   if (st) {
      std::memset(st, 0, sizeof(*st));
      st->n_traj = s->run_n;
      st->done = ctl.done;
      st->failed = ctl.failed;
      st->steps_done = ctl.steps_done;
      st->steps_executed = ctl.steps_executed;
      for (int k = 0; k < 8; k++) st->discards[k] = ctl.discards[k];
      double tmin, tmax;
      std::memcpy(&tmin, &ctl.tmin_bits, 8);
      std::memcpy(&tmax, &ctl.tmax_bits, 8);
      st->t_min = tmin; st->t_max = tmax;
      st->wall_s = wall;
      st->kernel_ms = kms;
      st->launches = launches;
      st->lanes = (int32_t)s->n_lanes;
      st->step_kernel_ms = s->step_ms;
      st->service_kernel_ms = s->service_ms;
   }
   return SPX_OK;
}

static int launch_many(spx_sim* s, int n, float* ms_out)
{
   Launch L = make_launch(s);
   HIPCHK(s, hipEventRecord(s->ev0, s->stream));
   if (s->graph_launches > 0) {
      if (!s->graph_exec) {
         HIPCHK(s, hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
         for (int i = 0; i < s->graph_launches; i++) {
            hipError_t e = launch_transport(s->plan, s->d_plan, L, s->d_launch, s->precision, s->stream, nullptr);
            if (e != hipSuccess) {
               hipGraph_t g;
               (void)hipStreamEndCapture(s->stream, &g);
               return fail(s, SPX_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
            }
         }
         HIPCHK(s, hipStreamEndCapture(s->stream, &s->graph));
         HIPCHK(s, hipGraphInstantiate(&s->graph_exec, s->graph, nullptr, nullptr, 0));
      }
      int reps = (n + s->graph_launches - 1) / s->graph_launches;
      for (int i = 0; i < reps; i++) HIPCHK(s, hipGraphLaunch(s->graph_exec, s->stream));
      n = reps * s->graph_launches;
   }
   else {
      while (s->tev.size() < (size_t)2 * n) {
         hipEvent_t e;
         HIPCHK(s, hipEventCreate(&e));
         s->tev.push_back(e);
      }
      for (int i = 0; i < n; i++) {
         HIPCHK(s, hipEventRecord(s->tev[2 * i], s->stream));
         hipError_t e = launch_transport(s->plan, s->d_plan, L, s->d_launch, s->precision, s->stream, s->tev[2 * i + 1]);
         if (e != hipSuccess) return fail(s, SPX_ERR_DEVICE, std::string("launch: ") + hipGetErrorString(e));
      }
   }
   HIPCHK(s, hipEventRecord(s->ev1, s->stream));
   HIPCHK(s, hipEventSynchronize(s->ev1));
   float ms = 0.f;
   HIPCHK(s, hipEventElapsedTime(&ms, s->ev0, s->ev1));
   if (s->graph_launches == 0) {
      for (int i = 0; i < n; i++) {
         float a = 0.f, b = 0.f;
         HIPCHK(s, hipEventElapsedTime(&a, s->tev[2 * i], s->tev[2 * i + 1]));
         HIPCHK(s, hipEventElapsedTime(&b, s->tev[2 * i + 1], i + 1 < n ? s->tev[2 * i + 2] : s->ev1));
         s->service_ms += a;
         s->step_ms += b;
      }
   }
   if (ms_out) *ms_out = ms;
   s->launches_total += n;
   return SPX_OK;
}

int spx_advance(spx_sim* s, int n_launches, spx_run_stats* st)
{
   if (!s || n_launches < 0) return SPX_ERR_ARG;
   if (!s->running) return fail(s, SPX_ERR_STATE, "advance without begin");
   auto w0 = std::chrono::steady_clock::now();
   s->step_ms = s->service_ms = 0.0;
   float ms = 0.f;
   int rc = launch_many(s, n_launches, &ms);
   if (rc) return rc;
   double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
   return read_stats(s, st, wall, ms, n_launches);
}

int spx_finish(spx_sim* s, spx_traj_record* records, spx_run_stats* st)
{
   if (!s) return SPX_ERR_ARG;
   if (!s->running) return fail(s, SPX_ERR_STATE, "finish without begin");
   auto w0 = std::chrono::steady_clock::now();
   s->step_ms = s->service_ms = 0.0;
   double kms = 0.0;
   int launches = 0;
   // Every launch advances each live slot by steps_per_launch steps; a trajectory needs at most
   // max_attempts * (max_steps + 1) steps, so this bound is never reached by a healthy run.
   const double guard = 64.0 + (double)s->plan.max_attempts * ((double)s->plan.max_steps + 1.0) / s->steps_per_launch *
                                   std::max(1.0, (double)s->run_n / (double)s->n_lanes + 1.0);
   int batch = 1;
   while (true) {
      RunCtl ctl;
      HIPCHK(s, hipMemcpyAsync(&ctl, s->d_ctl, sizeof(ctl), hipMemcpyDeviceToHost, s->stream));
      HIPCHK(s, hipStreamSynchronize(s->stream));
      if (ctl.done >= s->run_n) break;
      if (launches > guard) return fail(s, SPX_ERR_STATE, "run did not drain");
      float ms = 0.f;
      int rc = launch_many(s, batch, &ms);
      if (rc) return rc;
      kms += ms;
      launches += batch;
      if (batch < 8) batch *= 2;
   }
   double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
   int rc = read_stats(s, st, wall, kms, launches);
   if (rc) return rc;
   if (records && s->keep_records) {
      HIPCHK(s, hipMemcpyAsync(records, s->records.p, sizeof(spx_traj_record) * s->run_n, hipMemcpyDeviceToHost, s->stream));
      HIPCHK(s, hipStreamSynchronize(s->stream));
   }
   s->running = false;
   return SPX_OK;
}

int spx_run(spx_sim* s, uint64_t first, uint64_t n_traj, const spx_rng* rng, spx_traj_record* records, spx_run_stats* st)
{
   int rc = spx_begin(s, first, n_traj, rng, records != nullptr);
   if (rc) return rc;
   return spx_finish(s, records, st);
}

int spx_distro_info(const spx_sim* s, int d, int32_t* n_bins3, int32_t* width)
{
   if (!s || d < 0 || d >= s->plan.n_dist) return SPX_ERR_ARG;
   if (n_bins3) for (int k = 0; k < 3; k++) n_bins3[k] = s->plan.dist[d].n_bins[k];
   if (width) *width = s->plan.dist[d].width;
   return SPX_OK;
}

int spx_get_distro(const spx_sim* sc, int d, int64_t* counts, double* sums, int64_t* n_events)
{
   spx_sim* s = const_cast<spx_sim*>(sc);
   if (!s || d < 0 || d >= s->plan.n_dist) return SPX_ERR_ARG;
   const spx_ds_desc& D = s->plan.dist[d];
   const int64_t off = s->plan.hist_off[d];
   if (!s->hist_counts.p) {
      if (counts) std::memset(counts, 0, sizeof(int64_t) * D.nb);
      if (sums) std::memset(sums, 0, sizeof(double) * D.nb * D.width);
      if (n_events) *n_events = 0;
      return SPX_OK;
   }
   std::vector<unsigned long long> c(D.nb);
   std::vector<double> sm((size_t)D.nb * 9);
   unsigned long long ev[SPX_MAX_DIST];
   HIPCHK(s, hipMemcpyAsync(c.data(), s->hist_counts.p + off, sizeof(unsigned long long) * D.nb, hipMemcpyDeviceToHost, s->stream));
   HIPCHK(s, hipMemcpyAsync(sm.data(), s->hist_sums.p + off * 9, sizeof(double) * D.nb * 9, hipMemcpyDeviceToHost, s->stream));
   HIPCHK(s, hipMemcpyAsync(ev, s->hist_events.p, sizeof(ev), hipMemcpyDeviceToHost, s->stream));
   HIPCHK(s, hipStreamSynchronize(s->stream));
   if (counts) for (int b = 0; b < D.nb; b++) counts[b] = (int64_t)c[b];
   if (sums) for (int b = 0; b < D.nb; b++) for (int k = 0; k < D.width; k++) sums[(size_t)b * D.width + k] = sm[(size_t)b * 9 + k];
   if (n_events) *n_events = (int64_t)ev[d];
   return SPX_OK;
}

int spx_reset_distros(spx_sim* s)
{
   if (!s) return SPX_ERR_ARG;
   if (!s->hist_counts.p) return SPX_OK;
   HIPCHK(s, hipMemsetAsync(s->hist_counts.p, 0, sizeof(unsigned long long) * s->hist_counts.n, s->stream));
   HIPCHK(s, hipMemsetAsync(s->hist_sums.p, 0, sizeof(double) * s->hist_sums.n, s->stream));
   HIPCHK(s, hipMemsetAsync(s->hist_events.p, 0, sizeof(unsigned long long) * SPX_MAX_DIST, s->stream));
   HIPCHK(s, hipStreamSynchronize(s->stream));
   return SPX_OK;
}

const char* spx_last_error(const spx_sim* s) { return s ? s->err.c_str() : "null simulation"; }

int spx_device_info(char* buf, size_t len)
{
   int n = 0;
   if (!buf || !len) return SPX_ERR_ARG;
   buf[0] = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SPX_ERR_DEVICE;
   hipDeviceProp_t p;
   if (hipGetDeviceProperties(&p, 0) != hipSuccess) return SPX_ERR_DEVICE;
   std::snprintf(buf, len, "%s|%s|cu=%d|clock_khz=%d|mem=%zu", p.name, p.gcnArchName, p.multiProcessorCount, p.clockRate, (size_t)p.totalGlobalMem);
   return SPX_OK;
}

}  // extern "C"
