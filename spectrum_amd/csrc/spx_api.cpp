// spx_api.cpp -- host runtime behind the C ABI (include/spectrum_amd.h).
//
// Replaces the reference's MPI master/worker task farm (src/simulation.cc:253-305, 743-835) with a
// single-GPU scheduler: trajectories are queued on a device counter, resident lane slots are refilled
// by the persistent kernels, and distributions accumulate in device histograms.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "spx_kernels.h"
#include "spx_runtime.h"
#include "spx_static.h"

using namespace spx;

namespace {

constexpr double C_CGS_H = 2.99792458e+10;
constexpr double unit_velocity_h = 0.1 * C_CGS_H;
constexpr double c_code_h = C_CGS_H / unit_velocity_h;
constexpr double sp_tiny_h = 1.0E-15;

// ---- DataContainer reader (common/data_container.cc:214-238): [u8 len][len bytes] records -------------
// DataContainer::Read copies the record's own length into the value.  A record shorter than the value
// (an int where the class reads a double, as benchmarks/main_test_modulation_cartesian_parker.cc:190-200
// inserts its unused shock parameters) fills the low bytes; the rest, which the reference leaves as it
// was, is zero here.  A longer record would overrun the value and is rejected.
struct Reader {
   const uint8_t* p;
   const uint8_t* end;
   bool err = false;
   void raw(void* dst, size_t want)
   {
      if (err || p >= end) { err = true; return; }
      size_t len = *p;
      if (p + 1 + len > end || len > want || len == 0) { err = true; return; }
      std::memset(dst, 0, want);
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

// ---- gridded background (BackgroundServerCartesian) -------------------------------------------------
// Unit factors of the grid data (common/physics.hh:20-56 fluid units, common/specie.hh:30-64 code
// units), each formed as the reference's compile-time expressions are.
constexpr double MP_CGS_H = 1.67262192596E-24;
constexpr double E_CGS_H = 4.8032044E-10;
constexpr double AU_CGS_H = 1.495978707E+13;
constexpr double unit_length_h = 1.0E10;
inline void grid_units(spx_grid_desc& g)
{
   const double unit_momentum = MP_CGS_H * unit_velocity_h;
   const double unit_time = unit_length_h / unit_velocity_h;
   const double unit_magnetic = unit_momentum / (E_CGS_H * unit_time);
   const double fluid_magnetic = 1.0E7 * std::sqrt(MP_CGS_H * 1.0);
   g.f_len = AU_CGS_H / unit_length_h;
   g.f_vel = 1.0E7 / unit_velocity_h;
   g.f_mag = fluid_magnetic / unit_magnetic;
}

// Zone records of a grid on the host and, once a run needs them, in device memory (the FP64 zone array and
// the FP32 cell records, spx_grid_desc::z64 / c32).  Shared by a simulation and the queries served from it,
// so it lives as long as any of them.
struct GridData {
   spx_grid_desc desc{};
   std::vector<double> rec;            // [nz2][nz1][nz0][SPX_GRID_REC]
   DevBuf<double> z64;
   DevBuf<float> c32;                  // 8 corners per zone
   std::mutex mu;
   bool uploaded = false;
   size_t bytes() const { return rec.size() * (sizeof(double) + 8 * sizeof(float)); }
   // the zone records go up once and are expanded into the FP32 cell records on the device
   hipError_t upload()
   {
      std::lock_guard<std::mutex> lk(mu);
      if (uploaded) return hipSuccess;
      hipError_t e = z64.alloc(rec.size());
      if (e == hipSuccess) e = c32.alloc(rec.size() * 8);
      if (e == hipSuccess) e = hipMemcpy(z64.p, rec.data(), sizeof(double) * rec.size(), hipMemcpyHostToDevice);
      if (e == hipSuccess) e = grid_expand_cells(z64.p, desc.nz, c32.p, nullptr);
      if (e == hipSuccess) e = hipDeviceSynchronize();
      if (e != hipSuccess) return e;
      desc.z64 = z64.p;
      desc.c32 = c32.p;
      uploaded = true;
      return hipSuccess;
   }
   ~GridData() { z64.release(); c32.release(); }
};

// What the distribution file formats need beyond the plan: the class name, the configuration bytes
// (DistributionBase::Dump writes them), and the unit factors of Print1D / Print2D.
struct DistMeta {
   std::string cls;
   std::vector<uint8_t> container;
   uint32_t n_records = 0;
   double unit_distro[9] = {1, 1, 1, 1, 1, 1, 1, 1, 1};
   double unit_val[3] = {1, 1, 1};
};

}  // namespace

struct spx_sim {
   spx_plan_desc plan;
   int32_t soqlt_sc_idx = 0;            // DiffusionEmpiricalSOQLTandUNLT's solar_cycle_idx (checked against the background)
   bool have_bg = false, have_df = false, have_it = false, have_is = false, have_im = false;
   int precision = SPX_PREC_FP64;
   int block = 256;
   int64_t lanes_opt = 0;
   int steps_per_launch = 2000;
   int graph_launches = 0;
   std::string err;
   int64_t hist_bins = 0;
   // device resources: one slot pool with a single query, and histograms that accumulate over runs
   // like the reference's distributions until they are reset
   bool dev_ready = false;
   Pool pool;
   QueryRes q;
   DevBuf<unsigned long long> hist_counts, hist_events, hist_acc;   // counts, event totals, exact sums (spx_acc.h)
   std::shared_ptr<GridData> grid;     // BackgroundServerCartesian data
   std::vector<DistMeta> dmeta;        // per distribution, for the file formats
   // run state
   bool running = false;
   uint64_t run_first = 0, run_n = 0;
   bool keep_records = false;
   // SimulationMaster's tasks (src/simulation.hh:196-237): trajectory count, batch size, per-worker
   // maximum, the base name of the distribution files, the distributions to restore at the start
   int64_t task_n = 0, task_batch = 0, task_max_per_worker = 0;
   std::string distro_file_name = "distribution_";
   uint32_t restore_at_start = 0;
   uint64_t next_traj = 0;             // the main loop's trajectories continue from here
};

static int fail(spx_sim* s, int code, const std::string& msg)
{
   if (s) s->err = msg;
   return code;
}

static int ensure_hist(spx_sim* s);

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
      if (g.ep_diag && g.eprime[0][0] == 1.0 && g.eprime[1][1] == 1.0 && g.eprime[2][2] == 1.0) g.ep_diag = 2;   // x1 is exact: skipped
      g.ur0 = std::fabs(g.u0[0]);
      g.ur0_rcp = 1.0 / g.ur0;
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
   else if (cls == "BackgroundServerCartesian") {   // src/background_server.cc:62-75: the base parameters only
      g.kind = SPX_BG_GRID;
      s->grid.reset();                  // the grid comes with spx_set_background_grid / spx_add_background_file
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
   else if (cls == "DiffusionStraussEtAl2013" || cls == "DiffusionPotgieterEtAl2015") {   // src/diffusion_other.cc:869-880, :950-961
      d.kind = cls == "DiffusionStraussEtAl2013" ? SPX_DF_STRAUSS : SPX_DF_POTGIETER;
      d.region_idx = r.i(); d.lam_in = r.d(); d.lam_out = r.d(); d.R0 = r.d(); d.B0 = r.d();
      d.kap_rat_in = r.d(); d.kap_rat_out = r.d();
   }
   else if (cls == "DiffusionEmpiricalSOQLTandUNLT") {   // src/diffusion_other.cc:1034-1046
      d.kind = SPX_DF_SOQLT;
      d.lam_para = r.d(); d.lam_perp = r.d(); d.R0 = r.d(); d.B0 = r.d();
      d.region_idx = r.i(); d.kap_rat_red = r.d(); d.r_lim_red = r.d();
      s->soqlt_sc_idx = r.i();
      const double effect = r.d();
      // Kappa /= 1 + effect Sqr(cos(0.5 region[solar_cycle_idx])) (:1071): the indicator is +1 or -1 on the
      // backgrounds this model runs on and cos is even, so the divisor is one number
      const double c = std::cos(0.5 * 1.0);
      d.sc_den = 1.0 + effect * (c * c);
   }
   else if (cls == "DiffusionQLTConstant") {   // src/diffusion_other.cc:130-140
      d.kind = SPX_DF_QLT;
      d.A2A = r.d();
      const double l_max = r.d(), ps_index = r.d();
      d.k_min = (2.0 * M_PI) / l_max;
      d.ps_minus = ps_index - 1.0;
   }
   else if (cls == "DiffusionWNLTConstant" || cls == "DiffusionWNLTRampVLISM") {   // :130-140, :202-213, :281-296
      d.kind = cls == "DiffusionWNLTConstant" ? SPX_DF_WNLT : SPX_DF_WNLT_RAMP;
      d.A2A = r.d();
      const double l_max = r.d(), ps_index = r.d();
      d.k_min = (2.0 * M_PI) / l_max;
      d.ps_minus = ps_index - 1.0;
      d.A2T = r.d(); d.A2L = r.d();
      d.ps_plus = ps_index + 1.0;
      if (d.kind == SPX_DF_WNLT_RAMP) {
         d.l_max_HP = r.d(); d.z_nose = r.d(); d.z_sheath = r.d();
         d.k_min_ref = d.k_min; d.A2A_ref = d.A2A; d.A2T_ref = d.A2T; d.A2L_ref = d.A2L;
         d.dl_max = l_max - d.l_max_HP; d.dz = d.z_sheath - d.z_nose;
      }
      d.hyp_c = (5.0 + ps_index) / 4.0;
      d.hyp_g = tgamma(d.hyp_c) * tgamma(2.0 - d.hyp_c);
      d.hyp_r = (d.hyp_c - 1.0) / (d.hyp_c - 2.0);
   }
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
   else if (cls == "InitialSpaceFixed") { P.is_kind = SPX_IS_FIXED; r.v(P.is_pos); s->have_is = true; }   // src/initial_space.cc:49-59
   else if (cls == "InitialSpaceSphere") {                                                        // src/initial_space.cc:297-303
      P.is_kind = SPX_IS_SPHERE; r.v(P.is_pos); P.is_radius = r.d();
      P.is_ct1 = 1.0; P.is_ct2 = -1.0; P.is_phi1 = 0.0; P.is_phi2 = 2.0 * M_PI;
      s->have_is = true;
   }
   else if (cls == "InitialSpaceSphereSector") {                                                  // src/initial_space.cc:353-378
      P.is_kind = SPX_IS_SECTOR; r.v(P.is_pos); P.is_radius = r.d();
      double th1 = r.d(), th2 = r.d(), ph1 = r.d(), ph2 = r.d();
      if (th1 < 0.0) th1 = 0.0;
      if (th2 > M_PI) th2 = M_PI;
      if (th2 < th1) th2 = th1;
      if (ph2 > 2.0 * M_PI) ph2 = 2.0 * M_PI;
      if (ph2 < ph1) ph2 = ph1;
      P.is_ct1 = std::cos(th1); P.is_ct2 = std::cos(th2); P.is_phi1 = ph1; P.is_phi2 = ph2;
      s->have_is = true;
   }
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
   DistMeta M;
   r.mi(D.n_bins); r.v(minval); r.v(maxval); r.mi(D.log_bins); r.mi(D.bin_outside);
   // unit_distro has the distribution's value type: 1, 3 or 9 doubles (DistributionTemplated::ReadUnitDistro)
   r.raw(M.unit_distro, sizeof(double) * width);
   r.v(unit_val); (void)r.b();
   for (int k = 0; k < 3; k++) M.unit_val[k] = unit_val[k];
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
   P.acc_off[P.n_dist] = acc_components(P);
   s->hist_bins += D.nb;
   P.n_dist++;
   M.cls = cls;
   s->dmeta.push_back(M);
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
   HIPCHK(s, hipSetDevice(0));
   s->dev_ready = true;
   return SPX_OK;
}

// Device copy of the grid, and its addresses in the plan (both precision views).
static int bind_grid(spx_sim* s)
{
   if (s->plan.bg.kind != SPX_BG_GRID) return SPX_OK;
   HIPCHK(s, s->grid->upload());
   spx_grid_desc g = s->grid->desc;
   // TrajectoryParker::DiffusionCoeff (src/trajectory_parker.cc:100-119) evaluates the background at the
   // six derivative points.  Their fields enter div(kappa) only through (kappa_par - kappa_perp) bhat
   // and through kappa's own U or |B| dependence; when neither can matter, the only observable effects
   // of those evaluations are the domain check and GetFields' |B| < sp_tiny check.  The interpolated
   // |B| is a convex combination of zone magnitudes, so the latter cannot fire when the smallest zone
   // |B| clears sp_tiny with margin.  Then the kernel checks the points' positions and loads nothing.
   const spx_df_desc& d = s->plan.df;
   const bool k_equal = ((d.kind == SPX_DF_FLOWMOM || d.kind == SPX_DF_KINRAD || d.kind == SPX_DF_RIGBMAG) && d.kap_rat == 1.0) ||
                        (d.kind == SPX_DF_FULL && d.Dperp == d.Dpara);
   const bool need_u = d.kind == SPX_DF_FLOWMOM || (d.kind == SPX_DF_KINRAD && d.stream_dep_idx != 0);
   const bool need_b = (d.kind == SPX_DF_RIGBMAG && d.pow_law_B != 0.0) || d.kind == SPX_DF_STRAUSS || d.kind == SPX_DF_POTGIETER;
   const bool bmag_safe = g.bmag_min * g.f_mag * 0.999 > sp_tiny_h;
   g.fd_fields = (k_equal && !need_u && !need_b && bmag_safe) ? 0 : 1;
   s->plan.bg.grid = g;
   s->plan.bgf.grid = g;
   return SPX_OK;
}

// Resident slots per device: 8 waves per SIMD on every CU.  The register files hold 2 (FP64 Parker) to
// 5 (FP32 grid) of them at a time; the rest start as the first ones finish, which fills the lanes that
// finished trajectories leave idle until the next service pass (measured: 3-5% more steps/s than 4
// waves on the modulation workload, the autotuner's choice for the shock and gridded workloads).
static uint64_t auto_lanes()
{
   hipDeviceProp_t prop;
   if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 65536;
   return (uint64_t)prop.multiProcessorCount * 4 * 8 * 64;
}

// Float copies of the background / diffusion parameters for the mixed-precision kernel, and the
// position-dependence flag of kappa (pow(x, 0) == 1 exactly, so a zero exponent removes it bit for bit).
static void finalize_plan(spx_sim* s)
{
   spx_plan_desc& P = s->plan;
   // concentric spheres in consecutive slots share one |x - origin| per step
   for (int b = 0; b < P.n_bnd; b++) {
      const spx_bnd_desc& q = P.bnd[b > 0 ? b - 1 : 0];
      P.bnd[b].same_origin = b > 0 && P.bnd[b].kind == SPX_BK_SPHERE && q.kind == SPX_BK_SPHERE && q.origin[0] == P.bnd[b].origin[0] &&
                             q.origin[1] == P.bnd[b].origin[1] && q.origin[2] == P.bnd[b].origin[2];
   }
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
   f.ur0_rcp = (float)g.ur0_rcp;
   f.w0 = (float)g.w0; f.hp_rad = (float)g.hp_rad; f.v_shock = (float)g.v_shock; f.compression = (float)g.compression;
   f.width_shock = (float)g.width_shock;
   f.grid = g.grid;
   if (g.kind == SPX_BG_SOLARWIND) {
      f.sw_c1 = (float)(g.B0[0] * g.r_ref * g.r_ref);
      f.sw_k = (float)(g.w0 / g.ur0);
      // a diagonal basis S = diag(s0, s1, s2) (s0 s1 s2 = 1) maps the local spiral w(p) = (p0 + q p1, p1 - q p0, p2)
      // to S w(S d) = (d0 + s0 s1 q d1, d1 - s0 s1 q d0, d2) in the global frame
      f.sw_kg = g.ep_diag ? (float)(g.eprime[0][0] * g.eprime[1][1] * g.w0 / g.ur0) : 0.0f;
      // |B| = |B0| r_ref^2 |w| / r^3 with |w| >= r, so |B| >= |B0| r_ref^2 / r^2
      const double r2_ok = std::fabs(g.B0[0] * g.r_ref * g.r_ref) / sp_tiny_h;
      f.sw_r2_ok = r2_ok > (double)FLT_MAX ? INFINITY : (float)r2_ok;
   }
   spx_df_t<double>& d = P.df;
   d.pos_dep = 0;
   if (d.kind == SPX_DF_FLOWMOM) d.pos_dep = (d.pow_law_U != 0.0);
   if (d.kind == SPX_DF_KINRAD) d.pos_dep = 1;
   if (d.kind == SPX_DF_RIGBMAG) d.pos_dep = (d.pow_law_B != 0.0);
   if (d.kind == SPX_DF_STRAUSS || d.kind == SPX_DF_POTGIETER || d.kind == SPX_DF_SOQLT) d.pos_dep = 1;   // through |B| and the region
   spx_df_t<float>& e = P.dff;
   e.kind = d.kind; e.stream_dep_idx = d.stream_dep_idx; e.pos_dep = d.pos_dep; e.region_idx = d.region_idx;
   e.lam_in = (float)d.lam_in; e.lam_out = (float)d.lam_out; e.kap_rat_in = (float)d.kap_rat_in; e.kap_rat_out = (float)d.kap_rat_out;
   e.A2A = (float)d.A2A; e.k_min = (float)d.k_min; e.ps_minus = (float)d.ps_minus;
   e.A2T = (float)d.A2T; e.A2L = (float)d.A2L; e.ps_plus = (float)d.ps_plus;
   e.l_max_HP = (float)d.l_max_HP; e.z_nose = (float)d.z_nose; e.z_sheath = (float)d.z_sheath; e.k_min_ref = (float)d.k_min_ref;
   e.A2A_ref = (float)d.A2A_ref; e.A2T_ref = (float)d.A2T_ref; e.A2L_ref = (float)d.A2L_ref; e.dl_max = (float)d.dl_max; e.dz = (float)d.dz;
   e.hyp_c = (float)d.hyp_c; e.hyp_g = (float)d.hyp_g; e.hyp_r = (float)d.hyp_r;
   e.lam_para = (float)d.lam_para; e.lam_perp = (float)d.lam_perp; e.kap_rat_red = (float)d.kap_rat_red;
   e.r_lim_red = (float)d.r_lim_red; e.sc_den = (float)d.sc_den;
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
   // TrajectoryGuiding (no diffusion, src/trajectory_guiding.hh) integrates without a diffusion model
   if (!s->have_df && s->plan.traj_type != SPX_TRAJ_GUIDING && s->plan.traj_type != SPX_TRAJ_FIELDLINE &&
       s->plan.traj_type != SPX_TRAJ_LORENTZ && s->plan.traj_type != SPX_TRAJ_FOCUSED)
      return fail(s, SPX_ERR_NOT_READY, "no diffusion model");
   if (s->plan.traj_type == SPX_TRAJ_PARKER_SOURCE && s->plan.src_kind == SPX_SRC_NONE) return fail(s, SPX_ERR_NOT_READY, "no source");
   if (s->plan.bg.kind == SPX_BG_GRID && !s->grid) return fail(s, SPX_ERR_NOT_READY, "BackgroundServerCartesian has no grid data");
   return SPX_OK;
}

// Combinations the kernels implement (the reference rejects the others at compile time).
static int check_transport(spx_sim* s)
{
   if (s->plan.traj_type == SPX_TRAJ_LORENTZ) {
      if (s->plan.bg.kind == SPX_BG_SMOOTHSHOCK) return fail(s, SPX_ERR_ARG, "full orbit: SolarWind, Uniform, Shock and Cartesian-grid backgrounds");
      const int im = s->plan.im_kind;
      if (im != SPX_IM_FIXED && im != SPX_IM_SHELL && im != SPX_IM_THICKSHELL)
         return fail(s, SPX_ERR_ARG, "full orbit: initial momentum must be Fixed, Shell or ThickShell");
      if (s->plan.src_kind != SPX_SRC_NONE) return fail(s, SPX_ERR_ARG, "full orbit: no source term");
      // its pitch-angle values need bhat at the event (src/distribution_other.cc:284-287)
      for (int d = 0; d < s->plan.n_dist; d++) {
         const spx_ds_desc& D = s->plan.dist[d];
         if (D.kind == SPX_DS_POSPITCH || (D.kind == SPX_DS_MOMENTUM && D.val_coord != 0))
            return fail(s, SPX_ERR_ARG, "full orbit: pitch-angle distributions are not supported");
      }
   }
   else if (s->plan.traj_type == SPX_TRAJ_FOCUSED) {
      // every background kind: SolarWind, Uniform, Shock, SmoothShock and the Cartesian grid
      if (s->plan.src_kind != SPX_SRC_NONE) return fail(s, SPX_ERR_ARG, "focused transport: no source term");
      for (int d = 0; d < s->plan.n_dist; d++)
         if (s->plan.dist[d].kind == SPX_DS_POSPITCH) return fail(s, SPX_ERR_ARG, "focused transport: its momentum is (p, mu, 0); use DistributionMomentumUniform");
   }
   else if (s->plan.traj_type == SPX_TRAJ_FIELDLINE) {
      // BoundaryMomentum evaluates nothing for the field line (src/boundary_momentum.cc:69-76); its momentum is
      // (0, 0, p_par) (src/initial_momentum.cc:289-292, :390-391)
      for (int b = 0; b < s->plan.n_bnd; b++)
         if (s->plan.bnd[b].cls == SPX_BC_MOMENTUM) return fail(s, SPX_ERR_ARG, "field line: momentum boundaries are not defined");
      const int im = s->plan.im_kind;
      if (im == SPX_IM_FIXED ? (s->plan.im_mom[0] != 0.0 || s->plan.im_mom[1] != 0.0) : (im != SPX_IM_SHELL && im != SPX_IM_THICKSHELL))
         return fail(s, SPX_ERR_ARG, "field line: initial momentum must be Shell, ThickShell or Fixed (0, 0, p)");
      if (s->plan.src_kind != SPX_SRC_NONE) return fail(s, SPX_ERR_ARG, "field line: no source term");
   }
   else if (s->plan.traj_type >= SPX_TRAJ_GUIDING_DIFF_SCATT) {
      if (s->plan.time_flow != SPX_FLOW_FORWARD) return fail(s, SPX_ERR_ARG, "guiding-centre transport runs forward in time");
      // every background kind: the static SolarWind, Uniform and Cartesian grid, the discontinuous shock
      // and the moving smooth shock (with its dB/dt terms)
      if (s->plan.df.kind == SPX_DF_FLOWMOM || s->plan.df.kind == SPX_DF_KINRAD)
         return fail(s, SPX_ERR_ARG, "guiding-centre transport: diffusion model needs grad U, which its field mask does not compute");
      // Strauss / Potgieter set no component 2 (src/diffusion_other.cc:886-888), so with pitch-angle scattering
      // D_mumu would be whatever the class held before; their perpendicular drift steps over the solar wind's
      // numerical derivative steps (DiffusionBase::GetDirectionalDerivative, src/diffusion_base.cc:130-170)
      if ((s->plan.df.kind == SPX_DF_STRAUSS || s->plan.df.kind == SPX_DF_POTGIETER) &&
          (s->plan.traj_type != SPX_TRAJ_GUIDING_DIFF || s->plan.bg.kind != SPX_BG_SOLARWIND))
         return fail(s, SPX_ERR_ARG, "guiding-centre transport: the Strauss / Potgieter models run with GuidingDiff on the SolarWind background");
   }
   else if (s->plan.df.kind == SPX_DF_QLT)   // src/diffusion_other.cc:83 compiles it for the guiding-centre transports only
      return fail(s, SPX_ERR_ARG, "DiffusionQLTConstant needs a pitch-angle transport");
   else if (s->plan.df.kind == SPX_DF_WNLT || s->plan.df.kind == SPX_DF_WNLT_RAMP)   // :153 likewise
      return fail(s, SPX_ERR_ARG, "the WNLT models need a guiding-centre transport");
   if (s->plan.df.kind == SPX_DF_WNLT || s->plan.df.kind == SPX_DF_WNLT_RAMP) {
      const int tt = s->plan.traj_type;
      if (s->plan.bg.kind != SPX_BG_SOLARWIND && s->plan.bg.kind != SPX_BG_UNIFORM)
         return fail(s, SPX_ERR_ARG, "the WNLT models run on the SolarWind and Uniform backgrounds");
      // the perpendicular drift differentiates kappa_perp over the background's numerical derivative steps
      // (DiffusionBase::GetDirectionalDerivative, src/diffusion_base.cc:130-170), which only the solar wind sets
      if ((tt == SPX_TRAJ_GUIDING_DIFF || tt == SPX_TRAJ_GUIDING_DIFF_SCATT) && s->plan.bg.kind != SPX_BG_SOLARWIND)
         return fail(s, SPX_ERR_ARG, "the WNLT models with perpendicular diffusion need the SolarWind background");
      if (fabs(s->plan.df.hyp_c - 2.0) < 1e-3)
         return fail(s, SPX_ERR_ARG, "the WNLT models: (5 + ps_index) / 4 within 1e-3 of 2 (ps_index near 3) is not supported");
   }
   if (s->plan.df.kind == SPX_DF_SOQLT) {
      // the region indicators it reads (src/diffusion_other.cc:1062-1071): SolarWind sets region[0] (+1 inside
      // the heliopause, -1 beyond) and region[1] (-1) only (src/background_solarwind.cc:144-148), Uniform sets
      // all three to +1 (src/background_uniform.cc:68); a negative Bmix_idx reads region[-1] in the reference
      const int tt = s->plan.traj_type, bk = s->plan.bg.kind, bi = s->plan.df.region_idx, si = s->soqlt_sc_idx;
      if (bk != SPX_BG_SOLARWIND && bk != SPX_BG_UNIFORM)
         return fail(s, SPX_ERR_ARG, "DiffusionEmpiricalSOQLTandUNLT runs on the SolarWind and Uniform backgrounds");
      const int max_idx = bk == SPX_BG_SOLARWIND ? 1 : 2;
      if (bi < 0 || bi > max_idx || si < 0 || si > max_idx)
         return fail(s, SPX_ERR_ARG, "DiffusionEmpiricalSOQLTandUNLT: Bmix_idx / solar_cycle_idx name a region indicator this background does not set");
      // it sets no component 2 (:1055), so no pitch-angle transport; GuidingDiff's drift steps over the
      // background's numerical derivative steps, which only the solar wind sets
      if (tt != SPX_TRAJ_PARKER && tt != SPX_TRAJ_PARKER_SOURCE && !(tt == SPX_TRAJ_GUIDING_DIFF && bk == SPX_BG_SOLARWIND))
         return fail(s, SPX_ERR_ARG, "DiffusionEmpiricalSOQLTandUNLT runs with the Parker transports, or GuidingDiff on the SolarWind background");
   }
   if (s->plan.df.kind == SPX_DF_STRAUSS || s->plan.df.kind == SPX_DF_POTGIETER) {
      // the region indicators the backgrounds define: SolarWind sets region[0] (heliopause) and region[1] (-1) only
      // (src/background_solarwind.cc:144-148); the gridded server carries none here
      const int ri = s->plan.df.region_idx;
      if (ri > 2 || (ri == 2 && s->plan.bg.kind == SPX_BG_SOLARWIND) || (ri >= 0 && s->plan.bg.kind == SPX_BG_GRID))
         return fail(s, SPX_ERR_ARG, "LISM_idx names a region indicator this background does not set");
   }
   return SPX_OK;
}

// ---------------------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------------------
extern "C" {

int spx_create(int traj_type, int time_flow, spx_sim** out)
{
   if (!out || traj_type < 0 || traj_type > SPX_TRAJ_FOCUSED || (time_flow != 0 && time_flow != 1)) return SPX_ERR_ARG;
   spx_sim* s = new spx_sim();
   std::memset(&s->plan, 0, sizeof(s->plan));
   s->plan.traj_type = traj_type;
   s->plan.time_flow = time_flow;
   s->plan.max_steps = 10000000;
   s->plan.max_attempts = 16;
   s->plan.dt_scale = 1.0;
   *out = s;
   return SPX_OK;
}

void spx_destroy(spx_sim* s)
{
   if (!s) return;
   s->q.release();
   s->pool.release();
   s->hist_counts.release(); s->hist_events.release(); s->hist_acc.release();
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
   case SPX_DISTRIBUTION:
      rc = add_distribution(s, cls, r);
      if (rc == SPX_OK && r.err) {   // a malformed container: take the distribution back out
         s->plan.n_dist--;
         s->hist_bins -= s->plan.dist[s->plan.n_dist].nb;
         s->dmeta.pop_back();
      }
      else if (rc == SPX_OK) {
         DistMeta& M = s->dmeta.back();
         M.container.assign(container, container + len);
         for (size_t i = 0; i < len; i += 1 + container[i]) M.n_records++;   // DataContainer::size()
      }
      break;
   default: return fail(s, SPX_ERR_ARG, "unknown role");
   }
   if (rc == SPX_OK && r.err) return fail(s, SPX_ERR_CONTAINER, cls + ": container does not match the class parameters");
   return rc;
}

int spx_set_background_grid(spx_sim* s, const int32_t n_blocks[3], const double dom_min[3], const double dom_max[3], int32_t n_var,
                            const double* data)
{
   if (!s || !n_blocks || !dom_min || !dom_max || !data) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "the background cannot change during a run");
   if (s->plan.bg.kind != SPX_BG_GRID) return fail(s, SPX_ERR_STATE, "add a BackgroundServerCartesian background first");
   // BlockCartesian holds n_variables_cartesian = 6 values per zone: flow (0-2) and magnetic field (3-5)
   // (src/block_cartesian.hh:12, src/server_base.hh:32-35)
   if (n_var != 6) return fail(s, SPX_ERR_ARG, "a Cartesian grid holds 6 variables per zone (U, B)");
   auto G = std::make_shared<GridData>();
   spx_grid_desc& g = G->desc;
   int64_t nzone = 1;
   for (int a = 0; a < 3; a++) {
      if (n_blocks[a] < 1 || n_blocks[a] > (1 << 20) / SPX_GRID_BLOCK) return fail(s, SPX_ERR_ARG, "blocks per axis out of range");
      if (!(dom_max[a] > dom_min[a])) return fail(s, SPX_ERR_ARG, "empty domain");
      g.nb[a] = n_blocks[a];
      g.nz[a] = SPX_GRID_BLOCK * n_blocks[a];
      g.dom_min[a] = dom_min[a];
      g.dom_max[a] = dom_max[a];
      g.block_length[a] = (dom_max[a] - dom_min[a]) / (double)n_blocks[a];   // src/reader_cartesian.cc:88
      nzone *= g.nz[a];
   }
   grid_units(g);
   for (int a = 0; a < 3; a++) {
      const double zl = g.block_length[a] / SPX_GRID_BLOCK;
      const double lo = (g.dom_min[a] + 0.5 * zl) * g.f_len, hi = (g.dom_max[a] - 0.5 * zl) * g.f_len;
      const double margin = 1e-9 * (std::fabs(lo) + std::fabs(hi) + (hi - lo));
      g.inner_lo[a] = lo + margin;
      g.inner_hi[a] = hi - margin;
   }
   // the device holds the FP64 zone record and 8 FP32 corner records per zone (320 B): 2^28 zones are 86 GB
   if (nzone > ((int64_t)1 << 28)) return fail(s, SPX_ERR_LIMIT, "grid too large (more than 2^28 zones)");
   G->rec.assign((size_t)nzone * SPX_GRID_REC, 0.0);
   // ReadCartesianData's reordering (src/reader_cartesian.cc:103-140): blocks x fastest, zones x fastest
   const int B = SPX_GRID_BLOCK;
   const int64_t bsize = (int64_t)n_var * B * B * B;
   int64_t node = 0;
   for (int kb = 0; kb < g.nb[2]; kb++)
      for (int jb = 0; jb < g.nb[1]; jb++)
         for (int ib = 0; ib < g.nb[0]; ib++, node++)
            for (int kz = 0; kz < B; kz++)
               for (int jz = 0; jz < B; jz++)
                  for (int iz = 0; iz < B; iz++) {
                     const double* v = data + node * bsize + (int64_t)n_var * (iz + B * (jz + B * kz));
                     const int64_t zx = (int64_t)B * ib + iz, zy = (int64_t)B * jb + jz, zz = (int64_t)B * kb + kz;
                     double* r = G->rec.data() + (zx + g.nz[0] * (zy + (int64_t)g.nz[1] * zz)) * SPX_GRID_REC;
                     for (int m = 0; m < 6; m++) r[m] = v[m];
                     // |B| of the zone as GetVariablesInterp1 forms it (src/server_cartesian.cc:470-473)
                     double b2 = v[3] * v[3];
                     b2 = b2 + v[4] * v[4];
                     b2 = b2 + v[5] * v[5];
                     r[6] = std::sqrt(b2);
                  }
   double bmin = INFINITY;
   for (int64_t z = 0; z < nzone; z++) bmin = std::min(bmin, G->rec[(size_t)z * SPX_GRID_REC + 6]);
   g.bmag_min = bmin;
   s->grid = G;
   return SPX_OK;
}

int spx_add_background_file(spx_sim* s, const char* class_name, const uint8_t* container, size_t len, const char* fname_pattern)
{
   if (!s || !class_name || !fname_pattern) return SPX_ERR_ARG;
   int rc = spx_add_plugin(s, SPX_BACKGROUND, class_name, container, len);
   if (rc) return rc;
   if (s->plan.bg.kind != SPX_BG_GRID) return SPX_OK;   // analytic backgrounds have no files
   // ReadCartesianHeader (src/reader_cartesian.cc:25-95): blocks per axis, domain corners, variables
   const std::string pat(fname_pattern);
   std::ifstream hdr(pat + ".info");
   int32_t nb[3] = {0, 0, 0}, nvar = 0;
   double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
   hdr >> nb[0] >> nb[1] >> nb[2] >> lo[0] >> lo[1] >> lo[2] >> hi[0] >> hi[1] >> hi[2] >> nvar;
   if (!hdr) return fail(s, SPX_ERR_ARG, "cannot read grid header " + pat + ".info");
   if (nb[0] < 1 || nb[1] < 1 || nb[2] < 1 || nvar < 1 || nvar > 64) return fail(s, SPX_ERR_ARG, "bad grid header " + pat + ".info");
   // ReadCartesianData (src/reader_cartesian.cc:98-103): the binary doubles, block by block
   const size_t n = (size_t)nvar * SPX_GRID_BLOCK * SPX_GRID_BLOCK * SPX_GRID_BLOCK * (size_t)nb[0] * nb[1] * nb[2];
   std::vector<double> data(n);
   std::ifstream dat(pat + ".out", std::ios::binary);
   dat.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(n * sizeof(double)));
   if (!dat || (size_t)dat.gcount() != n * sizeof(double)) return fail(s, SPX_ERR_ARG, "grid data file " + pat + ".out is short");
   return spx_set_background_grid(s, nb, lo, hi, nvar, data.data());
}

int spx_set_option(spx_sim* s, const char* name, double v)
{
   if (!s || !name) return SPX_ERR_ARG;
   std::string n(name);
   if (n == "DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE") { if (v < 0 || v > 2) return SPX_ERR_ARG; s->plan.spectrum_type = (int)v; }
   else if (n == "max_trajectory_steps") { if (v < 1) return SPX_ERR_ARG; s->plan.max_steps = (int64_t)v; }
   else if (n == "max_attempts") { if (v < 1) return SPX_ERR_ARG; s->plan.max_attempts = (int)v; }
   else if (n == "step_scale") { if (!(v > 0.0 && v <= 1.0)) return SPX_ERR_ARG; s->plan.dt_scale = v; }
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
   if (rng->mode != 0 && rng->mode != 1) return fail(s, SPX_ERR_ARG, "rng mode");
   if (rng->mode == 1 && (!rng->normals || rng->normal_stride < 1 || rng->uniform_stride < 0)) return fail(s, SPX_ERR_ARG, "replay needs draws");
   int rc = check_ready(s);
   if (rc) return rc;
   finalize_plan(s);
   rc = check_transport(s);
   if (rc) return rc;
   rc = ensure_device(s);
   if (rc) return rc;
   rc = bind_grid(s);
   if (rc) return rc;
   uint64_t lanes = s->lanes_opt > 0 ? (uint64_t)s->lanes_opt : auto_lanes();
   lanes = (lanes + s->block - 1) / s->block * s->block;
   lanes = (lanes + 255) / 256 * 256;   // whole blocks of the kernels' block size
   Pool& P = s->pool;
   if (!P.matches(s->plan, s->precision) || P.n_lanes != lanes || P.steps_per_launch != s->steps_per_launch ||
       P.graph_launches != s->graph_launches) {
      HIPCHK(s, P.init(s->plan, s->precision, lanes, s->steps_per_launch, s->graph_launches, 1));
   }
   else HIPCHK(s, P.reset_slots());
   P.sig = s->plan;
   rc = ensure_hist(s);
   if (rc) return rc;
   s->q.release();
   s->keep_records = keep_records != 0;
   HIPCHK(s, s->q.create(s->plan, s->hist_bins, first, n_traj, rng, s->keep_records, s->hist_counts.p, s->hist_acc.p,
                         s->hist_events.p, P.stream));
   const int qi = P.alloc_index();
   if (qi < 0) return fail(s, SPX_ERR_STATE, "query table full");
   HIPCHK(s, P.set_query(qi, s->q.desc));
   HIPCHK(s, P.push_queue(qi));
   s->run_first = first;
   s->run_n = n_traj;
   s->running = true;
   return SPX_OK;
}

static int read_stats(spx_sim* s, spx_run_stats* st, double wall, double kms, int launches)
{
   RunCtl ctl;
   HIPCHK(s, s->q.read_ctl(ctl, s->pool.stream));
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
      st->lanes = (int32_t)s->pool.n_lanes;
      st->step_kernel_ms = s->pool.step_ms;
      st->service_kernel_ms = s->pool.service_ms;
   }
   return SPX_OK;
}

int spx_advance(spx_sim* s, int n_launches, spx_run_stats* st)
{
   if (!s || n_launches < 0) return SPX_ERR_ARG;
   if (!s->running) return fail(s, SPX_ERR_STATE, "advance without begin");
   auto w0 = std::chrono::steady_clock::now();
   s->pool.step_ms = s->pool.service_ms = 0.0;
   float ms = 0.f;
   HIPCHK(s, s->pool.launch(n_launches, &ms));
   double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
   return read_stats(s, st, wall, ms, n_launches);
}

int spx_finish(spx_sim* s, spx_traj_record* records, spx_run_stats* st)
{
   if (!s) return SPX_ERR_ARG;
   if (!s->running) return fail(s, SPX_ERR_STATE, "finish without begin");
   auto w0 = std::chrono::steady_clock::now();
   s->pool.step_ms = s->pool.service_ms = 0.0;
   double kms = 0.0;
   int launches = 0;
   // Every launch advances each live slot by steps_per_launch steps; a trajectory needs at most
   // max_attempts * (max_steps + 1) steps, so this bound is never reached by a healthy run.
   const double guard = 64.0 + (double)s->plan.max_attempts * ((double)s->plan.max_steps + 1.0) / s->steps_per_launch *
                                   std::max(1.0, (double)s->run_n / (double)s->pool.n_lanes + 1.0);
   int batch = 1;
   while (true) {
      RunCtl ctl;
      HIPCHK(s, s->q.read_ctl(ctl, s->pool.stream));
      if (ctl.done >= s->run_n) break;
      if (launches > guard) return fail(s, SPX_ERR_STATE, "run did not drain");
      float ms = 0.f;
      HIPCHK(s, s->pool.launch(batch, &ms));
      kms += ms;
      launches += batch;
      if (batch < 8) batch *= 2;
   }
   double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
   int rc = read_stats(s, st, wall, kms, launches);
   if (rc) return rc;
   if (records && s->keep_records) {
      HIPCHK(s, hipMemcpyAsync(records, s->q.records.p, sizeof(spx_traj_record) * s->run_n, hipMemcpyDeviceToHost, s->pool.stream));
      HIPCHK(s, hipStreamSynchronize(s->pool.stream));
   }
   HIPCHK(s, s->pool.remove_queue(0));
   s->pool.free_index(0);
   s->running = false;
   return SPX_OK;
}

int spx_run(spx_sim* s, uint64_t first, uint64_t n_traj, const spx_rng* rng, spx_traj_record* records, spx_run_stats* st)
{
   int rc = spx_begin(s, first, n_traj, rng, records != nullptr);
   if (rc) return rc;
   return spx_finish(s, records, st);
}

int spx_run_static(spx_sim* s, uint64_t first, uint64_t n_traj, uint64_t seed, int64_t* counts, double* sums, int64_t* n_events,
                   spx_run_stats* st)
{
   if (!s) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "static run during a run");
   int rc = check_ready(s);
   if (rc) return rc;
   finalize_plan(s);
   rc = check_transport(s);
   if (rc) return rc;
   if (s->plan.bg.kind == SPX_BG_GRID) return fail(s, SPX_ERR_ARG, "static run: a gridded background needs spx_run (its grid is allocated device memory)");
   if (s->hist_bins > SPX_STATIC_BINS || acc_components(s->plan) > SPX_STATIC_COMPONENTS)
      return fail(s, SPX_ERR_LIMIT, "static run: histograms beyond SPX_STATIC_BINS bins or SPX_STATIC_COMPONENTS bin components");
   rc = ensure_device(s);
   if (rc) return rc;
   StaticResult res;
   HIPCHK(s, static_run(s->plan, s->precision, s->steps_per_launch, first, n_traj, seed, counts, sums, n_events, res));
   if (st) {
      std::memset(st, 0, sizeof(*st));
      st->n_traj = n_traj;
      st->done = res.ctl.done;
      st->failed = res.ctl.failed;
      st->steps_done = res.ctl.steps_done;
      st->steps_executed = res.ctl.steps_executed;
      for (int k = 0; k < 8; k++) st->discards[k] = res.ctl.discards[k];
      std::memcpy(&st->t_min, &res.ctl.tmin_bits, 8);
      std::memcpy(&st->t_max, &res.ctl.tmax_bits, 8);
      st->wall_s = res.wall_s;
      st->launches = res.launches;
      st->lanes = (int32_t)res.lanes;
   }
   return SPX_OK;
}

int spx_autotune(spx_sim* s, uint64_t n_probe, uint64_t seed, spx_tune_entry* table, int max_entries, int* n_entries)
{
   if (!s || n_probe == 0) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "autotune during a run");
   int rc = ensure_device(s);
   if (rc) return rc;
   hipDeviceProp_t prop;
   HIPCHK(s, hipGetDeviceProperties(&prop, 0));
   const uint64_t cus = (uint64_t)prop.multiProcessorCount;
   const int waves_per_simd[] = {1, 2, 3, 4, 6, 8};
   const int spls[] = {250, 1000, 4000};
   const int64_t keep_lanes = s->lanes_opt;
   const int keep_spl = s->steps_per_launch, keep_graph = s->graph_launches;
   s->graph_launches = 0;
   double best = -1.0;
   int64_t best_lanes = 0;
   int best_spl = keep_spl;
   int n = 0;
   spx_rng rng{};
   rng.mode = 0;
   rng.seed = seed;
   for (int w : waves_per_simd) {
      for (int spl : spls) {
         s->lanes_opt = (int64_t)(cus * 4 * 64 * (uint64_t)w);
         s->steps_per_launch = spl;
         rc = spx_begin(s, 0, n_probe, &rng, 0);
         if (rc) break;
         spx_run_stats w0, w1;
         rc = spx_advance(s, 1, &w0);
         if (!rc) rc = spx_advance(s, 2, &w1);
         // drop the probe run: the slots are reset and the histograms zeroed below
         HIPCHK(s, s->pool.remove_queue(0));
         s->pool.free_index(0);
         s->running = false;
         if (rc) break;
         const double rate = (double)(w1.steps_executed - w0.steps_executed) / std::max(w1.wall_s, 1e-9);
         if (table && n < max_entries) {
            table[n].lanes = s->lanes_opt;
            table[n].steps_per_launch = spl;
            table[n].graph_launches = 0;
            table[n].steps_per_s = rate;
         }
         n++;
         if (rate > best) { best = rate; best_lanes = s->lanes_opt; best_spl = spl; }
      }
      if (rc) break;
   }
   // the best shape replayed as a HIP graph of two launches: kept when it runs more than 2% faster
   int use_graph = 0;
   if (!rc) {
      s->lanes_opt = best_lanes;
      s->steps_per_launch = best_spl;
      s->graph_launches = 2;
      rc = spx_begin(s, 0, n_probe, &rng, 0);
      if (!rc) {
         spx_run_stats w0, w1;
         rc = spx_advance(s, 2, &w0);   // captures the graph and replays it once
         if (!rc) rc = spx_advance(s, 2, &w1);
         HIPCHK(s, s->pool.remove_queue(0));
         s->pool.free_index(0);
         s->running = false;
         if (!rc) {
            const double rate = (double)(w1.steps_executed - w0.steps_executed) / std::max(w1.wall_s, 1e-9);
            if (table && n < max_entries) {
               table[n].lanes = best_lanes;
               table[n].steps_per_launch = best_spl;
               table[n].graph_launches = 2;
               table[n].steps_per_s = rate;
            }
            n++;
            use_graph = rate > 1.02 * best;
         }
      }
   }
   if (n_entries) *n_entries = n;
   s->graph_launches = use_graph ? 2 : 0;
   if (rc) {
      s->lanes_opt = keep_lanes;
      s->steps_per_launch = keep_spl;
      s->graph_launches = keep_graph;
      return rc;
   }
   return spx_reset_distros(s);
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
   if (!s->hist_counts.p) {
      if (counts) std::memset(counts, 0, sizeof(int64_t) * D.nb);
      if (sums) std::memset(sums, 0, sizeof(double) * D.nb * D.width);
      if (n_events) *n_events = 0;
      return SPX_OK;
   }
   HIPCHK(s, read_hist(s->plan, d, s->hist_counts.p, s->hist_acc.p, s->hist_events.p, counts, sums, n_events, s->pool.stream));
   return SPX_OK;
}

int spx_reset_distros(spx_sim* s)
{
   if (!s) return SPX_ERR_ARG;
   if (!s->hist_counts.p) return SPX_OK;
   HIPCHK(s, hipMemsetAsync(s->hist_counts.p, 0, sizeof(unsigned long long) * s->hist_counts.n, s->pool.stream));
   HIPCHK(s, hipMemsetAsync(s->hist_acc.p, 0, sizeof(unsigned long long) * s->hist_acc.n, s->pool.stream));
   HIPCHK(s, hipMemsetAsync(s->hist_events.p, 0, sizeof(unsigned long long) * SPX_MAX_DIST, s->pool.stream));
   HIPCHK(s, hipStreamSynchronize(s->pool.stream));
   return SPX_OK;
}

// ---- distribution files (src/distribution_templated.cc:394-632) --------------------------------------
static int ensure_hist(spx_sim* s)
{
   const size_t nb = (size_t)std::max<int64_t>(s->hist_bins, 1);
   const size_t nl = (size_t)std::max<int64_t>(acc_components(s->plan), 1) * SPX_ACC_LIMBS;
   if (s->hist_counts.p && s->hist_counts.n >= nb && s->hist_acc.n >= nl) return SPX_OK;
   int rc = ensure_device(s);
   if (rc) return rc;
   if (!s->pool.stream) HIPCHK(s, hipStreamCreateWithFlags(&s->pool.stream, hipStreamNonBlocking));
   HIPCHK(s, s->hist_counts.alloc(nb));
   HIPCHK(s, s->hist_acc.alloc(nl));
   HIPCHK(s, s->hist_events.alloc(SPX_MAX_DIST));
   HIPCHK(s, hipMemsetAsync(s->hist_counts.p, 0, sizeof(unsigned long long) * nb, s->pool.stream));
   HIPCHK(s, hipMemsetAsync(s->hist_acc.p, 0, sizeof(unsigned long long) * nl, s->pool.stream));
   HIPCHK(s, hipMemsetAsync(s->hist_events.p, 0, sizeof(unsigned long long) * SPX_MAX_DIST, s->pool.stream));
   HIPCHK(s, hipStreamSynchronize(s->pool.stream));
   return SPX_OK;
}

int spx_dump_distro(const spx_sim* sc, int d, const char* file_name)
{
   spx_sim* s = const_cast<spx_sim*>(sc);
   if (!s || !file_name || d < 0 || d >= s->plan.n_dist) return SPX_ERR_ARG;
   const spx_ds_desc& D = s->plan.dist[d];
   const DistMeta& M = s->dmeta[d];
   std::vector<int64_t> c((size_t)D.nb);
   std::vector<double> sm((size_t)D.nb * D.width);
   int64_t nev = 0;
   int rc = spx_get_distro(s, d, c.data(), sm.data(), &nev);
   if (rc) return rc;
   // the reference keeps int counts and an int event count
   if (nev > INT32_MAX) return fail(s, SPX_ERR_LIMIT, "event count exceeds the file format's 32-bit integer");
   std::vector<int32_t> c32((size_t)D.nb);
   for (size_t b = 0; b < c.size(); b++) {
      if (c[b] > INT32_MAX) return fail(s, SPX_ERR_LIMIT, "bin count exceeds the file format's 32-bit integer");
      c32[b] = (int32_t)c[b];
   }
   std::ofstream f(file_name, std::ios::binary);
   if (!f) return fail(s, SPX_ERR_ARG, std::string("cannot write ") + file_name);
   const uint32_t name_len = (uint32_t)M.cls.size(), size = M.n_records, length = (uint32_t)M.container.size(), zero = 0;
   const int32_t nev32 = (int32_t)nev;
   const uint8_t specie = 0;   // sizeof(Specie<...>) == 1: the empty specie object is one byte
   f.write((const char*)&name_len, 4);
   f.write(M.cls.data(), name_len);
   f.write((const char*)&specie, 1);
   f.write((const char*)&size, 4);
   f.write((const char*)&length, 4);
   f.write((const char*)M.container.data(), length);
   f.write((const char*)&nev32, 4);
   f.write((const char*)c32.data(), sizeof(int32_t) * c32.size());
   f.write((const char*)sm.data(), sizeof(double) * sm.size());
   f.write((const char*)&zero, 4);   // values_record: per-event records are not kept
   f.write((const char*)&zero, 4);   // weights_record
   if (!f) return fail(s, SPX_ERR_ARG, std::string("write failed: ") + file_name);
   return SPX_OK;
}

int spx_restore_distro(spx_sim* s, int d, const char* file_name)
{
   if (!s || !file_name || d < 0 || d >= s->plan.n_dist) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "restore during a run");
   const spx_ds_desc& D = s->plan.dist[d];
   const DistMeta& M = s->dmeta[d];
   std::ifstream f(file_name, std::ios::binary);
   if (!f) return fail(s, SPX_ERR_ARG, std::string("cannot read ") + file_name);
   uint32_t name_len = 0, size = 0, length = 0;
   f.read((char*)&name_len, 4);
   if (!f || name_len > 4096) return fail(s, SPX_ERR_ARG, "not a distribution file");
   std::string cls(name_len, '\0');
   f.read(&cls[0], name_len);
   // DistributionTemplated::Restore leaves the distribution alone when the class differs
   if (cls != M.cls) return fail(s, SPX_ERR_ARG, "distribution file holds " + cls + ", not " + M.cls);
   uint8_t specie = 0;
   f.read((char*)&specie, 1);
   f.read((char*)&size, 4);
   f.read((char*)&length, 4);
   if (!f || length > (1u << 20)) return fail(s, SPX_ERR_ARG, "bad container in distribution file");
   std::vector<uint8_t> cont(length);
   f.read((char*)cont.data(), length);
   // the binning must be this distribution's (the reference would re-read its setup from the file)
   Reader r{cont.data(), cont.data() + cont.size()};
   int32_t nb3[3];
   r.mi(nb3);
   int64_t nb = 1;
   for (int k = 0; k < 3; k++) nb *= std::max(nb3[k], 1);
   if (r.err || nb != D.nb) return fail(s, SPX_ERR_ARG, "distribution file has a different binning");
   int32_t nev32 = 0;
   std::vector<int32_t> c32((size_t)D.nb);
   std::vector<double> sm((size_t)D.nb * D.width);
   f.read((char*)&nev32, 4);
   f.read((char*)c32.data(), sizeof(int32_t) * c32.size());
   f.read((char*)sm.data(), sizeof(double) * sm.size());
   if (!f) return fail(s, SPX_ERR_ARG, "distribution file is short");
   int rc = ensure_hist(s);
   if (rc) return rc;
   std::vector<unsigned long long> c64((size_t)D.nb), acc(sm.size() * SPX_ACC_LIMBS);
   for (int64_t b = 0; b < D.nb; b++) c64[b] = (unsigned long long)std::max(c32[b], 0);
   for (size_t i = 0; i < sm.size(); i++) acc_from_double(sm[i], acc.data() + i * SPX_ACC_LIMBS);
   const unsigned long long ev = (unsigned long long)std::max(nev32, 0);
   const int64_t off = s->plan.hist_off[d];
   HIPCHK(s, hipMemcpyAsync(s->hist_counts.p + off, c64.data(), sizeof(unsigned long long) * D.nb, hipMemcpyHostToDevice, s->pool.stream));
   HIPCHK(s, hipMemcpyAsync(s->hist_acc.p + (size_t)s->plan.acc_off[d] * SPX_ACC_LIMBS, acc.data(), sizeof(unsigned long long) * acc.size(),
                            hipMemcpyHostToDevice, s->pool.stream));
   HIPCHK(s, hipMemcpyAsync(s->hist_events.p + d, &ev, sizeof(ev), hipMemcpyHostToDevice, s->pool.stream));
   HIPCHK(s, hipStreamSynchronize(s->pool.stream));
   return SPX_OK;
}

// Print1D / Print2D (src/distribution_templated.cc:546-632): the same columns, widths and precision.
namespace {
double bin_cent(const spx_ds_desc& D, int ijk, int bin)   // DistributionBase::BinCent (src/distribution_base.hh:267-272)
{
   if (bin < 0 || bin >= D.n_bins[ijk]) return 0.0;
   if (D.log_bins[ijk]) return std::pow(10.0, D.limits0[ijk] + (bin + 0.5) * D.bin_size[ijk]);
   return D.limits0[ijk] + (bin + 0.5) * D.bin_size[ijk];
}
void put_d(std::string& o, double v) { char b[64]; std::snprintf(b, sizeof b, "%20.10g", v); o += b; }
void put_l(std::string& o, long v) { char b[64]; std::snprintf(b, sizeof b, "%20ld", v); o += b; }
// PrintSumDistro (:477-537): mean and sum per component
void put_sums(std::string& o, const DistMeta& M, int width, long n, const double* sum, bool phys)
{
   if (width == 1) {
      put_d(o, (n ? sum[0] / n : 0.0) * (phys ? M.unit_distro[0] : 1.0));
      put_d(o, sum[0] * (phys ? M.unit_distro[0] : 1.0));
      return;
   }
   for (int pass = 0; pass < 2; pass++)
      for (int k = 0; k < width; k++) {
         if (!n) { put_d(o, 0.0); continue; }
         const double v = pass == 0 ? sum[k] / n : sum[k];
         put_d(o, phys ? v * M.unit_distro[k] : v);
      }
}
}  // namespace

static int print_distro(spx_sim* s, int d, int a1, int a2, const char* file_name, int phys)
{
   if (!s || !file_name || d < 0 || d >= s->plan.n_dist || a1 < 0 || a1 > 2 || a2 < 0 || a2 > 2) return SPX_ERR_ARG;
   const spx_ds_desc& D = s->plan.dist[d];
   const DistMeta& M = s->dmeta[d];
   std::vector<int64_t> c((size_t)D.nb);
   std::vector<double> sm((size_t)D.nb * D.width);
   int64_t nev = 0;
   int rc = spx_get_distro(s, d, c.data(), sm.data(), &nev);
   if (rc) return rc;
   const int* nb = D.n_bins;
   auto lin = [&](const int* bin) { return (int64_t)nb[2] * (nb[1] * bin[0] + bin[1]) + bin[2]; };   // MultiIndex::LinIdx
   std::string o = "# Total number of events: " + std::to_string(nev) + "\n\n";
   const bool one = (a1 == a2);
   const int c1 = one ? (a1 + 1) % 3 : 3 - a1 - a2, c2 = one ? (a1 + 2) % 3 : -1;   // collapsed dimensions
   int bin[3] = {0, 0, 0};
   double sum[9];
   for (bin[a1] = 0; bin[a1] < nb[a1]; bin[a1]++) {
      for (int j = 0; j < (one ? 1 : nb[a2]); j++) {
         if (!one) bin[a2] = j;
         long n = 0;
         for (int k = 0; k < D.width; k++) sum[k] = 0.0;
         if (one) {
            for (bin[c1] = 0; bin[c1] < nb[c1]; bin[c1]++)
               for (bin[c2] = 0; bin[c2] < nb[c2]; bin[c2]++) {
                  const int64_t l = lin(bin);
                  n += (long)c[l];
                  for (int k = 0; k < D.width; k++) sum[k] += sm[l * D.width + k];
               }
         }
         else {
            for (bin[c1] = 0; bin[c1] < nb[c1]; bin[c1]++) {
               const int64_t l = lin(bin);
               n += (long)c[l];
               for (int k = 0; k < D.width; k++) sum[k] += sm[l * D.width + k];
            }
         }
         put_d(o, bin_cent(D, a1, bin[a1]) * (phys ? M.unit_val[a1] : 1.0));
         if (!one) put_d(o, bin_cent(D, a2, bin[a2]) * (phys ? M.unit_val[a2] : 1.0));
         put_sums(o, M, D.width, n, sum, phys != 0);
         put_l(o, n);
         o += "\n";
      }
   }
   std::ofstream f(file_name);
   if (!f) return fail(s, SPX_ERR_ARG, std::string("cannot write ") + file_name);
   f << o;
   return f ? SPX_OK : fail(s, SPX_ERR_ARG, std::string("write failed: ") + file_name);
}

int spx_print_distro1d(const spx_sim* s, int d, int ijk, const char* file_name, int phys_units)
{
   return print_distro(const_cast<spx_sim*>(s), d, ijk, ijk, file_name, phys_units);
}

int spx_print_distro2d(const spx_sim* s, int d, int ijk1, int ijk2, const char* file_name, int phys_units)
{
   return print_distro(const_cast<spx_sim*>(s), d, ijk1, ijk2, file_name, phys_units);
}

// ---- SimulationMaster's task loop (src/simulation.cc:496-519, :558-643, :699-835), serial form ---------
int spx_set_tasks(spx_sim* s, int64_t n_traj, int64_t batch_size, int64_t max_traj_per_worker)
{
   if (!s) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "tasks cannot change during a run");
   if (n_traj < 0) n_traj = 0;
   if (batch_size < 1 || batch_size > n_traj) batch_size = std::min<int64_t>(1, n_traj);
   if (max_traj_per_worker * 1 < n_traj) max_traj_per_worker = 0;   // one worker: the device
   s->task_n = n_traj;
   s->task_batch = batch_size;
   s->task_max_per_worker = max_traj_per_worker;
   return SPX_OK;
}

int spx_distro_file_name(spx_sim* s, const char* base)
{
   if (!s || !base) return SPX_ERR_ARG;
   s->distro_file_name = base;
   return SPX_OK;
}

int spx_restore_distro_at_start(spx_sim* s, int d)
{
   if (!s || d < 0 || d >= s->plan.n_dist) return SPX_ERR_ARG;
   s->restore_at_start |= 1u << d;
   return SPX_OK;
}

static std::string distro_file(const spx_sim* s, int d) { return s->distro_file_name + std::to_string(d) + ".out"; }

static int dump_all(spx_sim* s)
{
   for (int d = 0; d < s->plan.n_dist; d++) {
      const int rc = spx_dump_distro(s, d, distro_file(s, d).c_str());
      if (rc) return rc;
   }
   return SPX_OK;
}

int spx_main_loop(spx_sim* s, const spx_rng* rng, spx_run_stats* st)
{
   if (!s || !rng) return SPX_ERR_ARG;
   if (s->running) return fail(s, SPX_ERR_STATE, "main loop during a run");
   const bool fwd = s->plan.time_flow == SPX_FLOW_FORWARD;
   const double unit_time = unit_length_h / unit_velocity_h;
   const auto t_start = std::chrono::system_clock::now();
   auto seconds_since_start = [&]() {
      return (long)std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - t_start).count();
   };
   // MasterStart (:699-737): fresh distributions, then the ones flagged by RestoreDistro from their files
   int rc = spx_reset_distros(s);
   if (rc) return rc;
   for (int d = 0; d < s->plan.n_dist; d++) {
      if (!((s->restore_at_start >> d) & 1u)) continue;
      // Restore leaves the distribution empty when there is no file yet (src/distribution_templated.cc:436-437)
      if (!std::ifstream(distro_file(s, d), std::ios::binary).is_open()) continue;
      rc = spx_restore_distro(s, d, distro_file(s, d).c_str());
      if (rc) return rc;
   }
   double shortest = fwd ? 1.0E300 : -1.0E300, longest = 0.0;
   double integ_ms = 0.0;   // the serial master's elapsed_time: time spent integrating, here the kernels' time
   const int64_t n_total = s->task_n;
   std::fprintf(stderr, "Trajectories left: %s\n", std::to_string(n_total).c_str());
   spx_run_stats run{};
   if (n_total > 0) {
      rc = spx_begin(s, s->next_traj, (uint64_t)n_total, rng, 0);
      if (rc) return rc;
      const auto w0 = std::chrono::steady_clock::now();
      s->pool.step_ms = s->pool.service_ms = 0.0;
      const double guard = 64.0 + (double)s->plan.max_attempts * ((double)s->plan.max_steps + 1.0) / s->steps_per_launch *
                                      std::max(1.0, (double)n_total / (double)s->pool.n_lanes + 1.0);
      int percentage_done = 0, launches = 0, burst = 1;
      while (true) {
         RunCtl ctl;
         HIPCHK(s, s->q.read_ctl(ctl, s->pool.stream));
         // DecrementTrajectoryCount (:558-643): the serial master counts whole batches as completed
         const int64_t done = (int64_t)ctl.done;
         const int64_t completed = done >= n_total ? n_total : done / s->task_batch * s->task_batch;
         const int percentage_new = (int)(100 * completed / n_total);
         if (percentage_new > percentage_done) {
            percentage_done = percentage_new;
            rc = dump_all(s);   // the partial distributions
            if (rc) return rc;
            const long elapsed = seconds_since_start();
            const long remaining = (long)(n_total - completed);
            const long time_left = remaining * (elapsed + 1) / (long)completed;
            const long integ_s = (long)(0.001 * integ_ms);
            std::ostringstream m;
            m << "\n";
            m << "Trajectories completed: " << std::setw(20) << percentage_done << " %" << "\n";
            m << "Est. time to finish:    " << std::setw(20) << time_left << " seconds" << "\n";
            m << "Time allocated for job: " << std::setw(20) << "inf" << " (No WM)" << "\n";   // no workload manager
            m << "SIMULATION EFFICIENCY: " << "\n" << "    " << 100 * integ_s / (elapsed + 1) << " %"
              << " of time was spent computing trajectories." << "\n";
            std::fputs(m.str().c_str(), stderr);
         }
         if (done >= n_total) break;
         if (launches > guard) return fail(s, SPX_ERR_STATE, "run did not drain");
         float ms = 0.f;
         HIPCHK(s, s->pool.launch(burst, &ms));
         integ_ms += ms;
         launches += burst;
         if (burst < 8) burst *= 2;
      }
      const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
      rc = spx_finish(s, nullptr, &run);
      if (rc) return rc;
      run.wall_s = wall;
      run.kernel_ms = integ_ms;
      run.launches = launches;
      s->next_traj += (uint64_t)n_total;
      // WorkerDuties (:266-273) keeps the extremes of ElapsedTime() = t - t0, negative for backward runs
      if (run.done > run.failed) {
         shortest = fwd ? run.t_min : 0.0 - run.t_min;
         longest = fwd ? run.t_max : 0.0 - run.t_max;
      }
   }
   if (st) *st = run;
   // MasterFinish (:776-811)
   std::fprintf(stderr, "%s:%d: info message: Simulation completed\n", __FILE__, __LINE__);
   rc = dump_all(s);
   if (rc) return rc;
   std::ostringstream m;
   m << "Shortest simulated trajectory time = " << shortest * unit_time << " s" << "\n";
   m << "Longest simulated trajectory time = " << longest * unit_time << " s" << "\n";
   m << "Time per trajectory integration = " << integ_ms / (double)n_total << " ms" << "\n";
   m << "Total time of this simulation: " << std::setw(20) << seconds_since_start() << " seconds." << "\n";
   std::fputs(m.str().c_str(), stderr);
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

// ===================================================================================================
// Query server: a persistent engine that runs many queries at once.
//
// Queries are submitted from any thread with a configured simulation as their prototype; they wait
// in a FIFO until their device footprint fits the memory budget and their pool has a free query-table
// entry, then join their pool's admission queue.  A worker thread keeps every pool with admitted
// queries launching; inside the kernels, each wave draws trajectories from its query and moves to the
// oldest admitted query with work left when it runs dry, so queries that arrive while others run
// start on the slots those free (continuous batching).  Each query has its own random-stream key,
// control block and histograms, so its result does not depend on what else runs beside it.
// ===================================================================================================
namespace {

struct SQuery {
   int64_t id = 0;
   int state = SPX_Q_QUEUED;
   spx_plan_desc plan;
   int64_t hist_bins = 0;
   uint64_t first = 0, n_traj = 0;
   spx_rng rng{};
   size_t footprint = 0;
   Pool* pool = nullptr;
   int qidx = -1;
   QueryRes res;
   double t_submit = 0.0, t_admit = -1.0, t_start = -1.0, t_done = -1.0;
   RunCtl ctl{};
   std::vector<std::vector<int64_t>> counts;
   std::vector<std::vector<double>> sums;
   std::vector<int64_t> nev;
   std::shared_ptr<GridData> grid;     // keeps a gridded background's device data alive
};

}  // namespace

struct spx_server {
   spx_server_opts o{};
   std::string err;
   std::mutex mu;
   std::condition_variable cv_work, cv_done;
   std::thread worker;
   bool stop = false;
   bool failed = false;
   std::deque<SQuery*> pending;
   std::vector<SQuery*> running;
   std::map<int64_t, SQuery*> all;
   std::vector<Pool*> pools;
   int64_t next_id = 1;
   size_t hbm_used = 0;
   std::chrono::steady_clock::time_point t0;
   uint64_t submitted = 0, completed = 0;
   double busy_s = 0.0;

   double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

static Pool* server_pool(spx_server* S, const spx_plan_desc& plan, hipError_t& e)
{
   for (Pool* p : S->pools)
      if (p->matches(plan, S->o.precision)) return p;
   Pool* p = new Pool();
   const uint64_t lanes = S->o.lanes > 0 ? (uint64_t)((S->o.lanes + 255) / 256 * 256) : auto_lanes();
   e = p->init(plan, S->o.precision, lanes, S->o.steps_per_launch, 0, S->o.max_queries);
   if (e != hipSuccess) {
      delete p;
      return nullptr;
   }
   S->pools.push_back(p);
   return p;
}

static void server_fail(spx_server* S, const std::string& msg)
{
   std::lock_guard<std::mutex> lk(S->mu);
   S->err = msg;
   S->failed = true;
   for (SQuery* q : S->running) { q->state = SPX_Q_FAILED; q->t_done = S->now(); }
   for (SQuery* q : S->pending) { q->state = SPX_Q_FAILED; q->t_done = S->now(); }
   S->running.clear();
   S->pending.clear();
   S->cv_done.notify_all();
}

static void server_loop(spx_server* S)
{
   if (hipSetDevice(0) != hipSuccess) { server_fail(S, "hipSetDevice failed"); return; }
   hipEvent_t e0 = nullptr, e1 = nullptr;
   (void)hipEventCreate(&e0);
   (void)hipEventCreate(&e1);
   while (true) {
      std::vector<SQuery*> admitted;
      {
         std::unique_lock<std::mutex> lk(S->mu);
         S->cv_work.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->running.empty(); });
         if (S->stop) break;
         // admission in arrival order, by device-memory footprint and free query-table entries
         while (!S->pending.empty()) {
            SQuery* q = S->pending.front();
            if (!S->running.empty() && S->hbm_used + q->footprint > (size_t)S->o.hbm_budget_bytes) break;
            hipError_t e = hipSuccess;
            Pool* pl = server_pool(S, q->plan, e);
            if (!pl) { lk.unlock(); server_fail(S, std::string("pool: ") + hipGetErrorString(e)); goto out; }
            const int qi = pl->alloc_index();
            if (qi < 0) break;
            S->pending.pop_front();
            q->pool = pl;
            q->qidx = qi;
            q->state = SPX_Q_RUNNING;
            q->t_admit = S->now();
            S->hbm_used += q->footprint;
            S->running.push_back(q);
            admitted.push_back(q);
         }
      }
      for (SQuery* q : admitted) {
         hipError_t e = q->res.create(q->plan, q->hist_bins, q->first, q->n_traj, &q->rng, false, nullptr, nullptr, nullptr, q->pool->stream);
         if (e == hipSuccess) e = q->pool->set_query(q->qidx, q->res.desc);
         if (e == hipSuccess) e = q->pool->push_queue(q->qidx);
         if (e != hipSuccess) { server_fail(S, std::string("admit: ") + hipGetErrorString(e)); goto out; }
      }
      // one round of launches on every pool that has admitted queries; pools run on their own streams
      {
         (void)hipEventRecord(e0, nullptr);
         bool any = false;
         for (Pool* pl : S->pools) {
            if (pl->host_queue.empty()) continue;
            hipError_t e = pl->enqueue(S->o.launches_per_poll);
            if (e != hipSuccess) { server_fail(S, std::string("launch: ") + hipGetErrorString(e)); goto out; }
            any = true;
         }
         for (Pool* pl : S->pools)
            if (!pl->host_queue.empty()) {
               hipError_t e = hipStreamSynchronize(pl->stream);
               if (e != hipSuccess) { server_fail(S, std::string("sync: ") + hipGetErrorString(e)); goto out; }
            }
         (void)hipEventRecord(e1, nullptr);
         (void)hipEventSynchronize(e1);
         float ms = 0.f;
         if (any && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) S->busy_s += 1e-3 * ms;
      }
      // completions
      {
         std::vector<SQuery*> snapshot;
         {
            std::lock_guard<std::mutex> lk(S->mu);
            snapshot = S->running;
         }
         for (SQuery* q : snapshot) {
            RunCtl c;
            hipError_t e = q->res.read_ctl(c, q->pool->stream);
            if (e != hipSuccess) { server_fail(S, std::string("poll: ") + hipGetErrorString(e)); goto out; }
            bool finished = c.done >= q->n_traj;
            if (finished) {
               q->counts.resize((size_t)q->plan.n_dist);
               q->sums.resize((size_t)q->plan.n_dist);
               q->nev.resize((size_t)q->plan.n_dist);
               for (int d = 0; d < q->plan.n_dist; d++) {
                  q->counts[(size_t)d].resize((size_t)q->plan.dist[d].nb);
                  q->sums[(size_t)d].resize((size_t)q->plan.dist[d].nb * q->plan.dist[d].width);
                  e = q->res.read_distro(q->plan, d, q->counts[(size_t)d].data(), q->sums[(size_t)d].data(), &q->nev[(size_t)d], q->pool->stream);
                  if (e != hipSuccess) { server_fail(S, std::string("result: ") + hipGetErrorString(e)); goto out; }
               }
               e = q->pool->remove_queue(q->qidx);
               if (e != hipSuccess) { server_fail(S, std::string("retire: ") + hipGetErrorString(e)); goto out; }
               q->pool->free_index(q->qidx);
               q->res.release();
            }
            std::lock_guard<std::mutex> lk(S->mu);
            q->ctl = c;
            // the round in which the query's first trajectory was claimed (its queue wait ends)
            if (q->t_start < 0.0 && c.next > q->first) q->t_start = S->now();
            if (finished) {
               q->state = SPX_Q_DONE;
               q->t_done = S->now();
               S->hbm_used -= q->footprint;
               S->running.erase(std::remove(S->running.begin(), S->running.end(), q), S->running.end());
               S->completed++;
               S->cv_done.notify_all();
            }
         }
      }
   }
out:
   if (e0) (void)hipEventDestroy(e0);
   if (e1) (void)hipEventDestroy(e1);
}

extern "C" {

int spx_server_create(const spx_server_opts* opts, spx_server** out)
{
   if (!out) return SPX_ERR_ARG;
   *out = nullptr;
   int n = 0;
   if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SPX_ERR_DEVICE;
   spx_server* S = new spx_server();
   if (opts) S->o = *opts;
   if (S->o.precision != SPX_PREC_FP64 && S->o.precision != SPX_PREC_MIXED) { delete S; return SPX_ERR_ARG; }
   if (S->o.steps_per_launch <= 0) S->o.steps_per_launch = 500;
   if (S->o.hbm_budget_bytes <= 0) S->o.hbm_budget_bytes = 4.0 * (1ull << 30);
   if (S->o.max_queries <= 0) S->o.max_queries = 64;
   if (S->o.launches_per_poll <= 0) S->o.launches_per_poll = 1;
   S->t0 = std::chrono::steady_clock::now();
   S->worker = std::thread(server_loop, S);
   *out = S;
   return SPX_OK;
}

int spx_server_submit(spx_server* S, const spx_sim* proto, uint64_t first, uint64_t n_traj, uint64_t seed, int64_t* query_id)
{
   if (!S || !proto || !query_id || n_traj == 0) return SPX_ERR_ARG;
   spx_sim* s = const_cast<spx_sim*>(proto);
   int rc = check_ready(s);
   if (rc) { S->err = s->err; return rc; }
   finalize_plan(s);
   rc = check_transport(s);
   if (rc) { S->err = s->err; return rc; }
   if (s->plan.bg.kind == SPX_BG_GRID) {
      rc = ensure_device(s);
      if (!rc) rc = bind_grid(s);
      if (rc) { S->err = s->err; return rc; }
   }
   SQuery* q = new SQuery();
   q->plan = s->plan;
   q->grid = s->grid;
   q->hist_bins = s->hist_bins;
   q->first = first;
   q->n_traj = n_traj;
   q->rng.mode = 0;
   q->rng.seed = seed;
   q->footprint = QueryRes::footprint(q->plan, q->hist_bins, n_traj, &q->rng, false);
   if (q->grid) q->footprint += q->grid->bytes();   // charged to every query that uses the grid
   std::lock_guard<std::mutex> lk(S->mu);
   if (S->failed) { delete q; return SPX_ERR_DEVICE; }
   q->id = S->next_id++;
   q->t_submit = S->now();
   S->all[q->id] = q;
   S->pending.push_back(q);
   S->submitted++;
   *query_id = q->id;
   S->cv_work.notify_all();
   return SPX_OK;
}

int spx_server_info(spx_server* S, int64_t id, spx_query_info* info)
{
   if (!S || !info) return SPX_ERR_ARG;
   std::lock_guard<std::mutex> lk(S->mu);
   auto it = S->all.find(id);
   if (it == S->all.end()) return SPX_ERR_ARG;
   const SQuery* q = it->second;
   std::memset(info, 0, sizeof(*info));
   info->state = q->state;
   info->n_traj = q->n_traj;
   info->done = q->ctl.done;
   info->failed = q->ctl.failed;
   info->steps_done = q->ctl.steps_done;
   info->steps_executed = q->ctl.steps_executed;
   info->t_submit = q->t_submit;
   info->t_admit = q->t_admit;
   info->t_start = q->t_start;
   info->t_done = q->t_done;
   info->latency_s = q->t_done >= 0.0 ? q->t_done - q->t_submit : -1.0;
   info->footprint_bytes = (int64_t)q->footprint;
   return SPX_OK;
}

int spx_server_wait(spx_server* S, int64_t id, double timeout_s)
{
   if (!S) return SPX_ERR_ARG;
   std::unique_lock<std::mutex> lk(S->mu);
   auto it = S->all.find(id);
   if (it == S->all.end()) return SPX_ERR_ARG;
   SQuery* q = it->second;
   auto pred = [&] { return q->state == SPX_Q_DONE || q->state == SPX_Q_FAILED; };
   if (timeout_s < 0) S->cv_done.wait(lk, pred);
   else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), pred)) return SPX_ERR_LIMIT;
   return q->state == SPX_Q_DONE ? SPX_OK : SPX_ERR_DEVICE;
}

int spx_server_result(spx_server* S, int64_t id, int d, int64_t* counts, double* sums, int64_t* n_events)
{
   if (!S) return SPX_ERR_ARG;
   std::lock_guard<std::mutex> lk(S->mu);
   auto it = S->all.find(id);
   if (it == S->all.end()) return SPX_ERR_ARG;
   const SQuery* q = it->second;
   if (q->state != SPX_Q_DONE) return SPX_ERR_STATE;
   if (d < 0 || d >= q->plan.n_dist) return SPX_ERR_ARG;
   if (counts) std::memcpy(counts, q->counts[(size_t)d].data(), sizeof(int64_t) * q->counts[(size_t)d].size());
   if (sums) std::memcpy(sums, q->sums[(size_t)d].data(), sizeof(double) * q->sums[(size_t)d].size());
   if (n_events) *n_events = q->nev[(size_t)d];
   return SPX_OK;
}

int spx_server_release(spx_server* S, int64_t id)
{
   if (!S) return SPX_ERR_ARG;
   std::lock_guard<std::mutex> lk(S->mu);
   auto it = S->all.find(id);
   if (it == S->all.end()) return SPX_ERR_ARG;
   if (it->second->state == SPX_Q_QUEUED || it->second->state == SPX_Q_RUNNING) return SPX_ERR_STATE;
   delete it->second;
   S->all.erase(it);
   return SPX_OK;
}

int spx_server_stats_get(spx_server* S, spx_server_stats* st)
{
   if (!S || !st) return SPX_ERR_ARG;
   std::lock_guard<std::mutex> lk(S->mu);
   std::memset(st, 0, sizeof(*st));
   st->submitted = S->submitted;
   st->completed = S->completed;
   for (Pool* p : S->pools) st->launches += (uint64_t)p->launches;
   for (auto& kv : S->all) st->steps_executed += kv.second->ctl.steps_executed;
   st->busy_s = S->busy_s;
   st->wall_s = S->now();
   st->pools = (int32_t)S->pools.size();
   st->running = (int32_t)S->running.size();
   st->queued = (int32_t)S->pending.size();
   return SPX_OK;
}

const char* spx_server_last_error(spx_server* S) { return S ? S->err.c_str() : "null server"; }

void spx_server_destroy(spx_server* S)
{
   if (!S) return;
   {
      std::lock_guard<std::mutex> lk(S->mu);
      S->stop = true;
   }
   S->cv_work.notify_all();
   if (S->worker.joinable()) S->worker.join();
   for (auto& kv : S->all) { kv.second->res.release(); delete kv.second; }
   for (Pool* p : S->pools) { p->release(); delete p; }
   delete S;
}

}  // extern "C"
