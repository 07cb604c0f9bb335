/*
 * spx_oracle.c -- FP64 CPU restatement of SPECTRUM's trajectory-integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only the test suite, the smoke check and
 * the CPU-baseline leg of the benchmark load this library; the HIP product path never
 * links, loads or calls it.
 *
 * Every routine restates the reference's algorithm and operation order; the file:line
 * citations point into the reference tree (vflorins/SPECTRUM).  Build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Scope: transports Parker / ParkerSource (src/trajectory_parker*.cc) and
 * GuidingDiffScatt (src/trajectory_guiding*.cc), Euler stepping (common/rk_config.hh:72-80);
 * backgrounds Uniform, SolarWind, Shock, SmoothShock; the diffusion models of
 * src/diffusion_other.cc used by the path; time/space/momentum boundaries; fixed and
 * shell initial conditions; the distributions of src/distribution_other.cc used by the
 * path; SourceMomentumInjection.
 */
#include "spx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* 2 pi (M_2PI of the reference) */
#define M_2PI_ORC (2.0 * M_PI)

/* the four guiding-centre transports */
#define ORC_IS_GC(t) ((t) >= ORC_TRAJ_GUIDING_DIFF_SCATT && (t) <= ORC_TRAJ_GUIDING_DIFF)

#ifdef ORC_PERTURB_LIBM
/* The last-bit-perturbed build of the power functions (a GPU's libm may round them one unit differently):
   test infrastructure for measuring how strongly a trajectory amplifies such differences. */
static double orc_pow_1ulp(double a, double b) { return pow(a, b) * (1.0 + 2.220446049250313e-16); }
static double orc_cbrt_1ulp(double a) { return cbrt(a) * (1.0 + 2.220446049250313e-16); }
#define pow orc_pow_1ulp
#define cbrt orc_cbrt_1ulp
#endif

/* ------------------------------------------------------------------------------------------------
 * Constants: common/physical_constants.hh:20-81 (in-tree values, USE_GSL off),
 * common/specie.hh:30-85 (code units), common/definitions.hh:54-66 (sp_*).
 * ---------------------------------------------------------------------------------------------- */
#define C_CGS 2.99792458e+10
#define MP_CGS 1.67262192596E-24
#define E_CGS 4.8032044E-10
#define EV_CGS 1.602176634E-12
#define AU_CGS 1.495978707E+13

static const double unit_length = 1.0E10;
static const double unit_velocity = 0.1 * C_CGS;
static const double unit_mass = MP_CGS;
static const double unit_charge = E_CGS;
#define unit_momentum (unit_mass * unit_velocity)
#define unit_energy (unit_mass * unit_velocity * unit_velocity)
#define unit_time (unit_length / unit_velocity)
#define unit_magnetic (unit_momentum / (unit_charge * unit_time))
#define unit_rigidity (unit_energy / unit_charge)
#define unit_diffusion (unit_velocity * unit_length)
#define c_code (C_CGS / unit_velocity)
/* proton_core: mass and charge in code units (common/specie.hh:150-200) */
#define sp_mass (MP_CGS / unit_mass)
#define sp_charge (E_CGS / unit_charge)

#define sp_large 1.0E+20
#define sp_little 1.0E-5
#define sp_small 1.0E-8
#define sp_tiny 1.0E-15

/* common/spatial_data.hh:22-48 */
#define BG_U 0x0001
#define BG_B 0x0002
#define BG_E 0x0004
#define BG_ALL 0x0007
#define BG_gradU 0x0010
#define BG_gradB 0x0020
#define BG_gradE 0x0040
#define BG_gradALL 0x0070
#define BG_dUdt 0x0100
#define BG_dBdt 0x0200
#define BG_dEdt 0x0400
#define BG_dALLdt 0x0700

/* src/trajectory_base.hh:45-57 */
#define TRAJ_FINISH 0x0010
#define TRAJ_TIME_CROSSED 0x0020
#define TRAJ_SPATIAL_CROSSED 0x0040
#define TRAJ_MOMENTUM_CROSSED 0x0080
#define TRAJ_DISCARD 0x0100

/* src/boundary_base.hh:22-40 */
#define BND_TIME 0x0010
#define BND_SPACE 0x0020
#define BND_MOMENTUM 0x0040
#define BND_REFLECT 0x0080
#define BND_TERMINAL 0x0100
#define BND_CROSSED 0x0200
#define BND_RECURRENT 0x0400

#define MAX_BND 8
#define MAX_DIST 4
#define MAX_ACT 8

static inline double Sqr(double x) { return x * x; }
static inline double Cube(double x) { return x * x * x; }   /* common/definitions.hh:198-201 */

/* ------------------------------------------------------------------------------------------------
 * GeoVector / GeoMatrix restated (common/vectors.hh, common/simple_array.hh, common/matrix.hh)
 * ---------------------------------------------------------------------------------------------- */
typedef struct { double v[3]; } vec3;
typedef struct { vec3 row[3]; } mat3;

static inline vec3 V(double x, double y, double z) { vec3 r = {{x, y, z}}; return r; }
static inline vec3 vadd(vec3 a, vec3 b) { return V(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static inline vec3 vsub(vec3 a, vec3 b) { return V(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
/* a * v (common/arithmetic.hh:140) */
static inline vec3 vscl(double a, vec3 b) { return V(a * b.v[0], a * b.v[1], a * b.v[2]); }
/* v / a (common/arithmetic.hh:188) */
static inline vec3 vdiv(vec3 b, double a) { return V(b.v[0] / a, b.v[1] / a, b.v[2] / a); }
static inline vec3 vneg(vec3 a) { return V(-a.v[0], -a.v[1], -a.v[2]); }
/* ScalarProd, common/simple_array.hh:460-465 */
static inline double vdot(vec3 a, vec3 b) { double s = a.v[0] * b.v[0]; s += a.v[1] * b.v[1]; s += a.v[2] * b.v[2]; return s; }
static inline double vnorm2(vec3 a) { double s = a.v[0] * a.v[0]; s += a.v[1] * a.v[1]; s += a.v[2] * a.v[2]; return s; }
static inline double vnorm(vec3 a) { return sqrt(vnorm2(a)); }
/* common/vectors.hh:210-221 */
static inline vec3 vunit(vec3 a) { double n = vnorm(a); return V(a.v[0] / n, a.v[1] / n, a.v[2] / n); }
/* operator ^=, common/vectors.hh:240-247 */
static inline vec3 vcross(vec3 a, vec3 b)
{
   return V(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
/* common/vectors.cc:58-65 */
static inline vec3 to_basis(vec3 a, const vec3* b)
{
   return V(a.v[0] * b[0].v[0] + a.v[1] * b[0].v[1] + a.v[2] * b[0].v[2],
            a.v[0] * b[1].v[0] + a.v[1] * b[1].v[1] + a.v[2] * b[1].v[2],
            a.v[0] * b[2].v[0] + a.v[1] * b[2].v[1] + a.v[2] * b[2].v[2]);
}
/* common/vectors.cc:86-93 */
static inline vec3 from_basis(vec3 a, const vec3* b)
{
   return V(a.v[0] * b[0].v[0] + a.v[1] * b[1].v[0] + a.v[2] * b[2].v[0],
            a.v[0] * b[0].v[1] + a.v[1] * b[1].v[1] + a.v[2] * b[2].v[1],
            a.v[0] * b[0].v[2] + a.v[1] * b[1].v[2] + a.v[2] * b[2].v[2]);
}
static const vec3 cart_unit[3] = {{{1.0, 0.0, 0.0}}, {{0.0, 1.0, 0.0}}, {{0.0, 0.0, 1.0}}};
/* GetSecondUnitVec, common/vectors.cc:453-468 */
static inline vec3 second_unit_vec(vec3 f)
{
   int k;
   if (fabs(f.v[0]) < fabs(f.v[1])) k = (fabs(f.v[0]) < fabs(f.v[2])) ? 0 : 2;
   else k = (fabs(f.v[1]) < fabs(f.v[2])) ? 1 : 2;
   return vunit(vcross(f, cart_unit[k]));
}
/* GeoMatrix * GeoVector, common/matrix.hh:480-483 */
static inline vec3 mvmul(const mat3* m, vec3 a) { return V(vdot(m->row[0], a), vdot(m->row[1], a), vdot(m->row[2], a)); }
/* GeoMatrix * GeoMatrix, common/matrix.hh:509-519 */
static inline mat3 mmmul(const mat3* l, const mat3* r)
{
   mat3 t, o; int i, j;
   for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) t.row[i].v[j] = r->row[j].v[i];
   for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) o.row[i].v[j] = vdot(l->row[i], t.row[j]);
   return o;
}
static inline mat3 mtranspose(const mat3* m)
{
   mat3 t; int i, j;
   for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) t.row[i].v[j] = m->row[j].v[i];
   return t;
}
/* Trace, common/matrix.hh:186-189 */
static inline double mtrace(const mat3* m) { return m->row[0].v[0] + m->row[1].v[1] + m->row[2].v[2]; }

/* ------------------------------------------------------------------------------------------------
 * Relativistic kinematics, common/physics.hh:246-456 (proton_core)
 * ---------------------------------------------------------------------------------------------- */
static inline double EnrTot(double p) { return sqrt(Sqr(p) + Sqr(sp_mass * c_code)) * c_code; }
static inline double EnrKin(double p) { return EnrTot(p) - sp_mass * Sqr(c_code); }
static inline double VelP(double p) { return p * Sqr(c_code) / EnrTot(p); }
static inline double MomT(double T) { return sqrt(T * (T + 2.0 * sp_mass * Sqr(c_code))) / c_code; }
static inline double Rigidity(double p) { return p * c_code / fabs(sp_charge); }
static inline double RelFactor(double v) { return 1.0 / sqrt(1.0 - Sqr(v / c_code)); }
static inline double CyclotronFrequency(double v, double B) { return sp_charge * B / (RelFactor(v) * sp_mass * c_code); }
static inline double LarmorRadius(double p, double B) { return p * c_code / (fabs(sp_charge) * B); }
/* Vel<specie, Cartesian>(mom), common/physics.hh:290-297 with Metric<Cartesian> (common/coordinates.hh:314-345) */
static inline vec3 VelCart(vec3 m) { double mm = vnorm(m); return vscl(VelP(mm) / mm, m); }
/* Vel<specie, Cylindrical>: components 0 and 2 rescaled (common/coordinates.hh:327-357) */
static inline vec3 VelCyl(vec3 m)
{
   double mm = sqrt(Sqr(m.v[0]) + Sqr(m.v[2]));
   double a = VelP(mm) / mm;
   return V(m.v[0] * a, m.v[1], m.v[2] * a);
}

double orc_const(const char* name)
{
   if (!strcmp(name, "c_code")) return c_code;
   if (!strcmp(name, "unit_length")) return unit_length;
   if (!strcmp(name, "unit_velocity")) return unit_velocity;
   if (!strcmp(name, "unit_time")) return unit_time;
   if (!strcmp(name, "unit_energy")) return unit_energy;
   if (!strcmp(name, "unit_magnetic")) return unit_magnetic;
   if (!strcmp(name, "unit_rigidity")) return unit_rigidity;
   if (!strcmp(name, "unit_diffusion")) return unit_diffusion;
   if (!strcmp(name, "one_au")) return AU_CGS / unit_length;
   if (!strcmp(name, "GeV")) return 1.0E9 * EV_CGS / unit_energy;
   if (!strcmp(name, "MeV")) return 1.0E6 * EV_CGS / unit_energy;
   if (!strcmp(name, "keV")) return 1.0E3 * EV_CGS / unit_energy;
   if (!strcmp(name, "mass")) return sp_mass;
   if (!strcmp(name, "charge")) return sp_charge;
   return NAN;
}

/* ------------------------------------------------------------------------------------------------
 * Counter-based RNG: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller.
 * Stream key = (seed), counter = (draw, draw_hi | kind, traj_lo, traj_hi).
 * ---------------------------------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
   uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
   uint32_t k0 = key_in[0], k1 = key_in[1];
   int r;
   for (r = 0; r < 10; r++) {
      uint64_t p0 = (uint64_t)0xD2511F53u * c0;
      uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
      uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
      uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
      uint32_t n0 = hi1 ^ c1 ^ k0;
      uint32_t n1 = lo1;
      uint32_t n2 = hi0 ^ c3 ^ k1;
      uint32_t n3 = lo0;
      c0 = n0; c1 = n1; c2 = n2; c3 = n3;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
   }
   out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_block(uint64_t seed, uint64_t traj, uint64_t draw, uint32_t kind, uint32_t out[4])
{
   uint32_t ctr[4] = {(uint32_t)draw, ((uint32_t)(draw >> 32) & 0x0FFFFFFFu) | (kind << 28), (uint32_t)traj, (uint32_t)(traj >> 32)};
   uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
   orc_philox4x32_10(ctr, key, out);
}

/* 53-bit uniform in (0,1]: ((x >> 11) + 1) * 2^-53 */
static inline double u53(uint32_t lo, uint32_t hi)
{
   uint64_t x = ((uint64_t)hi << 32) | lo;
   return (double)((x >> 11) + 1) * 0x1.0p-53;
}

double orc_counter_uniform(uint64_t seed, uint64_t traj, uint64_t draw)
{
   uint32_t o[4];
   philox_block(seed, traj, draw, 1u, o);
   return u53(o[0], o[1]);
}

/* Box-Muller pairs: normal draw k uses Philox block k/2 of the normal stream; the even draw takes the
   cosine branch and the odd draw the sine branch of the same radius and angle 2*pi*u2. */
double orc_counter_normal(uint64_t seed, uint64_t traj, uint64_t draw)
{
   uint32_t o[4];
   philox_block(seed, traj, draw >> 1, 2u, o);
   double u1 = u53(o[0], o[1]);
   double u2 = u53(o[2], o[3]);
   double r = sqrt(-2.0 * log(u1));
   double a = 3.14159265358979323846 * (2.0 * u2);
   return (draw & 1) ? r * sin(a) : r * cos(a);
}

/* ------------------------------------------------------------------------------------------------
 * DataContainer reader (common/data_container.cc:214-238): records are [u8 len][len bytes].
 * ---------------------------------------------------------------------------------------------- */
typedef struct { const uint8_t* p; const uint8_t* end; int err; } dcread;

/* DataContainer::Read (common/data_container.cc:214-222) copies the record's own length; a shorter
   record fills the low bytes of a zeroed value, a longer one is rejected. */
static void rd_raw(dcread* r, void* dst, size_t want)
{
   if (r->err || r->p >= r->end) { r->err = 1; return; }
   size_t len = *r->p;
   if (r->p + 1 + len > r->end || len > want || len == 0) { r->err = 1; return; }
   memset(dst, 0, want);
   memcpy(dst, r->p + 1, len);
   r->p += len + 1;
}
static double rd_d(dcread* r) { double x = 0.0; rd_raw(r, &x, 8); return x; }
static int rd_i(dcread* r) { int32_t x = 0; rd_raw(r, &x, 4); return x; }
static int rd_b(dcread* r) { uint8_t x = 0; rd_raw(r, &x, 1); return x; }
static vec3 rd_v(dcread* r) { vec3 x = {{0, 0, 0}}; rd_raw(r, x.v, 24); return x; }
static void rd_mi(dcread* r, int32_t out[3]) { out[0] = out[1] = out[2] = 0; rd_raw(r, out, 12); }
static int rd_vi(dcread* r, int32_t* out, int maxn)
{
   if (r->err || r->p >= r->end) { r->err = 1; return 0; }
   size_t len = *r->p;
   if (r->p + 1 + len > r->end || (len % 4) || (int)(len / 4) > maxn) { r->err = 1; return 0; }
   memcpy(out, r->p + 1, len);
   r->p += len + 1;
   return (int)(len / 4);
}
static void rd_skip(dcread* r)
{
   if (r->err || r->p >= r->end) { r->err = 1; return; }
   size_t len = *r->p;
   if (r->p + 1 + len > r->end) { r->err = 1; return; }
   r->p += len + 1;
}

/* ------------------------------------------------------------------------------------------------
 * Plugin parameter blocks
 * ---------------------------------------------------------------------------------------------- */
enum { BG_NONE = 0, BG_UNIFORM, BG_SOLARWIND, BG_SHOCK, BG_SMOOTHSHOCK, BG_GRID };

/* BackgroundServerCartesian data, stored as the reference's reader keeps it
   (src/reader_cartesian.cc:60-103): blocks of 4x4x4 zones, block index x fastest, each block's
   zones x fastest with 6 doubles (flow, magnetic field) per zone, fluid units. */
#define GRID_BS 4
#define GRID_NVAR 6
typedef struct {
   int nb[3];
   double dom_min[3], dom_max[3], block_length[3];
   double* data;
} grid_t;
enum { DF_NONE = 0, DF_ISO, DF_PARA, DF_PERP, DF_FULL, DF_FLOWMOM, DF_KINRAD, DF_RIGBMAG, DF_PERPISO, DF_STRAUSS, DF_POTGIETER, DF_QLT,
       DF_WNLT, DF_WNLT_RAMP, DF_SOQLT };
enum { BK_TIME = 0, BK_SPHERE, BK_PLANE, BK_MOMENTUM };
enum { IT_FIXED = 0 };
enum { IS_FIXED = 0, IS_SPHERE = 1, IS_SECTOR = 2 };
enum { IM_FIXED = 0, IM_SHELL, IM_THICKSHELL, IM_SHELLGRID, IM_RING };
enum { DS_SPECTRUM = 0, DS_TIME, DS_MOMENTUM, DS_POSMOM, DS_CUMO1, DS_CUMO2, DS_POSPITCH };
enum { SRC_NONE = 0, SRC_MOMINJ };

typedef struct {
   int kind;
   double t0; vec3 r0, u0, B0; double dmax0;
   /* SolarWind (src/background_solarwind.cc:58-84) */
   vec3 Omega; double r_ref, dmax_fraction; vec3 eprime[3]; double ur0, w0;
   /* Shock (src/background_shock.cc:58-86) */
   vec3 n_shock; double v_shock, compression; vec3 u1, B1, E0, E1;
   /* SmoothShock (src/background_smooth_shock.cc:139-146) */
   double width_shock;
   /* Uniform (src/background_uniform.cc:58-64) */
   vec3 Eu;
   /* ServerCartesian */
   const grid_t* grid;
} bg_t;

typedef struct {
   int kind;
   double D0, Dperp, Dpara;
   double kap0, U0, pow_law_U, p0, pow_law_p, kap_rat;
   double T0, r0, pow_law_T, pow_law_r; int stream_dep_idx; double u_up, w_sh, s_sh, dn_up_rat;
   double lam0, R0, B0, pow_law_R, pow_law_B;
   int LISM_idx; double lam_in, lam_out, kap_rat_in, kap_rat_out;   /* Strauss (lam) / Potgieter (kappa) */
   double A2A, l_max, ps_index, k_min, ps_minus;                     /* QLTConstant */
   double A2T, A2L, ps_plus;                                         /* WNLTConstant */
   double l_max_HP, z_nose, z_sheath, k_min_ref, A2A_ref, A2T_ref, A2L_ref, dl_max, dz;   /* WNLTRampVLISM */
   double hyp_c, hyp_g, hyp_r;                                       /* 2F1(1, 1; c; z): c, Gamma(c) Gamma(2 - c), (c - 1) / (c - 2) */
   double lam_para, lam_perp, kap_rat_red, radial_limit_perp_red, solar_cycle_effect;   /* EmpiricalSOQLTandUNLT */
   int Bmix_idx, solar_cycle_idx;
} df_t;

typedef struct {
   int kind; uint16_t status0; int max_crossings; int n_act; int32_t actions[MAX_ACT];
   double timemark; vec3 origin; double radius; vec3 norm; double momentum;
} bnd_t;

typedef struct {
   int kind;
   int n_bins[3]; vec3 minval, maxval; int log_bins[3], bin_outside[3];
   vec3 unit_val; int keep_records; uint8_t dims; int width;
   vec3 limits0, limits1, bin_size;
   double val_hot, val_cold; int val_time, val_coord, pos_idx, mom_idx;
   double J0, T0, pow_law;
   int nb;
   int64_t* counts; double* sums; int64_t n_events;
} dist_t;

struct orc_sim {
   int traj_type, time_flow;
   bg_t bg; int have_bg;
   df_t df; int have_df;
   int src_kind; double src_p_inj, src_rate;
   int it_kind; double it_t;
   int is_kind; vec3 is_pos; double is_radius, is_ct1, is_ct2, is_phi1, is_phi2;
   int im_kind; double im_p0, im_p1, im_p2; int im_log; vec3 im_mom; double im_T1, im_T2; int im_n; double im_mu0, im_st0;
   int n_t, n_s, n_m; bnd_t bt[MAX_BND], bs[MAX_BND], bm[MAX_BND];
   int n_dist; dist_t dist[MAX_DIST];
   int spectrum_type;          /* DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE, src/distribution_other.hh:325 */
   int64_t max_steps;          /* max_trajectory_steps, src/trajectory_base.hh:38 */
   double dt_scale;            /* "step_scale": factor on the physical step of the Parker and guiding-centre transports (1 = the reference) */
   int max_attempts;
   int64_t discards[8];
   grid_t grid; int have_grid;
};

/* ------------------------------------------------------------------------------------------------
 * Per-lane transient state: the SpatialData members the path reads (common/spatial_data.hh:55-134)
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
   uint16_t mask;
   vec3 U, B, E; double Bmag; vec3 bhat;
   mat3 gradU, gradB; vec3 gradBmag;
   vec3 dUdt, dBdt; double dBmagdt;
   double region[3]; double dmax;
   double dr[3];   /* the spatial derivative steps (2 r_g) of NumericalDerivatives */
} spd_t;

typedef struct {
   uint64_t traj; int attempt;
   uint64_t n_norm, n_unif;       /* draws consumed by this trajectory */
   const orc_rng* rng; int exhausted;
   double smooth_ds;              /* hidden "ds_shock" of BackgroundSmoothShock */
   double src_del_old;            /* hidden "del_mom_old" of SourceMomentumInjection */
} lane_t;

static double draw_normal(lane_t* L)
{
   uint64_t k = L->n_norm++;
   if (L->rng->mode == ORC_RNG_REPLAY) {
      if ((int64_t)k >= L->rng->normal_stride) { L->exhausted = 1; return 0.0; }
      return L->rng->normals[L->traj * (uint64_t)L->rng->normal_stride + k];
   }
   return orc_counter_normal(L->rng->seed, L->traj, k);
}
static double draw_uniform(lane_t* L)
{
   uint64_t k = L->n_unif++;
   if (L->rng->mode == ORC_RNG_REPLAY) {
      if ((int64_t)k >= L->rng->uniform_stride) { L->exhausted = 1; return 0.5; }
      return L->rng->uniforms[L->traj * (uint64_t)L->rng->uniform_stride + k];
   }
   return orc_counter_uniform(L->rng->seed, L->traj, k);
}

/* ------------------------------------------------------------------------------------------------
 * Backgrounds: EvaluateBackground for each class, writes U/B/E by mask (3 lowest bits)
 * ---------------------------------------------------------------------------------------------- */
/* src/background_smooth_shock.cc:46-75, SMOOTH_SHOCK_ORDER 4 (src/background_smooth_shock.hh:17) */
static inline double shock_transition(double x) { return 0.5 * (1.0 + tanh(8.0 * x)); }
static inline double shock_transition_deriv(double x) { return 0.5 * 8.0 * (1.0 - Sqr(tanh(8.0 * x))); }

static void bg_evaluate(const bg_t* bg, lane_t* L, double t, vec3 pos, uint16_t mask, spd_t* sp)
{
   switch (bg->kind) {
   case BG_UNIFORM:  /* src/background_uniform.cc:70-77 */
      if (mask & BG_U) sp->U = bg->u0;
      if (mask & BG_B) sp->B = bg->B0;
      if (mask & BG_E) sp->E = bg->Eu;
      sp->region[0] = sp->region[1] = sp->region[2] = 1.0;
      break;
   case BG_SOLARWIND: {  /* src/background_solarwind.cc:111-255 (all switches 0) */
      vec3 pp = to_basis(vsub(pos, bg->r0), bg->eprime);
      double r = vnorm(pp);
      double r_mns = r - bg->r_ref;
      double costheta = pp.v[2] / r;
      double ur = bg->ur0;
      sp->region[0] = (r < 117.0 * AU_CGS / unit_length ? 1.0 : -1.0);
      sp->region[1] = -1.0;
      if (mask & BG_U) sp->U = from_basis(vscl(ur, vunit(pp)), bg->eprime);
      if (mask & BG_B) {
         double sintheta = sqrt(fmax(1.0 - Sqr(costheta), 0.0));
         double s = sqrt(Sqr(pp.v[0]) + Sqr(pp.v[1]));
         double cosphi, sinphi;
         if (s < r * sp_tiny) { cosphi = 0.0; sinphi = 0.0; }
         else { cosphi = pp.v[0] / s; sinphi = pp.v[1] / s; }
         double Br = bg->B0.v[0] * Sqr(bg->r_ref / r);
         double Bp = -Br * sintheta * r_mns * bg->w0 / ur;
         vec3 Bl = V(Br * sintheta * cosphi - Bp * sinphi, Br * sintheta * sinphi + Bp * cosphi, Br * costheta);
         sp->B = from_basis(Bl, bg->eprime);
      }
      if (mask & BG_E) sp->E = vdiv(vneg(vcross(sp->U, sp->B)), c_code);
      break;
   }
   case BG_SHOCK:  /* src/background_shock.cc:91-111 */
      if (vdot(vsub(pos, bg->r0), bg->n_shock) - bg->v_shock * t > 0) {
         if (mask & BG_U) sp->U = bg->u0;
         if (mask & BG_B) sp->B = bg->B0;
         if (mask & BG_E) sp->E = bg->E0;
         sp->region[0] = sp->region[1] = sp->region[2] = 1.0;
      }
      else {
         if (mask & BG_U) sp->U = bg->u1;
         if (mask & BG_B) sp->B = bg->B1;
         if (mask & BG_E) sp->E = bg->E1;
         sp->region[0] = sp->region[1] = sp->region[2] = 2.0;
      }
      break;
   case BG_SMOOTHSHOCK: {  /* src/background_smooth_shock.cc:139-154 */
      double ds = (vdot(vsub(pos, bg->r0), bg->n_shock) - bg->v_shock * t) / bg->width_shock;
      L->smooth_ds = ds;
      double a1 = shock_transition(ds), a2 = 1.0 - a1;
      if ((mask & BG_U) || (mask & BG_E)) sp->U = vadd(vscl(a1, bg->u0), vscl(a2, bg->u1));
      if (mask & BG_B) sp->B = vadd(vscl(a1, bg->B0), vscl(a2, bg->B1));
      if (mask & BG_E) sp->E = vdiv(vneg(vcross(sp->U, sp->B)), c_code);
      double reg = 1.0 * a1 + 2.0 * a2;
      sp->region[0] = sp->region[1] = sp->region[2] = reg;
      break;
   }
   default: break;
   }
}

/* EvaluateDmax per class */
static double bg_dmax(const bg_t* bg, lane_t* L, vec3 pos)
{
   switch (bg->kind) {
   case BG_SOLARWIND: return fmin(bg->dmax_fraction * vnorm(vsub(pos, bg->r0)), bg->dmax0);  /* src/background_solarwind.cc:310-314 */
   case BG_SMOOTHSHOCK: return fmin(bg->dmax_fraction * bg->width_shock * fmax(1.0, fabs(L->smooth_ds)), bg->dmax0);  /* src/background_smooth_shock.cc:192-196 */
   default: return bg->dmax0;  /* src/background_base.cc:378-382 */
   }
}

/* BackgroundBase::DirectionalDerivative for spatial directions (src/background_base.cc:70-149), no failures. */
static void numerical_derivatives(const bg_t* bg, lane_t* L, double t, vec3 pos, double pmag, spd_t* sp, uint16_t full_mask)
{
   uint16_t m1 = (uint16_t)(full_mask >> 4);
   if (m1 & BG_ALL) {
      double r_g = fmin(LarmorRadius(pmag, sp->Bmag), sp->dmax);
      vec3 fa[3];
      fa[2] = sp->bhat;
      fa[0] = second_unit_vec(sp->bhat);
      fa[1] = vcross(fa[2], fa[0]);
      mat3 gU, gB, gE; vec3 gBm;
      int xyz;
      for (xyz = 0; xyz < 3; xyz++) {
         spd_t f, b;
         double dr = r_g;
         vec3 pf = vadd(pos, vscl(dr, fa[xyz]));
         bg_evaluate(bg, L, t, pf, m1, &f);
         dr *= 2.0;
         vec3 pb = vsub(pf, vscl(dr, fa[xyz]));
         bg_evaluate(bg, L, t, pb, m1, &b);
         if (m1 & BG_U) gU.row[xyz] = vdiv(vsub(f.U, b.U), dr);
         if (m1 & BG_B) {
            double bmf = vnorm(f.B), bmb = vnorm(b.B);
            gB.row[xyz] = vdiv(vsub(f.B, b.B), dr);
            gBm.v[xyz] = (bmf - bmb) / dr;
         }
         if (m1 & BG_E) gE.row[xyz] = vdiv(vsub(f.E, b.E), dr);
      }
      sp->dr[0] = sp->dr[1] = sp->dr[2] = 2.0 * r_g;
      mat3 rot = mtranspose((const mat3*)fa);
      if (m1 & BG_U) sp->gradU = mmmul(&rot, &gU);
      if (m1 & BG_B) { sp->gradB = mmmul(&rot, &gB); sp->gradBmag = mvmul(&rot, gBm); }
      (void)gE;
   }
   uint16_t m2 = (uint16_t)(m1 >> 4);
   if (m2 & BG_ALL) {
      /* Time derivatives (src/background_base.cc:151-220).  All analytic backgrounds of this path
         with a time dependence are evaluated here; static ones give exact zeros. */
      double w_g = fmin(CyclotronFrequency(VelP(pmag), sp->Bmag), VelP(pmag) / sp->dmax);
      spd_t f, b;
      double dtt = 1.0 / w_g;
      double tf = t + dtt;
      bg_evaluate(bg, L, tf, pos, m2, &f);
      dtt *= 2.0;
      double tb = tf - dtt;
      bg_evaluate(bg, L, tb, pos, m2, &b);
      if (m2 & BG_U) sp->dUdt = vdiv(vsub(f.U, b.U), dtt);
      if (m2 & BG_B) { sp->dBdt = vdiv(vsub(f.B, b.B), dtt); sp->dBmagdt = (vnorm(f.B) - vnorm(b.B)) / dtt; }
   }
}

/* BackgroundSmoothShock analytic derivatives (src/background_smooth_shock.cc:159-187). */
static void smooth_shock_derivatives(const bg_t* bg, lane_t* L, spd_t* sp, uint16_t mask)
{
   double ds = L->smooth_ds;
   int i, j;
   if ((mask & BG_gradU) || (mask & BG_gradE)) {
      vec3 du = vsub(bg->u0, bg->u1);
      for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) sp->gradU.row[i].v[j] = bg->n_shock.v[i] * du.v[j];
      double f = shock_transition_deriv(ds) / bg->width_shock;
      for (i = 0; i < 3; i++) sp->gradU.row[i] = V(sp->gradU.row[i].v[0] * f, sp->gradU.row[i].v[1] * f, sp->gradU.row[i].v[2] * f);
   }
   if (mask & BG_gradB) {
      vec3 db = vsub(bg->B0, bg->B1);
      for (i = 0; i < 3; i++) for (j = 0; j < 3; j++) sp->gradB.row[i].v[j] = bg->n_shock.v[i] * db.v[j];
      double f = shock_transition_deriv(ds) / bg->width_shock;
      for (i = 0; i < 3; i++) sp->gradB.row[i] = V(sp->gradB.row[i].v[0] * f, sp->gradB.row[i].v[1] * f, sp->gradB.row[i].v[2] * f);
      sp->gradBmag = mvmul(&sp->gradB, sp->bhat);
   }
   if ((mask & BG_dUdt) || (mask & BG_dEdt)) sp->dUdt = vscl(shock_transition_deriv(ds) * bg->v_shock / bg->width_shock, vsub(bg->u1, bg->u0));
   if (mask & BG_dBdt) {
      sp->dBdt = vscl(shock_transition_deriv(ds) * bg->v_shock / bg->width_shock, vsub(bg->B1, bg->B0));
      /* the class sets no d|B|/dt (the guiding centre reads it, src/trajectory_guiding.cc:115 and
         common/spatial_data.hh:340); this is its value, bhat . dB/dt */
      sp->dBmagdt = vdot(sp->dBdt, sp->bhat);
   }
}

/* ServerCartesianFront::GetVariables + GetGradients with SERVER_INTERP_ORDER 1 and no ghost cells
   (src/server_cartesian.cc:570-610, :715-740), on the reference's block storage.  Returns nonzero where
   the reference fails: a position outside the domain (ExServerError, src/server_cartesian.cc:1058) or
   a stencil that needs a block beyond the domain (which the reference does not guard). */
static int grid_fields(const grid_t* G, vec3 pos, spd_t* sp)
{
   /* the block of the position: ReadCartesianGetNode (src/reader_cartesian.cc:255-270), fluid units */
   const double ulen_ratio = AU_CGS / unit_length;                    /* Fluid / Particle length */
   const double uvel_ratio = 1.0E7 / unit_velocity;                   /* Fluid / Particle velocity */
   const double umag_ratio = (1.0E7 * sqrt(MP_CGS * 1.0)) / unit_magnetic;
   const double c_fluid = C_CGS / 1.0E7;
   int a, iz, v, uvw, node[3];
   for (a = 0; a < 3; a++) {
      double pf = pos.v[a] / AU_CGS * unit_length;
      if (pf < G->dom_min[a] || pf > G->dom_max[a]) return 1;
      node[a] = (int)((pf - G->dom_min[a]) / G->block_length[a]);
      if (node[a] >= G->nb[a]) return 1;
   }
   /* its geometry: ReadCartesianGetBlockCorners (src/reader_cartesian.cc:277-290), BlockBase::LoadDimensions
      and ConfigureProperties (src/block_base.cc:110-146) */
   vec3 face_min, face_max, zone_length, cent_min, offset_lo, offset_hi;
   for (a = 0; a < 3; a++) {
      face_min.v[a] = G->dom_min[a] + G->block_length[a] * node[a];
      face_max.v[a] = face_min.v[a] + G->block_length[a];
      face_min.v[a] *= ulen_ratio;
      face_max.v[a] *= ulen_ratio;
      zone_length.v[a] = (face_max.v[a] - face_min.v[a]) / GRID_BS;
      cent_min.v[a] = face_min.v[a] + 0.5 * zone_length.v[a];
   }
   /* BlockBase::GetZoneOffset (src/block_base.cc:164-176) */
   int zone_lo[3];
   for (a = 0; a < 3; a++) {
      offset_lo.v[a] = (pos.v[a] - cent_min.v[a]) / zone_length.v[a];
      zone_lo[a] = (int)offset_lo.v[a];
      if (offset_lo.v[a] < 0.0) zone_lo[a]--;
      offset_lo.v[a] = offset_lo.v[a] - zone_lo[a];
      offset_hi.v[a] = 1.0 - offset_lo.v[a];
   }
   /* InteriorInterpolationStencil (src/server_cartesian.cc:135-183): corner iz sits at zone_lo + (i, j, k)
      with i = iz & 1, j = (iz >> 1) & 1, k = iz >> 2 */
   static const int sh[8][3] = {{0,0,0},{1,0,0},{0,1,0},{1,1,0},{0,0,1},{1,0,1},{0,1,1},{1,1,1}};
   double w[8], der[24];
   const double* d = zone_length.v;
   const double *lo = offset_lo.v, *hi = offset_hi.v;
   w[0] = hi[0] * hi[1] * hi[2]; der[0] = -hi[1] * hi[2] / d[0]; der[1] = -hi[0] * hi[2] / d[1]; der[2] = -hi[0] * hi[1] / d[2];
   w[1] = lo[0] * hi[1] * hi[2]; der[3] = hi[1] * hi[2] / d[0]; der[4] = -lo[0] * hi[2] / d[1]; der[5] = -lo[0] * hi[1] / d[2];
   w[2] = hi[0] * lo[1] * hi[2]; der[6] = -lo[1] * hi[2] / d[0]; der[7] = hi[0] * hi[2] / d[1]; der[8] = -hi[0] * lo[1] / d[2];
   w[3] = lo[0] * lo[1] * hi[2]; der[9] = lo[1] * hi[2] / d[0]; der[10] = lo[0] * hi[2] / d[1]; der[11] = -lo[0] * lo[1] / d[2];
   w[4] = hi[0] * hi[1] * lo[2]; der[12] = -hi[1] * lo[2] / d[0]; der[13] = -hi[0] * lo[2] / d[1]; der[14] = hi[0] * hi[1] / d[2];
   w[5] = lo[0] * hi[1] * lo[2]; der[15] = hi[1] * lo[2] / d[0]; der[16] = -lo[0] * lo[2] / d[1]; der[17] = lo[0] * hi[1] / d[2];
   w[6] = hi[0] * lo[1] * lo[2]; der[18] = -lo[1] * lo[2] / d[0]; der[19] = hi[0] * lo[2] / d[1]; der[20] = hi[0] * lo[1] / d[2];
   w[7] = lo[0] * lo[1] * lo[2]; der[21] = lo[1] * lo[2] / d[0]; der[22] = lo[0] * lo[2] / d[1]; der[23] = lo[0] * lo[1] / d[2];
   /* BuildInterpolationStencil with no ghost cells (src/server_cartesian.cc:215-246): a zone index of -1 or
      4 moves the corner into the previous / next block along that axis, the weights unchanged */
   double vars[GRID_NVAR], grads[GRID_NVAR][3], Bmag = 0.0, gBm[3] = {0.0, 0.0, 0.0};
   for (v = 0; v < GRID_NVAR; v++) { vars[v] = 0.0; grads[v][0] = grads[v][1] = grads[v][2] = 0.0; }
   for (iz = 0; iz < 8; iz++) {
      int zc[3], bc[3];
      for (a = 0; a < 3; a++) {
         zc[a] = zone_lo[a] + sh[iz][a];
         bc[a] = node[a];
         if (zc[a] == -1) { zc[a] = GRID_BS - 1; bc[a]--; }
         else if (zc[a] == GRID_BS) { zc[a] = 0; bc[a]++; }
         if (bc[a] < 0 || bc[a] >= G->nb[a]) return 1;
      }
      const int64_t blk = bc[0] + (int64_t)G->nb[0] * (bc[1] + (int64_t)G->nb[1] * bc[2]);
      /* BlockBase::DataLoc (src/block_base.cc:192-196) */
      const double* val = G->data + blk * (GRID_NVAR * GRID_BS * GRID_BS * GRID_BS) + GRID_NVAR * (zc[0] + GRID_BS * (zc[1] + GRID_BS * zc[2]));
      for (v = 0; v < GRID_NVAR; v++) vars[v] += w[iz] * val[v];
      const double bm = sqrt(Sqr(val[3]) + Sqr(val[4]) + Sqr(val[5]));
      Bmag += w[iz] * bm;
      for (v = 0; v < GRID_NVAR; v++)
         for (uvw = 0; uvw < 3; uvw++) grads[v][uvw] += der[3 * iz + uvw] * val[v];
      for (uvw = 0; uvw < 3; uvw++) gBm[uvw] += der[3 * iz + uvw] * bm;
   }
   /* SERVER_VAR_INDEX_FLO 0, SERVER_VAR_INDEX_MAG 3 (src/server_base.hh:32-35); E = -(U x B) / c in fluid units */
   vec3 U = V(vars[0], vars[1], vars[2]), B = V(vars[3], vars[4], vars[5]);
   vec3 E = vdiv(vneg(vcross(U, B)), c_fluid);
   sp->U = vscl(uvel_ratio, U);
   sp->B = vscl(umag_ratio, B);
   sp->E = vscl(umag_ratio, E);
   sp->Bmag = Bmag * umag_ratio;
   for (uvw = 0; uvw < 3; uvw++) {
      for (a = 0; a < 3; a++) {
         sp->gradU.row[uvw].v[a] = grads[a][uvw] * uvel_ratio;
         sp->gradB.row[uvw].v[a] = grads[3 + a][uvw] * umag_ratio;
      }
      sp->gradBmag.v[uvw] = gBm[uvw] * umag_ratio;
   }
   sp->dmax = fmin(sp->dmax, fmin(fmin(zone_length.v[0], zone_length.v[1]), zone_length.v[2]));
   return 0;
}

/* BackgroundBase::GetFields (src/background_base.cc:425-456).  Returns nonzero on ExFieldError. */
static int get_fields(const bg_t* bg, lane_t* L, double t, vec3 pos, double pmag, uint16_t mask, spd_t* sp)
{
   sp->mask = mask;
   sp->dmax = bg_dmax(bg, L, pos);
   if (bg->kind == BG_GRID) {
      /* BackgroundServer (src/background_server.cc:80-110): |B| is interpolated, not recomputed; the grid is
         static, so its time derivatives are zero */
      if (!bg->grid || grid_fields(bg->grid, pos, sp)) return 1;
      if (sp->Bmag < sp_tiny) return 1;
      sp->bhat = vdiv(sp->B, sp->Bmag);
      sp->dUdt = V(0, 0, 0); sp->dBdt = V(0, 0, 0); sp->dBmagdt = 0.0;
      return 0;
   }
   bg_evaluate(bg, L, t, pos, mask, sp);
   if (mask & BG_B) {
      sp->Bmag = vnorm(sp->B);
      if (sp->Bmag < sp_tiny) return 1;
      sp->bhat = vdiv(sp->B, sp->Bmag);
   }
   switch (bg->kind) {
   case BG_SOLARWIND: numerical_derivatives(bg, L, t, pos, pmag, sp, mask); break;  /* SOLARWIND_DERIVATIVE_METHOD 1 */
   case BG_SMOOTHSHOCK: smooth_shock_derivatives(bg, L, sp, mask); break;
   default: {  /* Uniform / Shock: zero derivatives */
      mat3 z; memset(&z, 0, sizeof(z));
      if (mask & BG_gradU) sp->gradU = z;
      if (mask & BG_gradB) { sp->gradB = z; }
      if (mask & BG_dUdt) sp->dUdt = V(0, 0, 0);
      if (mask & BG_dBdt) sp->dBdt = V(0, 0, 0);
      break;
   }
   }
   return 0;
}

/* SpatialData helpers (common/spatial_data.hh:231-341) */
static inline double divU(const spd_t* s) { return mtrace(&s->gradU); }
static inline vec3 curlB(const spd_t* s)
{
   return V(s->gradB.row[1].v[2] - s->gradB.row[2].v[1], s->gradB.row[2].v[0] - s->gradB.row[0].v[2], s->gradB.row[0].v[1] - s->gradB.row[1].v[0]);
}
static inline vec3 curlbhat(const spd_t* s) { return vdiv(vsub(curlB(s), vcross(s->gradBmag, s->bhat)), s->Bmag); }
static inline vec3 dbhatdt(const spd_t* s) { return vdiv(vsub(s->dBdt, vscl(s->dBmagdt, s->bhat)), s->Bmag); }

/* ------------------------------------------------------------------------------------------------
 * Diffusion: DiffusionBase::GetComponent (src/diffusion_base.cc:105-121) + models
 * ---------------------------------------------------------------------------------------------- */
typedef struct { double K[3]; } kap_t;

/* Gauss's hypergeometric function 2F1(1, 1; c; z) for 0 <= z < 1, the function GSL's gsl_sf_hyperg_2F1 gives
   DiffusionWNLTConstant (src/diffusion_other.cc:236-241).  Written from the series (Abramowitz & Stegun
   15.1.1: term ratio (n + 1) z / (c + n) for a = b = 1) below z = 1/2, and above it from the linear
   transformation to 1 - z (A&S 15.3.6 with a = b = 1, where 2F1(c-1, c-1; c-1; 1-z) = z^(1-c)):
     2F1(1, 1; c; z) = (c-1)/(c-2) 2F1(1, 1; 3-c; 1-z) + Gamma(c) Gamma(2-c) (1-z)^(c-2) z^(1-c),
   valid for c - 2 not an integer.  g = Gamma(c) Gamma(2 - c), r = (c - 1)/(c - 2). */
static double hyp_series_11(double c, double z)
{
   double sum = 1.0, term = 1.0;
   int n;
   for (n = 0; n < 400; n++) {
      term *= (n + 1.0) * z / (c + n);
      sum += term;
      if (fabs(term) < 1e-17 * fabs(sum)) break;
   }
   return sum;
}

static double hyp2f1_11(double c, double g, double r, double z)
{
   if (z < 0.5) return hyp_series_11(c, z);
   double w = 1.0 - z;
   return r * hyp_series_11(3.0 - c, w) + g * pow(w, c - 2.0) * pow(z, 1.0 - c);
}

double orc_hyp2f1_11(double c, double z)
{
   return hyp2f1_11(c, tgamma(c) * tgamma(2.0 - c), (c - 1.0) / (c - 2.0), z);
}

/* DiffusionWNLTConstant::EvaluateDiffusion (src/diffusion_other.cc:215-245) with DiffusionQLTConstant's
   component 2 (:145-149), and DiffusionWNLTRampVLISM's position-dependent k_min and amplitudes
   (:301-328): components 0 (and 2 when comp == 2) into kp. */
static void wnlt_eval(const df_t* df, kap_t* kp, int comp, vec3 pos, double vmag, double mu, double Bmag)
{
   double k_min = df->k_min, A2A = df->A2A, A2T = df->A2T, A2L = df->A2L;
   if (df->kind == DF_WNLT_RAMP) {
      double r0 = vnorm(pos);
      double z0 = sqrt(0.5 * r0 * (r0 + pos.v[2]));
      if (z0 > df->z_sheath) { k_min = df->k_min_ref; A2A = df->A2A_ref; A2T = df->A2T_ref; A2L = df->A2L_ref; }
      else {
         k_min = M_2PI_ORC / (df->l_max_HP + df->dl_max * (z0 - df->z_nose) / df->dz);
         double k_ratio = pow(df->k_min_ref / k_min, df->ps_minus);
         A2A = df->A2A_ref * k_ratio; A2T = df->A2T_ref * k_ratio; A2L = df->A2L_ref * k_ratio;
      }
   }
   double st2 = 1.0 - Sqr(mu);
   double Omega = CyclotronFrequency(vmag, Bmag);
   double CT = 0.5 * df->ps_minus / df->ps_plus * Sqr(mu / k_min) * A2T;
   double CL = 0.125 * Sqr(vmag * st2 / Omega) * A2L;
   kp->K[0] = vmag * sqrt(CT + CL);
   if (comp == 0) return;
   kp->K[2] = 0.25 * M_PI * df->ps_minus * fabs(Omega) * st2 * pow(vmag * k_min * fabs(mu / Omega), df->ps_minus) * A2A;
   if (A2T > sp_tiny) {
      double xi1 = vmag * k_min * sqrt(st2) / M_SQRT2 / fabs(Omega);
      double F21 = hyp2f1_11(df->hyp_c, df->hyp_g, df->hyp_r, 1.0 / (1.0 + Sqr(xi1)));
      double DT1 = 1.0 / (1.0 + Sqr(xi1)) * F21;
      double xi2 = kp->K[0] * k_min * k_min / fabs(Omega);
      F21 = hyp2f1_11(df->hyp_c, df->hyp_g, df->hyp_r, 1.0 / (1.0 + Sqr(xi2)));
      double DT2 = 1.0 / (1.0 + Sqr(xi2)) * F21;
      kp->K[2] += (df->ps_minus / df->ps_plus) * Omega * st2 * A2T * DT1 * DT2 / sqrt(Sqr(DT1) + Sqr(DT2));
   }
}

static double df_component(const df_t* df, kap_t* kp, int comp, vec3 pos, double p, double mu, const spd_t* sp)
{
   double vmag = VelP(p);
   double st2 = 1.0 - Sqr(mu);
   (void)st2;
   switch (df->kind) {
   case DF_ISO:  /* src/diffusion_other.cc:64-68 */
      if (comp == 0 || comp == 1) break;
      kp->K[2] = df->D0 * (1.0 - Sqr(mu));
      break;
   case DF_PARA:  /* :381-385 */
      if (comp == 0 || comp == 2) break;
      kp->K[1] = df->D0;
      break;
   case DF_PERP:  /* :459-463 */
      if (comp == 1 || comp == 2) break;
      kp->K[0] = df->D0;
      break;
   case DF_FULL:  /* :538-543 */
      if (comp == 2) break;
      kp->K[0] = df->Dperp; kp->K[1] = df->Dpara;
      break;
   case DF_FLOWMOM:  /* :622-627 */
      if (comp == 2) break;
      kp->K[1] = df->kap0 * pow(vnorm(sp->U) / df->U0, df->pow_law_U) * pow(p / df->p0, df->pow_law_p);
      kp->K[0] = df->kap_rat * kp->K[1];
      break;
   case DF_KINRAD: {  /* :713-723 */
      if (comp == 2) break;
      double r = vnorm(pos);
      double kap_mom = df->kap0 * pow(EnrKin(p) / df->T0, df->pow_law_T);
      if (df->stream_dep_idx == 0 || r < df->r0) kp->K[1] = kap_mom * pow(r / df->r0, df->pow_law_r);
      else if (r < df->r0 + df->w_sh) kp->K[1] = kap_mom * Sqr(vnorm(sp->U) / df->u_up);
      else kp->K[1] = kap_mom * df->dn_up_rat;
      kp->K[0] = df->kap_rat * kp->K[1];
      break;
   }
   case DF_RIGBMAG:  /* :804-809 */
      if (comp == 2) break;
      kp->K[1] = (df->lam0 * vmag / 3.0) * pow(Rigidity(p) / df->R0, df->pow_law_R) * pow(sp->Bmag / df->B0, df->pow_law_B);
      kp->K[0] = df->kap_rat * kp->K[1];
      break;
   case DF_PERPISO:  /* composite: PerpConstant (comp 0) + IsotropicConstant (comp 2) */
      if (comp == 0) kp->K[0] = df->Dperp;
      else if (comp == 2) kp->K[2] = df->D0 * (1.0 - Sqr(mu));
      break;
   case DF_STRAUSS: case DF_POTGIETER: {  /* :886-906, :968-988 */
      if (comp == 2) break;
      double LISM_ind = (df->LISM_idx < 0) ? 0.0 : (sp->region[df->LISM_idx] > 0.0 ? 0.0 : 1.0);
      double lam_para = LISM_ind * df->lam_out + (1.0 - LISM_ind) * df->lam_in;
      double B0_eff = LISM_ind * sp->Bmag + (1.0 - LISM_ind) * df->B0;
      double rig = Rigidity(p);
      if (df->kind == DF_STRAUSS)
         kp->K[1] = (lam_para * vmag / 3.0) * (rig < df->R0 ? cbrt(rig / df->R0) : rig / df->R0) * (B0_eff / sp->Bmag);
      else
         kp->K[1] = lam_para * (vmag / c_code) * (rig < df->R0 ? 1.0 : sqrt(Cube(rig / df->R0))) * (B0_eff / sp->Bmag);
      if (comp == 0) {
         double kap_rat = LISM_ind * df->kap_rat_out + (1.0 - LISM_ind) * df->kap_rat_in;
         kp->K[0] = kap_rat * kp->K[1];
      }
      break;
   }
   case DF_SOQLT: {  /* src/diffusion_other.cc:1053-1072: bent power laws in rigidity, one component per call */
      if (comp == 2) break;
      double lam = 0.0, rig_dep = 0.0;
      double rig = Rigidity(p);
      if (comp == 1) {
         rig_dep = cbrt((rig / df->R0) * (1.0 + Sqr(Sqr(rig / df->R0))));
         lam = df->lam_para;
      }
      else {
         rig_dep = cbrt((rig / df->R0) * (1.0 + Sqr(rig / df->R0)));
         double Bmix_ind = (sp->region[df->Bmix_idx] < 0.0 ? 0.0 : 1.0);
         if (vnorm(pos) < df->radial_limit_perp_red) lam = df->lam_perp * (Bmix_ind + (1.0 - Bmix_ind) * df->kap_rat_red);
         else lam = df->lam_perp;
      }
      kp->K[comp] = (lam * vmag / 3.0) * rig_dep * (df->B0 / sp->Bmag);
      kp->K[comp] /= 1.0 + df->solar_cycle_effect * Sqr(cos(0.5 * sp->region[df->solar_cycle_idx]));
      break;
   }
   case DF_WNLT: case DF_WNLT_RAMP:
      if (comp == 1) break;
      wnlt_eval(df, kp, comp, pos, vmag, mu, sp->Bmag);
      break;
   case DF_QLT: {  /* :145-149 */
      if (comp == 0 || comp == 1) break;
      double Omega = CyclotronFrequency(vmag, sp->Bmag);
      kp->K[2] = 0.25 * M_PI * df->ps_minus * fabs(Omega) * st2 * pow(vmag * df->k_min * fabs(mu / Omega), df->ps_minus) * df->A2A;
      break;
   }
   default: break;
   }
   return kp->K[comp];
}

/* GetDirectionalDerivative overrides (analytic per model) for xyz in 0..2 */
/* DiffusionBase::GetDirectionalDerivative (src/diffusion_base.cc:130-170): the component at the position
   moved by half the background's derivative step along the Cartesian axis xyz, B and |B| carried there by
   their gradients, and back by the full step */
static double df_generic_dir_deriv(const df_t* df, kap_t* kp, int comp, int xyz, vec3 pos, double p, double mu, const spd_t* sp)
{
   kap_t saved = *kp;
   spd_t f = *sp;
   double dr = 0.5 * sp->dr[xyz];
   vec3 pf = pos;
   pf.v[xyz] += dr;
   f.Bmag += sp->gradBmag.v[xyz] * dr;
   double Kf = df_component(df, kp, comp, pf, p, mu, &f);
   dr *= 2.0;
   pf.v[xyz] -= dr;
   f.Bmag -= sp->gradBmag.v[xyz] * dr;
   double Kb = df_component(df, kp, comp, pf, p, mu, &f);
   *kp = saved;
   return (Kf - Kb) / dr;
}

static double df_dir_deriv(const df_t* df, kap_t* kp, int comp, int xyz, vec3 pos, double p, double mu, const spd_t* sp)
{
   switch (df->kind) {
   case DF_WNLT: case DF_WNLT_RAMP: case DF_STRAUSS: case DF_POTGIETER: case DF_SOQLT:
      return df_generic_dir_deriv(df, kp, comp, xyz, pos, p, mu, sp);
   case DF_PARA: case DF_PERP: case DF_FULL: case DF_PERPISO: return 0.0;
   case DF_FLOWMOM: return kp->K[comp] * df->pow_law_U * vdot(sp->gradU.row[xyz], sp->U) / Sqr(vnorm(sp->U));  /* :637-642 */
   case DF_KINRAD: {  /* :732-742 */
      double r = vnorm(pos);
      if (df->stream_dep_idx == 0 || r < df->r0) return kp->K[comp] * df->pow_law_r * pos.v[xyz] / Sqr(r);
      else if (r < df->r0 + df->w_sh) return kp->K[comp] * 2.0 * vdot(sp->gradU.row[xyz], sp->U) / Sqr(vnorm(sp->U));
      return 0.0;
   }
   case DF_RIGBMAG: return kp->K[comp] * df->pow_law_B * sp->gradBmag.v[xyz] / sp->Bmag;  /* :818-823 */
   default: return 0.0;  /* DiffusionIsotropicConstant has no spatial dependence in comp 0 (=0) */
   }
}

/* GetMuDerivative: the overrides, and DiffusionBase's forward difference (src/diffusion_base.cc:226-255)
   for the classes without one (QLTConstant); kp holds the component 2 just evaluated at mu */
static double df_mu_deriv(const df_t* df, kap_t* kp, vec3 pos, double p, double mu, const spd_t* sp)
{
   switch (df->kind) {
   case DF_ISO: case DF_PERPISO: return -2.0 * df->D0 * mu;  /* src/diffusion_other.cc:76-79 */
   case DF_QLT: case DF_WNLT: case DF_WNLT_RAMP: {
      kap_t saved = *kp;
      double dmu = sp_small * (mu + sp_small < 1.0 ? 1.0 : -1.0);
      double derivative = (df_component(df, kp, 2, pos, p, mu + dmu, sp) - saved.K[2]) / dmu;
      *kp = saved;
      return derivative;
   }
   default: return 0.0;
   }
}

/* ------------------------------------------------------------------------------------------------
 * Boundaries (src/boundary_base.cc:105-191, src/boundary_{time,space,momentum}.cc)
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
   uint16_t status; int crossings_left; double delta, delta_old, scale; vec3 normal;
} bst_t;

static double bnd_eval(const bnd_t* b, bst_t* st, int traj_type, double t, vec3 pos, vec3 mom)
{
   switch (b->kind) {
   case BK_TIME:
      if (b->status0 & BND_RECURRENT) return t - (b->max_crossings - st->crossings_left + 1) * b->timemark;  /* src/boundary_time.cc:205-208 */
      return t - b->timemark;  /* :70-73 */
   case BK_SPHERE: {  /* src/boundary_space.cc:395-411 */
      vec3 d = vsub(pos, b->origin);
      st->normal = vunit(d);
      return vnorm(d) - b->radius;
   }
   case BK_PLANE:  /* src/boundary_space.cc:71-84 */
      st->normal = b->norm;
      return vdot(vsub(pos, b->origin), b->norm);
   case BK_MOMENTUM:  /* src/boundary_momentum.cc:69-76 */
      if (ORC_IS_GC(traj_type) || traj_type == ORC_TRAJ_LORENTZ) return vnorm(mom) - b->momentum;
      return mom.v[0] - b->momentum;
   }
   return 0.0;
}

static void bnd_reset(const bnd_t* b, bst_t* st, int traj_type, double t, vec3 pos, vec3 mom)
{
   st->status = b->status0;
   st->crossings_left = b->max_crossings;
   if (st->crossings_left == 1) st->status |= BND_TERMINAL; else st->status &= ~BND_TERMINAL;
   st->delta = bnd_eval(b, st, traj_type, t, pos, mom);
   st->delta_old = st->delta;
   if (st->delta_old == 0.0) st->delta_old = -sp_tiny * st->scale;
   st->status &= ~BND_CROSSED;
}

static void bnd_compute(const bnd_t* b, bst_t* st, int traj_type, double t, vec3 pos, vec3 mom)
{
   st->delta = bnd_eval(b, st, traj_type, t, pos, mom);
   if (st->delta * st->delta_old <= 0.0) {
      st->status |= BND_CROSSED;
      if (st->delta * st->delta_old == 0.0) st->delta = -st->delta_old;
   }
   else st->status &= ~BND_CROSSED;
}

static void bnd_record(const bnd_t* b, bst_t* st, int traj_type, double t, vec3 pos, vec3 mom)
{
   if (st->status & BND_CROSSED) {
      st->crossings_left--;
      if (st->status & BND_RECURRENT) st->delta = bnd_eval(b, st, traj_type, t, pos, mom);
      else st->delta_old = st->delta;
      st->status &= ~BND_CROSSED;
   }
   else st->delta_old = st->delta;
   if (st->crossings_left == 1) st->status |= BND_TERMINAL; else st->status &= ~BND_TERMINAL;
}

/* ------------------------------------------------------------------------------------------------
 * Distributions: DistributionBase::ProcessTrajectory + DistributionTemplated::AddEvent
 * (src/distribution_base.cc:174-194, src/distribution_templated.cc:145-180)
 * ---------------------------------------------------------------------------------------------- */
typedef struct { int bin; double w[9]; int valid; } event_t;

static void dist_event(const orc_sim* s, const dist_t* d, int action,
                       double t1, vec3 pos1, vec3 mom1, double t2, vec3 pos2, vec3 mom2,
                       double amp, double logwgt, event_t* ev)
{
   vec3 value = V(0, 0, 0);
   double w[9] = {0};
   int traj_type = s->traj_type;
   int parker = (traj_type < ORC_TRAJ_GUIDING_DIFF_SCATT || traj_type == ORC_TRAJ_FOCUSED);
   ev->valid = 0;
   /* EvaluateValue */
   switch (d->kind) {
   case DS_SPECTRUM: value.v[0] = EnrKin(parker ? mom1.v[0] : vnorm(mom1)); break;  /* src/distribution_other.cc:561-570 */
   case DS_TIME: value.v[0] = (d->val_time == 0 ? t1 : t2); break;  /* :143-148 */
   case DS_MOMENTUM: {  /* :262-290 */
      vec3 m = (d->val_time == 0 ? mom1 : mom2);
      if (d->val_coord == 0 || parker) value = m;
      else if (traj_type == ORC_TRAJ_FIELDLINE) value = V(m.v[2], 0.0, 0.0);
      else { value.v[0] = vnorm(m); value.v[1] = m.v[2] / value.v[0]; value.v[2] = 0.0; }
      break;
   }
   case DS_POSMOM: {  /* :342-353 */
      vec3 p = (d->val_time == 0 ? pos1 : pos2), m = (d->val_time == 0 ? mom1 : mom2);
      value.v[0] = (d->pos_idx == 3 ? vnorm(p) : p.v[d->pos_idx]);
      value.v[1] = m.v[d->mom_idx];
      break;
   }
   case DS_POSPITCH: {  /* build extension: (|x|, p_para/|p|) */
      vec3 p = (d->val_time == 0 ? pos1 : pos2), m = (d->val_time == 0 ? mom1 : mom2);
      value.v[0] = vnorm(p);
      value.v[1] = m.v[2] / vnorm(m);
      break;
   }
   case DS_CUMO1: case DS_CUMO2: value.v[0] = t2; break;  /* :730-733, :777-780 */
   }
   /* EvaluateWeight(action) */
   switch (d->kind) {
   case DS_SPECTRUM:
      if (action == 0) {  /* :575-594 */
         double m2 = parker ? mom2.v[0] : vnorm(mom2);
         double T = EnrKin(m2);
         if (s->spectrum_type == 0) w[0] = d->J0 * VelP(m2) * pow(T / d->T0, d->pow_law) / Sqr(m2);
         else if (s->spectrum_type == 1) w[0] = d->J0 * pow(T / d->T0, d->pow_law) / Sqr(m2);
         else w[0] = d->J0 * pow(T / d->T0, d->pow_law);
      }
      else if (action == 1) w[0] = d->val_cold;
      else return;
      break;
   case DS_TIME: case DS_MOMENTUM: case DS_POSMOM: case DS_POSPITCH:
      if (action == 0) w[0] = d->val_hot; else if (action == 1) w[0] = d->val_cold; else return;  /* :79-93 */
      break;
   case DS_CUMO1: {  /* :738-741 */
      if (action != 0) return;
      vec3 dx = vsub(pos2, pos1);
      w[0] = dx.v[0]; w[1] = dx.v[1]; w[2] = dx.v[2];
      break;
   }
   case DS_CUMO2: {  /* :785-788, Dyadic common/matrix.cc:23-34 */
      if (action != 0) return;
      vec3 dx = vsub(pos2, pos1);
      int i, j;
      for (i = 0; i < 3; i++) for (j = i; j < 3; j++) w[3 * i + j] = dx.v[i] * dx.v[j];
      w[3] = w[1]; w[6] = w[2]; w[7] = w[5];
      break;
   }
   }
   /* AddEvent */
   int bin[3] = {0, 0, 0}, ijk;
   for (ijk = 0; ijk < 3; ijk++) {
      if (!(d->dims & (1 << ijk))) continue;
      double val = (d->log_bins[ijk] ? log10(value.v[ijk]) : value.v[ijk]) - d->limits0.v[ijk];
      if (val < 0.0) {
         if (!d->bin_outside[ijk]) return;
         /* the reference then truncates val/bin_size toward zero (src/distribution_templated.cc:162-167);
            indices below zero are clamped to the first bin here */
      }
      bin[ijk] = (int)(val / d->bin_size.v[ijk]);
      if (bin[ijk] < 0) bin[ijk] = 0;
      if (bin[ijk] >= d->n_bins[ijk]) {
         if (d->bin_outside[ijk]) bin[ijk] = d->n_bins[ijk] - 1;
         else return;
      }
   }
   ev->bin = d->n_bins[2] * (d->n_bins[1] * bin[0] + bin[1]) + bin[2];
   double e = exp(logwgt);
   int k;
   for (k = 0; k < d->width; k++) ev->w[k] = w[k] * e + amp;
   ev->valid = 1;
}

/* ------------------------------------------------------------------------------------------------
 * Trajectory integration
 * ---------------------------------------------------------------------------------------------- */
typedef struct {
   double t; vec3 pos, mom; double amp, wgt;
   double t0; vec3 pos0, mom0;
   spd_t sp;
   double dt, dt_adaptive, dt_physical;
   vec3 U0;   /* full orbit: the flow the background keeps from SetStart's evaluation with U in the mask */
   uint16_t status;
   int64_t segs;
   bst_t bt[MAX_BND], bs[MAX_BND], bm[MAX_BND];
   int n_refl, n_mirr;
   /* per trajectory event record */
   int n_events; int ev_bin[MAX_DIST]; double ev_w[MAX_DIST][9]; int ev_has[MAX_DIST];
   /* histogram sinks of the calling thread: events are binned when they happen */
   int64_t** hc; double** hs; int64_t* hev;
   double ct, cp; vec3 cpos;   /* ORC_COMPENSATED: low-order bits lost by the sums of t, p and x */
} traj_t;

/* x += inc.  The reference adds in plain FP64 (src/trajectory_base.cc:378-389); the ORC_COMPENSATED
   build (liboracle_comp.so) keeps Kahan's compensation term instead, which measures what the plain sums
   lose over a whole trajectory. */
static inline void acc_add(double* x, double* c, double inc)
{
#ifdef ORC_COMPENSATED
   double y = inc - *c, t = *x + y;
   *c = (t - *x) - y;
   *x = t;
#else
   (void)c;
   *x += inc;
#endif
}
/* x -= dec, likewise */
static inline void acc_sub(double* x, double* c, double dec)
{
#ifdef ORC_COMPENSATED
   acc_add(x, c, -dec);
#else
   (void)c;
   *x -= dec;
#endif
}

#define EX(code) do { return (code); } while (0)

/* ConvertMomentum (src/trajectory_guiding.hh:112-115) */
static inline vec3 gc_convert(vec3 m) { return V(sqrt(Sqr(m.v[0]) + Sqr(m.v[2])), m.v[2] / vnorm(m), 0.0); }

static int common_fields(const orc_sim* s, lane_t* L, traj_t* T, uint16_t mask)
{
   /* ConvertMomentum: (|p|, mu) for the guiding centre, (|p_par|, 0, 0) on a field line (src/trajectory_fieldline.hh:86-89) */
   double pm = ORC_IS_GC(s->traj_type) ? gc_convert(T->mom).v[0] : (s->traj_type == ORC_TRAJ_FIELDLINE ? fabs(T->mom.v[2]) : T->mom.v[0]);
   return get_fields(&s->bg, L, T->t, T->pos, pm, mask, &T->sp);
}

static void process_event(const orc_sim* s, traj_t* T, const bnd_t* b)
{
   int d;
   for (d = 0; d < s->n_dist; d++) {
      int action = (d < b->n_act) ? b->actions[d] : -1;
      if (action < 0) continue;
      event_t ev;
      dist_event(s, &s->dist[d], action, T->t0, T->pos0, T->mom0, T->t, T->pos, T->mom, T->amp, T->wgt, &ev);
      T->n_events++;
      if (ev.valid) {
         int k, w = s->dist[d].width;
         T->ev_bin[d] = ev.bin;
         memcpy(T->ev_w[d], ev.w, sizeof(ev.w));
         T->ev_has[d] = 1;
         T->hc[d][ev.bin]++;
         for (k = 0; k < w; k++) T->hs[d][(size_t)ev.bin * w + k] += ev.w[k];
         T->hev[d]++;
      }
      else T->ev_bin[d] = -1;
   }
}

/* TrajectoryBase::HandleBoundaries (src/trajectory_base.cc:413-535) */
static void handle_boundaries(const orc_sim* s, traj_t* T)
{
   int b, active;
   for (b = 0; b < s->n_t; b++) bnd_compute(&s->bt[b], &T->bt[b], s->traj_type, T->t, T->pos, T->mom);
   for (b = 0; b < s->n_s; b++) bnd_compute(&s->bs[b], &T->bs[b], s->traj_type, T->t, T->pos, T->mom);
   for (b = 0; b < s->n_m; b++) bnd_compute(&s->bm[b], &T->bm[b], s->traj_type, T->t, T->pos, T->mom);

   active = -1;
   for (b = 0; b < s->n_m; b++) {
      uint16_t st = T->bm[b].status;
      if (st & BND_CROSSED) {
         active = b;
         if (st & BND_REFLECT) T->n_mirr++;
         else if (st & BND_TERMINAL) T->status |= TRAJ_MOMENTUM_CROSSED | TRAJ_FINISH;
      }
   }
   if (active >= 0) process_event(s, T, &s->bm[active]);

   active = -1;
   for (b = 0; b < s->n_s; b++) {
      uint16_t st = T->bs[b].status;
      if (st & BND_CROSSED) {
         active = b;
         if (st & BND_REFLECT) {
            T->pos = vsub(T->pos, vscl(2.0 * T->bs[b].delta, T->bs[b].normal));
            if (ORC_IS_GC(s->traj_type)) T->mom.v[2] = -T->mom.v[2];  /* src/trajectory_guiding.hh:121-124 */
            else if (s->traj_type == ORC_TRAJ_FIELDLINE || s->traj_type == ORC_TRAJ_LORENTZ) T->mom = vneg(T->mom);  /* TrajectoryBase's, src/trajectory_base.cc:139-142 */
            else if (s->traj_type == ORC_TRAJ_FOCUSED) T->mom.v[1] = -T->mom.v[1];  /* src/trajectory_focused.hh:126-129 */
            T->n_refl++;
            bnd_compute(&s->bs[b], &T->bs[b], s->traj_type, T->t, T->pos, T->mom);
            T->bs[b].crossings_left--;
            T->n_mirr++;
         }
         else if (st & BND_TERMINAL) T->status |= TRAJ_SPATIAL_CROSSED | TRAJ_FINISH;
      }
   }
   if (active >= 0) process_event(s, T, &s->bs[active]);

   active = -1;
   for (b = 0; b < s->n_t; b++) {
      uint16_t st = T->bt[b].status;
      if (st & BND_CROSSED) {
         active = b;
         if (st & BND_TERMINAL) T->status |= TRAJ_TIME_CROSSED | TRAJ_FINISH;
      }
   }
   if (active >= 0) process_event(s, T, &s->bt[active]);

   for (b = 0; b < s->n_t; b++) bnd_record(&s->bt[b], &T->bt[b], s->traj_type, T->t, T->pos, T->mom);
   for (b = 0; b < s->n_s; b++) bnd_record(&s->bs[b], &T->bs[b], s->traj_type, T->t, T->pos, T->mom);
   for (b = 0; b < s->n_m; b++) bnd_record(&s->bm[b], &T->bm[b], s->traj_type, T->t, T->pos, T->mom);
}

/* TrajectoryBase::TimeBoundaryProximityCheck (src/trajectory_base.cc:151-179) */
static void time_boundary_proximity(const orc_sim* s, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   double dn = fwd ? -sp_large * T->sp.dmax / c_code : sp_large * T->sp.dmax / c_code;
   int b;
   for (b = 0; b < s->n_t; b++) {
      double d = T->bt[b].delta;
      if (fwd) { if ((d <= 0.0) && (d > dn)) dn = d; }
      else { if ((d >= 0.0) && (d < dn)) dn = d; }
   }
   if (fwd) { if (T->dt >= -dn) T->dt = fmax(-(1.0 + sp_little) * dn, sp_small * T->sp.dmax / c_code); }
   else { if (T->dt >= dn) T->dt = fmax((1.0 + sp_little) * dn, sp_small * T->sp.dmax / c_code); }
}

/* SourceMomentumInjection::EvaluateSource (src/source_other.cc:113-119) */
static double source_term(const orc_sim* s, lane_t* L, double p, double dt)
{
   if (s->src_kind != SRC_MOMINJ) return 0.0;
   double del = p - s->src_p_inj;
   double st = (del * L->src_del_old < 0.0) ? s->src_rate / dt : 0.0;
   L->src_del_old = del;
   return st;
}

/* Parker step: TrajectoryParker::Advance (src/trajectory_parker.cc:255-306).  Returns exception code. */
static int parker_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   const df_t* df = &s->df;
   spd_t* sp = &T->sp;
   kap_t kp = {{0, 0, 0}};
   int i, j;
   double p = T->mom.v[0];

   /* Slopes -> DriftCoeff + DiffusionCoeff (src/trajectory_parker.cc:189-246, 83-153) */
   vec3 drift = sp->U;
   double delta = fmin(LarmorRadius(p, sp->Bmag), sp->dmax);
   double Kperp = df_component(df, &kp, 0, T->pos, p, 0.0, sp);
   double Kpara = df_component(df, &kp, 1, T->pos, p, 0.0, sp);
   vec3 divK = V(0, 0, 0);
   for (j = 0; j < 3; j++) {
      spd_t f, b;
      vec3 pt = vadd(T->pos, vscl(delta, cart_unit[j]));
      if (get_fields(&s->bg, L, T->t, pt, p, BG_U | BG_B, &f)) EX(ORC_EX_FIELD);
      double Kperp_f = df_component(df, &kp, 0, pt, p, 0.0, &f);
      double Kpara_f = df_component(df, &kp, 1, pt, p, 0.0, &f);
      pt.v[j] -= 2.0 * delta;
      if (get_fields(&s->bg, L, T->t, pt, p, BG_U | BG_B, &b)) EX(ORC_EX_FIELD);
      double Kperp_b = df_component(df, &kp, 0, pt, p, 0.0, &b);
      double Kpara_b = df_component(df, &kp, 1, pt, p, 0.0, &b);
      for (i = 0; i < 3; i++) {
         double Kf = Kperp_f * (i == j ? 1.0 : 0.0) + (Kpara_f - Kperp_f) * f.bhat.v[j] * f.bhat.v[i];
         double Kb = Kperp_b * (i == j ? 1.0 : 0.0) + (Kpara_b - Kperp_b) * b.bhat.v[j] * b.bhat.v[i];
         divK.v[i] += 0.5 * (Kf - Kb) / delta;
      }
   }
   vec3 slope_pos = fwd ? vadd(drift, divK) : vsub(drift, divK);
   double slope_mom = -p * divU(sp) / 3.0;
   double slope_amp = 0.0;
   if (s->traj_type == ORC_TRAJ_PARKER_SOURCE) slope_amp = source_term(s, L, p, T->dt);  /* src/trajectory_parker_source.cc:58-65 */

   /* PhysicalStep (src/trajectory_parker.cc:212-221) */
   T->dt_physical = 0.5 * sp->dmax / (fwd ? vnorm(vadd(drift, divK)) : vnorm(vsub(drift, divK)));
   T->dt_physical = fmin(T->dt_physical, 0.5 * Sqr(sp->dmax) / fmax(Kperp, Kpara));
   T->dt_physical = fmin(T->dt_physical, 0.5 * 3.0 * 0.01 / fabs(divU(sp)));
   T->dt = fmin(s->dt_scale == 1.0 ? T->dt_physical : T->dt_physical * s->dt_scale, T->dt_adaptive);
   time_boundary_proximity(s, T);

   /* EulerDiffSlopes (src/trajectory_parker.cc:159-180) */
   vec3 fa[3];
   fa[2] = sp->bhat;
   fa[0] = second_unit_vec(sp->bhat);
   fa[1] = vcross(fa[2], fa[0]);
   double dWx = sqrt(T->dt) * draw_normal(L);
   double dWy = sqrt(T->dt) * draw_normal(L);
   double dWz = sqrt(T->dt) * draw_normal(L);
   if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
   Kperp = df_component(df, &kp, 0, T->pos, p, 0.0, sp);
   Kpara = df_component(df, &kp, 1, T->pos, p, 0.0, sp);
   vec3 dr = V(sqrt(2.0 * Kperp) * dWx, sqrt(2.0 * Kperp) * dWy, sqrt(2.0 * Kpara) * dWz);
   dr = from_basis(dr, fa);

   /* RKStep, Euler (src/trajectory_base.cc:362-389) */
   double dtv = T->dt * 1.0;
   if (fwd) {
      acc_add(&T->t, &T->ct, T->dt);
      for (i = 0; i < 3; i++) acc_add(&T->pos.v[i], &T->cpos.v[i], dtv * slope_pos.v[i]);
      acc_add(&T->mom.v[0], &T->cp, dtv * slope_mom);
   }
   else {
      acc_sub(&T->t, &T->ct, T->dt);
      for (i = 0; i < 3; i++) acc_sub(&T->pos.v[i], &T->cpos.v[i], dtv * slope_pos.v[i]);
      acc_sub(&T->mom.v[0], &T->cp, dtv * slope_mom);
   }
   T->amp += dtv * slope_amp * exp(T->wgt);
   T->wgt += dtv * 0.0;
   for (i = 0; i < 3; i++) acc_add(&T->pos.v[i], &T->cpos.v[i], dr.v[i]);

   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (common_fields(s, L, T, BG_U | BG_B | BG_gradU | BG_gradB)) EX(ORC_EX_FIELD);
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* Guiding center drift pieces (src/trajectory_guiding.cc:65-122) */
static void gc_drift(const traj_t* T, vec3 vel, vec3* drift, vec3* Estar, vec3* Bstar)
{
   const spd_t* sp = &T->sp;
   double rL = LarmorRadius(T->mom.v[0], sp->Bmag);
   double rR = LarmorRadius(T->mom.v[2], sp->Bmag);
   *Estar = vsub(vsub(sp->E, vdiv(vscl(rR * sp->Bmag, dbhatdt(sp)), c_code)), vdiv(vscl(rL * vel.v[0], sp->gradBmag), 2.0 * c_code));
   *Bstar = vadd(sp->B, vscl(rR * sp->Bmag, curlbhat(sp)));
   *drift = vdiv(vadd(vscl(vel.v[2], *Bstar), vscl(c_code, vcross(*Estar, sp->bhat))), vdot(*Bstar, sp->bhat));
}

/* TrajectoryGuidingDiff::DiffusionCoeff (src/trajectory_guiding_diff.cc:72-98) */
static void gcd_diffcoeff(const orc_sim* s, traj_t* T, kap_t* kp, vec3* fa, double* Dperp, vec3* Vperp)
{
   const spd_t* sp = &T->sp;
   fa[2] = sp->bhat;
   fa[0] = second_unit_vec(sp->bhat);
   fa[1] = vcross(fa[2], fa[0]);
   vec3 mc = gc_convert(T->mom);
   *Dperp = df_component(&s->df, kp, 0, T->pos, mc.v[0], mc.v[1], sp);
   vec3 g;
   int k;
   for (k = 0; k < 3; k++) g.v[k] = df_dir_deriv(&s->df, kp, 0, k, T->pos, mc.v[0], mc.v[1], sp);
   vec3 vp = V(vdot(g, fa[0]), vdot(g, fa[1]), 0.0);
   *Vperp = from_basis(vp, fa);
}

/* TrajectoryGuidingDiffScatt::Advance (src/trajectory_guiding_diff_scatt.cc:76-174), forward flow */
static int gcds_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   spd_t* sp = &T->sp;
   kap_t kp = {{0, 0, 0}};
   vec3 fa[3], Vperp, drift, Es, Bs;
   double Dperp, Dmumu, Vmu;
   vec3 vel = VelCart(T->mom);  /* Load() */

   /* DiffusionCoeff (src/trajectory_guiding_diff_scatt.cc:53-62) */
   gcd_diffcoeff(s, T, &kp, fa, &Dperp, &Vperp);
   vec3 mc = gc_convert(T->mom);
   Dmumu = df_component(&s->df, &kp, 2, T->pos, mc.v[0], mc.v[1], sp);
   Vmu = df_mu_deriv(&s->df, &kp, T->pos, mc.v[0], mc.v[1], sp);
   /* DriftCoeff */
   gc_drift(T, vel, &drift, &Es, &Bs);
   /* PhysicalStep: GuidingDiff (src/trajectory_guiding_diff.cc:324-332) then GuidingScatt (src/trajectory_guiding_scatt.cc:255-262) */
   T->dt_physical = 0.5 * sp->dmax / ((fwd ? vnorm(vadd(drift, Vperp)) : vnorm(vsub(drift, Vperp))) + 0.5 * vnorm(vel));
   T->dt_physical = fmin(T->dt_physical, 0.5 * Sqr(sp->dmax) / Dperp);
   {
      const double dthetamax = 2.0 * M_PI / 180.0;
      double amu = fabs(T->mom.v[2] / vnorm(T->mom));
      double dmumax = sqrt(1 - amu) * dthetamax + 0.5 * amu * Sqr(dthetamax);
      T->dt_physical = fmin(T->dt_physical, 0.5 * Sqr(dmumax) / Dmumu);
      T->dt_physical = fmin(T->dt_physical, 0.5 * dmumax / fabs(Vmu));
   }
   T->dt = fmin(s->dt_scale == 1.0 ? T->dt_physical : T->dt_physical * s->dt_scale, T->dt_adaptive);
   time_boundary_proximity(s, T);

   /* EulerPitchAngleScatt (src/trajectory_guiding_scatt.cc:81-106) */
   {
      double mu_new;
      vec3 m = gc_convert(T->mom);
      int guard = 0;
      do {
         mu_new = m.v[1] + Vmu * T->dt + sqrt(2.0 * Dmumu * T->dt) * draw_normal(L);
         if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
         if (++guard > 100000) EX(ORC_EX_STEP_NAN);
      } while (fabs(mu_new) > 1.0);
      T->mom.v[0] = m.v[0] * sqrt(1 - Sqr(mu_new));
      T->mom.v[2] = m.v[0] * mu_new;
      vel = VelCyl(T->mom);
   }
   /* Slopes: TrajectoryGuiding::Slopes + GuidingDiff::DiffusionCoeff (src/trajectory_guiding_diff.cc:290-314) */
   gc_drift(T, vel, &drift, &Es, &Bs);
   vec3 slope_pos = drift;
   vec3 slope_mom;
   slope_mom.v[0] = 0.5 * T->mom.v[0] / sp->Bmag * (sp->dBmagdt + vdot(drift, sp->gradBmag));
   slope_mom.v[1] = 0.0;
   slope_mom.v[2] = sp_charge * vdot(Es, Bs) / vdot(Bs, sp->bhat);
   gcd_diffcoeff(s, T, &kp, fa, &Dperp, &Vperp);
   slope_pos = fwd ? vadd(slope_pos, Vperp) : vsub(slope_pos, Vperp);
   /* EulerPerpDiffSlopes (src/trajectory_guiding_diff.cc:123-140) */
   vec3 mconv = gc_convert(T->mom);
   double dWx = sqrt(T->dt) * draw_normal(L);
   double dWy = sqrt(T->dt) * draw_normal(L);
   if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
   Dperp = df_component(&s->df, &kp, 0, T->pos, mconv.v[0], mconv.v[1], sp);
   vec3 dr = V(sqrt(2.0 * Dperp) * dWx, sqrt(2.0 * Dperp) * dWy, 0.0);
   dr = from_basis(dr, fa);
   /* RKStep (src/trajectory_base.cc:362-389) */
   double dtv = T->dt * 1.0;
   if (fwd) {
      T->t += T->dt;
      T->pos = vadd(T->pos, vscl(dtv, slope_pos));
      T->mom = vadd(T->mom, vscl(dtv, slope_mom));
   }
   else {
      T->t -= T->dt;
      T->pos = vsub(T->pos, vscl(dtv, slope_pos));
      T->mom = vsub(T->mom, vscl(dtv, slope_mom));
   }
   T->amp += dtv * 0.0 * exp(T->wgt);
   T->wgt += dtv * 0.0;
   T->pos = vadd(T->pos, dr);

   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (common_fields(s, L, T, BG_ALL | BG_gradB | BG_dBdt)) EX(ORC_EX_FIELD);
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* TrajectoryGuiding::Slopes (src/trajectory_guiding.cc:109-124, PPERP_METHOD 1) with the drift of the
   present momentum */
static void gcg_slopes(traj_t* T, vec3 vel, vec3* drift, vec3* slope_mom)
{
   const spd_t* sp = &T->sp;
   vec3 Es, Bs;
   gc_drift(T, vel, drift, &Es, &Bs);
   slope_mom->v[0] = 0.5 * T->mom.v[0] / sp->Bmag * (sp->dBmagdt + vdot(*drift, sp->gradBmag));
   slope_mom->v[1] = 0.0;
   slope_mom->v[2] = sp_charge * vdot(Es, Bs) / vdot(Bs, sp->bhat);
}

/* RKStep with the Euler table (src/trajectory_base.cc:362-389), forward flow, then HandleBoundaries and
   the end-of-step fields (src/trajectory_base.cc:566-573) */
static int gc_finish_step(const orc_sim* s, lane_t* L, traj_t* T, vec3 slope_pos, vec3 slope_mom)
{
   double dtv = T->dt * 1.0;
   T->t += T->dt;
   T->pos = vadd(T->pos, vscl(dtv, slope_pos));
   T->mom = vadd(T->mom, vscl(dtv, slope_mom));
   T->amp += dtv * 0.0 * exp(T->wgt);
   T->wgt += dtv * 0.0;
   return 0;
}

static int gc_end_step(const orc_sim* s, lane_t* L, traj_t* T)
{
   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (common_fields(s, L, T, BG_ALL | BG_gradB | BG_dBdt)) EX(ORC_EX_FIELD);
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* TrajectoryGuiding::Advance = RKAdvance (src/trajectory_guiding.cc:131-134, src/trajectory_base.cc:545-579):
   Slopes, PhysicalStep (src/trajectory_guiding.cc:96-100), the time step, RKStep; no random numbers */
static int gcg_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   spd_t* sp = &T->sp;
   vec3 vel = VelCart(T->mom);
   vec3 drift, slope_mom;
   gcg_slopes(T, vel, &drift, &slope_mom);
   T->dt_physical = 0.5 * sp->dmax / (vnorm(drift) + 0.5 * vnorm(vel));
   T->dt = fmin(s->dt_scale == 1.0 ? T->dt_physical : T->dt_physical * s->dt_scale, T->dt_adaptive);
   time_boundary_proximity(s, T);
   gc_finish_step(s, L, T, drift, slope_mom);
   return gc_end_step(s, L, T);
}

/* TrajectoryGuidingScatt::Advance (src/trajectory_guiding_scatt.cc:272-340): DiffusionCoeff (D_mumu and
   its mu derivative), DriftCoeff, TrajectoryGuiding::PhysicalStep and the pitch-angle limits (:255-262),
   EulerPitchAngleScatt (:81-106), Slopes with the scattered momentum, RKStep */
static int gcs_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   spd_t* sp = &T->sp;
   kap_t kp = {{0, 0, 0}};
   vec3 vel = VelCart(T->mom);
   vec3 mc = gc_convert(T->mom);
   double Dmumu = df_component(&s->df, &kp, 2, T->pos, mc.v[0], mc.v[1], sp);
   double Vmu = df_mu_deriv(&s->df, &kp, T->pos, mc.v[0], mc.v[1], sp);
   vec3 drift, Es, Bs;
   gc_drift(T, vel, &drift, &Es, &Bs);
   T->dt_physical = 0.5 * sp->dmax / (vnorm(drift) + 0.5 * vnorm(vel));
   {
      const double dthetamax = 2.0 * M_PI / 180.0;
      double amu = fabs(T->mom.v[2] / vnorm(T->mom));
      double dmumax = sqrt(1 - amu) * dthetamax + 0.5 * amu * Sqr(dthetamax);
      T->dt_physical = fmin(T->dt_physical, 0.5 * Sqr(dmumax) / Dmumu);
      T->dt_physical = fmin(T->dt_physical, 0.5 * dmumax / fabs(Vmu));
   }
   T->dt = fmin(s->dt_scale == 1.0 ? T->dt_physical : T->dt_physical * s->dt_scale, T->dt_adaptive);
   time_boundary_proximity(s, T);
   {
      double mu_new;
      vec3 m = gc_convert(T->mom);
      int guard = 0;
      do {
         mu_new = m.v[1] + Vmu * T->dt + sqrt(2.0 * Dmumu * T->dt) * draw_normal(L);
         if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
         if (++guard > 100000) EX(ORC_EX_STEP_NAN);
      } while (fabs(mu_new) > 1.0);
      T->mom.v[0] = m.v[0] * sqrt(1 - Sqr(mu_new));
      T->mom.v[2] = m.v[0] * mu_new;
      vel = VelCyl(T->mom);
   }
   vec3 slope_mom;
   gcg_slopes(T, vel, &drift, &slope_mom);
   gc_finish_step(s, L, T, drift, slope_mom);
   return gc_end_step(s, L, T);
}

/* TrajectoryGuidingDiff::Advance (src/trajectory_guiding_diff.cc:324-370), forward flow: Slopes with
   V_perp (:290-299), PhysicalStep (:306-314), EulerPerpDiffSlopes (:123-140), RKStep, the perpendicular
   displacement */
static int gcd_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   spd_t* sp = &T->sp;
   kap_t kp = {{0, 0, 0}};
   vec3 fa[3], Vperp, drift, slope_mom;
   double Dperp;
   vec3 vel = VelCart(T->mom);
   gcg_slopes(T, vel, &drift, &slope_mom);
   gcd_diffcoeff(s, T, &kp, fa, &Dperp, &Vperp);
   vec3 slope_pos = vadd(drift, Vperp);
   T->dt_physical = 0.5 * sp->dmax / (vnorm(vadd(drift, Vperp)) + 0.5 * vnorm(vel));
   T->dt_physical = fmin(T->dt_physical, 0.5 * Sqr(sp->dmax) / Dperp);
   T->dt = fmin(s->dt_scale == 1.0 ? T->dt_physical : T->dt_physical * s->dt_scale, T->dt_adaptive);
   time_boundary_proximity(s, T);
   vec3 mconv = gc_convert(T->mom);
   double dWx = sqrt(T->dt) * draw_normal(L);
   double dWy = sqrt(T->dt) * draw_normal(L);
   if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
   Dperp = df_component(&s->df, &kp, 0, T->pos, mconv.v[0], mconv.v[1], sp);
   vec3 dr = from_basis(V(sqrt(2.0 * Dperp) * dWx, sqrt(2.0 * Dperp) * dWy, 0.0), fa);
   gc_finish_step(s, L, T, slope_pos, slope_mom);
   T->pos = vadd(T->pos, dr);
   return gc_end_step(s, L, T);
}

/* TrajectoryFieldline::Advance = RKAdvance (src/trajectory_fieldline.cc:41-77, src/trajectory_base.cc:545-579):
   Slopes follow the magnetic field at the velocity's third component, PhysicalStep 0.5 dmax / |v_z|,
   the momentum is constant */
static int fl_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   spd_t* sp = &T->sp;
   vec3 vel = VelCart(T->mom);
   vec3 slope_pos = vscl(vel.v[2], sp->bhat);
   T->dt_physical = 0.5 * sp->dmax / fabs(vel.v[2]);
   T->dt = fmin(T->dt_physical, T->dt_adaptive);
   time_boundary_proximity(s, T);
   double dtv = T->dt * 1.0;
   if (fwd) { T->t += T->dt; T->pos = vadd(T->pos, vscl(dtv, slope_pos)); }
   else { T->t -= T->dt; T->pos = vsub(T->pos, vscl(dtv, slope_pos)); }
   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (common_fields(s, L, T, BG_B)) EX(ORC_EX_FIELD);
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* The full orbit's fields after SetStart: mask E | B (src/trajectory_lorentz.cc:36-37); the solar wind's E is
   formed from the flow the background still holds, that of SetStart's evaluation (src/background_solarwind.cc:252,
   src/background_base.cc:425-456) */
static int lz_fields(const orc_sim* s, lane_t* L, traj_t* T)
{
   /* BackgroundShock sets E0 / E1 by side with the mask E (src/background_shock.cc:91-111); the Cartesian
      server interpolates U and B and forms E from them (grid_fields) */
   if (common_fields(s, L, T, s->bg.kind == BG_SHOCK ? (BG_B | BG_E) : BG_B)) return 1;
   if (s->bg.kind == BG_UNIFORM) T->sp.E = s->bg.Eu;
   else if (s->bg.kind != BG_SHOCK && s->bg.kind != BG_GRID) T->sp.E = vdiv(vneg(vcross(T->U0, T->sp.B)), c_code);
   return 0;
}

/* TrajectoryLorentz::Advance = RKAdvance (src/trajectory_lorentz.cc:51-83, src/trajectory_base.cc:545-579):
   Slopes pos' = v, p' = LorentzForce (common/physics.hh:409-412), PhysicalStep 2 pi / Omega / 100 and
   0.1 dmax / |v|, the Euler step */
static int lz_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   spd_t* sp = &T->sp;
   vec3 vel = VelCart(T->mom);
   double Omega = fmax(CyclotronFrequency(vnorm(vel), sp->Bmag), sp_tiny);
   T->dt_physical = M_2PI_ORC / Omega / 100;
   T->dt_physical = fmin(T->dt_physical, 0.1 * sp->dmax / vnorm(vel));
   T->dt = fmin(T->dt_physical, T->dt_adaptive);
   time_boundary_proximity(s, T);
   vec3 force = vscl(sp_charge, vadd(sp->E, vdiv(vcross(vel, sp->B), c_code)));
   double dtv = T->dt * 1.0;
   if (fwd) { T->t += T->dt; T->pos = vadd(T->pos, vscl(dtv, vel)); T->mom = vadd(T->mom, vscl(dtv, force)); }
   else { T->t -= T->dt; T->pos = vsub(T->pos, vscl(dtv, vel)); T->mom = vsub(T->mom, vscl(dtv, force)); }
   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (lz_fields(s, L, T)) EX(ORC_EX_FIELD);
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* TrajectoryFocused (src/trajectory_focused.cc:31-158, src/trajectory_focused.hh:17-32): momentum (p, mu, 0),
   velocity (p (v(p) / p), mu, 0) (Particle::Vel with the SphericalRMP metric, common/physics.hh:330-336,
   common/coordinates.hh:48,379-382); fields with mask ALL | gradB | gradU | dUdt (:36-37) */
#define FC_MASK (BG_ALL | BG_gradB | BG_gradU | BG_dUdt)

/* V * M with V a row vector: (V * M)_j = sum_i V_i M_ij (common/matrix.hh) */
static inline vec3 vmmul(vec3 a, const mat3* m)
{
   return vadd(vadd(vscl(a.v[0], m->row[0]), vscl(a.v[1], m->row[1])), vscl(a.v[2], m->row[2]));
}

/* DriftCoeff with TRAJ_FOCUSED_USE_B_DRIFTS (:55-79), Slopes with PPERP_METHOD 1 (:98-125), PhysicalStep
   (:87-91), the Euler RKStep, HandleBoundaries, then CommonFields and MomentumCorrection (:142-152) in
   RKAdvance's order (src/trajectory_base.cc:545-579); no random numbers */
static int fc_advance(const orc_sim* s, lane_t* L, traj_t* T)
{
   int fwd = (s->time_flow == ORC_FLOW_FORWARD);
   const spd_t* sp = &T->sp;
   double p = T->mom.v[0], mu = T->mom.v[1];
   double v = p * (VelP(p) / p);
   double ct2 = Sqr(mu), st2 = 1.0 - ct2;
   vec3 drift = vdiv(vscl(0.5 * st2, vcross(sp->bhat, sp->gradBmag)), sp->Bmag);
   vec3 BgB = vmmul(sp->B, &sp->gradB);
   drift = vadd(drift, vdiv(vadd(vscl(vdot(sp->B, curlB(sp)), vscl(0.5 * st2, sp->bhat)), vscl(ct2, vcross(sp->bhat, BgB))), Sqr(sp->Bmag)));
   drift = vscl(LarmorRadius(p, sp->Bmag) * v, drift);
   drift = vadd(drift, vadd(sp->U, vscl(v * mu, sp->bhat)));
   /* bb : grad U (common/matrix.hh:556-561), 2 b . (dU/dt + U . grad U) / v */
   double bbgU = 0.0;
   int i;
   for (i = 0; i < 3; i++) bbgU += vdot(vscl(sp->bhat.v[i], sp->bhat), sp->gradU.row[i]);
   vec3 cdU = vadd(sp->dUdt, vmmul(sp->U, &sp->gradU));
   double bcdU = vdot(vscl(2.0, sp->bhat), cdU) / v;
   double divu = divU(sp);
   double divbhat = (mtrace(&sp->gradB) - vdot(sp->gradBmag, sp->bhat)) / sp->Bmag;   /* common/spatial_data.hh:304-307 */
   vec3 slope_mom = V(0.5 * p * ((3.0 * st2 - 2.0) * bbgU - st2 * divu - mu * bcdU),
                      0.5 * st2 * (v * divbhat - bcdU + mu * (divu - 3.0 * bbgU)), 0.0);
   T->dt_physical = 0.5 * sp->dmax / (vnorm(drift) + 0.5 * v);
   T->dt = fmin(T->dt_physical, T->dt_adaptive);
   time_boundary_proximity(s, T);
   double dtv = T->dt * 1.0;
   if (fwd) { T->t += T->dt; T->pos = vadd(T->pos, vscl(dtv, drift)); T->mom = vadd(T->mom, vscl(dtv, slope_mom)); }
   else { T->t -= T->dt; T->pos = vsub(T->pos, vscl(dtv, drift)); T->mom = vsub(T->mom, vscl(dtv, slope_mom)); }
   handle_boundaries(s, T);
   if (!(T->status & TRAJ_FINISH)) {
      if (common_fields(s, L, T, FC_MASK)) EX(ORC_EX_FIELD);
      if (T->mom.v[1] > 1.0) T->mom.v[1] = 1.0;
      else if (T->mom.v[1] < -1.0) T->mom.v[1] = -1.0;
   }
   T->segs++;
   return ORC_EX_NONE;
}

/* Initial momentum sample (src/initial_momentum.cc:283-410) */
/* InitialSpaceFixed / InitialSpaceSphere / InitialSpaceSphereSector::EvaluateInitial (src/initial_space.cc:49-66,
   :308-316, :384-392): the polar cosine uniform between cos(theta1) and cos(theta2), then the azimuth uniform
   between phi1 and phi2 (the full sphere: -1 + 2 u and 2 pi u) */
static vec3 pos_sample(const orc_sim* s, lane_t* L)
{
   if (s->is_kind == IS_FIXED) return s->is_pos;
   double ct, st, phi;
   if (s->is_kind == IS_SPHERE) ct = -1.0 + 2.0 * draw_uniform(L);
   else ct = s->is_ct1 + (s->is_ct2 - s->is_ct1) * draw_uniform(L);
   st = sqrt(1.0 - Sqr(ct));
   if (s->is_kind == IS_SPHERE) phi = M_2PI_ORC * draw_uniform(L);
   else phi = s->is_phi1 + (s->is_phi2 - s->is_phi1) * draw_uniform(L);
   return vadd(s->is_pos, V(s->is_radius * st * cos(phi), s->is_radius * st * sin(phi), s->is_radius * ct));
}

static vec3 mom_sample(const orc_sim* s, lane_t* L)
{
   int gc = ORC_IS_GC(s->traj_type), fc = (s->traj_type == ORC_TRAJ_FOCUSED);
   double p;
   switch (s->im_kind) {
   case IM_FIXED: return s->im_mom;
   default: break;
   }
   if (s->traj_type == ORC_TRAJ_LORENTZ) {  /* src/initial_momentum.cc:300-325, :381-413 with TRAJ_LORENTZ */
      if (s->im_kind == IM_THICKSHELL) {
         if (s->im_log) p = pow(10.0, s->im_p1 + (s->im_p2 - s->im_p1) * draw_uniform(L));
         else p = s->im_p1 + (s->im_p2 - s->im_p1) * draw_uniform(L);
      }
      else p = s->im_p0;
      double mu = -1.0 + 2.0 * draw_uniform(L);
      double st = sqrt(1.0 - Sqr(mu));
      double phi = M_2PI_ORC * draw_uniform(L);
      return V(p * st * cos(phi), p * st * sin(phi), p * mu);
   }
   switch (s->im_kind) {
   case IM_RING:  /* src/initial_momentum.cc:203-220: fixed (p sin theta0, 0, p cos theta0), no draw */
      if (gc) return V(s->im_p0 * s->im_st0, 0.0, s->im_p0 * s->im_mu0);
      if (fc) return V(s->im_p0, s->im_mu0, 0.0);
      return V(s->im_p0, 0.0, 0.0);
   case IM_SHELL: p = s->im_p0; break;
   case IM_THICKSHELL:
      if (s->im_log) p = pow(10.0, s->im_p1 + (s->im_p2 - s->im_p1) * draw_uniform(L));
      else p = s->im_p1 + (s->im_p2 - s->im_p1) * draw_uniform(L);
      break;
   case IM_SHELLGRID: {
      int k = (int)(L->traj % (uint64_t)s->im_n);
      double lT1 = log10(s->im_T1), lT2 = log10(s->im_T2);
      double Tk = (s->im_n > 1) ? pow(10.0, lT1 + (lT2 - lT1) * k / (double)(s->im_n - 1)) : s->im_T1;
      p = MomT(Tk);
      break;
   }
   default: p = 0.0;
   }
   if (s->traj_type == ORC_TRAJ_FIELDLINE) return V(0.0, 0.0, p);  /* src/initial_momentum.cc:289-292, :390-391 */
   if (!gc && !fc) return V(p, 0.0, 0.0);
   double mu = -1.0 + 2.0 * draw_uniform(L);
   double st = sqrt(1.0 - Sqr(mu));
   if (fc) return V(p, mu, 0.0);  /* src/initial_momentum.cc:311-312, :399-400 */
   return V(p * st, 0.0, p * mu);
}

/* TrajectoryBase::SetStart (src/trajectory_base.cc:867-929) + per-transport SetStart; then
   TrajectoryBase::Integrate (src/trajectory_base.cc:946-995).  Returns exception code. */
static int integrate_one(const orc_sim* s, lane_t* L, traj_t* T)
{
   int b;
   int64_t** hc = T->hc; double** hs = T->hs; int64_t* hev = T->hev;
   memset(T, 0, sizeof(*T));
   T->hc = hc; T->hs = hs; T->hev = hev;
   T->t = s->it_t;
   T->pos = pos_sample(s, L);
   T->mom = mom_sample(s, L);
   /* hidden state of BackgroundSmoothShock::ds_shock: taken at the starting point */
   if (s->bg.kind == BG_SMOOTHSHOCK)
      L->smooth_ds = (vdot(vsub(T->pos, s->bg.r0), s->bg.n_shock) - s->bg.v_shock * T->t) / s->bg.width_shock;
   if (common_fields(s, L, T, BG_ALL | BG_gradALL | BG_dALLdt)) EX(ORC_EX_FIELD);
   T->U0 = T->sp.U;
   T->mom = mom_sample(s, L);
   if (L->exhausted) EX(ORC_EX_REPLAY_EXHAUSTED);
   T->dt_adaptive = sp_large * T->sp.dmax / c_code;
   T->t0 = T->t; T->pos0 = T->pos; T->mom0 = T->mom;
   T->status = 0;
   for (b = 0; b < s->n_t; b++) { T->bt[b].scale = fabs(T->sp.dmax / c_code); bnd_reset(&s->bt[b], &T->bt[b], s->traj_type, T->t, T->pos, T->mom); }
   for (b = 0; b < s->n_s; b++) { T->bs[b].scale = fabs(T->sp.dmax); bnd_reset(&s->bs[b], &T->bs[b], s->traj_type, T->t, T->pos, T->mom); }
   for (b = 0; b < s->n_m; b++) { T->bm[b].scale = fabs(vnorm(T->mom)); bnd_reset(&s->bm[b], &T->bm[b], s->traj_type, T->t, T->pos, T->mom); }
   T->amp = 0.0; T->wgt = 0.0;
   for (b = 0; b < MAX_DIST; b++) T->ev_bin[b] = -1;
   /* TrajectoryParkerSource::SetStart: source->GetSource(..., 1.0) primes del_mom_old */
   if (s->traj_type == ORC_TRAJ_PARKER_SOURCE) { L->src_del_old = 0.0; (void)source_term(s, L, T->mom.v[0], 1.0); }

   while (!(T->status & (TRAJ_FINISH | TRAJ_DISCARD))) {
      int ex;
      switch (s->traj_type) {
      case ORC_TRAJ_GUIDING_DIFF_SCATT: ex = gcds_advance(s, L, T); break;
      case ORC_TRAJ_FIELDLINE: ex = fl_advance(s, L, T); break;
      case ORC_TRAJ_LORENTZ: ex = lz_advance(s, L, T); break;
      case ORC_TRAJ_FOCUSED: ex = fc_advance(s, L, T); break;
      case ORC_TRAJ_GUIDING: ex = gcg_advance(s, L, T); break;
      case ORC_TRAJ_GUIDING_SCATT: ex = gcs_advance(s, L, T); break;
      case ORC_TRAJ_GUIDING_DIFF: ex = gcd_advance(s, L, T); break;
      default: ex = parker_advance(s, L, T);
      }
      if (ex) return ex;
      if (T->segs > s->max_steps) EX(ORC_EX_MAX_STEPS);
      if (T->dt < sp_tiny * T->sp.dmax / c_code) EX(ORC_EX_STEP_TOO_SMALL);
      else if (!isnormal(T->dt)) EX(ORC_EX_STEP_NAN);
   }
   return ORC_EX_NONE;
}

/* ------------------------------------------------------------------------------------------------
 * Public API
 * ---------------------------------------------------------------------------------------------- */
orc_sim* orc_create(int traj_type, int time_flow)
{
   orc_sim* s = (orc_sim*)calloc(1, sizeof(orc_sim));
   s->traj_type = traj_type;
   s->time_flow = time_flow;
   s->spectrum_type = 0;
   s->max_steps = 10000000;
   s->dt_scale = 1.0;
   s->max_attempts = 16;
   return s;
}

int orc_set_grid(orc_sim* s, const int32_t nb[3], const double dom_min[3], const double dom_max[3], int32_t n_var, const double* data)
{
   int a;
   if (!s || n_var != GRID_NVAR) return -1;
   for (a = 0; a < 3; a++) if (nb[a] < 1 || !(dom_max[a] > dom_min[a])) return -1;
   const size_t n = (size_t)GRID_NVAR * GRID_BS * GRID_BS * GRID_BS * (size_t)nb[0] * nb[1] * nb[2];
   free(s->grid.data);
   s->grid.data = (double*)malloc(n * sizeof(double));
   if (!s->grid.data) return -1;
   memcpy(s->grid.data, data, n * sizeof(double));
   for (a = 0; a < 3; a++) {
      s->grid.nb[a] = nb[a];
      s->grid.dom_min[a] = dom_min[a];
      s->grid.dom_max[a] = dom_max[a];
      s->grid.block_length[a] = (dom_max[a] - dom_min[a]) / nb[a];   /* src/reader_cartesian.cc:88 */
   }
   s->have_grid = 1;
   if (s->bg.kind == BG_GRID) s->bg.grid = &s->grid;
   return 0;
}

int orc_grid_fields(orc_sim* s, const double pos[3], double out[32])
{
   spd_t sp; memset(&sp, 0, sizeof(sp));
   int k, j;
   if (!s || s->bg.kind != BG_GRID || !s->bg.grid) return -1;
   sp.dmax = s->bg.dmax0;
   if (grid_fields(s->bg.grid, V(pos[0], pos[1], pos[2]), &sp)) return 1;
   for (k = 0; k < 3; k++) { out[k] = sp.U.v[k]; out[3 + k] = sp.B.v[k]; out[25 + k] = sp.gradBmag.v[k]; }
   out[6] = sp.Bmag;
   for (k = 0; k < 3; k++)
      for (j = 0; j < 3; j++) { out[7 + 3 * k + j] = sp.gradU.row[k].v[j]; out[16 + 3 * k + j] = sp.gradB.row[k].v[j]; }
   out[28] = sp.dmax;
   for (k = 0; k < 3; k++) out[29 + k] = sp.E.v[k];
   return 0;
}

void orc_destroy(orc_sim* s)
{
   int d;
   if (!s) return;
   free(s->grid.data);
   for (d = 0; d < s->n_dist; d++) { free(s->dist[d].counts); free(s->dist[d].sums); }
   free(s);
}

int orc_set_option(orc_sim* s, const char* name, double value)
{
   if (!strcmp(name, "DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE")) { s->spectrum_type = (int)value; return 0; }
   if (!strcmp(name, "max_trajectory_steps")) { s->max_steps = (int64_t)value; return 0; }
   if (!strcmp(name, "step_scale")) { if (!(value > 0.0 && value <= 1.0)) return -1; s->dt_scale = value; return 0; }
   if (!strcmp(name, "max_attempts")) { s->max_attempts = (int)value; return 0; }
   return -1;
}

static int setup_background(orc_sim* s, const char* cls, dcread* r)
{
   bg_t* bg = &s->bg;
   memset(bg, 0, sizeof(*bg));
   /* BackgroundBase::SetupBackground (src/background_base.cc:334-352) */
   bg->t0 = rd_d(r); bg->r0 = rd_v(r); bg->u0 = rd_v(r); bg->B0 = rd_v(r); bg->dmax0 = rd_d(r);
   if (!strcmp(cls, "BackgroundUniform")) {
      bg->kind = BG_UNIFORM;
      bg->Eu = vdiv(vneg(vcross(bg->u0, bg->B0)), c_code);
   }
   else if (!strcmp(cls, "BackgroundSolarWind")) {
      bg->kind = BG_SOLARWIND;
      bg->Omega = rd_v(r); bg->r_ref = rd_d(r); bg->dmax_fraction = rd_d(r);
      bg->w0 = vnorm(bg->Omega);
      if (bg->w0 < sp_tiny) bg->eprime[2] = V(0, 0, 1);
      else bg->eprime[2] = vunit(bg->Omega);
      bg->eprime[0] = second_unit_vec(bg->eprime[2]);
      bg->eprime[1] = vcross(bg->eprime[2], bg->eprime[0]);
      bg->ur0 = fabs(bg->u0.v[0]);
   }
   else if (!strcmp(cls, "BackgroundShock") || !strcmp(cls, "BackgroundSmoothShock")) {
      bg->kind = strcmp(cls, "BackgroundShock") ? BG_SMOOTHSHOCK : BG_SHOCK;
      bg->n_shock = rd_v(r); bg->v_shock = rd_d(r); bg->compression = rd_d(r);
      bg->n_shock = vunit(bg->n_shock);
      vec3 u0n = vscl(vdot(bg->u0, bg->n_shock) - bg->v_shock, bg->n_shock);
      vec3 u0t = vsub(bg->u0, vscl(vdot(bg->u0, bg->n_shock), bg->n_shock));
      vec3 B0n = vscl(vdot(bg->B0, bg->n_shock), bg->n_shock);
      vec3 B0t = vsub(bg->B0, B0n);
      bg->u1 = vadd(vadd(vdiv(u0n, bg->compression), vscl(bg->v_shock, bg->n_shock)), u0t);
      bg->B1 = vadd(B0n, vscl(bg->compression, B0t));
      bg->E0 = vdiv(vneg(vcross(bg->u0, bg->B0)), c_code);
      bg->E1 = vdiv(vneg(vcross(bg->u1, bg->B1)), c_code);
      if (bg->kind == BG_SMOOTHSHOCK) { bg->width_shock = rd_d(r); bg->dmax_fraction = rd_d(r); }
   }
   else if (!strcmp(cls, "BackgroundServerCartesian")) {  /* src/background_server.cc:62-75: base parameters only */
      bg->kind = BG_GRID;
      bg->grid = s->have_grid ? &s->grid : NULL;
   }
   else return -1;
   s->have_bg = 1;
   return 0;
}

static int setup_diffusion(orc_sim* s, const char* cls, dcread* r)
{
   df_t* df = &s->df;
   memset(df, 0, sizeof(*df));
   if (!strcmp(cls, "DiffusionIsotropicConstant")) { df->kind = DF_ISO; df->D0 = rd_d(r); }
   else if (!strcmp(cls, "DiffusionParaConstant")) { df->kind = DF_PARA; df->D0 = rd_d(r); }
   else if (!strcmp(cls, "DiffusionPerpConstant")) { df->kind = DF_PERP; df->D0 = rd_d(r); }
   else if (!strcmp(cls, "DiffusionFullConstant")) { df->kind = DF_FULL; df->Dperp = rd_d(r); df->Dpara = rd_d(r); }
   else if (!strcmp(cls, "DiffusionFlowMomentumPowerLaw")) {
      df->kind = DF_FLOWMOM;
      df->kap0 = rd_d(r); df->U0 = rd_d(r); df->pow_law_U = rd_d(r); df->p0 = rd_d(r); df->pow_law_p = rd_d(r); df->kap_rat = rd_d(r);
   }
   else if (!strcmp(cls, "DiffusionKineticEnergyRadialDistancePowerLaw")) {
      df->kind = DF_KINRAD;
      df->kap0 = rd_d(r); df->T0 = rd_d(r); df->r0 = rd_d(r); df->pow_law_T = rd_d(r); df->pow_law_r = rd_d(r); df->kap_rat = rd_d(r);
      df->stream_dep_idx = rd_i(r); df->u_up = rd_d(r); df->w_sh = rd_d(r); df->s_sh = rd_d(r);
      df->dn_up_rat = Sqr(1.0 / df->s_sh);
   }
   else if (!strcmp(cls, "DiffusionRigidityMagneticFieldPowerLaw")) {
      df->kind = DF_RIGBMAG;
      df->lam0 = rd_d(r); df->R0 = rd_d(r); df->B0 = rd_d(r); df->pow_law_R = rd_d(r); df->pow_law_B = rd_d(r); df->kap_rat = rd_d(r);
   }
   else if (!strcmp(cls, "DiffusionPerpConstantPitchIsotropic")) { df->kind = DF_PERPISO; df->Dperp = rd_d(r); df->D0 = rd_d(r); }
   else if (!strcmp(cls, "DiffusionStraussEtAl2013") || !strcmp(cls, "DiffusionPotgieterEtAl2015")) {
      /* src/diffusion_other.cc:869-880, :950-961 */
      df->kind = !strcmp(cls, "DiffusionStraussEtAl2013") ? DF_STRAUSS : DF_POTGIETER;
      df->LISM_idx = rd_i(r); df->lam_in = rd_d(r); df->lam_out = rd_d(r); df->R0 = rd_d(r); df->B0 = rd_d(r);
      df->kap_rat_in = rd_d(r); df->kap_rat_out = rd_d(r);
   }
   else if (!strcmp(cls, "DiffusionEmpiricalSOQLTandUNLT")) {  /* src/diffusion_other.cc:1034-1046 */
      df->kind = DF_SOQLT;
      df->lam_para = rd_d(r); df->lam_perp = rd_d(r); df->R0 = rd_d(r); df->B0 = rd_d(r);
      df->Bmix_idx = rd_i(r); df->kap_rat_red = rd_d(r); df->radial_limit_perp_red = rd_d(r);
      df->solar_cycle_idx = rd_i(r); df->solar_cycle_effect = rd_d(r);
   }
   else if (!strcmp(cls, "DiffusionQLTConstant")) {  /* src/diffusion_other.cc:130-140 */
      df->kind = DF_QLT;
      df->A2A = rd_d(r); df->l_max = rd_d(r); df->ps_index = rd_d(r);
      df->k_min = 2.0 * M_PI / df->l_max;
      df->ps_minus = df->ps_index - 1.0;
   }
   else if (!strcmp(cls, "DiffusionWNLTConstant") || !strcmp(cls, "DiffusionWNLTRampVLISM")) {  /* :202-213, :281-296 */
      df->kind = !strcmp(cls, "DiffusionWNLTConstant") ? DF_WNLT : DF_WNLT_RAMP;
      df->A2A = rd_d(r); df->l_max = rd_d(r); df->ps_index = rd_d(r);
      df->k_min = 2.0 * M_PI / df->l_max;
      df->ps_minus = df->ps_index - 1.0;
      df->A2T = rd_d(r); df->A2L = rd_d(r);
      df->ps_plus = df->ps_index + 1.0;
      if (df->kind == DF_WNLT_RAMP) {
         df->l_max_HP = rd_d(r); df->z_nose = rd_d(r); df->z_sheath = rd_d(r);
         df->k_min_ref = df->k_min; df->A2A_ref = df->A2A; df->A2T_ref = df->A2T; df->A2L_ref = df->A2L;
         df->dl_max = df->l_max - df->l_max_HP; df->dz = df->z_sheath - df->z_nose;
      }
      df->hyp_c = (5.0 + df->ps_index) / 4.0;
      df->hyp_g = tgamma(df->hyp_c) * tgamma(2.0 - df->hyp_c);
      df->hyp_r = (df->hyp_c - 1.0) / (df->hyp_c - 2.0);
   }
   else return -1;
   s->have_df = 1;
   return 0;
}

static int setup_boundary(orc_sim* s, const char* cls, dcread* r)
{
   bnd_t b;
   memset(&b, 0, sizeof(b));
   /* BoundaryBase::SetupBoundary (src/boundary_base.cc:73-83) */
   b.max_crossings = rd_i(r);
   b.n_act = rd_vi(r, b.actions, MAX_ACT);
   int cls_t = -1;
   if (!strcmp(cls, "BoundaryTimeExpire")) { b.kind = BK_TIME; b.timemark = rd_d(r); b.max_crossings = 1; b.status0 = BND_TIME; cls_t = 0; }
   else if (!strcmp(cls, "BoundaryTimePass")) { b.kind = BK_TIME; b.timemark = rd_d(r); b.max_crossings = -1; b.status0 = BND_TIME; cls_t = 0; }
   else if (!strcmp(cls, "BoundaryTimeRecurrent")) { b.kind = BK_TIME; b.timemark = rd_d(r); b.status0 = BND_TIME | BND_RECURRENT; cls_t = 0; }
   else if (!strcmp(cls, "BoundarySphereAbsorb")) { b.kind = BK_SPHERE; b.origin = rd_v(r); b.radius = rd_d(r); b.max_crossings = 1; b.status0 = BND_SPACE; cls_t = 1; }
   else if (!strcmp(cls, "BoundarySphereReflect")) { b.kind = BK_SPHERE; b.origin = rd_v(r); b.radius = rd_d(r); b.status0 = BND_SPACE | BND_REFLECT; cls_t = 1; }
   else if (!strcmp(cls, "BoundaryPlaneAbsorb")) { b.kind = BK_PLANE; b.origin = rd_v(r); b.norm = vunit(rd_v(r)); b.max_crossings = 1; b.status0 = BND_SPACE; cls_t = 1; }
   else if (!strcmp(cls, "BoundaryPlaneReflect")) { b.kind = BK_PLANE; b.origin = rd_v(r); b.norm = vunit(rd_v(r)); b.status0 = BND_SPACE | BND_REFLECT; cls_t = 1; }
   else if (!strcmp(cls, "BoundaryPlanePass")) { b.kind = BK_PLANE; b.origin = rd_v(r); b.norm = vunit(rd_v(r)); b.status0 = BND_SPACE; cls_t = 1; }
   else if (!strcmp(cls, "BoundaryMomentumInject")) { b.kind = BK_MOMENTUM; b.momentum = rd_d(r); b.max_crossings = 1; b.status0 = BND_MOMENTUM; cls_t = 2; }
   else if (!strcmp(cls, "BoundaryMomentumPass")) { b.kind = BK_MOMENTUM; b.momentum = rd_d(r); b.status0 = BND_MOMENTUM; cls_t = 2; }
   else return -1;
   if (cls_t == 0) { if (s->n_t >= MAX_BND) return -2; s->bt[s->n_t++] = b; }
   else if (cls_t == 1) { if (s->n_s >= MAX_BND) return -2; s->bs[s->n_s++] = b; }
   else { if (s->n_m >= MAX_BND) return -2; s->bm[s->n_m++] = b; }
   return 0;
}

static int setup_initial(orc_sim* s, const char* cls, dcread* r)
{
   if (!strcmp(cls, "InitialTimeFixed")) { s->it_kind = IT_FIXED; s->it_t = rd_d(r); }
   else if (!strcmp(cls, "InitialSpaceFixed")) { s->is_kind = IS_FIXED; s->is_pos = rd_v(r); }
   else if (!strcmp(cls, "InitialSpaceSphere")) {   /* src/initial_space.cc:297-316 */
      s->is_kind = IS_SPHERE; s->is_pos = rd_v(r); s->is_radius = rd_d(r);
      s->is_ct1 = 1.0; s->is_ct2 = -1.0; s->is_phi1 = 0.0; s->is_phi2 = M_2PI_ORC;
   }
   else if (!strcmp(cls, "InitialSpaceSphereSector")) {   /* src/initial_space.cc:353-378 */
      double th1, th2, ph1, ph2;
      s->is_kind = IS_SECTOR; s->is_pos = rd_v(r); s->is_radius = rd_d(r);
      th1 = rd_d(r); if (th1 < 0.0) th1 = 0.0;
      th2 = rd_d(r); if (th2 > M_PI) th2 = M_PI; if (th2 < th1) th2 = th1;
      ph1 = rd_d(r);
      ph2 = rd_d(r); if (ph2 > M_2PI_ORC) ph2 = M_2PI_ORC; if (ph2 < ph1) ph2 = ph1;
      s->is_ct1 = cos(th1); s->is_ct2 = cos(th2); s->is_phi1 = ph1; s->is_phi2 = ph2;
   }
   else if (!strcmp(cls, "InitialMomentumFixed")) { s->im_kind = IM_FIXED; s->im_mom = rd_v(r); }
   else if (!strcmp(cls, "InitialMomentumShell")) { s->im_kind = IM_SHELL; s->im_p0 = rd_d(r); }
   else if (!strcmp(cls, "InitialMomentumRing")) {
      double th;
      s->im_kind = IM_RING; s->im_p0 = rd_d(r); th = rd_d(r);
      s->im_mu0 = cos(th); s->im_st0 = sin(th);
   }
   else if (!strcmp(cls, "InitialMomentumThickShell")) {
      s->im_kind = IM_THICKSHELL; s->im_p1 = rd_d(r); s->im_p2 = rd_d(r); s->im_log = rd_b(r);
      if (s->im_log) { s->im_p1 = log10(s->im_p1); s->im_p2 = log10(s->im_p2); }
   }
   else if (!strcmp(cls, "InitialMomentumShellGrid")) { s->im_kind = IM_SHELLGRID; s->im_T1 = rd_d(r); s->im_T2 = rd_d(r); s->im_n = rd_i(r); if (s->im_n < 1) return -2; }
   else return -1;
   return 0;
}

static int setup_distribution(orc_sim* s, const char* cls, dcread* r)
{
   if (s->n_dist >= MAX_DIST) return -2;
   dist_t* d = &s->dist[s->n_dist];
   memset(d, 0, sizeof(*d));
   int width = 1;
   if (!strcmp(cls, "DistributionPositionCumulativeOrder1")) width = 3;
   else if (!strcmp(cls, "DistributionPositionCumulativeOrder2")) width = 9;
   /* DistributionTemplated::SetupDistribution (src/distribution_templated.cc:92-139) */
   rd_mi(r, d->n_bins); d->minval = rd_v(r); d->maxval = rd_v(r); rd_mi(r, d->log_bins); rd_mi(r, d->bin_outside);
   rd_skip(r);   /* unit_distro (distroClass) */
   d->unit_val = rd_v(r); d->keep_records = rd_b(r);
   int ijk;
   d->dims = 0;
   for (ijk = 0; ijk < 3; ijk++) {
      if (d->n_bins[ijk] > 0) d->dims |= (uint8_t)(1 << ijk);
      if (d->n_bins[ijk] < 1) d->n_bins[ijk] = 1;
   }
   for (ijk = 0; ijk < 3; ijk++) {
      if (d->log_bins[ijk]) { d->limits0.v[ijk] = log10(d->minval.v[ijk]); d->limits1.v[ijk] = log10(d->maxval.v[ijk]); }
      else { d->limits0.v[ijk] = d->minval.v[ijk]; d->limits1.v[ijk] = d->maxval.v[ijk]; }
      d->bin_size.v[ijk] = (d->limits1.v[ijk] - d->limits0.v[ijk]) / d->n_bins[ijk];
   }
   int need_dims = 1;
   if (!strcmp(cls, "DistributionSpectrumKineticEnergyPowerLaw")) {
      d->kind = DS_SPECTRUM; d->J0 = rd_d(r); d->T0 = rd_d(r); d->pow_law = rd_d(r); d->val_cold = rd_d(r);
   }
   else if (!strcmp(cls, "DistributionTimeUniform")) { d->kind = DS_TIME; d->val_hot = rd_d(r); d->val_cold = rd_d(r); d->val_time = rd_i(r); }
   else if (!strcmp(cls, "DistributionMomentumUniform")) {
      d->kind = DS_MOMENTUM; d->val_hot = rd_d(r); d->val_cold = rd_d(r); d->val_time = rd_i(r); d->val_coord = rd_i(r); need_dims = 7;
   }
   else if (!strcmp(cls, "DistributionPositionMomentumUniform")) {
      d->kind = DS_POSMOM; d->val_hot = rd_d(r); d->val_cold = rd_d(r); d->val_time = rd_i(r); d->pos_idx = rd_i(r); d->mom_idx = rd_i(r); need_dims = 5;
   }
   else if (!strcmp(cls, "DistributionPositionPitchUniform")) {
      d->kind = DS_POSPITCH; d->val_hot = rd_d(r); d->val_cold = rd_d(r); d->val_time = rd_i(r); need_dims = 3;
   }
   else if (!strcmp(cls, "DistributionPositionCumulativeOrder1")) d->kind = DS_CUMO1;
   else if (!strcmp(cls, "DistributionPositionCumulativeOrder2")) d->kind = DS_CUMO2;
   else return -1;
   if (d->dims != need_dims) return -3;  /* STATE_SETUP_COMPLETE would be lowered in the reference */
   d->width = width;
   d->nb = d->n_bins[0] * d->n_bins[1] * d->n_bins[2];
   d->counts = (int64_t*)calloc((size_t)d->nb, sizeof(int64_t));
   d->sums = (double*)calloc((size_t)d->nb * (size_t)width, sizeof(double));
   s->n_dist++;
   return 0;
}

int orc_add(orc_sim* s, int role, const char* class_name, const uint8_t* bytes, size_t len)
{
   dcread r = {bytes, bytes + len, 0};
   int rc;
   switch (role) {
   case ORC_BACKGROUND: rc = setup_background(s, class_name, &r); break;
   case ORC_DIFFUSION: rc = setup_diffusion(s, class_name, &r); break;
   case ORC_SOURCE:
      if (!strcmp(class_name, "SourceMomentumInjection")) { s->src_kind = SRC_MOMINJ; s->src_p_inj = rd_d(&r); s->src_rate = rd_d(&r); rc = 0; }
      else rc = -1;
      break;
   case ORC_BOUNDARY: rc = setup_boundary(s, class_name, &r); break;
   case ORC_INITIAL: rc = setup_initial(s, class_name, &r); break;
   case ORC_DISTRIBUTION: rc = setup_distribution(s, class_name, &r); break;
   default: rc = -1;
   }
   if (rc == 0 && r.err) return -4;
   return rc;
}

int64_t orc_run(orc_sim* s, uint64_t first_traj, uint64_t n_traj, const orc_rng* rng, int n_threads, orc_traj* per_traj)
{
   if (!s->have_bg || s->n_t < 1) return -1;
   if (s->traj_type < ORC_TRAJ_GUIDING_DIFF_SCATT && !s->have_df) return -1;
   int nthr = n_threads > 0 ? n_threads : 1;
   int64_t total_steps = 0;
   int d;
   /* per-thread histograms, merged in thread order */
   int64_t** tcounts = (int64_t**)calloc((size_t)nthr * MAX_DIST, sizeof(int64_t*));
   double** tsums = (double**)calloc((size_t)nthr * MAX_DIST, sizeof(double*));
   int64_t* tev = (int64_t*)calloc((size_t)nthr * MAX_DIST, sizeof(int64_t));
   int64_t* tdisc = (int64_t*)calloc((size_t)nthr * 8, sizeof(int64_t));
   int th;
   for (th = 0; th < nthr; th++) for (d = 0; d < s->n_dist; d++) {
      tcounts[th * MAX_DIST + d] = (int64_t*)calloc((size_t)s->dist[d].nb, sizeof(int64_t));
      tsums[th * MAX_DIST + d] = (double*)calloc((size_t)s->dist[d].nb * (size_t)s->dist[d].width, sizeof(double));
   }
   int64_t i;
#ifdef _OPENMP
#pragma omp parallel for num_threads(nthr) schedule(dynamic, 16) reduction(+:total_steps)
#endif
   for (i = 0; i < (int64_t)n_traj; i++) {
#ifdef _OPENMP
      int tid = omp_get_thread_num();
#else
      int tid = 0;
#endif
      traj_t T;
      lane_t L;
      memset(&L, 0, sizeof(L));
      T.hc = &tcounts[tid * MAX_DIST];
      T.hs = &tsums[tid * MAX_DIST];
      T.hev = &tev[tid * MAX_DIST];
      L.traj = first_traj + (uint64_t)i;
      L.rng = rng;
      int ex = ORC_EX_NONE, att;
      for (att = 0; att < s->max_attempts; att++) {
         L.attempt = att;
         L.exhausted = 0;
         ex = integrate_one(s, &L, &T);
         if (ex == ORC_EX_NONE) break;
         tdisc[tid * 8 + ex]++;
         if (ex == ORC_EX_REPLAY_EXHAUSTED) break;
      }
      if (ex == ORC_EX_NONE) total_steps += T.segs;
      if (per_traj) {
         orc_traj* o = &per_traj[i];
         memset(o, 0, sizeof(*o));
         o->t = T.t; o->amp = T.amp; o->wgt = T.wgt; o->t0 = T.t0;
         int k;
         for (k = 0; k < 3; k++) { o->pos[k] = T.pos.v[k]; o->mom[k] = T.mom.v[k]; o->pos0[k] = T.pos0.v[k]; o->mom0[k] = T.mom0.v[k]; }
         o->steps = T.segs;
         o->status = (ex == ORC_EX_NONE) ? ORC_ST_OK : (ORC_ST_DISCARDED | (ex << 8));
         o->attempts = att + (ex == ORC_EX_NONE ? 1 : 0);
         o->n_events = T.n_events;
         for (d = 0; d < 4; d++) {
            o->ev_bin[d] = (d < s->n_dist && ex == ORC_EX_NONE && T.ev_has[d]) ? T.ev_bin[d] : -1;
            o->ev_weight[d] = (d < s->n_dist && ex == ORC_EX_NONE && T.ev_has[d]) ? T.ev_w[d][0] : 0.0;
         }
      }
   }
   for (th = 0; th < nthr; th++) {
      for (d = 0; d < s->n_dist; d++) {
         dist_t* D = &s->dist[d];
         int b, k;
         for (b = 0; b < D->nb; b++) {
            D->counts[b] += tcounts[th * MAX_DIST + d][b];
            for (k = 0; k < D->width; k++) D->sums[(size_t)b * D->width + k] += tsums[th * MAX_DIST + d][(size_t)b * D->width + k];
         }
         D->n_events += tev[th * MAX_DIST + d];
         free(tcounts[th * MAX_DIST + d]); free(tsums[th * MAX_DIST + d]);
      }
      for (d = 0; d < 8; d++) s->discards[d] += tdisc[th * 8 + d];
   }
   free(tcounts); free(tsums); free(tev); free(tdisc);
   return total_steps;
}

int orc_n_distros(const orc_sim* s) { return s->n_dist; }
int orc_distro_nbins(const orc_sim* s, int d) { return (d >= 0 && d < s->n_dist) ? s->dist[d].nb : -1; }
int orc_distro_width(const orc_sim* s, int d) { return (d >= 0 && d < s->n_dist) ? s->dist[d].width : -1; }

int orc_get_distro(const orc_sim* s, int d, int64_t* counts, double* sums, int64_t* n_events)
{
   if (d < 0 || d >= s->n_dist) return -1;
   const dist_t* D = &s->dist[d];
   if (counts) memcpy(counts, D->counts, sizeof(int64_t) * (size_t)D->nb);
   if (sums) memcpy(sums, D->sums, sizeof(double) * (size_t)D->nb * (size_t)D->width);
   if (n_events) *n_events = D->n_events;
   return 0;
}

void orc_reset_distros(orc_sim* s)
{
   int d;
   for (d = 0; d < s->n_dist; d++) {
      memset(s->dist[d].counts, 0, sizeof(int64_t) * (size_t)s->dist[d].nb);
      memset(s->dist[d].sums, 0, sizeof(double) * (size_t)s->dist[d].nb * (size_t)s->dist[d].width);
      s->dist[d].n_events = 0;
   }
   memset(s->discards, 0, sizeof(s->discards));
}

int64_t orc_discards(const orc_sim* s, int reason) { return (reason >= 0 && reason < 8) ? s->discards[reason] : -1; }

void orc_background_fields(orc_sim* s, double t, const double pos[3], double p, double out_U[3], double out_B[3], double* out_dmax, double* out_divU)
{
   lane_t L; memset(&L, 0, sizeof(L));
   spd_t sp; memset(&sp, 0, sizeof(sp));
   vec3 x = V(pos[0], pos[1], pos[2]);
   if (s->bg.kind == BG_SMOOTHSHOCK) L.smooth_ds = (vdot(vsub(x, s->bg.r0), s->bg.n_shock) - s->bg.v_shock * t) / s->bg.width_shock;
   get_fields(&s->bg, &L, t, x, p, BG_U | BG_B | BG_gradU | BG_gradB, &sp);
   int k;
   for (k = 0; k < 3; k++) { out_U[k] = sp.U.v[k]; out_B[k] = sp.B.v[k]; }
   if (out_dmax) *out_dmax = sp.dmax;
   if (out_divU) *out_divU = divU(&sp);
}

/* One diffusion component with the background's fields (incl. its region indicators) at pos */
double orc_kappa_at(orc_sim* s, int comp, double t, const double pos[3], double p, double mu)
{
   kap_t kp = {{0, 0, 0}};
   spd_t sp; memset(&sp, 0, sizeof(sp));
   lane_t L; memset(&L, 0, sizeof(L));
   vec3 x = V(pos[0], pos[1], pos[2]);
   bg_evaluate(&s->bg, &L, t, x, BG_ALL, &sp);
   sp.Bmag = vnorm(sp.B);
   return df_component(&s->df, &kp, comp, x, p, mu, &sp);
}

/* TrajectoryGuidingDiff::DiffusionCoeff at one point: kappa_perp and the drift V_perp = (grad kappa_perp)_perp
   for momentum magnitude p and pitch cosine mu, with the guiding centre's fields (mask ALL | gradB | dBdt). */
int orc_gc_vperp(orc_sim* s, double t, const double pos[3], double p, double mu, double out[4])
{
   traj_t T; memset(&T, 0, sizeof(T));
   lane_t L; memset(&L, 0, sizeof(L));
   kap_t kp = {{0, 0, 0}};
   vec3 fa[3], Vperp;
   double Dperp;
   T.t = t; T.pos = V(pos[0], pos[1], pos[2]);
   T.mom = V(p * sqrt(1.0 - Sqr(mu)), 0.0, p * mu);
   if (get_fields(&s->bg, &L, t, T.pos, p, BG_ALL | BG_gradB | BG_dBdt, &T.sp)) return 1;
   gcd_diffcoeff(s, &T, &kp, fa, &Dperp, &Vperp);
   out[0] = Dperp; out[1] = Vperp.v[0]; out[2] = Vperp.v[1]; out[3] = Vperp.v[2];
   return 0;
}

double orc_kappa(orc_sim* s, int comp, double t, const double pos[3], double p, double mu, double Bmag)
{
   kap_t kp = {{0, 0, 0}};
   spd_t sp; memset(&sp, 0, sizeof(sp));
   lane_t L; memset(&L, 0, sizeof(L));
   /* the background at the point (its region indicators), with |B| given */
   if (s->bg.kind != BG_GRID) bg_evaluate(&s->bg, &L, t, V(pos[0], pos[1], pos[2]), BG_ALL, &sp);
   sp.Bmag = Bmag;
   return df_component(&s->df, &kp, comp, V(pos[0], pos[1], pos[2]), p, mu, &sp);
}
