// spx_plan.h -- plain-data descriptors of a configured trajectory simulation.
//
// The host lowers each plugin (class name + DataContainer bytes, the reference's
// configuration format) into these POD blocks; the HIP kernels read them from device
// memory with uniform (scalar) loads.  Shared by host and device code.
#pragma once

#include <stdint.h>

// Host and device C++ (hipcc); the C ABI of include/spectrum_amd.h does not expose these types.

#define SPX_MAX_BND 4     // boundary slots per simulation (time + space + momentum)
#define SPX_MAX_DIST 4    // distributions per simulation
#define SPX_MAX_ACT 4     // one action per distribution

enum spx_bg_kind { SPX_BG_NONE = 0, SPX_BG_UNIFORM = 1, SPX_BG_SOLARWIND = 2, SPX_BG_SHOCK = 3, SPX_BG_SMOOTHSHOCK = 4, SPX_BG_GRID = 5 };
enum spx_df_kind {
   SPX_DF_NONE = 0, SPX_DF_ISO, SPX_DF_PARA, SPX_DF_PERP, SPX_DF_FULL, SPX_DF_FLOWMOM, SPX_DF_KINRAD, SPX_DF_RIGBMAG, SPX_DF_PERPISO,
   SPX_DF_STRAUSS, SPX_DF_POTGIETER, SPX_DF_QLT, SPX_DF_WNLT, SPX_DF_WNLT_RAMP, SPX_DF_SOQLT
};
enum spx_bnd_class { SPX_BC_TIME = 0, SPX_BC_SPACE = 1, SPX_BC_MOMENTUM = 2 };
enum spx_bnd_kind { SPX_BK_TIME = 0, SPX_BK_SPHERE = 1, SPX_BK_PLANE = 2, SPX_BK_MOMENTUM = 3 };
enum spx_is_kind { SPX_IS_FIXED = 0, SPX_IS_SPHERE = 1, SPX_IS_SECTOR = 2 };
enum spx_im_kind { SPX_IM_FIXED = 0, SPX_IM_SHELL = 1, SPX_IM_THICKSHELL = 2, SPX_IM_SHELLGRID = 3, SPX_IM_RING = 4 };
enum spx_ds_kind { SPX_DS_SPECTRUM = 0, SPX_DS_TIME, SPX_DS_MOMENTUM, SPX_DS_POSMOM, SPX_DS_CUMO1, SPX_DS_CUMO2, SPX_DS_POSPITCH };
enum spx_src_kind { SPX_SRC_NONE = 0, SPX_SRC_MOMINJ = 1 };

// Status bits of a boundary slot (src/boundary_base.hh:22-40 of the reference).
#define SPX_BND_REFLECT 0x0080
#define SPX_BND_TERMINAL 0x0100
#define SPX_BND_CROSSED 0x0200
#define SPX_BND_RECURRENT 0x0400

// Uniform Cartesian grid of BackgroundServerCartesian (src/reader_cartesian.cc:25-110): blocks of
// 4x4x4 cell-centred zones, stored in HBM as one zone array, x fastest, one record per zone of
// U(3), B(3), |B| and a pad in fluid units (64 B in double, 32 B in float).
#define SPX_GRID_BLOCK 4
#define SPX_GRID_REC 8
typedef struct {
   int32_t nb[3], nz[3];                // blocks and zones per axis
   double dom_min[3], dom_max[3];       // domain corners, fluid units (au)
   double block_length[3];              // (dom_max - dom_min) / nb, fluid units
   double f_len, f_vel, f_mag;          // Fluid / Particle units of length, velocity, magnetic field
   double bmag_min;                     // smallest zone |B| of the grid, fluid units
   // a box in code units, a relative 1e-9 inside the first and last zone centres: a position in it has
   // a stencil within the grid, whatever the rounding of the exact block arithmetic
   double inner_lo[3], inner_hi[3];
   // 0 when the derivative points of div(kappa) need no field values: kappa_perp == kappa_par
   // identically, kappa reads neither U nor |B|, and no interpolated |B| can fall below sp_tiny
   int32_t fd_fields, pad_;
   // device records.  FP64: the zone array itself (64 B per zone, x fastest; a stencil reads four rows of two
   // x-neighbours).  FP32: cell records, for each zone the 8 zone records of the trilinear stencil whose
   // lowest corner it is, in the interpolation's corner order c = i + 2j + 4k (256 B: one stencil is two
   // whole 128-B lines instead of four scattered rows).  The FP64 kernel keeps the zone array: its cell
   // records would be 512 B per zone, 8x the grid, and on grids beyond a few GB the scattered gathers then
   // outrun the address-translation caches (measured: half the rate on the 64^3-block grid).
   const double* z64;
   const float* c32;
} spx_grid_desc;

// Background parameters in the step type T (double for the drop-in kernel, float for the mixed one).
template <typename T>
struct spx_bg_t {
   int32_t kind;
   int32_t ep_diag;                    // SolarWind: eprime is diagonal with entries +-1 (rotation axis along +-z); 2: the identity
   double t0, r0[3];                   // always double: positions are accumulated in double
   T u0[3], B0[3], dmax0;
   // BackgroundSolarWind (src/background_solarwind.cc:58-84)
   T r_ref, dmax_fraction, eprime[3][3], ur0, w0, hp_rad;
   T ur0_rcp;                          // 1 / ur0, correctly rounded (the FP64 step's quotients by ur0)
   T sw_c1, sw_k;                      // FP32 step: B0 r_ref^2 and w0 / ur0 of the closed-form spiral
   T sw_kg;                            // FP32 step: w0 / ur0 times eprime[0][0] eprime[1][1], the winding in the global frame (diagonal bases)
   T sw_r2_ok;                         // FP32 step: |B| >= sp_tiny wherever |x - r0|^2 <= this (|B0| r_ref^2 / sp_tiny)
   // BackgroundShock / BackgroundSmoothShock (src/background_shock.cc:58-86, src/background_smooth_shock.cc:139-146)
   T n_shock[3], v_shock, compression, u1[3], B1[3], E0[3], E1[3], width_shock;
   // BackgroundUniform (src/background_uniform.cc:58-64)
   T Eu[3];
   // BackgroundServerCartesian (src/background_server_cartesian.cc, src/server_cartesian.cc)
   spx_grid_desc grid;
};
typedef spx_bg_t<double> spx_bg_desc;

template <typename T>
struct spx_df_t {
   int32_t kind, stream_dep_idx, pos_dep;
   int32_t region_idx;                  // Strauss / Potgieter: LISM_idx, the region indicator read (-1: none); SOQLT: Bmix_idx
   T D0, Dperp, Dpara;
   T kap0, U0, pow_law_U, p0, pow_law_p, kap_rat;
   T T0, r0, pow_law_T, pow_law_r, u_up, w_sh, s_sh, dn_up_rat;
   T lam0, R0, B0, pow_law_R, pow_law_B;
   // DiffusionStraussEtAl2013 (lam_in / lam_out) and DiffusionPotgieterEtAl2015 (kappa_in / kappa_out)
   T lam_in, lam_out, kap_rat_in, kap_rat_out;
   // DiffusionQLTConstant: A2A, k_min = 2 pi / l_max, ps_minus = ps_index - 1
   T A2A, k_min, ps_minus;
   // DiffusionWNLTConstant: A2T, A2L, ps_plus = ps_index + 1; DiffusionWNLTRampVLISM: the ramp of l_max
   // between z_nose and z_sheath and the reference k_min and amplitudes; 2F1(1, 1; c; z) with
   // c = (5 + ps_index) / 4, hyp_g = Gamma(c) Gamma(2 - c), hyp_r = (c - 1) / (c - 2)
   T A2T, A2L, ps_plus;
   T l_max_HP, z_nose, z_sheath, k_min_ref, A2A_ref, A2T_ref, A2L_ref, dl_max, dz;
   T hyp_c, hyp_g, hyp_r;
   // DiffusionEmpiricalSOQLTandUNLT: mean free paths, the perpendicular reduction factor and radius, and the
   // solar-cycle divisor 1 + solar_cycle_effect cos^2(region / 2), constant on the backgrounds it runs on
   T lam_para, lam_perp, kap_rat_red, r_lim_red, sc_den;
};
typedef spx_df_t<double> spx_df_desc;

typedef struct {
   int32_t cls, kind;           // spx_bnd_class, spx_bnd_kind
   int32_t status0;             // REFLECT / RECURRENT bits
   int32_t max_crossings;
   int32_t n_act;
   int32_t same_origin;         // a sphere whose previous slot is a sphere with the same origin (its |x - origin| is reused)
   int32_t actions[SPX_MAX_ACT];
   double timemark, origin[3], radius, norm[3], momentum;
} spx_bnd_desc;

typedef struct {
   int32_t kind, n_bins[3], log_bins[3], bin_outside[3], dims, width, nb;
   int32_t val_time, val_coord, pos_idx, mom_idx, pad_;
   double limits0[3], bin_size[3];
   double val_hot, val_cold, J0, T0, pow_law;
} spx_ds_desc;

typedef struct {
   int32_t traj_type, time_flow;        // 0 Parker, 1 ParkerSource, 2 GuidingDiffScatt, 3 Guiding, 4 GuidingScatt,
                                        // 5 GuidingDiff; 0 forward, 1 backward
   int32_t n_bnd, n_dist;
   int32_t spectrum_type;               // DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE
   int32_t max_attempts;
   int64_t max_steps;                   // max_trajectory_steps
   double dt_scale;                     // Parker and guiding-centre transports: factor on the physical time step (1: the reference's)
   spx_bg_desc bg;
   spx_df_desc df;
   spx_bg_t<float> bgf;                 // the same parameters rounded to float (mixed-precision kernel)
   spx_df_t<float> dff;
   int32_t src_kind, pad_;
   double src_p_inj, src_rate;
   // initial conditions (src/initial_{time,space,momentum}.cc)
   int32_t im_kind, im_log, im_n, pad2_;
   double it_t, is_pos[3], im_mom[3], im_p0, im_p1, im_p2, im_T1, im_T2, im_mu0, im_st0;
   int32_t is_kind, pad4_;              // spx_is_kind; is_pos is the fixed point or the sphere's origin
   double is_radius, is_ct1, is_ct2, is_phi1, is_phi2;
   spx_bnd_desc bnd[SPX_MAX_BND];       // ordered: all time, then space, then momentum (ComputeAllBoundaries order)
   int32_t n_bt, n_bs, n_bm, pad3_;
   spx_ds_desc dist[SPX_MAX_DIST];
   int64_t hist_off[SPX_MAX_DIST];      // offsets of each distribution in the count array (bins)
   int64_t acc_off[SPX_MAX_DIST];       // offsets in the exact-sum array (bin components: bins x width)
} spx_plan_desc;

// RNG description passed to a run (counter-based streams or replayed draws).
typedef struct {
   int32_t mode;                       // 0 counter, 1 replay
   int32_t pad_;
   uint64_t seed;
   const double* normals;              // device pointers in a launch; host pointers at the C ABI
   const double* uniforms;
   int64_t normal_stride, uniform_stride;
} spx_rng_desc;
