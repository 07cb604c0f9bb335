/*
 * spx_oracle.h -- CPU restatement of SPECTRUM's trajectory-integration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it.  The product never links or calls it.
 *
 * Plain C, FP64, the reference operation order, built with -ffp-contract=off.
 * Plugins are configured exactly as in the reference: a class name plus the
 * bytes of a DataContainer (common/data_container.cc:133-240 of the reference).
 */
#ifndef SPX_ORACLE_H
#define SPX_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Plugin roles, matching the Simulation::Add* families (src/simulation.hh:37-140). */
enum { ORC_BACKGROUND = 0, ORC_DIFFUSION = 1, ORC_SOURCE = 2, ORC_BOUNDARY = 3, ORC_INITIAL = 4, ORC_DISTRIBUTION = 5 };

/* Transport families (configure.ac:237-261 TRAJ_TYPE) and time flow (configure.ac:264-280). */
enum { ORC_TRAJ_PARKER = 0, ORC_TRAJ_PARKER_SOURCE = 1, ORC_TRAJ_GUIDING_DIFF_SCATT = 2, ORC_TRAJ_GUIDING = 3,
       ORC_TRAJ_GUIDING_SCATT = 4, ORC_TRAJ_GUIDING_DIFF = 5, ORC_TRAJ_FIELDLINE = 6,
       ORC_TRAJ_LORENTZ = 7, ORC_TRAJ_FOCUSED = 8 };
enum { ORC_FLOW_FORWARD = 0, ORC_FLOW_BACKWARD = 1 };

/* RNG modes: counter-based Philox4x32-10 streams, or replay of supplied draws. */
enum { ORC_RNG_COUNTER = 0, ORC_RNG_REPLAY = 1 };

typedef struct {
   int32_t mode;
   uint64_t seed;
   const double* normals;   /* replay: [n_traj][normal_stride] N(0,1) draws */
   const double* uniforms;  /* replay: [n_traj][uniform_stride] U(0,1) draws */
   int64_t normal_stride;
   int64_t uniform_stride;
} orc_rng;

/* Per-trajectory endpoint record (what the parity tests compare). */
typedef struct {
   double t, pos[3], mom[3], amp, wgt;
   double t0, pos0[3], mom0[3];
   int64_t steps;
   int32_t status;      /* ORC_ST_* */
   int32_t attempts;    /* attempts used (1 = no discard) */
   int32_t n_events;    /* distribution events recorded (all distributions) */
   int32_t ev_bin[4];   /* linear bin of the last event per distribution (-1 if none or outlier) */
   double ev_weight[4]; /* weight*exp(logwgt)+amp of that event */
} orc_traj;

enum { ORC_ST_OK = 0, ORC_ST_DISCARDED = 1 };
/* discard reasons mirror the reference exceptions */
enum {
   ORC_EX_NONE = 0, ORC_EX_FIELD = 1, ORC_EX_STEP_TOO_SMALL = 2, ORC_EX_STEP_NAN = 3,
   ORC_EX_MAX_STEPS = 4, ORC_EX_REPLAY_EXHAUSTED = 5, ORC_EX_BOUNDARY = 6
};

typedef struct orc_sim orc_sim;

orc_sim* orc_create(int traj_type, int time_flow);
void orc_destroy(orc_sim* s);
/* Returns 0 on success, negative on unknown class or malformed container. */
int orc_add(orc_sim* s, int role, const char* class_name, const uint8_t* bytes, size_t len);
/* Runtime replacements of compile-time switches: name/value pairs. */
int orc_set_option(orc_sim* s, const char* name, double value);
/* Integrate trajectories first_traj .. first_traj+n_traj-1.  per_traj may be NULL.
   Returns total number of accepted steps, or negative on error. */
int64_t orc_run(orc_sim* s, uint64_t first_traj, uint64_t n_traj, const orc_rng* rng, int n_threads, orc_traj* per_traj);
int orc_n_distros(const orc_sim* s);
int orc_distro_nbins(const orc_sim* s, int d);
/* Element width of the distro array (1 double, 3 GeoVector, 9 GeoMatrix). */
int orc_distro_width(const orc_sim* s, int d);
int orc_get_distro(const orc_sim* s, int d, int64_t* counts, double* sums, int64_t* n_events);
void orc_reset_distros(orc_sim* s);
int64_t orc_discards(const orc_sim* s, int reason);

/* BackgroundServerCartesian data (blocks of 4x4x4 zones, 6 values per zone, the .out order of
   src/reader_cartesian.cc:98-140).  Returns 0, or -1 on a malformed grid. */
int orc_set_grid(orc_sim* s, const int32_t nb[3], const double dom_min[3], const double dom_max[3], int32_t n_var, const double* data);
/* The gridded fields at one point: U(3), B(3), |B|, grad U (9, row = derivative direction),
   grad B (9), grad |B| (3), dmax, E(3).  Returns 0, 1 where the reference fails (field error), -1 without a grid. */
int orc_grid_fields(orc_sim* s, const double pos[3], double out[32]);
double orc_kappa_at(orc_sim* s, int comp, double t, const double pos[3], double p, double mu);
double orc_hyp2f1_11(double c, double z);   /* 2F1(1, 1; c; z) of DiffusionWNLTConstant */
/* guiding-centre kappa_perp and V_perp at a point: out = (kappa_perp, V_perp); 1 on a field error */
int orc_gc_vperp(orc_sim* s, double t, const double pos[3], double p, double mu, double out[4]);

/* Building blocks exposed for closed-form tests. */
void orc_background_fields(orc_sim* s, double t, const double pos[3], double p, double out_U[3], double out_B[3], double* out_dmax, double* out_divU);
double orc_kappa(orc_sim* s, int comp, double t, const double pos[3], double p, double mu, double Bmag);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_counter_normal(uint64_t seed, uint64_t traj, uint64_t draw);
double orc_counter_uniform(uint64_t seed, uint64_t traj, uint64_t draw);
/* Code-unit constants of common/specie.hh, evaluated as the reference does. */
double orc_const(const char* name);

#ifdef __cplusplus
}
#endif

#endif
