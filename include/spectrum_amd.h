/*
 * spectrum_amd.h -- C ABI of the MI355X trajectory-integration engine.
 *
 * A drop-in for the per-trajectory loop of SPECTRUM's Simulation: plugins are configured
 * with the reference's own class names and DataContainer bytes, trajectories are integrated
 * by HIP kernels on one GPU, and the distributions come back as the reference's counts and
 * sums.  Plain C types only; every call returns an SPX_* status (no exceptions cross the ABI).
 * Citations name the reference interface (vflorins/SPECTRUM) each entry point replaces.
 */
#ifndef SPECTRUM_AMD_H
#define SPECTRUM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes: the reference's exception set mapped to return values. */
enum {
   SPX_OK = 0,
   SPX_ERR_ARG = -1,          /* bad argument */
   SPX_ERR_CLASS = -2,        /* unknown plugin class name */
   SPX_ERR_CONTAINER = -3,    /* DataContainer does not match the class (ExUninitialized analogue) */
   SPX_ERR_NOT_READY = -4,    /* TrajectoryBase::IsSimmulationReady false (src/trajectory_base.cc:608-632) */
   SPX_ERR_DEVICE = -5,       /* HIP runtime failure */
   SPX_ERR_STATE = -6,        /* call out of order (e.g. advance without begin) */
   SPX_ERR_LIMIT = -7         /* request beyond a documented limit */
};

/* Plugin families (SimulationWorker::Add*, src/simulation.hh:37-140). */
enum { SPX_BACKGROUND = 0, SPX_DIFFUSION = 1, SPX_SOURCE = 2, SPX_BOUNDARY = 3, SPX_INITIAL = 4, SPX_DISTRIBUTION = 5 };

/* Transport (configure.ac:237-261 TRAJ_TYPE) and time flow (configure.ac:264-280 TRAJ_TIME_FLOW). */
enum { SPX_TRAJ_PARKER = 0, SPX_TRAJ_PARKER_SOURCE = 1, SPX_TRAJ_GUIDING_DIFF_SCATT = 2, SPX_TRAJ_GUIDING = 3,
       SPX_TRAJ_GUIDING_SCATT = 4, SPX_TRAJ_GUIDING_DIFF = 5, SPX_TRAJ_FIELDLINE = 6, SPX_TRAJ_LORENTZ = 7,
       SPX_TRAJ_FOCUSED = 8 };
enum { SPX_FLOW_FORWARD = 0, SPX_FLOW_BACKWARD = 1 };

/* Step precision: FP64 drop-in, or FP32 step arithmetic with FP64 accumulation. */
enum { SPX_PREC_FP64 = 0, SPX_PREC_MIXED = 1 };

/* Discard reasons (the reference's exceptions, src/trajectory_base.hh:70-152, src/background_base.hh:62). */
enum {
   SPX_EX_FIELD = 1, SPX_EX_STEP_TOO_SMALL = 2, SPX_EX_STEP_NAN = 3, SPX_EX_MAX_STEPS = 4,
   SPX_EX_REPLAY_EXHAUSTED = 5, SPX_EX_BOUNDARY = 6
};

typedef struct spx_sim spx_sim;

/* Random numbers: counter-based Philox4x32-10 streams keyed by (seed, trajectory id), or replay of
   supplied N(0,1) / U(0,1] draws, one row per trajectory of the run (host pointers). */
typedef struct {
   int32_t mode;               /* 0 counter, 1 replay */
   int32_t pad_;
   uint64_t seed;
   const double* normals;      /* [n_traj][normal_stride] */
   const double* uniforms;     /* [n_traj][uniform_stride] */
   int64_t normal_stride;
   int64_t uniform_stride;
} spx_rng;

/* Endpoint of one trajectory (what the parity tests compare). */
typedef struct {
   double t, pos[3], mom[3], amp, wgt;
   double t0, pos0[3], mom0[3];
   int64_t steps;
   int32_t status;             /* 0 accepted; 1 | (reason << 8) when every attempt was discarded */
   int32_t attempts;
   int32_t n_events;
   int32_t ev_bin[4];          /* linear bin of the last event of each distribution, -1 if none */
   double ev_weight[4];        /* weight*exp(logwgt)+amp of that event (first component) */
} spx_traj_record;

typedef struct {
   uint64_t n_traj, done, failed;
   uint64_t steps_done;        /* accepted steps of finished trajectories */
   uint64_t steps_executed;    /* every executed step, the throughput counter */
   uint64_t discards[8];
   double t_min, t_max;        /* shortest / longest elapsed trajectory time (src/simulation.cc:266-273) */
   double wall_s;              /* host wall time of the call */
   double kernel_ms;           /* HIP-event time of the kernels of the call */
   int32_t launches, lanes;
   double step_kernel_ms;      /* HIP-event time of the step kernels alone (0 under graph replay) */
   double service_kernel_ms;   /* HIP-event time of the service kernels alone (0 under graph replay) */
} spx_run_stats;

/* CreateSimulation (src/simulation.cc:899-907) with TRAJ_TYPE / TRAJ_TIME_FLOW chosen at run time. */
int spx_create(int traj_type, int time_flow, spx_sim** out);
void spx_destroy(spx_sim* sim);

/* AddBackground / AddDiffusion / AddSource / AddBoundary / AddInitial / AddDistribution
   (src/simulation.hh:37-140 -> TrajectoryBase::Add*, src/trajectory_base.cc:644-741):
   `container` holds the DataContainer bytes of common/data_container.cc:133-240. */
int spx_add_plugin(spx_sim* sim, int role, const char* class_name, const uint8_t* container, size_t len);

/* Gridded background, BackgroundServerCartesian with first-order interpolation and no ghost cells
   (src/background_server_cartesian.cc, src/server_cartesian.cc; benchmarks/run_tests.sh:383).
   spx_add_background_file is SimulationWorker::AddBackground(proto, container, fname_pattern)
   (src/simulation.hh:111, src/simulation.cc:387-396): it adds the background and, for the gridded class,
   reads `fname_pattern`.info (blocks per axis, domain corners in au, variable count) and
   `fname_pattern`.out (binary doubles, block by block) as ReadCartesianHeader / ReadCartesianData do
   (src/reader_cartesian.cc:25-140).  spx_set_background_grid takes the same data from memory: blocks of
   4x4x4 zones, `n_var` (= 6: flow, magnetic field, fluid units) doubles per zone, the .out order.  The
   grid is kept in device memory once a run needs it; a position outside it discards the attempt. */
int spx_add_background_file(spx_sim* sim, const char* class_name, const uint8_t* container, size_t len, const char* fname_pattern);
int spx_set_background_grid(spx_sim* sim, const int32_t n_blocks[3], const double dom_min[3], const double dom_max[3], int32_t n_var,
                            const double* data);

/* Run-time replacement of compile-time switches and engine knobs:
   "DISTRO_KINETIC_ENERGY_POWER_LAW_TYPE" (src/distribution_other.hh:325), "max_trajectory_steps"
   (src/trajectory_base.hh:38), "max_attempts", "precision" (SPX_PREC_*), "block", "lanes",
   "steps_per_launch", "graph_launches" (0 = no graph capture), and "step_scale" in (0, 1]: a factor on
   the physical time step of the Parker and guiding-centre transports (TrajectoryParker::PhysicalStep's
   CFL and dlnpmax constants, src/trajectory_parker.hh:32-41; the guiding-centre, perpendicular and
   pitch-angle limits of src/trajectory_guiding*.cc; 1 = the reference's step), for time-step
   convergence studies of the Euler-Maruyama estimate. */
int spx_set_option(spx_sim* sim, const char* name, double value);

/* SetTasks + MainLoop (src/simulation.cc:509-519, :818-835): integrate trajectories
   first_traj .. first_traj+n_traj-1 to completion.  `records` (host, n_traj entries) may be NULL. */
int spx_run(spx_sim* sim, uint64_t first_traj, uint64_t n_traj, const spx_rng* rng, spx_traj_record* records, spx_run_stats* stats);

/* The same run split into steps of the persistent engine: begin queues the trajectories, advance
   executes `n_launches` launches of `steps_per_launch` steps per resident slot, finish drains. */
int spx_begin(spx_sim* sim, uint64_t first_traj, uint64_t n_traj, const spx_rng* rng, int keep_records);
int spx_advance(spx_sim* sim, int n_launches, spx_run_stats* stats);
int spx_finish(spx_sim* sim, spx_traj_record* records, spx_run_stats* stats);

/* Persistent-engine autotuner (no counterpart in the reference, whose MPI workers have no launch
   shape): times a short probe run of `n_probe` trajectories of this configuration for resident slots of
   1, 2, 3, 4, 6 and 8 waves per SIMD and 250 / 1000 / 4000 steps per launch, keeps the fastest as the
   simulation's "lanes" and "steps_per_launch" options, then times that shape replayed as a HIP graph and
   sets "graph_launches" to 2 when the graph is more than 2% faster (0 otherwise); resets the
   distributions.  `table` (may be NULL) receives up to `max_entries` measurements (19 are made). */
typedef struct {
   int64_t lanes;
   int32_t steps_per_launch;
   int32_t graph_launches;     /* 0: plain launches; 2: the shape replayed as a HIP graph of two launches */
   double steps_per_s;
} spx_tune_entry;
int spx_autotune(spx_sim* sim, uint64_t n_probe, uint64_t seed, spx_tune_entry* table, int max_entries, int* n_entries);

/* DistributionBase::NBins / GetCountsAddress / GetDistroAddress (src/distribution_base.hh,
   src/distribution_templated.cc).  counts: n_bins entries; sums: n_bins * width doubles. */
int spx_distro_info(const spx_sim* sim, int d, int32_t* n_bins3, int32_t* width);
int spx_get_distro(const spx_sim* sim, int d, int64_t* counts, double* sums, int64_t* n_events);
/* DistributionTemplated::ResetDistribution (src/distribution_templated.cc:212-235) for every distribution. */
int spx_reset_distros(spx_sim* sim);

/* Distribution files, byte for byte the reference's formats:
   spx_dump_distro is DistributionTemplated::Dump (src/distribution_templated.cc:394-422), the binary
   file the reference writes per distribution at the end of MainLoop (src/simulation.cc:790-793): class
   name, the configuration DataContainer, int event count, int counts and the bin sums (per-event
   records are not kept: both record counts are 0).  spx_restore_distro is Restore (:430-467), the
   checkpoint/resume path (SimulationWorker::RestoreDistro, src/simulation.cc:323): the file's counts,
   sums and event count replace distribution d's, and later runs add to them; the class and the number
   of bins must match.  spx_print_distro1d / 2d are Print1D / Print2D (:546-632) behind
   SimulationMaster::PrintDistro1D / 2D (src/simulation.cc:856-874): one text row per bin of dimension
   ijk (or per pair ijk1, ijk2), the other dimensions summed, with bin centre, mean, sum and count,
   in physical units when phys_units != 0. */
int spx_dump_distro(const spx_sim* sim, int d, const char* file_name);
int spx_restore_distro(spx_sim* sim, int d, const char* file_name);
int spx_print_distro1d(const spx_sim* sim, int d, int ijk, const char* file_name, int phys_units);
int spx_print_distro2d(const spx_sim* sim, int d, int ijk1, int ijk2, const char* file_name, int phys_units);

/* SimulationMaster's task loop in its serial form (one worker: the device).
   spx_set_tasks is SetTasks (src/simulation.cc:509-519) with its clamps: a negative count becomes 0, a
   batch size outside [1, n_traj] becomes min(1, n_traj), and a per-worker maximum below n_traj becomes
   0 (none).  spx_distro_file_name is DistroFileName (:496-499, default "distribution_"), and
   spx_restore_distro_at_start is SimulationMaster::RestoreDistro (:842-845): the main loop restores
   distribution d from <base><d>.out before it starts.
   spx_main_loop is MainLoop (:818-835): MasterStart (:699-737) resets the distributions, restores the
   flagged ones and prints the trajectory count; the run integrates n_traj trajectories (ids continue
   from the previous main loop's); whenever the completed percentage, counted in whole batches as
   DecrementTrajectoryCount counts them (:558-643), goes up, every distribution is dumped to
   <base><d>.out and the progress block is printed to stderr; MasterFinish (:776-811) dumps the final
   distributions and prints the shortest and longest elapsed trajectory time (s), the integration
   time per trajectory (ms, kernel time here) and the total time.  No workload manager is detected. */
int spx_set_tasks(spx_sim* sim, int64_t n_traj, int64_t batch_size, int64_t max_traj_per_worker);
int spx_distro_file_name(spx_sim* sim, const char* base);
int spx_restore_distro_at_start(spx_sim* sim, int d);
int spx_main_loop(spx_sim* sim, const spx_rng* rng, spx_run_stats* stats);

/* Run without allocating (no counterpart in the reference): for callers that cannot allocate device
   memory -- code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  The
   same integration as spx_run (counter streams keyed by `seed`, the sim's precision and steps per
   launch), but every piece of device state -- launch block, query entry, plan, control block,
   histograms and SPX_STATIC_LANES resident slots, 40.7 MiB in all -- lives in one static device array per GPU,
   held until the process ends; no hipMalloc, hipFree or other allocating call is made.  Results go to
   the caller's host arrays, all distributions in order: `counts` (n_bins of each, spx_distro_info),
   `sums` (n_bins x width of each) and `n_events` (one per distribution); any may be NULL.  They equal
   spx_get_distro's after spx_reset_distros + spx_run of the same trajectories, bit for bit; the sim's
   own distributions are untouched.  Rejected: a gridded background (SPX_ERR_ARG; its grid is device
   memory the library allocates), histograms beyond SPX_STATIC_BINS bins or SPX_STATIC_COMPONENTS
   bin components (SPX_ERR_LIMIT).  Every call on one GPU shares the array and clears what must start
   at zero; calls that overlap on one GPU race. */
#define SPX_STATIC_LANES 65536
#define SPX_STATIC_BINS 4096
#define SPX_STATIC_COMPONENTS 4096
int spx_run_static(spx_sim* sim, uint64_t first_traj, uint64_t n_traj, uint64_t seed, int64_t* counts, double* sums,
                   int64_t* n_events, spx_run_stats* stats);

const char* spx_last_error(const spx_sim* sim);
/* Name, compute units and architecture of device 0 ("" and SPX_ERR_DEVICE without a GPU). */
int spx_device_info(char* buf, size_t len);

/* ---- Query server ---------------------------------------------------------------------------
   Replaces the MPI master/worker request loop (SimulationMaster::MainLoop / WorkerDuties,
   src/simulation.cc:253-305, 743-835) for many concurrent simulations: each submitted query is a
   prototype simulation (its plugins and options) plus a trajectory range and a random-stream key.
   Queries wait in arrival order until their device memory fits `hbm_budget_bytes`, then run beside
   the others on shared resident slots (continuous batching); each keeps its own counters and
   histograms on the device, so its result is the same alone and in a mix. */
typedef struct spx_server spx_server;

enum { SPX_Q_QUEUED = 0, SPX_Q_RUNNING = 1, SPX_Q_DONE = 2, SPX_Q_FAILED = 3 };

typedef struct {
   int32_t precision;          /* SPX_PREC_* of every query */
   int32_t steps_per_launch;   /* steps per slot per engine launch (0: 500) */
   int64_t lanes;              /* resident slots per kernel pool (0: 8 waves per SIMD on every CU) */
   double hbm_budget_bytes;    /* device memory the admitted queries may hold (0: 4 GiB) */
   int32_t max_queries;        /* queries admitted at once per pool (0: 64) */
   int32_t launches_per_poll;  /* engine launches between completion checks (0: 1) */
} spx_server_opts;

typedef struct {
   int32_t state;              /* SPX_Q_* */
   int32_t pad_;
   uint64_t n_traj, done, failed, steps_done, steps_executed;
   double t_submit, t_admit, t_done;   /* seconds since the server started (-1: not yet) */
   double t_start;             /* end of the launch round in which its first trajectory started (-1: not yet) */
   double latency_s;           /* t_done - t_submit (-1 until done) */
   int64_t footprint_bytes;    /* device memory charged against the budget */
} spx_query_info;

typedef struct {
   uint64_t submitted, completed, launches, steps_executed;
   double busy_s;              /* engine time with admitted work */
   double wall_s;              /* since the server started */
   int32_t pools, running, queued, pad_;
} spx_server_stats;

int spx_server_create(const spx_server_opts* opts, spx_server** out);
/* Snapshot `proto`'s configuration and queue trajectories first_traj .. first_traj+n_traj-1 with
   counter-based streams keyed by `seed`. */
int spx_server_submit(spx_server* srv, const spx_sim* proto, uint64_t first_traj, uint64_t n_traj, uint64_t seed, int64_t* query_id);
int spx_server_info(spx_server* srv, int64_t query_id, spx_query_info* info);
/* Block until the query is done (SPX_OK), failed (SPX_ERR_DEVICE) or `timeout_s` passed (SPX_ERR_LIMIT; < 0 waits forever). */
int spx_server_wait(spx_server* srv, int64_t query_id, double timeout_s);
/* Histograms of a finished query, laid out as spx_get_distro's. */
int spx_server_result(spx_server* srv, int64_t query_id, int d, int64_t* counts, double* sums, int64_t* n_events);
int spx_server_release(spx_server* srv, int64_t query_id);
int spx_server_stats_get(spx_server* srv, spx_server_stats* stats);
const char* spx_server_last_error(spx_server* srv);
void spx_server_destroy(spx_server* srv);

#ifdef __cplusplus
}
#endif

#endif
