/*
 * spx_simulation_gpu.hpp -- C++ adaptor over the C ABI (spectrum_amd.h) for SPECTRUM drivers.
 *
 * A driver in the style of runs/gcr_modulation_example.cc configures a SimulationWorker with
 * AddBackground(proto, container) and friends (src/simulation.hh:37-140), then SetTasks / MainLoop
 * (src/simulation.cc:509-519, :818-845).  SimulationGPU offers the same calls; a plugin prototype is any
 * object with GetName() (the reference's class name) and a container any object with data() / length()
 * (common/data_container.hh:68-77 gives the DataContainer both), so the reference's own classes plug in
 * unchanged.  Errors become std::runtime_error with spx_last_error's message.
 */
#ifndef SPX_SIMULATION_GPU_HPP
#define SPX_SIMULATION_GPU_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "spectrum_amd.h"

class SimulationGPU {
   spx_sim* sim_ = nullptr;

   void check(int rc) const
   {
      if (rc != SPX_OK) throw std::runtime_error(std::string("spectrum_amd: ") + spx_last_error(sim_));
   }
   template <class P, class C>
   void add(int role, const P& proto, const C& c)
   {
      check(spx_add_plugin(sim_, role, proto.GetName().c_str(), reinterpret_cast<const uint8_t*>(c.data()), (size_t)c.length()));
   }

public:
   // CreateSimulation (src/simulation.cc:899-907); TRAJ_TYPE / TRAJ_TIME_FLOW as arguments
   SimulationGPU(int traj_type, int time_flow)
   {
      if (spx_create(traj_type, time_flow, &sim_) != SPX_OK) throw std::runtime_error("spectrum_amd: spx_create");
   }
   ~SimulationGPU() { spx_destroy(sim_); }
   SimulationGPU(const SimulationGPU&) = delete;
   SimulationGPU& operator=(const SimulationGPU&) = delete;
   spx_sim* handle() const { return sim_; }

   // AddBackground(proto, container, fname_pattern): a BackgroundServerCartesian reads its grid from
   // fname_pattern.info / fname_pattern.out into device memory; analytic backgrounds read no files
   template <class P, class C>
   void AddBackground(const P& b, const C& c, const std::string& fname_pattern = "")
   {
      check(spx_add_background_file(sim_, b.GetName().c_str(), reinterpret_cast<const uint8_t*>(c.data()), (size_t)c.length(),
                                    fname_pattern.c_str()));
   }
   // the same grid from memory (blocks of 4x4x4 zones in the .out order, 6 doubles per zone)
   void SetBackgroundGrid(const int32_t nb[3], const double lo[3], const double hi[3], const std::vector<double>& blocks)
   {
      check(spx_set_background_grid(sim_, nb, lo, hi, 6, blocks.data()));
   }
   template <class P, class C> void AddDiffusion(const P& d, const C& c) { add(SPX_DIFFUSION, d, c); }
   template <class P, class C> void AddSource(const P& s, const C& c) { add(SPX_SOURCE, s, c); }
   template <class P, class C> void AddBoundary(const P& b, const C& c) { add(SPX_BOUNDARY, b, c); }
   template <class P, class C> void AddInitial(const P& i, const C& c) { add(SPX_INITIAL, i, c); }
   template <class P, class C> void AddDistribution(const P& d, const C& c) { add(SPX_DISTRIBUTION, d, c); }
   // compile-time switches of the reference as run-time options (spx_set_option)
   void SetOption(const char* name, double v) { check(spx_set_option(sim_, name, v)); }

   // every trajectory on the GPU with counter-based streams keyed by `seed`
   spx_run_stats Run(uint64_t first_traj, uint64_t n_traj, uint64_t seed)
   {
      spx_rng rng{};
      rng.mode = 0;
      rng.seed = seed;
      spx_run_stats st;
      check(spx_run(sim_, first_traj, n_traj, &rng, nullptr, &st));
      return st;
   }

   // SimulationMaster's task loop, serial form (the device is the one worker)
   void SetTasks(int64_t n_traj, int64_t batch_size, int64_t max_traj_per_worker = 0)
   {
      check(spx_set_tasks(sim_, n_traj, batch_size, max_traj_per_worker));
   }
   void DistroFileName(const std::string& base) { check(spx_distro_file_name(sim_, base.c_str())); }
   void RestoreDistro(int d) { check(spx_restore_distro_at_start(sim_, d)); }
   spx_run_stats MainLoop(uint64_t seed)
   {
      spx_rng rng{};
      rng.mode = 0;
      rng.seed = seed;
      spx_run_stats st;
      check(spx_main_loop(sim_, &rng, &st));
      return st;
   }

   // DistributionBase::NBins / GetCountsAddress / GetDistroAddress: counts and width sums per bin
   void GetDistro(int d, std::vector<int64_t>& counts, std::vector<double>& sums, int64_t& n_events) const
   {
      int32_t nb[3], w;
      check(spx_distro_info(sim_, d, nb, &w));
      const size_t n = (size_t)nb[0] * nb[1] * nb[2];
      counts.resize(n);
      sums.resize(n * (size_t)w);
      check(spx_get_distro(sim_, d, counts.data(), sums.data(), &n_events));
   }
   // SimulationMaster::PrintDistro1D / 2D, the end-of-run Dump and RestoreDistro
   void PrintDistro1D(int d, int ijk, const std::string& f, bool phys) const { check(spx_print_distro1d(sim_, d, ijk, f.c_str(), phys)); }
   void PrintDistro2D(int d, int i, int j, const std::string& f, bool phys) const { check(spx_print_distro2d(sim_, d, i, j, f.c_str(), phys)); }
   void Dump(int d, const std::string& f) const { check(spx_dump_distro(sim_, d, f.c_str())); }
   void Restore(int d, const std::string& f) { check(spx_restore_distro(sim_, d, f.c_str())); }
};

#endif
