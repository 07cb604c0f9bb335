// Test harness: a C++ driver that configures the engine through SimulationGPU (include/spx_simulation_gpu.hpp)
// the way a reference driver would, from a plugin list written by the Python side (role, class name,
// DataContainer bytes), then either reports the distributions' shapes ("info") or runs the trajectories
// and prints the distributions ("run").
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "spx_simulation_gpu.hpp"

namespace {

struct Proto {   // a plugin prototype: the reference's classes answer GetName() with their class name
   std::string name;
   const std::string& GetName() const { return name; }
};
struct Container {   // a DataContainer's wire bytes
   std::vector<uint8_t> bytes;
   const uint8_t* data() const { return bytes.data(); }
   size_t length() const { return bytes.size(); }
};

template <class T>
bool rd(std::ifstream& f, T& v) { return (bool)f.read(reinterpret_cast<char*>(&v), sizeof(T)); }

}  // namespace

int main(int argc, char** argv)
{
   if (argc < 3) { std::fprintf(stderr, "usage: driver plugins.bin info|run [n_traj seed]\n"); return 2; }
   std::ifstream f(argv[1], std::ios::binary);
   int32_t traj_type = 0, time_flow = 0, n_plugins = 0, n_options = 0;
   if (!rd(f, traj_type) || !rd(f, time_flow) || !rd(f, n_plugins) || !rd(f, n_options)) return 2;
   try {
      SimulationGPU sim(traj_type, time_flow);
      int n_dist = 0;
      for (int i = 0; i < n_plugins; i++) {
         int32_t role = 0;
         uint32_t nl = 0, bl = 0;
         rd(f, role); rd(f, nl);
         Proto p{std::string(nl, '\0')};
         f.read(&p.name[0], nl);
         rd(f, bl);
         Container c{std::vector<uint8_t>(bl)};
         f.read(reinterpret_cast<char*>(c.bytes.data()), bl);
         switch (role) {
         case SPX_BACKGROUND: sim.AddBackground(p, c); break;
         case SPX_DIFFUSION: sim.AddDiffusion(p, c); break;
         case SPX_SOURCE: sim.AddSource(p, c); break;
         case SPX_BOUNDARY: sim.AddBoundary(p, c); break;
         case SPX_INITIAL: sim.AddInitial(p, c); break;
         default: sim.AddDistribution(p, c); n_dist++; break;
         }
      }
      for (int i = 0; i < n_options; i++) {
         uint32_t nl = 0;
         double v = 0.0;
         rd(f, nl);
         std::string name(nl, '\0');
         f.read(&name[0], nl);
         rd(f, v);
         sim.SetOption(name.c_str(), v);
      }
      if (!std::strcmp(argv[2], "run")) {
         const uint64_t n = argc > 3 ? std::stoull(argv[3]) : 1000, seed = argc > 4 ? std::stoull(argv[4]) : 1;
         const spx_run_stats st = sim.Run(0, n, seed);
         std::printf("done %llu failed %llu steps %llu\n", (unsigned long long)st.done, (unsigned long long)st.failed,
                     (unsigned long long)st.steps_done);
      }
      for (int d = 0; d < n_dist; d++) {
         std::vector<int64_t> counts;
         std::vector<double> sums;
         int64_t nev = 0;
         sim.GetDistro(d, counts, sums, nev);
         std::printf("distro %d bins %zu width %zu events %lld\n", d, counts.size(), counts.empty() ? (size_t)0 : sums.size() / counts.size(),
                     (long long)nev);
         for (size_t b = 0; b < counts.size(); b++) {
            std::printf("%lld", (long long)counts[b]);
            for (size_t k = b * (sums.size() / counts.size()); k < (b + 1) * (sums.size() / counts.size()); k++) std::printf(" %.17g", sums[k]);
            std::printf("\n");
         }
      }
   }
   catch (const std::exception& e) {
      std::printf("error %s\n", e.what());
      return 3;
   }
   return 0;
}
