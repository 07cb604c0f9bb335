// spx_static.hip -- the run that allocates nothing (spx_run_static in include/spectrum_amd.h).
//
// Everything a run keeps on the device beyond its inputs and outputs -- the launch block, the query
// table entry, the plan, the control block, the histograms and the state of SPX_STATIC_LANES resident
// slots -- is carved out of one file-scope device array, `scratch`, instead of hipMalloc'd buffers.
// The array exists once per GPU for the life of the process.  Host code never passes the array's name
// to a kernel or a copy: at the first call on a device, with that device current, hipGetSymbolAddress
// gives its device address, kept per device, and every kernel argument and copy below uses that.
// Every call on one GPU shares the array: the call clears what must start at zero, and calls that
// overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <mutex>

#include "spx_kernels.h"
#include "spx_static.h"

namespace spx {
namespace {

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Bytes of the per-lane arrays of a slot pool (the fields of Launch, sized as Pool::init sizes them).
constexpr size_t lane_bytes(size_t n)
{
   return 6 * align256(4 * n) + 2 * align256(4 * n * SPX_MAX_BND) + align256(4 * n * SPX_MAX_DIST) + align256(4 * n * 3) +
          22 * align256(8 * n) + 2 * align256(8 * n * SPX_MAX_BND) + align256(8 * n * SPX_MAX_DIST) + align256(8 * n * 27) +
          align256(8 * n) + 2 * align256(4 * n) + align256(8 * n) + align256(4 * n * SPX_KC);
}

constexpr size_t header_bytes()
{
   return align256(sizeof(Launch)) + align256(sizeof(QueryDesc)) + align256(sizeof(int32_t)) + align256(sizeof(spx_plan_desc)) +
          align256(sizeof(RunCtl)) + align256(8 * SPX_MAX_DIST) + align256(8 * (size_t)SPX_STATIC_BINS) +
          align256(8 * (size_t)SPX_ACC_LIMBS * SPX_STATIC_COMPONENTS);
}

constexpr size_t kScratchBytes = header_bytes() + lane_bytes(SPX_STATIC_LANES);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on each GPU, looked up once per device with that device current.
std::mutex g_mutex;
unsigned char* g_base[64] = {};

hipError_t scratch_base(int dev, unsigned char** out)
{
   if (dev < 0 || dev >= 64) return hipErrorInvalidDevice;
   std::lock_guard<std::mutex> lock(g_mutex);
   if (!g_base[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return e;
      g_base[dev] = (unsigned char*)p;
   }
   *out = g_base[dev];
   return hipSuccess;
}

struct Carver {
   unsigned char* base;
   size_t off = 0;
   template <typename T>
   T* take(size_t count)
   {
      off = align256(off);
      T* p = (T*)(base + off);
      off += sizeof(T) * count;
      return p;
   }
};

// The device pieces of one static run.
struct StaticRun {
   Launch* launch;
   QueryDesc* query;
   int32_t* queue;
   spx_plan_desc* plan;
   RunCtl* ctl;
   unsigned long long *events, *counts, *acc;
   Launch host{};
};

size_t carve(unsigned char* base, uint64_t n, StaticRun& R)
{
   Carver c{base};
   R.launch = c.take<Launch>(1);
   R.query = c.take<QueryDesc>(1);
   R.queue = c.take<int32_t>(1);
   R.plan = c.take<spx_plan_desc>(1);
   R.ctl = c.take<RunCtl>(1);
   R.events = c.take<unsigned long long>(SPX_MAX_DIST);
   R.counts = c.take<unsigned long long>(SPX_STATIC_BINS);
   R.acc = c.take<unsigned long long>((size_t)SPX_ACC_LIMBS * SPX_STATIC_COMPONENTS);
   Launch& L = R.host;
   L.slot_q = c.take<int32_t>(n); L.active = c.take<int32_t>(n); L.attempt = c.take<int32_t>(n);
   L.n_events = c.take<int32_t>(n); L.ev_pending = c.take<int32_t>(n); L.ex_code = c.take<int32_t>(n);
   L.bcross = c.take<int32_t>(n * SPX_MAX_BND); L.bstat = c.take<int32_t>(n * SPX_MAX_BND);
   L.ev_bin = c.take<int32_t>(n * SPX_MAX_DIST); L.ev_bnd = c.take<int32_t>(n * 3);
   double** dv[] = {&L.t, &L.x, &L.y, &L.z, &L.m0, &L.m1, &L.m2, &L.amp, &L.wgt, &L.dt_prev, &L.src_old,
                    &L.ds_prev, &L.dmax, &L.dt_adapt, &L.pf, &L.t0, &L.x0, &L.y0, &L.z0, &L.m00, &L.m01, &L.m02};
   for (double** d : dv) *d = c.take<double>(n);
   L.bdelta = c.take<double>(n * SPX_MAX_BND); L.bdelta_old = c.take<double>(n * SPX_MAX_BND);
   L.ev_w = c.take<double>(n * SPX_MAX_DIST); L.evs = c.take<double>(n * 27);
   L.traj = c.take<uint64_t>(n); L.nn = c.take<uint32_t>(n); L.nu = c.take<uint32_t>(n); L.steps = c.take<int64_t>(n);
   L.kc = c.take<float>(n * SPX_KC);
   L.queries = R.query;
   L.queue = R.queue;
   L.n_queue = 1;
   L.n_lanes = n;
   return c.off;
}

__global__ void fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
   for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void fill_u64(unsigned long long* p, uint64_t n, unsigned long long v)
{
   for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

hipError_t fill32(int32_t* p, uint64_t n, uint32_t v)
{
   if (!n) return hipSuccess;
   hipLaunchKernelGGL(fill_u32, dim3((unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)), dim3(256), 0, nullptr, (uint32_t*)p, n, v);
   return hipGetLastError();
}
hipError_t fill64(unsigned long long* p, uint64_t n, unsigned long long v)
{
   if (!n) return hipSuccess;
   hipLaunchKernelGGL(fill_u64, dim3((unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096)), dim3(256), 0, nullptr, p, n, v);
   return hipGetLastError();
}

}  // namespace

#define TRY(expr)                          \
   do {                                    \
      hipError_t e_ = (expr);              \
      if (e_ != hipSuccess) return e_;     \
   } while (0)

hipError_t static_run(const spx_plan_desc& plan, int precision, int steps_per_launch, uint64_t first, uint64_t n_traj, uint64_t seed,
                      int64_t* counts, double* sums, int64_t* n_events, StaticResult& res)
{
   int dev = 0;
   TRY(hipGetDevice(&dev));
   unsigned char* base = nullptr;
   TRY(scratch_base(dev, &base));
   StaticRun R;
   const uint64_t n = SPX_STATIC_LANES;
   if (carve(base, n, R) > kScratchBytes) return hipErrorInvalidValue;
   int64_t bins = 0, comps = 0;
   for (int d = 0; d < plan.n_dist; d++) {
      bins += plan.dist[d].nb;
      comps += (int64_t)plan.dist[d].nb * plan.dist[d].width;
   }
   if (bins > SPX_STATIC_BINS || comps > SPX_STATIC_COMPONENTS) return hipErrorInvalidValue;
   R.host.steps_per_launch = steps_per_launch;

   QueryDesc q;
   std::memset(&q, 0, sizeof(q));
   q.plan = R.plan;
   q.ctl = R.ctl;
   q.rng.mode = 0;
   q.rng.seed = seed;
   q.first = first;
   q.end = first + n_traj;
   q.n_dist = plan.n_dist;
   for (int d = 0; d < SPX_MAX_DIST; d++) { q.hist_off[d] = plan.hist_off[d]; q.acc_off[d] = plan.acc_off[d]; }
   q.hist_counts = R.counts;
   q.hist_acc = R.acc;
   q.hist_events = R.events;
   RunCtl ctl;
   std::memset(&ctl, 0, sizeof(ctl));
   ctl.next = first;
   ctl.tmin_bits = 0x7FF0000000000000ull;   // +inf
   const int32_t q0 = 0;
   // uploads (synchronous copies on the null stream, ordered before the kernels below)
   TRY(hipMemcpy(R.plan, &plan, sizeof(plan), hipMemcpyHostToDevice));
   TRY(hipMemcpy(R.query, &q, sizeof(q), hipMemcpyHostToDevice));
   TRY(hipMemcpy(R.queue, &q0, sizeof(q0), hipMemcpyHostToDevice));
   TRY(hipMemcpy(R.ctl, &ctl, sizeof(ctl), hipMemcpyHostToDevice));
   TRY(hipMemcpy(R.launch, &R.host, sizeof(Launch), hipMemcpyHostToDevice));
   // the previous call's contents: idle every slot, zero the histograms this plan uses
   TRY(fill32(R.host.active, n, 0u));            // SLOT_IDLE
   TRY(fill32(R.host.slot_q, n, 0xffffffffu));   // -1: no query
   TRY(fill32(R.host.ev_pending, n, 0u));
   TRY(fill64(R.events, SPX_MAX_DIST, 0ull));
   TRY(fill64(R.counts, (uint64_t)bins, 0ull));
   TRY(fill64(R.acc, (uint64_t)comps * SPX_ACC_LIMBS, 0ull));

   auto w0 = std::chrono::steady_clock::now();
   // a trajectory needs at most max_attempts * (max_steps + 1) steps; a healthy run never reaches this
   const double guard = 64.0 + (double)plan.max_attempts * ((double)plan.max_steps + 1.0) / steps_per_launch *
                                   (1.0 + (double)n_traj / (double)n);
   int launches = 0, batch = 1;
   while (true) {
      TRY(hipMemcpy(&ctl, R.ctl, sizeof(ctl), hipMemcpyDeviceToHost));
      if (ctl.done >= n_traj) break;
      if (launches > guard) return hipErrorLaunchTimeOut;
      for (int i = 0; i < batch; i++) TRY(launch_transport(plan, n, R.launch, precision, nullptr, nullptr));
      launches += batch;
      if (batch < 8) batch *= 2;
   }
   TRY(hipDeviceSynchronize());
   res.wall_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
   res.launches = launches;
   res.lanes = (int64_t)n;
   res.ctl = ctl;

   // outputs into the caller's host arrays; exact sums converted a few components at a time
   for (int d = 0; d < plan.n_dist; d++) {
      const spx_ds_desc& D = plan.dist[d];
      if (counts)
         TRY(hipMemcpy(counts + plan.hist_off[d], R.counts + plan.hist_off[d], sizeof(unsigned long long) * D.nb, hipMemcpyDeviceToHost));
      if (sums) {
         const int64_t nc = (int64_t)D.nb * D.width;
         constexpr int64_t chunk = 32;
         unsigned long long limbs[chunk * SPX_ACC_LIMBS];
         for (int64_t c0 = 0; c0 < nc; c0 += chunk) {
            const int64_t m = nc - c0 < chunk ? nc - c0 : chunk;
            TRY(hipMemcpy(limbs, R.acc + (size_t)(plan.acc_off[d] + c0) * SPX_ACC_LIMBS, sizeof(unsigned long long) * SPX_ACC_LIMBS * m,
                          hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m; i++) sums[plan.acc_off[d] + c0 + i] = acc_value(limbs + i * SPX_ACC_LIMBS);
         }
      }
   }
   if (n_events && plan.n_dist > 0) {
      unsigned long long ev[SPX_MAX_DIST];
      TRY(hipMemcpy(ev, R.events, sizeof(ev), hipMemcpyDeviceToHost));
      for (int d = 0; d < plan.n_dist; d++) n_events[d] = (int64_t)ev[d];
   }
   return hipSuccess;
}

}  // namespace spx
