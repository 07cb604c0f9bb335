// spx_runtime.cpp -- slot pools and query resources (see spx_runtime.h).
#include "spx_runtime.h"

#include <algorithm>
#include <cstring>

namespace spx {

#define TRY(expr)                          \
   do {                                    \
      hipError_t e_ = (expr);              \
      if (e_ != hipSuccess) return e_;     \
   } while (0)

// ---- QueryRes ---------------------------------------------------------------------------------
size_t QueryRes::footprint(const spx_plan_desc& plan, int64_t hist_bins, uint64_t n_traj, const spx_rng* rng, bool keep_records)
{
   size_t b = sizeof(spx_plan_desc) + sizeof(RunCtl);
   b += (size_t)std::max<int64_t>(hist_bins, 1) * sizeof(unsigned long long) + SPX_MAX_DIST * sizeof(unsigned long long);
   b += (size_t)std::max<int64_t>(acc_components(plan), 1) * SPX_ACC_LIMBS * sizeof(unsigned long long);
   if (keep_records) b += (size_t)n_traj * sizeof(spx_traj_record);
   if (rng && rng->mode == 1)
      b += (size_t)n_traj * sizeof(double) * ((size_t)std::max<int64_t>(rng->normal_stride, 0) + (size_t)std::max<int64_t>(rng->uniform_stride, 1));
   return b;
}

int64_t acc_components(const spx_plan_desc& p)
{
   int64_t n = 0;
   for (int d = 0; d < p.n_dist; d++) n += (int64_t)p.dist[d].nb * p.dist[d].width;
   return n;
}

hipError_t QueryRes::create(const spx_plan_desc& p, int64_t bins, uint64_t first, uint64_t n, const spx_rng* rng, bool keep_records,
                            unsigned long long* hist_counts, unsigned long long* hist_acc, unsigned long long* hist_events, hipStream_t st)
{
   n_traj = n;
   hist_bins = bins;
   std::memset(&desc, 0, sizeof(desc));
   TRY(plan.alloc(1));
   TRY(ctl.alloc(1));
   TRY(hipMemcpyAsync(plan.p, &p, sizeof(spx_plan_desc), hipMemcpyHostToDevice, st));
   RunCtl c;
   std::memset(&c, 0, sizeof(c));
   c.next = first;
   c.tmin_bits = 0x7FF0000000000000ull;   // +inf
   TRY(hipMemcpyAsync(ctl.p, &c, sizeof(c), hipMemcpyHostToDevice, st));
   if (!hist_counts) {
      const size_t nb = (size_t)std::max<int64_t>(bins, 1);
      const size_t nl = (size_t)std::max<int64_t>(acc_components(p), 1) * SPX_ACC_LIMBS;
      TRY(own_counts.alloc(nb));
      TRY(own_acc.alloc(nl));
      TRY(own_events.alloc(SPX_MAX_DIST));
      TRY(hipMemsetAsync(own_counts.p, 0, sizeof(unsigned long long) * nb, st));
      TRY(hipMemsetAsync(own_acc.p, 0, sizeof(unsigned long long) * nl, st));
      TRY(hipMemsetAsync(own_events.p, 0, sizeof(unsigned long long) * SPX_MAX_DIST, st));
      hist_counts = own_counts.p;
      hist_acc = own_acc.p;
      hist_events = own_events.p;
   }
   desc.plan = plan.p;
   desc.ctl = ctl.p;
   desc.rng.mode = rng ? rng->mode : 0;
   desc.rng.seed = rng ? rng->seed : 0;
   if (rng && rng->mode == 1) {
      const size_t nn = (size_t)n * (size_t)rng->normal_stride;
      TRY(rep_normals.alloc(nn ? nn : 1));
      TRY(hipMemcpyAsync(rep_normals.p, rng->normals, sizeof(double) * nn, hipMemcpyHostToDevice, st));
      desc.rng.normals = rep_normals.p;
      desc.rng.normal_stride = rng->normal_stride;
      if (rng->uniforms && rng->uniform_stride > 0) {
         const size_t nu = (size_t)n * (size_t)rng->uniform_stride;
         TRY(rep_uniforms.alloc(nu));
         TRY(hipMemcpyAsync(rep_uniforms.p, rng->uniforms, sizeof(double) * nu, hipMemcpyHostToDevice, st));
         desc.rng.uniforms = rep_uniforms.p;
         desc.rng.uniform_stride = rng->uniform_stride;
      }
   }
   desc.first = first;
   desc.end = first + n;
   desc.n_dist = p.n_dist;
   for (int d = 0; d < SPX_MAX_DIST; d++) { desc.hist_off[d] = p.hist_off[d]; desc.acc_off[d] = p.acc_off[d]; }
   desc.hist_counts = hist_counts;
   desc.hist_acc = hist_acc;
   desc.hist_events = hist_events;
   if (keep_records) {
      TRY(records.alloc(n ? n : 1));
      TRY(hipMemsetAsync(records.p, 0, sizeof(spx_traj_record) * (n ? n : 1), st));
      desc.records = records.p;
      desc.rec_first = first;
      desc.rec_n = n;
   }
   return hipSuccess;
}

hipError_t QueryRes::read_ctl(RunCtl& out, hipStream_t st) const
{
   TRY(hipMemcpyAsync(&out, ctl.p, sizeof(RunCtl), hipMemcpyDeviceToHost, st));
   return hipStreamSynchronize(st);
}

hipError_t read_hist(const spx_plan_desc& p, int d, const unsigned long long* dev_counts, const unsigned long long* dev_acc,
                     const unsigned long long* dev_events, int64_t* counts, double* sums, int64_t* n_events, hipStream_t st)
{
   const spx_ds_desc& D = p.dist[d];
   const size_t nc = (size_t)D.nb * D.width;
   std::vector<unsigned long long> c((size_t)D.nb), acc(nc * SPX_ACC_LIMBS);
   unsigned long long ev[SPX_MAX_DIST];
   TRY(hipMemcpyAsync(c.data(), dev_counts + p.hist_off[d], sizeof(unsigned long long) * D.nb, hipMemcpyDeviceToHost, st));
   TRY(hipMemcpyAsync(acc.data(), dev_acc + (size_t)p.acc_off[d] * SPX_ACC_LIMBS, sizeof(unsigned long long) * acc.size(),
                      hipMemcpyDeviceToHost, st));
   TRY(hipMemcpyAsync(ev, dev_events, sizeof(ev), hipMemcpyDeviceToHost, st));
   TRY(hipStreamSynchronize(st));
   if (counts) for (int b = 0; b < D.nb; b++) counts[b] = (int64_t)c[b];
   if (sums) for (size_t i = 0; i < nc; i++) sums[i] = acc_value(acc.data() + i * SPX_ACC_LIMBS);
   if (n_events) *n_events = (int64_t)ev[d];
   return hipSuccess;
}

hipError_t QueryRes::read_distro(const spx_plan_desc& p, int d, int64_t* counts, double* sums, int64_t* n_events, hipStream_t st) const
{
   return read_hist(p, d, desc.hist_counts, desc.hist_acc, desc.hist_events, counts, sums, n_events, st);
}

void QueryRes::release()
{
   plan.release(); ctl.release(); own_counts.release(); own_events.release(); own_acc.release();
   rep_normals.release(); rep_uniforms.release(); records.release();
   std::memset(&desc, 0, sizeof(desc));
}

// ---- Pool -------------------------------------------------------------------------------------
bool Pool::matches(const spx_plan_desc& p, int prec) const
{
   return n_lanes && prec == precision && p.traj_type == sig.traj_type && p.time_flow == sig.time_flow && p.bg.kind == sig.bg.kind &&
          p.df.kind == sig.df.kind;
}

hipError_t Pool::init(const spx_plan_desc& s, int prec, uint64_t lanes, int spl, int graphs, int cap)
{
   release();
   sig = s;
   precision = prec;
   n_lanes = lanes;
   steps_per_launch = spl;
   graph_launches = graphs;
   qcap = cap;
   TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
   TRY(hipEventCreate(&ev0));
   TRY(hipEventCreate(&ev1));
   const uint64_t n = lanes;
   DevBuf<int32_t>* di[] = {&active, &slot_q, &attempt, &n_events, &ev_pending, &ex_code};
   for (auto* b : di) TRY(b->alloc(n));
   TRY(bcross.alloc(n * SPX_MAX_BND)); TRY(bstat.alloc(n * SPX_MAX_BND));
   TRY(ev_bin.alloc(n * SPX_MAX_DIST)); TRY(ev_w.alloc(n * SPX_MAX_DIST));
   TRY(evs.alloc(n * 3 * 9)); TRY(ev_bnd.alloc(n * 3));
   DevBuf<double>* dv[] = {&t, &x, &y, &z, &m0, &m1, &m2, &amp, &wgt, &dt_prev, &src_old, &ds_prev, &dmax, &dt_adapt, &pf,
                           &t0, &x0, &y0, &z0, &m00, &m01, &m02};
   for (auto* b : dv) TRY(b->alloc(n));
   TRY(bdelta.alloc(n * SPX_MAX_BND)); TRY(bdelta_old.alloc(n * SPX_MAX_BND));
   TRY(traj.alloc(n)); TRY(nn.alloc(n)); TRY(nu.alloc(n)); TRY(steps.alloc(n)); TRY(kc.alloc(n * SPX_KC));
   TRY(queries.alloc((size_t)cap));
   TRY(queue.alloc((size_t)cap));
   TRY(d_launch.alloc(1));
   qused.assign((size_t)cap, 0);
   host_queue.clear();
   std::memset(&host, 0, sizeof(host));
   host.queries = queries.p;
   host.queue = queue.p;
   host.n_queue = 0;
   host.steps_per_launch = spl;
   host.n_lanes = n;
   host.slot_q = slot_q.p;
   host.active = active.p;
   host.t = t.p; host.x = x.p; host.y = y.p; host.z = z.p; host.m0 = m0.p; host.m1 = m1.p; host.m2 = m2.p;
   host.amp = amp.p; host.wgt = wgt.p; host.dt_prev = dt_prev.p; host.src_old = src_old.p; host.ds_prev = ds_prev.p;
   host.dmax = dmax.p; host.dt_adapt = dt_adapt.p; host.pf = pf.p; host.kc = kc.p;
   host.bdelta = bdelta.p; host.bdelta_old = bdelta_old.p; host.bcross = bcross.p; host.bstat = bstat.p;
   host.traj = traj.p; host.nn = nn.p; host.nu = nu.p; host.attempt = attempt.p; host.steps = steps.p; host.n_events = n_events.p;
   host.t0 = t0.p; host.x0 = x0.p; host.y0 = y0.p; host.z0 = z0.p; host.m00 = m00.p; host.m01 = m01.p; host.m02 = m02.p;
   host.ev_bin = ev_bin.p; host.ev_w = ev_w.p;
   host.evs = evs.p; host.ev_bnd = ev_bnd.p; host.ev_pending = ev_pending.p; host.ex_code = ex_code.p;
   TRY(reset_slots());
   return hipSuccess;
}

hipError_t Pool::reset_slots()
{
   TRY(hipMemsetAsync(active.p, 0, sizeof(int32_t) * n_lanes, stream));          // SLOT_IDLE
   TRY(hipMemsetAsync(slot_q.p, 0xff, sizeof(int32_t) * n_lanes, stream));       // -1: no query
   TRY(hipMemsetAsync(ev_pending.p, 0, sizeof(int32_t) * n_lanes, stream));
   host_queue.clear();
   std::fill(qused.begin(), qused.end(), 0);
   return upload_launch();
}

hipError_t Pool::upload_launch()
{
   host.n_queue = (int32_t)host_queue.size();
   if (!host_queue.empty())
      TRY(hipMemcpyAsync(queue.p, host_queue.data(), sizeof(int32_t) * host_queue.size(), hipMemcpyHostToDevice, stream));
   TRY(hipMemcpyAsync(d_launch.p, &host, sizeof(Launch), hipMemcpyHostToDevice, stream));
   // the host vectors are pageable: the copies above have been staged before this returns
   return hipStreamSynchronize(stream);
}

int Pool::alloc_index()
{
   for (int q = 0; q < qcap; q++)
      if (!qused[(size_t)q]) { qused[(size_t)q] = 1; return q; }
   return -1;
}

void Pool::free_index(int q)
{
   if (q >= 0 && q < qcap) qused[(size_t)q] = 0;
}

hipError_t Pool::set_query(int q, const QueryDesc& d)
{
   TRY(hipMemcpyAsync(queries.p + q, &d, sizeof(QueryDesc), hipMemcpyHostToDevice, stream));
   return hipStreamSynchronize(stream);
}

hipError_t Pool::push_queue(int q)
{
   host_queue.push_back(q);
   return upload_launch();
}

hipError_t Pool::remove_queue(int q)
{
   host_queue.erase(std::remove(host_queue.begin(), host_queue.end(), q), host_queue.end());
   return upload_launch();
}

hipError_t Pool::enqueue(int n)
{
   if (graph_launches > 0) {
      if (!graph_exec) {
         TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
         for (int i = 0; i < graph_launches; i++) {
            hipError_t e = launch_transport(sig, n_lanes, d_launch.p, precision, stream, nullptr);
            if (e != hipSuccess) {
               hipGraph_t g;
               (void)hipStreamEndCapture(stream, &g);
               return e;
            }
         }
         TRY(hipStreamEndCapture(stream, &graph));
         TRY(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
      }
      const int reps = (n + graph_launches - 1) / graph_launches;
      for (int i = 0; i < reps; i++) TRY(hipGraphLaunch(graph_exec, stream));
      launches += (int64_t)reps * graph_launches;
      return hipSuccess;
   }
   for (int i = 0; i < n; i++) TRY(launch_transport(sig, n_lanes, d_launch.p, precision, stream, nullptr));
   launches += n;
   return hipSuccess;
}

hipError_t Pool::launch(int n, float* ms_out)
{
   TRY(hipEventRecord(ev0, stream));
   if (graph_launches > 0) {
      TRY(enqueue(n));
   }
   else {
      while (tev.size() < (size_t)2 * n) {
         hipEvent_t e;
         TRY(hipEventCreate(&e));
         tev.push_back(e);
      }
      for (int i = 0; i < n; i++) {
         TRY(hipEventRecord(tev[2 * i], stream));
         TRY(launch_transport(sig, n_lanes, d_launch.p, precision, stream, tev[2 * i + 1]));
      }
      launches += n;
   }
   TRY(hipEventRecord(ev1, stream));
   TRY(hipEventSynchronize(ev1));
   float ms = 0.f;
   TRY(hipEventElapsedTime(&ms, ev0, ev1));
   if (ms_out) *ms_out = ms;
   if (graph_launches == 0) {
      for (int i = 0; i < n; i++) {
         float a = 0.f, b = 0.f;
         TRY(hipEventElapsedTime(&a, tev[2 * i], tev[2 * i + 1]));
         TRY(hipEventElapsedTime(&b, tev[2 * i + 1], i + 1 < n ? tev[2 * i + 2] : ev1));
         service_ms += a;
         step_ms += b;
      }
   }
   return hipSuccess;
}

void Pool::release()
{
   if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
   if (graph) (void)hipGraphDestroy(graph);
   graph_exec = nullptr;
   graph = nullptr;
   DevBuf<int32_t>* di[] = {&active, &slot_q, &bcross, &bstat, &attempt, &n_events, &ev_bin, &ev_bnd, &ev_pending, &ex_code, &queue};
   for (auto* b : di) b->release();
   DevBuf<double>* dv[] = {&t, &x, &y, &z, &m0, &m1, &m2, &amp, &wgt, &dt_prev, &src_old, &ds_prev, &dmax, &dt_adapt, &pf, &bdelta,
                           &bdelta_old, &t0, &x0, &y0, &z0, &m00, &m01, &m02, &ev_w, &evs};
   for (auto* b : dv) b->release();
   traj.release(); nn.release(); nu.release(); steps.release(); kc.release(); queries.release(); d_launch.release();
   for (hipEvent_t e : tev) (void)hipEventDestroy(e);
   tev.clear();
   if (ev0) (void)hipEventDestroy(ev0);
   if (ev1) (void)hipEventDestroy(ev1);
   if (stream) (void)hipStreamDestroy(stream);
   ev0 = ev1 = nullptr;
   stream = nullptr;
   n_lanes = 0;
   launches = 0;
   step_ms = service_ms = 0.0;
}

}  // namespace spx
