// spx_runtime.h -- host-side engine objects behind the C ABI (internal to the library).
//
// A Pool is a set of resident trajectory slots that runs one kernel instantiation (transport,
// background, diffusion class, time flow, precision).  A QueryRes is the device state of one run or
// one served query: its plan, control block, histograms, optional records and replay buffers.  The
// pool's query table maps a small index to each admitted query; the kernels bind every wave to one
// query at a time, so queries of the same kind share the slots and run concurrently.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "spx_kernels.h"

namespace spx {

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

// Device state of one run or served query.
struct QueryRes {
   DevBuf<spx_plan_desc> plan;
   DevBuf<RunCtl> ctl;
   DevBuf<unsigned long long> own_counts, own_events, own_acc;
   DevBuf<double> rep_normals, rep_uniforms;
   DevBuf<spx_traj_record> records;
   QueryDesc desc{};
   uint64_t n_traj = 0;
   int64_t hist_bins = 0;

   // Bytes of device memory the query holds (histograms, records, replay buffers, plan, control block).
   static size_t footprint(const spx_plan_desc& plan, int64_t hist_bins, uint64_t n_traj, const spx_rng* rng, bool keep_records);

   // Upload the plan, zero the control block (next = first) and, when `hist_counts` is null, allocate
   // and zero the query's own histograms; otherwise bin into the given (persistent) arrays.
   hipError_t create(const spx_plan_desc& plan, int64_t hist_bins, uint64_t first, uint64_t n_traj, const spx_rng* rng,
                     bool keep_records, unsigned long long* hist_counts, unsigned long long* hist_acc, unsigned long long* hist_events,
                     hipStream_t st);
   hipError_t read_ctl(RunCtl& out, hipStream_t st) const;
   // counts (n_bins) and sums (n_bins * width) of distribution d, plus its event count
   hipError_t read_distro(const spx_plan_desc& plan, int d, int64_t* counts, double* sums, int64_t* n_events, hipStream_t st) const;
   void release();
};

// Bin components (bins x width, all distributions) of a plan's exact sums; SPX_ACC_LIMBS limbs each.
int64_t acc_components(const spx_plan_desc& plan);
// Host copy of distribution d of a query's histograms: counts, sums (n_bins x width, each the exact
// sum rounded once) and the event count.
hipError_t read_hist(const spx_plan_desc& p, int d, const unsigned long long* dev_counts, const unsigned long long* dev_acc,
                     const unsigned long long* dev_events, int64_t* counts, double* sums, int64_t* n_events, hipStream_t st);

// Resident slots running one kernel instantiation.
struct Pool {
   spx_plan_desc sig{};                 // host plan whose kinds select the kernels
   int precision = SPX_PREC_FP64;
   uint64_t n_lanes = 0;
   int steps_per_launch = 2000;
   int graph_launches = 0;
   int qcap = 0;                        // size of the query table
   hipStream_t stream = nullptr;
   hipEvent_t ev0 = nullptr, ev1 = nullptr;
   std::vector<hipEvent_t> tev;         // [2i] before launch i, [2i+1] between its two kernels
   DevBuf<int32_t> active, slot_q, bcross, bstat, attempt, n_events, ev_bin, ev_bnd, ev_pending, ex_code;
   DevBuf<double> t, x, y, z, m0, m1, m2, amp, wgt, dt_prev, src_old, ds_prev, dmax, dt_adapt, pf, bdelta, bdelta_old;
   DevBuf<double> t0, x0, y0, z0, m00, m01, m02, ev_w, evs;
   DevBuf<uint64_t> traj;
   DevBuf<uint32_t> nn, nu;
   DevBuf<float> kc;
   DevBuf<int64_t> steps;
   DevBuf<QueryDesc> queries;
   DevBuf<int32_t> queue;
   DevBuf<Launch> d_launch;
   Launch host{};
   std::vector<int32_t> host_queue;     // admitted query indices, oldest first
   std::vector<char> qused;             // query-table entries in use
   hipGraph_t graph = nullptr;
   hipGraphExec_t graph_exec = nullptr;
   int64_t launches = 0;
   double step_ms = 0.0, service_ms = 0.0;   // accumulated by launch() when not replaying a graph

   hipError_t init(const spx_plan_desc& sig, int precision, uint64_t lanes, int steps_per_launch, int graph_launches, int qcap);
   bool matches(const spx_plan_desc& p, int prec) const;
   // idle every slot and forget the wave-to-query binding
   hipError_t reset_slots();
   int alloc_index();                   // free query-table entry, or -1
   void free_index(int q);
   hipError_t set_query(int q, const QueryDesc& d);
   hipError_t push_queue(int q);
   hipError_t remove_queue(int q);
   // enqueue n engine launches on the pool's stream (graph replays when graph_launches > 0)
   hipError_t enqueue(int n);
   // enqueue, wait, and account kernel time (ms of the call, step / service kernel split)
   hipError_t launch(int n, float* ms);
   void release();

private:
   hipError_t upload_launch();
};

}  // namespace spx
