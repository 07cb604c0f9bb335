// spx_engine.h -- device pieces shared by the transports' service and step kernels.
//
// Slot states, wave-level work claiming, random streams, boundary evaluation and handling,
// distribution binning and trajectory records.  Included by the kernel translation units only.
#pragma once

#include <hip/hip_runtime.h>

#include "spx_physics.h"
#include "spx_kernels.h"

namespace spx {

typedef const SPX_CP Launch* LaunchPtr;
typedef const SPX_CP QueryDesc* QueryPtr;

// The query of a wave (uniform by construction) and its plan, as constant-address-space pointers.
__device__ __forceinline__ QueryPtr query_of(LaunchPtr L, int q) { return (QueryPtr)L->queries + q; }
__device__ __forceinline__ PlanPtr plan_of(QueryPtr Q) { return (PlanPtr)Q->plan; }

// Member-wise copy of a query's random-stream description out of the constant query table.
__device__ __forceinline__ spx_rng_desc load_rng(QueryPtr Q)
{
   spx_rng_desc r;
   r.mode = Q->rng.mode;
   r.pad_ = 0;
   r.seed = Q->rng.seed;
   r.normals = Q->rng.normals;
   r.uniforms = Q->rng.uniforms;
   r.normal_stride = Q->rng.normal_stride;
   r.uniform_stride = Q->rng.uniform_stride;
   return r;
}

// Slot states
enum { SLOT_IDLE = 0, SLOT_RUN = 1, SLOT_START = 2, SLOT_EVENT = 3, SLOT_DONE = 4, SLOT_DISCARD = 5 };

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------------------------
// Wave helpers (64 lanes)
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }

// Hand out trajectory ids to the requesting lanes of this wave with a single atomic.
__device__ __forceinline__ uint64_t wave_claim(RunCtl* ctl, bool need)
{
   const uint64_t mask = __ballot(need);
   if (mask == 0) return ~0ull;
   const int leader = __ffsll((unsigned long long)mask) - 1;
   unsigned long long base = 0;
   if (lane_id() == leader) base = atomicAdd(&ctl->next, (unsigned long long)__popcll(mask));
   base = __shfl(base, leader);
   return need ? base + (uint64_t)__popcll(mask & lanes_below()) : ~0ull;
}

__device__ __forceinline__ void atomic_min_pos(unsigned long long* a, double v)
{
   if (v >= 0.0) atomicMin(a, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void atomic_max_pos(unsigned long long* a, double v)
{
   if (v >= 0.0) atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------------------------
// Random draws: counter streams (Philox4x32-10 + Box-Muller pairs) or replayed buffers
// ---------------------------------------------------------------------------------------------
template <typename R>
struct NormalStream {
   uint64_t traj;
   uint32_t nn;          // normal draws consumed
   uint32_t ns;          // spares: draws nn .. nn+ns-1 in s0, s1, s2, always the tail of a Philox block
   R s0, s1, s2;
   bool exhausted;
};

// The stream: FP64 normal k is the cosine (even k) or sine (odd k) branch of Philox block k/2
// (normal_pair_f64); FP32 normal k is word k%4 of block k/4 (normal_quad_f32).  The spares left by a
// draw are valid within one launch; a launch that loads a slot whose stream stops inside a block
// regenerates that block's tail (prime_normals), so the same draws come out whatever the launch shape.
template <typename R>
__device__ __forceinline__ void prime_normals(const spx_rng_desc& rng, NormalStream<R>& S)
{
   S.ns = 0;
   S.s0 = S.s1 = S.s2 = R(0);
   if (rng.mode == 1) return;
   if constexpr (is_f32<R>()) {
      const uint32_t j = S.nn & 3u;
      if (j != 0u) {
         float n0, n1, n2, n3;
         normal_quad_f32(rng.seed, S.traj, S.nn >> 2, n0, n1, n2, n3);
         S.s0 = j == 1u ? n1 : (j == 2u ? n2 : n3);
         S.s1 = j == 1u ? n2 : n3;
         S.s2 = n3;
         S.ns = 4u - j;
      }
   }
   else if (S.nn & 1u) {
      double c, s;
      normal_pair_f64(rng.seed, S.traj, S.nn >> 1, c, s);
      S.s0 = s;
      S.ns = 1u;
   }
}

// The next N (1..3) normal draws of the stream, in order.  A lane generates at most one Philox block
// per call (FP32: four normals) or two (FP64: two each), at one place in the code: the lanes of a wave
// sit at different offsets in their blocks, and a generator inlined once per draw would run, masked,
// once per draw for the whole wave.  New normals are appended behind the spares (at most two of them
// when a block is needed, so six places suffice), and the spares left over end at a block boundary.
template <typename R, int N>
__device__ __forceinline__ void next_normals(const spx_rng_desc& rng, uint64_t first, NormalStream<R>& S, R (&out)[N])
{
   static_assert(N >= 1 && N <= 3, "one to three draws per call");
   if (rng.mode == 1) {
      const uint64_t base = (S.traj - first) * (uint64_t)rng.normal_stride;
#pragma unroll
      for (int i = 0; i < N; i++) {
         const uint32_t k = S.nn + (uint32_t)i;
         if ((int64_t)k >= rng.normal_stride) { S.exhausted = true; out[i] = R(0); }
         else out[i] = R(rng.normals[base + k]);
      }
      S.nn += N;
      return;
   }
   uint32_t ns = S.ns;
   R b0 = S.s0, b1 = S.s1, b2 = S.s2, b3 = R(0), b4 = R(0), b5 = R(0);
   if constexpr (is_f32<R>()) {
      if (ns < (uint32_t)N) {   // ns is 0, 1 or 2 here
         float w0, w1, w2, w3;
         normal_quad_f32(rng.seed, S.traj, (S.nn + ns) >> 2, w0, w1, w2, w3);
         b5 = w3;
         b4 = ns == 1u ? w3 : w2;
         b3 = ns == 0u ? w3 : (ns == 1u ? w2 : w1);
         b2 = ns == 0u ? w2 : (ns == 1u ? w1 : w0);
         b1 = ns == 0u ? w1 : (ns == 1u ? w0 : b1);
         b0 = ns == 0u ? w0 : b0;
         ns += 4u;
      }
   }
   else {
      // ns is 0 or 1 on entry: the first pair is appended at ns, a second one (N = 3 from ns = 0) at 2
      if (ns < (uint32_t)N) {
         double c, s;
         normal_pair_f64(rng.seed, S.traj, (S.nn + ns) >> 1, c, s);
         b2 = ns == 1u ? s : b2;
         b1 = ns == 1u ? c : s;
         b0 = ns == 1u ? b0 : c;
         ns += 2u;
      }
      if (N == 3 && ns < 3u) {
         double c, s;
         normal_pair_f64(rng.seed, S.traj, (S.nn + ns) >> 1, c, s);
         b2 = c;
         b3 = s;
         ns += 2u;
      }
   }
   out[0] = b0;
   if (N > 1) out[1] = b1;
   if (N > 2) out[2] = b2;
   S.s0 = N == 1 ? b1 : (N == 2 ? b2 : b3);
   S.s1 = N == 1 ? b2 : (N == 2 ? b3 : b4);
   S.s2 = N == 1 ? b3 : (N == 2 ? b4 : b5);
   S.ns = ns - (uint32_t)N;
   S.nn += N;
}

__device__ __forceinline__ double next_uniform(const spx_rng_desc& rng, uint64_t first, uint64_t traj, uint32_t& nu, bool& exhausted)
{
   const uint32_t k = nu++;
   if (rng.mode == 1) {
      if ((int64_t)k >= rng.uniform_stride) { exhausted = true; return 0.5; }
      return rng.uniforms[(traj - first) * (uint64_t)rng.uniform_stride + k];
   }
   return counter_uniform(rng.seed, traj, k);
}

// Initial position: InitialSpaceFixed, InitialSpaceSphere and InitialSpaceSphereSector::EvaluateInitial
// (src/initial_space.cc:49-66, :308-316, :384-392).  The polar cosine is uniform between cos(theta1) and
// cos(theta2) (the sphere: -1 + 2 u), then the azimuth between phi1 and phi2 (the sphere: 2 pi u).
__device__ __forceinline__ void pos_sample(PlanPtr P, const spx_rng_desc& rng, uint64_t first, uint64_t traj, uint32_t& nu, bool& ex,
                                           double& x, double& y, double& z)
{
   x = P->is_pos[0]; y = P->is_pos[1]; z = P->is_pos[2];
   if (P->is_kind == SPX_IS_FIXED) return;
   const bool sphere = P->is_kind == SPX_IS_SPHERE;
   const double u1 = next_uniform(rng, first, traj, nu, ex);
   const double ct = sphere ? -1.0 + 2.0 * u1 : P->is_ct1 + (P->is_ct2 - P->is_ct1) * u1;
   const double st = sqrt(1.0 - ct * ct);
   const double u2 = next_uniform(rng, first, traj, nu, ex);
   const double phi = sphere ? 6.283185307179586 * u2 : P->is_phi1 + (P->is_phi2 - P->is_phi1) * u2;
   double sn, cs;
   sincos(phi, &sn, &cs);
   x = x + P->is_radius * st * cs;
   y = y + P->is_radius * st * sn;
   z = z + P->is_radius * ct;
}

// ---------------------------------------------------------------------------------------------
// Boundaries (src/boundary_base.cc:105-191, src/boundary_{time,space,momentum}.cc)
// ---------------------------------------------------------------------------------------------
// |x - origin| of a sphere boundary in the step type (FP32 from the FP64 offset): the distance the sphere's
// delta subtracts the radius from
template <typename R>
__device__ __forceinline__ R sphere_dist(const SPX_CP spx_bnd_desc& b, double x, double y, double z)
{
   if constexpr (is_f32<R>()) {
      const v3<float> d = mk<float>((float)(x - b.origin[0]), (float)(y - b.origin[1]), (float)(z - b.origin[2]));
      return __builtin_sqrtf(norm2(d));   // one v_sqrt_f32 in the FP32 unit (approximate functions allowed); 0 at the origin
   }
   else return norm(mk<double>(x - b.origin[0], y - b.origin[1], z - b.origin[2]));
}

// With R = float the distance to a sphere or plane is formed in FP32 from the FP64 offset (the FP32
// step's crossing test); time and momentum boundaries and every normal stay in FP64.
template <typename R = double>
__device__ __forceinline__ double bnd_eval(const SPX_CP spx_bnd_desc& b, int cross_left, bool gc, double t, double x, double y, double z,
                                           double m0, double m1, double m2, double* nrm)
{
   if constexpr (is_f32<R>()) {
      if (!nrm && b.kind == SPX_BK_SPHERE) return (double)(sphere_dist<float>(b, x, y, z) - (float)b.radius);
      if (!nrm && b.kind == SPX_BK_PLANE) {
         const v3<float> d = mk<float>((float)(x - b.origin[0]), (float)(y - b.origin[1]), (float)(z - b.origin[2]));
         return (double)dot(d, mk<float>((float)b.norm[0], (float)b.norm[1], (float)b.norm[2]));
      }
   }
   switch (b.kind) {
   case SPX_BK_TIME:
      if (b.status0 & SPX_BND_RECURRENT) return t - (double)(b.max_crossings - cross_left + 1) * b.timemark;  // src/boundary_time.cc:205-208
      return t - b.timemark;                                                                                   // :70-73
   case SPX_BK_SPHERE: {                                                                                        // src/boundary_space.cc:395-411
      const v3<double> d = mk<double>(x - b.origin[0], y - b.origin[1], z - b.origin[2]);
      if (nrm) { const v3<double> u = unit(d); nrm[0] = u.x; nrm[1] = u.y; nrm[2] = u.z; }
      return norm(d) - b.radius;
   }
   case SPX_BK_PLANE: {                                                                                         // src/boundary_space.cc:71-84
      if (nrm) { nrm[0] = b.norm[0]; nrm[1] = b.norm[1]; nrm[2] = b.norm[2]; }
      const v3<double> d = mk<double>(x - b.origin[0], y - b.origin[1], z - b.origin[2]);
      return dot(d, mk<double>(b.norm[0], b.norm[1], b.norm[2]));
   }
   default:                                                                                                     // src/boundary_momentum.cc:69-76
      if (gc) return norm(mk<double>(m0, m1, m2)) - b.momentum;
      return m0 - b.momentum;
   }
}

// Does boundary b carry an action >= 0 for any distribution?
__device__ __forceinline__ bool bnd_has_action(PlanPtr P, int b)
{
   const SPX_CP spx_bnd_desc& B = P->bnd[b];
   bool any = false;
   for (int d = 0; d < P->n_dist; d++) any = any || (d < B.n_act && B.actions[d] >= 0);
   return any;
}

// Pending-event buffer: 3 entries per slot (momentum, space, time classes), 9 doubles each.
enum { EV_T = 0, EV_X, EV_Y, EV_Z, EV_M0, EV_M1, EV_M2, EV_AMP, EV_WGT, EV_N };
__device__ __forceinline__ double& ev_field(LaunchPtr L, int cls, int f, uint64_t slot)
{
   return L->evs[((uint64_t)cls * EV_N + f) * L->n_lanes + slot];
}

// ---------------------------------------------------------------------------------------------
// Distribution events (src/distribution_base.cc:174-194, src/distribution_templated.cc:145-180)
// Called by every lane of a wave together; `has` marks the lanes with an event at boundary `bnd`.
// Each lane forms the value and weight of every distribution and its bin; the wave then merges the
// lanes that hit the same bin and issues one count atomic per bin; the weights go to the bin's exact
// fixed-point sum (spx_acc.h), digits of lanes with the same bin and limb added as integers first.
// Counts and sums are therefore independent of event order; also the per-slot last-event record.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ long long wave_sum_i64(long long v)
{
#pragma unroll
   for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
   return v;
}
__device__ __forceinline__ int wave_or(int v)
{
#pragma unroll
   for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off);
   return v;
}

static __device__ int bin_events(PlanPtr P, QueryPtr Q, LaunchPtr L, uint64_t slot, bool has, int bnd,
                                 double t2, double x2, double y2, double z2, double m20, double m21, double m22, double amp, double logwgt)
{
   const bool parker = (P->traj_type < 2 || P->traj_type == SPX_TRAJ_FOCUSED);   // momentum (p, ., .)
   const double t1 = L->t0[slot], x1 = L->x0[slot], y1 = L->y0[slot], z1 = L->z0[slot];
   const double m10 = L->m00[slot], m11 = L->m01[slot], m12 = L->m02[slot];
   int n_ev = 0;
   for (int d = 0; d < P->n_dist; d++) {
      const SPX_CP spx_ds_desc& D = P->dist[d];
      int action = -1;
      if (has) {
         const SPX_CP spx_bnd_desc& B = P->bnd[bnd];
         action = d < B.n_act ? B.actions[d] : -1;
      }
      const bool ev = action >= 0;
      n_ev += ev ? 1 : 0;
      bool keep = ev;
      double w[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      int lin = 0;
      if (ev) {
         const bool first = D.val_time == 0;
         const double px = first ? x1 : x2, py = first ? y1 : y2, pz = first ? z1 : z2;
         const double ma = first ? m10 : m20, mb = first ? m11 : m21, mc = first ? m12 : m22;
         double v0 = 0.0, v1 = 0.0, v2 = 0.0;
         switch (D.kind) {
         case SPX_DS_SPECTRUM: v0 = EnrKin<double>(parker ? m10 : norm(mk<double>(m10, m11, m12))); break;   // src/distribution_other.cc:561-570
         case SPX_DS_TIME: v0 = first ? t1 : t2; break;                                                      // :143-148
         case SPX_DS_MOMENTUM:                                                                                // :262-290
            if (D.val_coord == 0 || parker) { v0 = ma; v1 = mb; v2 = mc; }
            else if (P->traj_type == SPX_TRAJ_FIELDLINE) { v0 = mc; v1 = 0.0; v2 = 0.0; }
            else { v0 = norm(mk<double>(ma, mb, mc)); v1 = mc / v0; v2 = 0.0; }
            break;
         case SPX_DS_POSMOM:                                                                                  // :342-353
            v0 = D.pos_idx == 3 ? norm(mk<double>(px, py, pz)) : (D.pos_idx == 0 ? px : (D.pos_idx == 1 ? py : pz));
            v1 = D.mom_idx == 0 ? ma : (D.mom_idx == 1 ? mb : mc);
            break;
         case SPX_DS_POSPITCH:                                                                                // (|x|, p_par/|p|)
            v0 = norm(mk<double>(px, py, pz));
            v1 = mc / norm(mk<double>(ma, mb, mc));
            break;
         default: v0 = t2; break;                                                                             // :730-733, :777-780
         }
         // EvaluateWeight through the action table (w and keep = true are declared above)
         switch (D.kind) {
         case SPX_DS_SPECTRUM:                                                                                // :575-594
            if (action == 0) {
               const double m2 = parker ? m20 : norm(mk<double>(m20, m21, m22));
               const double T = EnrKin<double>(m2);
               if (P->spectrum_type == 0) w[0] = D.J0 * VelP<double>(m2) * pow(T / D.T0, D.pow_law) / sqr(m2);
               else if (P->spectrum_type == 1) w[0] = D.J0 * pow(T / D.T0, D.pow_law) / sqr(m2);
               else w[0] = D.J0 * pow(T / D.T0, D.pow_law);
            }
            else if (action == 1) w[0] = D.val_cold;
            else keep = false;
            break;
         case SPX_DS_CUMO1:                                                                                   // :738-741
            if (action != 0) { keep = false; break; }
            w[0] = x2 - x1; w[1] = y2 - y1; w[2] = z2 - z1;
            break;
         case SPX_DS_CUMO2: {                                                                                 // :785-788
            if (action != 0) { keep = false; break; }
            const double dx = x2 - x1, dy = y2 - y1, dz = z2 - z1;
            w[0] = dx * dx; w[1] = dx * dy; w[2] = dx * dz; w[4] = dy * dy; w[5] = dy * dz; w[8] = dz * dz;
            w[3] = w[1]; w[6] = w[2]; w[7] = w[5];
            break;
         }
         default:                                                                                             // :79-93
            if (action == 0) w[0] = D.val_hot;
            else if (action == 1) w[0] = D.val_cold;
            else keep = false;
         }
         // AddEvent: bin index per active dimension
         int bin[3] = {0, 0, 0};
#pragma unroll
         for (int ijk = 0; ijk < 3; ijk++) {
            if (!keep || !(D.dims & (1 << ijk))) continue;
            const double val = ijk == 0 ? v0 : (ijk == 1 ? v1 : v2);
            const double v = (D.log_bins[ijk] ? log10(val) : val) - D.limits0[ijk];
            if (v < 0.0 && !D.bin_outside[ijk]) { keep = false; continue; }
            int b = (int)(v / D.bin_size[ijk]);
            if (b < 0) b = 0;   // the reference can index below zero here (src/distribution_templated.cc:162-167)
            if (b >= D.n_bins[ijk]) {
               if (D.bin_outside[ijk]) b = D.n_bins[ijk] - 1;
               else keep = false;
            }
            bin[ijk] = b;
         }
         if (keep) {
            lin = D.n_bins[2] * (D.n_bins[1] * bin[0] + bin[1]) + bin[2];
            const double ew = exp(logwgt);
#pragma unroll
            for (int k = 0; k < 9; k++) w[k] = w[k] * ew + amp;
            L->ev_bin[(uint64_t)d * L->n_lanes + slot] = lin;
            L->ev_w[(uint64_t)d * L->n_lanes + slot] = w[0];
         }
         else L->ev_bin[(uint64_t)d * L->n_lanes + slot] = -1;
      }
      // one count atomic per distinct bin of the wave
      const uint64_t kept = __ballot(keep);
      if (kept == 0) continue;
      if (lane_id() == __ffsll((unsigned long long)kept) - 1) atomicAdd(&Q->hist_events[d], (unsigned long long)__popcll(kept));
      uint64_t todo = kept;
      while (todo) {
         const int leader = __ffsll((unsigned long long)todo) - 1;
         const int key = __shfl(lin, leader);
         const uint64_t grp = __ballot(keep && lin == key);
         if (lane_id() == leader) atomicAdd(&Q->hist_counts[Q->hist_off[d] + key], (unsigned long long)__popcll(grp));
         todo &= ~grp;
      }
      // exact sums: lanes with the same bin and the same lowest limb add their digits as integers
      // (associative), then one atomic per touched limb
#pragma unroll
      for (int k = 0; k < 9; k++) {
         if (k >= D.width) continue;
         const AccParts a = acc_split(keep ? w[k] : 0.0);
         const bool act = keep && (a.special || a.c0 || a.c1 || a.c2);
         todo = __ballot(act);
         while (todo) {
            const int leader = __ffsll((unsigned long long)todo) - 1;
            const int kb = __shfl(lin, leader), ki = __shfl(a.idx, leader);
            const bool mine = act && lin == kb && a.idx == ki;
            const uint64_t grp = __ballot(mine);
            const long long s0 = wave_sum_i64(mine ? a.c0 : 0ll), s1 = wave_sum_i64(mine ? a.c1 : 0ll), s2 = wave_sum_i64(mine ? a.c2 : 0ll);
            const int sp = wave_or(mine ? a.special : 0);
            if (lane_id() == leader) {
               unsigned long long* limb = Q->hist_acc + (uint64_t)(Q->acc_off[d] + (int64_t)kb * D.width + k) * SPX_ACC_LIMBS;
               if (s0) atomicAdd(limb + ki, (unsigned long long)s0);
               if (s1) atomicAdd(limb + ki + 1, (unsigned long long)s1);
               if (s2) atomicAdd(limb + ki + 2, (unsigned long long)s2);
               if (sp) atomicOr(limb + SPX_ACC_FLAGS, (unsigned long long)sp);
            }
            todo &= ~grp;
         }
      }
   }
   return n_ev;
}

// Record a finished (or failed) trajectory: run statistics and, when asked, its endpoint.
static __device__ void finish_record(QueryPtr Q, LaunchPtr Lp, uint64_t slot, uint64_t traj, int ex, int attempts, int64_t steps,
                                     int n_events, double t, double x, double y, double z, double m0, double m1, double m2, double amp,
                                     double wgt)
{
   const SPX_CP Launch& L = *Lp;
   if (ex == 0) {
      const double el = fabs(t - L.t0[slot]);
      atomic_min_pos(&Q->ctl->tmin_bits, el);
      atomic_max_pos(&Q->ctl->tmax_bits, el);
      atomicAdd(&Q->ctl->steps_done, (unsigned long long)steps);
   }
   else atomicAdd(&Q->ctl->failed, 1ull);
   atomicAdd(&Q->ctl->done, 1ull);
   if (Q->records && traj >= Q->rec_first && traj < Q->rec_first + Q->rec_n) {
      spx_traj_record& r = Q->records[traj - Q->rec_first];
      r.t = t; r.pos[0] = x; r.pos[1] = y; r.pos[2] = z; r.mom[0] = m0; r.mom[1] = m1; r.mom[2] = m2;
      r.amp = amp; r.wgt = wgt;
      r.t0 = L.t0[slot]; r.pos0[0] = L.x0[slot]; r.pos0[1] = L.y0[slot]; r.pos0[2] = L.z0[slot];
      r.mom0[0] = L.m00[slot]; r.mom0[1] = L.m01[slot]; r.mom0[2] = L.m02[slot];
      r.steps = steps;
      r.status = ex == 0 ? 0 : (1 | (ex << 8));
      r.attempts = attempts;
      r.n_events = n_events;
      for (int d = 0; d < 4; d++) {
         const bool has = ex == 0 && d < Q->n_dist && L.ev_bin[(uint64_t)d * L.n_lanes + slot] >= 0;
         r.ev_bin[d] = has ? L.ev_bin[(uint64_t)d * L.n_lanes + slot] : -1;
         r.ev_weight[d] = has ? L.ev_w[(uint64_t)d * L.n_lanes + slot] : 0.0;
      }
   }
}

// ---------------------------------------------------------------------------------------------
// Step-kernel helpers
// ---------------------------------------------------------------------------------------------
// Per-lane state that is rarely touched lives in LDS for the duration of a launch.
struct ColdLDS {
   double amp[BLOCK], wgt[BLOCK], dt_adapt[BLOCK], dt_prev[BLOCK], src_old[BLOCK];
   int32_t cross[SPX_MAX_BND][BLOCK];
   float kc[SPX_KC][BLOCK];             // FP32 step: compensation terms of the FP64 sums (kahan_add)
};


// TrajectoryBase::TimeBoundaryProximityCheck (src/trajectory_base.cc:151-179)
template <int FWD, typename HOT>
__device__ __forceinline__ double time_proximity(PlanPtr P, const HOT& H, double dt, double dmax)
{
   typedef typename HOT::real R;
   double dn = FWD ? -divc_d<R>(sp_large * dmax, c_code) : divc_d<R>(sp_large * dmax, c_code);
   // time boundaries fill the first n_bt slots; one is the common case, and the loop's guards would
   // otherwise become selects over every slot
   const int nbt = P->n_bt;
   if (nbt == 1) {
      const double d = H.bd[0];
      if (FWD) { if ((d <= 0.0) && (d > dn)) dn = d; }
      else { if ((d >= 0.0) && (d < dn)) dn = d; }
   }
   else if (nbt > 1) {
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         if (b < nbt) {
            const double d = H.bd[b];
            if (FWD) { if ((d <= 0.0) && (d > dn)) dn = d; }
            else { if ((d >= 0.0) && (d < dn)) dn = d; }
         }
      }
   }
   if (FWD) { if (dt >= -dn) dt = fmax(-(1.0 + sp_little) * dn, divc_d<R>(sp_small * dmax, c_code)); }
   else { if (dt >= dn) dt = fmax((1.0 + sp_little) * dn, divc_d<R>(sp_small * dmax, c_code)); }
   return dt;
}

// TrajectoryBase::HandleBoundaries (src/trajectory_base.cc:413-535) + UpdateAllBoundaries.  Events are
// queued in the slot's event buffer for the service kernel.  Returns a bit set: 1 = TRAJ_FINISH,
// 2 = an event was queued.  HOT is the transport's register state: t, x, y, z, the boundary arrays
// bd / bdo / bst and the momentum accessors m0() / m1() / m2() (Parker keeps (p, 0, 0), the guiding centre
// (p_perp, 0, p_par), the field line (0, 0, p_par), the full orbit (px, py, pz)) and reflect_mom() (a reflecting wall reverses p_par of a guiding centre,
// src/trajectory_guiding.hh:121-124).
template <bool GC, typename HOT>
__device__ __forceinline__ int handle_boundaries(PlanPtr P, LaunchPtr L, uint64_t slot,
                                                 HOT& H, ColdLDS& C, int tid)
{
   typedef typename HOT::real R;
   bool any = false;
   R sdist = R(0);   // |x - origin| of the last sphere, reused by a concentric sphere in the next slot
#pragma unroll
   for (int b = 0; b < SPX_MAX_BND; b++) {
      if (b < P->n_bnd) {
         // ComputeBoundary (src/boundary_base.cc:105-127); a slot's recurrent bit is its boundary's (status0)
         const SPX_CP spx_bnd_desc& B = P->bnd[b];
         double d;
         if (B.kind == SPX_BK_SPHERE) {
            if (!B.same_origin) sdist = sphere_dist<R>(B, H.x, H.y, H.z);
            if constexpr (is_f32<R>()) d = (double)(sdist - (float)B.radius);
            else d = sdist - B.radius;
         }
         else {
            const int left = (B.status0 & SPX_BND_RECURRENT) ? C.cross[b][tid] : 0;
            d = bnd_eval<R>(B, left, GC, H.t, H.x, H.y, H.z, H.m0(), H.m1(), H.m2(), nullptr);
         }
         const double prod = d * H.bdo[b];
         H.bd[b] = (prod == 0.0) ? -H.bdo[b] : d;
         H.bst[b] = (prod <= 0.0) ? (H.bst[b] | SPX_BND_CROSSED) : (H.bst[b] & ~SPX_BND_CROSSED);
         any = any || (prod <= 0.0);
      }
   }
   if (!any) {
      // RecordBoundary without crossings: _delta_old = _delta
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) H.bdo[b] = H.bd[b];
      return 0;
   }
   int out = 0;
   // momentum, then space, then time
   for (int cls = 2; cls >= 0; cls--) {
      int active = -1;
#pragma unroll
      for (int b = 0; b < SPX_MAX_BND; b++) {
         if (b >= P->n_bnd || P->bnd[b].cls != cls || !(H.bst[b] & SPX_BND_CROSSED)) continue;
         active = b;
         if (cls == SPX_BC_SPACE && (H.bst[b] & SPX_BND_REFLECT)) {
            double nrm[3];
            (void)bnd_eval(P->bnd[b], 0, GC, H.t, H.x, H.y, H.z, H.m0(), H.m1(), H.m2(), nrm);
            H.x -= 2.0 * H.bd[b] * nrm[0];
            H.y -= 2.0 * H.bd[b] * nrm[1];
            H.z -= 2.0 * H.bd[b] * nrm[2];
            H.reflect_mom();
            const double d = bnd_eval(P->bnd[b], 0, GC, H.t, H.x, H.y, H.z, H.m0(), H.m1(), H.m2(), nullptr);
            const double prod = d * H.bdo[b];
            H.bd[b] = (prod == 0.0) ? -H.bdo[b] : d;
            H.bst[b] = (prod <= 0.0) ? (H.bst[b] | SPX_BND_CROSSED) : (H.bst[b] & ~SPX_BND_CROSSED);
            C.cross[b][tid] -= 1;
         }
         else if (cls != SPX_BC_MOMENTUM || !(H.bst[b] & SPX_BND_REFLECT)) {
            if (H.bst[b] & SPX_BND_TERMINAL) out |= 1;
         }
      }
      if (active >= 0 && bnd_has_action(P, active)) {
         ev_field(L, cls, EV_T, slot) = H.t; ev_field(L, cls, EV_X, slot) = H.x; ev_field(L, cls, EV_Y, slot) = H.y;
         ev_field(L, cls, EV_Z, slot) = H.z; ev_field(L, cls, EV_M0, slot) = H.m0(); ev_field(L, cls, EV_M1, slot) = H.m1();
         ev_field(L, cls, EV_M2, slot) = H.m2(); ev_field(L, cls, EV_AMP, slot) = C.amp[tid]; ev_field(L, cls, EV_WGT, slot) = C.wgt[tid];
         L->ev_bnd[cls * L->n_lanes + slot] = active;
         L->ev_pending[slot] |= (1 << cls);
         out |= 2;
      }
   }
   // RecordBoundary (src/boundary_base.cc:168-191)
#pragma unroll
   for (int b = 0; b < SPX_MAX_BND; b++) {
      if (b < P->n_bnd) {
         const bool crossed = (H.bst[b] & SPX_BND_CROSSED) != 0;
         const bool recur = crossed && (H.bst[b] & SPX_BND_RECURRENT) != 0;
         const int left = C.cross[b][tid] - (crossed ? 1 : 0);
         C.cross[b][tid] = left;
         const double d_now = H.bd[b];
         const double d_rec = recur ? bnd_eval(P->bnd[b], left, GC, H.t, H.x, H.y, H.z, H.m0(), H.m1(), H.m2(), nullptr) : d_now;
         H.bdo[b] = recur ? H.bdo[b] : d_now;
         H.bd[b] = d_rec;
         const int s = H.bst[b] & ~(SPX_BND_CROSSED | SPX_BND_TERMINAL);
         H.bst[b] = (left == 1) ? (s | SPX_BND_TERMINAL) : s;
      }
   }
   return out;
}

// ---------------------------------------------------------------------------------------------
// Service pass, one thread per slot: bin the events the step kernel queued, finish or retry
// trajectories, refill idle slots from their query's work counter (moving a fully idle wave to the
// oldest admitted query with trajectories left) and run SetStart through `set_start`, a functor
// (plan, query, rng, slot, traj, nu&) -> exception code of the transport.
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long load_next(const RunCtl* ctl)
{
   return __hip_atomic_load(&ctl->next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename StartFn>
__device__ __forceinline__ void service_slot(LaunchPtr L, StartFn&& set_start)
{
   const uint64_t slot = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
   int st = L->active[slot];
   int q = __builtin_amdgcn_readfirstlane(L->slot_q[slot]);
   if (q >= 0) {
      const QueryPtr Q = query_of(L, q);
      const PlanPtr P = plan_of(Q);
      // 1. pending distribution events of the last step kernel (binned by the whole wave together)
      const bool evl = (st == SLOT_EVENT || st == SLOT_DONE);
      const int pend = evl ? L->ev_pending[slot] : 0;
      int n_ev = evl ? L->n_events[slot] : 0;
      for (int cls = 2; cls >= 0; cls--) {
         const bool has = (pend >> cls) & 1;
         if (__ballot(has) == 0) continue;
         const int b = has ? L->ev_bnd[cls * L->n_lanes + slot] : 0;
         const double f[EV_N] = {has ? ev_field(L, cls, EV_T, slot) : 0.0, has ? ev_field(L, cls, EV_X, slot) : 0.0,
                                 has ? ev_field(L, cls, EV_Y, slot) : 0.0, has ? ev_field(L, cls, EV_Z, slot) : 0.0,
                                 has ? ev_field(L, cls, EV_M0, slot) : 0.0, has ? ev_field(L, cls, EV_M1, slot) : 0.0,
                                 has ? ev_field(L, cls, EV_M2, slot) : 0.0, has ? ev_field(L, cls, EV_AMP, slot) : 0.0,
                                 has ? ev_field(L, cls, EV_WGT, slot) : 0.0};
         n_ev += bin_events(P, Q, L, slot, has, b, f[EV_T], f[EV_X], f[EV_Y], f[EV_Z], f[EV_M0], f[EV_M1], f[EV_M2], f[EV_AMP], f[EV_WGT]);
      }
      if (evl) {
         L->ev_pending[slot] = 0;
         L->n_events[slot] = n_ev;
         if (st == SLOT_DONE) {
            finish_record(Q, L, slot, L->traj[slot], 0, L->attempt[slot] + 1, L->steps[slot], n_ev, L->t[slot], L->x[slot], L->y[slot],
                          L->z[slot], L->m0[slot], L->m1[slot], L->m2[slot], L->amp[slot], L->wgt[slot]);
            st = SLOT_IDLE;
         }
         else st = SLOT_RUN;
      }
      // 2. discarded attempts: retry with continuing draws (a process-wide stream in the reference), or give up
      if (st == SLOT_DISCARD) {
         const int exc = L->ex_code[slot];
         atomicAdd(&Q->ctl->discards[exc], 1ull);
         L->ev_pending[slot] = 0;
         const int att = L->attempt[slot] + 1;
         L->attempt[slot] = att;
         if (att < P->max_attempts && exc != 5) st = SLOT_START;
         else {
            finish_record(Q, L, slot, L->traj[slot], exc, att, L->steps[slot], L->n_events[slot], L->t[slot], L->x[slot], L->y[slot],
                          L->z[slot], L->m0[slot], L->m1[slot], L->m2[slot], L->amp[slot], L->wgt[slot]);
            st = SLOT_IDLE;
         }
      }
   }
   // 3. refill idle slots from the wave's query
   if (q >= 0) {
      const bool need = st == SLOT_IDLE;
      if (__ballot(need) != 0) {
         const QueryPtr Q = query_of(L, q);
         const uint64_t id = wave_claim(Q->ctl, need);
         if (need && id < Q->end) {
            L->traj[slot] = id;
            L->nu[slot] = 0;
            L->nn[slot] = 0;
            L->attempt[slot] = 0;
            st = SLOT_START;
         }
      }
   }
   // a wave with nothing running and nothing left in its query moves to the oldest admitted query with work
   if (__ballot(st != SLOT_IDLE) == 0) {
      int nq = -1;
      if (lane_id() == 0) {
         for (int i = 0; i < L->n_queue; i++) {
            const int c = L->queue[i];
            const QueryPtr C = query_of(L, c);
            if (load_next(C->ctl) < C->end) { nq = c; break; }
         }
      }
      q = __builtin_amdgcn_readfirstlane(__shfl(nq, 0));
      if (q >= 0) {
         const QueryPtr Q = query_of(L, q);
         const uint64_t id = wave_claim(Q->ctl, true);
         if (id < Q->end) {
            L->traj[slot] = id;
            L->nu[slot] = 0;
            L->nn[slot] = 0;
            L->attempt[slot] = 0;
            st = SLOT_START;
         }
      }
   }
   L->slot_q[slot] = q;
   // 4. SetStart, repeated with continuing draws when an attempt is discarded
   if (st == SLOT_START) {
      const QueryPtr Q = query_of(L, q);
      const PlanPtr P = plan_of(Q);
      const spx_rng_desc rng = load_rng(Q);
      const uint64_t traj = L->traj[slot];
      uint32_t nu = L->nu[slot];
      int att = L->attempt[slot];
      int exc = 1;
      while (att < P->max_attempts) {
         exc = set_start(P, Q, rng, slot, traj, nu);
         if (exc == 0) break;
         atomicAdd(&Q->ctl->discards[exc], 1ull);
         att++;
         if (exc == 5) break;
      }
      L->nu[slot] = nu;
      L->attempt[slot] = att;
      if (exc == 0) {
         st = SLOT_RUN;
#pragma unroll
         for (int k = 0; k < SPX_KC; k++) L->kc[k * L->n_lanes + slot] = 0.0f;
      }
      else {
         finish_record(Q, L, slot, traj, exc, att, 0, 0, P->it_t, P->is_pos[0], P->is_pos[1], P->is_pos[2], 0.0, 0.0, 0.0, 0.0, 0.0);
         st = SLOT_IDLE;
      }
   }
   L->active[slot] = st;
}

}  // namespace spx
