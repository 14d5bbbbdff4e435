// shp_solver.hip - ordered FP64 (drop-in) solve: dataflow over per-body
// counters.
//
// The reference applies the cached lagrangians in descending key order
// (applyCachedSlns, shapes/src/Physics/Solvers/Contact.hs:74-121) and then
// sweeps the contacts `iterations` times in the same order (improveWorld,
// :146-157).  Written as one sequence, item (p, k) is pass p of contact k
// (pass 0 is the warm start) and the items run in (p, k) order.  An item
// reads and writes the velocities of its two bodies and the lagrangians of
// its own contact only, so it depends on nothing but the latest earlier
// items touching one of its dynamic bodies.  For a dynamic body a touched
// by d(a) contacts, with contact k at rank r_a(k) among them in key order,
// the items on a are numbered p * d(a) + r_a(k) in sequence order.  A
// per-body counter holds how many items on the body have completed: item
// (p, k) runs when counter[a] == p * d(a) + r_a(k) and the same holds for
// b, and on completion it raises both counters by one (a release store
// after its velocity stores; the waiting side loads with acquire).  Every
// read therefore sees exactly the values of the sequential sweep and the
// result is the sequential result bit for bit, whatever the timing.  Static
// bodies (both inverse masses 0) are read-only and carry no counter.
//
// Lane t of the workgroup owns contacts t, t + T, t + 2T, ... and runs its
// items in sequence order.  The earliest unfinished item of the world is
// always ready (every earlier item is done) and it is its lane's current
// item, and lanes wait on nothing but counters inside the loop, so the loop
// always progresses.  As a bound that holds even if this argument were
// broken by a defect, a lane that sees no item of the workgroup complete
// for WAIT_LIMIT consecutive polls stops and raises SHP_FLAG_SOLVER_STALL.
//
// Which lane runs which item only changes the speed.  Sequence order per
// lane (lane t: contacts t, t + T, ...) is ideal while a world has no more
// contacts than lanes; beyond that a lane's second contact holds up its
// next pass.  That still pays while velocities sit in LDS.  A world whose
// velocities live in global memory and that has more contacts than lanes
// runs level-synchronously instead: the dataflow computes, integer work
// only, every contact's level in one pass (1 + the largest level of the
// earlier contacts on its dynamic bodies; contacts of one level share no
// dynamic body), the contacts are sorted by level, and each pass runs
// level by level with a barrier between levels, the next level's records
// loaded before the barrier.  Both give the sequential result bit for bit.
//
// One workgroup per world:
//   1. degrees d(a) by atomics, prefix sum -> per-body segments; every
//      (contact, side) is listed in its body's segment; rank r_a(k) = number
//      of entries of the segment that precede it in key order
//   2. counters are cleared, velocities staged in LDS when they fit
//   3. level-synchronous worlds: levels, counting sort by level
//   4. each contact's constants are packed into a compact record (the
//      jacobian rows are rebuilt from the normal and the two cross terms,
//      exactly: every other component is a sign change of the normal), in
//      contact order or in level order
//   5. dataflow loop over all items, or the passes level by level
//   6. lagrangians and velocities are written back
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_solve_common.h"

using namespace shp;

#define WAIT_LIMIT (1u << 20)

// compact record fields (doubles), stride contact_cap
enum {
  FL_UX = 0,  // jn[3], jn[4]; jn[0..1] = -u, jt[3..4] = clockwise(u), jt[0..1] = -jt[3..4]
  FL_UY,
  FL_JN2,
  FL_JN5,
  FL_JT2,
  FL_JT5,
  FL_BN,
  FL_RAX,
  FL_RAY,
  FL_RBX,
  FL_RBY,
  FL_MCN,
  FL_MCT,
  FL_MU,
  FL_E,
  FL_IMLA,
  FL_IMRA,
  FL_IMLB,
  FL_IMRB,
  FL_LN,
  FL_LT,
  FL_FIELDS
};

struct FRec {
  double ux, uy, jn2, jn5, jt2, jt5, bn, rax, ray, rbx, rby, mcn, mct, mu, e;
  double ima, ira, imb, irb, ln, lt;
  int ga, gb, flags;  // flags: 1 a static, 2 b static, 4 warm
};

__device__ __forceinline__ void load_frec(const Ctx& c, int k, FRec& r) {
  const double* P = c.pk;
  const long CC = c.contact_cap;
  r.ux = P[FL_UX * CC + k];
  r.uy = P[FL_UY * CC + k];
  r.jn2 = P[FL_JN2 * CC + k];
  r.jn5 = P[FL_JN5 * CC + k];
  r.jt2 = P[FL_JT2 * CC + k];
  r.jt5 = P[FL_JT5 * CC + k];
  r.bn = P[FL_BN * CC + k];
  r.rax = P[FL_RAX * CC + k];
  r.ray = P[FL_RAY * CC + k];
  r.rbx = P[FL_RBX * CC + k];
  r.rby = P[FL_RBY * CC + k];
  r.mcn = P[FL_MCN * CC + k];
  r.mct = P[FL_MCT * CC + k];
  r.mu = P[FL_MU * CC + k];
  r.e = P[FL_E * CC + k];
  r.ima = P[FL_IMLA * CC + k];
  r.ira = P[FL_IMRA * CC + k];
  r.imb = P[FL_IMLB * CC + k];
  r.irb = P[FL_IMRB * CC + k];
  r.ln = P[FL_LN * CC + k];
  r.lt = P[FL_LT * CC + k];
  r.ga = c.pk_g[k];
  r.gb = c.pk_g[CC + k];
  r.flags = c.pk_f[k];
}

// The full rows of contact k: non-flipped, jn = (-n, xa_n, n, xb_n) and
// jt = (-cw(n), xa_t, cw(n), xb_t); flipped, the halves are swapped back
// (shp_engine.hip, constraint generation), which gives the same shape with
// u = -n.  So jn[3..4] = u, jn[0..1] = -u, jt[3..4] = clockwise(u) and
// jt[0..1] = -jt[3..4] in both cases, and negation is exact.
__device__ __forceinline__ void pack_flow(const Ctx& c, int k, int dst) {
  const long CC = c.contact_cap;
  const int ga = c.ga[k], gb = c.gb[k];
  // every value is loaded before the first store, so the loads are not
  // serialised behind stores the compiler cannot prove disjoint
  double f[FL_FIELDS];
  const d2 u = row_normal(c, k);
  const double4 x = c.jx[k];
  f[FL_UX] = u.x;
  f[FL_UY] = u.y;
  f[FL_JN2] = x.x;
  f[FL_JN5] = x.y;
  f[FL_JT2] = x.z;
  f[FL_JT5] = x.w;
  f[FL_BN] = c.bn[k];
  f[FL_RAX] = c.rax[k];
  f[FL_RAY] = c.ray[k];
  f[FL_RBX] = c.rbx[k];
  f[FL_RBY] = c.rby[k];
  f[FL_MCN] = c.mcn[k];
  f[FL_MCT] = c.mct[k];
  f[FL_MU] = (c.mu[ga] + c.mu[gb]) / 2.0;  // pairMu (Friction.hs:46-48)
  f[FL_E] = hs_min(c.e[ga], c.e[gb]);       // min bounciness
  f[FL_IMLA] = c.iml[ga];
  f[FL_IMRA] = c.imr[ga];
  f[FL_IMLB] = c.iml[gb];
  f[FL_IMRB] = c.imr[gb];
  f[FL_LN] = c.ln[k];
  f[FL_LT] = c.lt[k];
  const uint8_t fl = (uint8_t)((c.is_static[ga] ? 1 : 0) | (c.is_static[gb] ? 2 : 0) | (c.warm[k] ? 4 : 0));
  double* P = c.pk;
#pragma unroll
  for (int q = 0; q < FL_FIELDS; ++q) P[q * CC + dst] = f[q];
  c.pk_g[dst] = ga;
  c.pk_g[CC + dst] = gb;
  c.pk_f[dst] = fl;
}

__device__ __forceinline__ void rows(const FRec& r, double* jn, double* jt, double* im) {
  d2 t = clockwise(mk(r.ux, r.uy));
  jn[0] = -r.ux;
  jn[1] = -r.uy;
  jn[2] = r.jn2;
  jn[3] = r.ux;
  jn[4] = r.uy;
  jn[5] = r.jn5;
  jt[0] = -t.x;
  jt[1] = -t.y;
  jt[2] = r.jt2;
  jt[3] = t.x;
  jt[4] = t.y;
  jt[5] = r.jt5;
  im[0] = r.ima;
  im[1] = r.ima;
  im[2] = r.ira;
  im[3] = r.imb;
  im[4] = r.imb;
  im[5] = r.irb;
}

// Velocity access: LDS (3 doubles per local body) or the global arrays.
template <bool VLDS>
struct Vel {
  double* lds;
  const Ctx* c;
  int b0;
  __device__ __forceinline__ void load(int g, double* v) const {
    if (VLDS) {
      const double* p = lds + 3 * (g - b0);
      v[0] = p[0];
      v[1] = p[1];
      v[2] = p[2];
    } else {
      v[0] = c->vx[g];
      v[1] = c->vy[g];
      v[2] = c->w[g];
    }
  }
  __device__ __forceinline__ void store(int g, const double* v) const {
    if (VLDS) {
      double* p = lds + 3 * (g - b0);
      p[0] = v[0];
      p[1] = v[1];
      p[2] = v[2];
    } else {
      c->vx[g] = v[0];
      c->vy[g] = v[1];
      c->w[g] = v[2];
    }
  }
};

// useCache (Solvers/Contact.hs:100-114): applySln with the cached lambdas
template <bool VLDS>
__device__ __forceinline__ void warm_frec(const FRec& r, const Vel<VLDS>& V) {
  double v[6], jn[6], jt[6], im[6];
  rows(r, jn, jt, im);
  V.load(r.ga, v);
  V.load(r.gb, v + 3);
  apply_lag(r.ln, jn, im, v);
  apply_lag(r.lt, jt, im, v);
  if (!(r.flags & 1)) V.store(r.ga, v);
  if (!(r.flags & 2)) V.store(r.gb, v + 3);
}

// improveContactSln (Solvers/Contact.hs:124-143): contactLagrangian with
// bounceB (Constraints/Contact.hs:74-95, Restitution.hs:34-47), then
// solutionProcessor = positive / clampAbs (Constraints/Contact.hs:101-110,
// SolutionProcessors.hs:33-53), then applySln, on a compact record.  The
// new lagrangians go to the record and to the packed array.
template <bool VLDS>
__device__ __forceinline__ void improve_frec(const Ctx& c, int k, FRec& r, const Vel<VLDS>& V) {
  double v[6], jn[6], jt[6], im[6];
  rows(r, jn, jt, im);
  V.load(r.ga, v);
  V.load(r.gb, v + 3);
  d2 ra = mk(r.rax, r.ray), rb = mk(r.rbx, r.rby), nr = mk(r.ux, r.uy);
  d2 closing = add(add(add(neg(mk(v[0], v[1])), zcross(-v[2], ra)), mk(v[3], v[4])), zcross(v[5], rb));
  double bb = hs_min(0.0, r.e * dot(closing, nr));
  double bn = r.bn + bb;
  double new_n = (-(dot6(jn, v) + bn)) / r.mcn;
  double new_t = (-(dot6(jt, v) + 0.0)) / r.mct;
  double app_n = hs_max(new_n, -r.ln);
  double acc_n = r.ln + app_n;
  double thr = acc_n * r.mu;
  double acc_t = r.lt + new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - r.lt;
  apply_lag(app_n, jn, im, v);
  apply_lag(app_t, jt, im, v);
  if (!(r.flags & 1)) V.store(r.ga, v);
  if (!(r.flags & 2)) V.store(r.gb, v + 3);
  const long CC = c.contact_cap;
  r.ln = acc_n;
  r.lt = accp;
  c.pk[FL_LN * CC + k] = acc_n;
  c.pk[FL_LT * CC + k] = accp;
}

__device__ __forceinline__ int ld_acquire(int* p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void st_release(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Exclusive prefix sum of a[0, n) in place by the T threads of the
// workgroup; wsum holds T / 64 ints of LDS.
template <int T>
__device__ void block_scan_exclusive(int* a, int n, int* wsum) {
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int chunk = (n + T - 1) / T;
  const int lo = min(n, tid * chunk), hi = min(n, lo + chunk);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += a[i];
  int x = s;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[wid] = x;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int q = 0; q < T / 64; ++q) {
      int t = wsum[q];
      wsum[q] = acc;
      acc += t;
    }
  }
  __syncthreads();
  int run = wsum[wid] + x - s;
  for (int i = lo; i < hi; ++i) {
    int t = a[i];
    a[i] = run;
    run += t;
  }
  __syncthreads();
}

// Ranks of every (contact, dynamic side) of world contacts [c0, c0 + n):
// flow_rk[k] = {r_a, d(a), r_b, d(b)}, with r = -1 for a static side.
// off and cur hold nb ints each (LDS or global, local body index).
template <int T>
__device__ void compute_ranks(const Ctx& c, int c0, int n, int b0, int nb, int* off, int* cur,
                              int* wsum) {
  const int tid = threadIdx.x;
  int* list = c.flow_list + 2L * c0;
  for (int i = tid; i < nb; i += T) off[i] = 0;
  __syncthreads();
  for (int i = tid; i < n; i += T) {
    int ga = c.ga[c0 + i], gb = c.gb[c0 + i];
    if (!c.is_static[ga]) atomicAdd(&off[ga - b0], 1);
    if (!c.is_static[gb]) atomicAdd(&off[gb - b0], 1);
  }
  __syncthreads();
  block_scan_exclusive<T>(off, nb, wsum);
  for (int i = tid; i < nb; i += T) cur[i] = off[i];
  __syncthreads();
  for (int i = tid; i < n; i += T) {
    int ga = c.ga[c0 + i], gb = c.gb[c0 + i];
    if (!c.is_static[ga]) list[atomicAdd(&cur[ga - b0], 1)] = 2 * i;
    if (!c.is_static[gb]) list[atomicAdd(&cur[gb - b0], 1)] = 2 * i + 1;
  }
  __syncthreads();
  // segment of body x: [off[x], cur[x]); entries are 2 * contact + side,
  // so ascending entries are ascending key order
  for (int i = tid; i < n; i += T) {
    int4 rk = make_int4(-1, 0, -1, 0);
    int g[2] = {c.ga[c0 + i], c.gb[c0 + i]};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      if (c.is_static[g[s]]) continue;
      int x = g[s] - b0, lo = off[x], hi = cur[x], me = 2 * i + s, r = 0;
      for (int j = lo; j < hi; ++j) r += list[j] < me;
      if (s == 0) {
        rk.x = r;
        rk.y = hi - lo;
      } else {
        rk.z = r;
        rk.w = hi - lo;
      }
    }
    c.flow_rk[c0 + i] = rk;
  }
  __syncthreads();
}

// Bounded waiting shared by the dataflow loops: every wave iteration in
// which some lane completed an item adds to the workgroup's progress count;
// a lane that sees the count unchanged for WAIT_LIMIT consecutive polls
// stops (SHP_FLAG_SOLVER_STALL).  Returns false when the lane must stop.
struct Waiter {
  unsigned spins = 0;
  int seen = -1;
  __device__ __forceinline__ bool after_poll(const Ctx& c, bool ready, int* progress) {
    const unsigned long long m = __ballot(ready);
    if (m) {
      if ((int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(progress, __popcll(m));
    } else {
      __builtin_amdgcn_s_sleep(1);
    }
    if (ready) {
      spins = 0;
      return true;
    }
    int pr = __hip_atomic_load(progress, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (pr != seen) {
      seen = pr;
      spins = 0;
    } else if (++spins > WAIT_LIMIT) {
      atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_SOLVER_STALL);
      return false;
    }
    return true;
  }
};

// Whether item (p, contact with ranks rk) may run: both counters at its
// rank for pass p.  la/lb: local body indices.
__device__ __forceinline__ bool item_ready(int* cnt, int la, int lb, int4 rk, int p) {
  bool ready = true;
  if (rk.x >= 0) ready = ld_acquire(cnt + la) == p * rk.y + rk.x;
  if (ready && rk.z >= 0) ready = ld_acquire(cnt + lb) == p * rk.w + rk.z;
  return ready;
}

__device__ __forceinline__ void item_done(int* cnt, int la, int lb, int4 rk, int p) {
  if (rk.x >= 0) st_release(cnt + la, p * rk.y + rk.x + 1);
  if (rk.z >= 0) st_release(cnt + lb, p * rk.w + rk.z + 1);
}

// Lane order "cyclic": lane t runs (p, t), (p, t + T), ... for p = 0, 1,
// ..., which is sequence order.  With at most one contact per lane this is
// the best order (a lane's items are then one contact's passes, which
// depend on each other anyway).  cnt: per-body counters, cleared.
template <bool VLDS, int T>
__device__ void run_flow(const Ctx& c, int c0, int n, int b0, int* cnt, const Vel<VLDS>& V,
                         int* progress) {
  const int tid = threadIdx.x;
  const int P = c.iterations;
  int i = tid, p = 0;
  bool live = i < n;
  FRec r;
  int4 rk = make_int4(-1, 0, -1, 0);
  if (live) {
    load_frec(c, c0 + i, r);
    rk = c.flow_rk[c0 + i];
  }
  Waiter wt;
  while (live) {
    const int la = r.ga - b0, lb = r.gb - b0;
    const bool ready = item_ready(cnt, la, lb, rk, p);
    if (ready) {
      if (p == 0) {
        if (r.flags & 4) warm_frec<VLDS>(r, V);
      } else {
        improve_frec<VLDS>(c, c0 + i, r, V);
      }
      item_done(cnt, la, lb, rk, p);
      int ni = i + T, np = p;
      if (ni >= n) {
        ni = tid;
        ++np;
      }
      if (np > P) {
        live = false;
      } else {
        if (ni != i) {
          load_frec(c, c0 + ni, r);
          rk = c.flow_rk[c0 + ni];
        }
        i = ni;
        p = np;
      }
    }
    if (!wt.after_poll(c, ready, progress)) live = false;
  }
}

// Level of every contact in one pass (the same counters, integer work
// only, cyclic lane order): lvl[k] = 1 + max(level of the previous contact
// on a, on b), with last[] holding per body the level of its latest
// contact (cleared).  Returns the lane's largest level.
template <int T>
__device__ int run_levels(const Ctx& c, int c0, int n, int b0, int* cnt, int* last, int* lvl,
                          int* progress) {
  const int tid = threadIdx.x;
  const int P = 0;
  int i = tid, p = 0, maxl = 0;
  bool live = i < n;
  int la = 0, lb = 0;
  int4 rk = make_int4(-1, 0, -1, 0);
  if (live) {
    la = c.ga[c0 + i] - b0;
    lb = c.gb[c0 + i] - b0;
    rk = c.flow_rk[c0 + i];
  }
  Waiter wt;
  while (live) {
    const bool ready = item_ready(cnt, la, lb, rk, p);
    if (ready) {
      int l = 0;
      if (rk.x >= 0) l = last[la];
      if (rk.z >= 0) l = max(l, last[lb]);
      ++l;
      if (rk.x >= 0) last[la] = l;
      if (rk.z >= 0) last[lb] = l;
      lvl[i] = l;
      maxl = max(maxl, l);
      item_done(cnt, la, lb, rk, p);
      int ni = i + T, np = p;
      if (ni >= n) {
        ni = tid;
        ++np;
      }
      if (np > P) {
        live = false;
      } else {
        if (ni != i) {
          la = c.ga[c0 + ni] - b0;
          lb = c.gb[c0 + ni] - b0;
          rk = c.flow_rk[c0 + ni];
        }
        i = ni;
        p = np;
      }
    }
    if (!wt.after_poll(c, ready, progress)) live = false;
  }
  return maxl;
}

// Counting sort of the n contacts by level (levels 1..maxl): order[s]
// gets the contact of slot s, and bins[l] (LDS, maxl + 1 ints) ends as the
// end of level l's slots (level l = [bins[l - 1], bins[l]), bins[0] = 0).
template <int T>
__device__ void sort_by_level(const int* lvl, int n, int maxl, int* bins, int* order, int* wsum) {
  const int tid = threadIdx.x;
  for (int l = tid; l <= maxl; l += T) bins[l] = 0;
  __syncthreads();
  for (int q = tid; q < n; q += T) atomicAdd(&bins[lvl[q]], 1);
  __syncthreads();
  block_scan_exclusive<T>(bins, maxl + 1, wsum);
  for (int q = tid; q < n; q += T) order[atomicAdd(&bins[lvl[q]], 1)] = q;
  __syncthreads();
}

// Pass p over slot-ordered records, level by level.  Thread t takes slots
// start(l) + t, + T, ... of every level, so a slot keeps its thread in
// every pass and the lagrangians it wrote last pass are its own; the
// record of its first slot of the next level (lagrangians included) is
// loaded before the barrier, the velocities after it.
template <bool VLDS, int T>
__device__ void run_pass_levels(const Ctx& c, int c0, int p, int maxl, const int* bins,
                                const Vel<VLDS>& V) {
  const int tid = threadIdx.x;
  FRec r, rn;
  int s = tid;  // level 1 starts at slot 0
  bool has = maxl >= 1 && s < bins[1];
  if (has) load_frec(c, c0 + s, r);
  for (int l = 1; l <= maxl; ++l) {
    const int end = bins[l];
    const int sn = end + tid;
    const bool hasn = l < maxl && sn < bins[l + 1];
    if (hasn) load_frec(c, c0 + sn, rn);
    if (has) {
      if (p == 0) {
        if (r.flags & 4) warm_frec<VLDS>(r, V);
      } else {
        improve_frec<VLDS>(c, c0 + s, r, V);
      }
    }
    for (int s2 = s + T; s2 < end; s2 += T) {
      FRec r2;
      load_frec(c, c0 + s2, r2);
      if (p == 0) {
        if (r2.flags & 4) warm_frec<VLDS>(r2, V);
      } else {
        improve_frec<VLDS>(c, c0 + s2, r2, V);
      }
    }
    __syncthreads();
    r = rn;
    has = hasn;
    s = sn;
  }
}

// CLDS: per-body offsets/counters in LDS (else in flow_off/flow_cur);
// VLDS: velocities in LDS.
template <bool CLDS, bool VLDS, int T>
__global__ __launch_bounds__(T) void k_solve_flow(Ctx c) {
  extern __shared__ double lds_d[];
  __shared__ int wsum[T / 64];
  __shared__ int progress;
  __shared__ int s_maxl;
  const int w = c.world_order[blockIdx.x];
  const int tid = threadIdx.x;
  const int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  const int n = c1 - c0;
  const int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  double* vel = lds_d;
  int* ip = (int*)(lds_d + (VLDS ? 3 * nb : 0));
  int* bins = ip;  // LEVEL_CAP + 1
  int* off = CLDS ? ip + (LEVEL_CAP + 1) : c.flow_off + b0;
  int* cur = CLDS ? off + nb : c.flow_cur + b0;
  tmark(c, 0);
  compute_ranks<T>(c, c0, n, b0, nb, off, cur, wsum);
  tmark(c, 1);
  int* cnt = off;
  int* last = cur;
  const bool levels = !VLDS && n > T;
  for (int i = tid; i < nb; i += T) {
    cnt[i] = 0;
    if (levels) last[i] = 0;
  }
  if (tid == 0) {
    progress = 0;
    s_maxl = 0;
  }
  Vel<VLDS> V;
  V.lds = vel;
  V.c = &c;
  V.b0 = b0;
  if (VLDS) {
    for (int i = tid; i < nb; i += T) {
      vel[3 * i] = c.vx[b0 + i];
      vel[3 * i + 1] = c.vy[b0 + i];
      vel[3 * i + 2] = c.w[b0 + i];
    }
  }
  __syncthreads();
  tmark(c, 2);
  int maxl = 0;
  int* order = c.flow_order + c0;
  if (levels) {
    int m = run_levels<T>(c, c0, n, b0, cnt, last, c.flow_lvl + c0, &progress);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if ((tid & 63) == 0) atomicMax(&s_maxl, m);
    __syncthreads();
    maxl = s_maxl;
    if (maxl <= LEVEL_CAP) {
      sort_by_level<T>(c.flow_lvl + c0, n, maxl, bins, order, wsum);  // ends with a barrier
      for (int s = tid; s < n; s += T) pack_flow(c, c0 + order[s], c0 + s);
    } else {  // too deep for the bins: the dataflow with its counters
      for (int i = tid; i < nb; i += T) cnt[i] = 0;
      if (tid == 0) progress = 0;
      for (int k = c0 + tid; k < c1; k += T) pack_flow(c, k, k);
    }
  } else {
    for (int k = c0 + tid; k < c1; k += T) pack_flow(c, k, k);
  }
  __syncthreads();
  tmark(c, 3);
  const bool sync = levels && maxl <= LEVEL_CAP;
  if (sync) {
    for (int p = 0; p <= c.iterations; ++p) run_pass_levels<VLDS, T>(c, c0, p, maxl, bins, V);
  } else {
    run_flow<VLDS, T>(c, c0, n, b0, cnt, V, &progress);
  }
  __syncthreads();
  tmark(c, 4);
  const long CC = c.contact_cap;
  for (int s = tid; s < n; s += T) {
    const int k = sync ? c0 + order[s] : c0 + s;
    c.ln[k] = c.pk[FL_LN * CC + c0 + s];
    c.lt[k] = c.pk[FL_LT * CC + c0 + s];
  }
  if (VLDS) {
    for (int i = tid; i < nb; i += T) {
      if (c.is_static[b0 + i]) continue;
      c.vx[b0 + i] = vel[3 * i];
      c.vy[b0 + i] = vel[3 * i + 1];
      c.w[b0 + i] = vel[3 * i + 2];
    }
  }
  __syncthreads();
  tmark(c, 5);
}

template <bool CLDS, bool VLDS>
static const void* flow_kernel(int threads) {
  return threads == 512 ? (const void*)k_solve_flow<CLDS, VLDS, 512>
                        : (const void*)k_solve_flow<CLDS, VLDS, 256>;
}

static_assert(FL_FIELDS <= FLOW_RECORD_DOUBLES, "compact record exceeds the packed buffer");

// LDS per world: level bins, then velocities (24 B/body) and offsets +
// cursors (8 B/body) when they fit; offsets + cursors alone; else both in
// global memory.  4 waves per world; a large world gets 8 (a
// level-synchronous pass keeps two records per lane, so 8 waves leave it
// 256 registers per lane).
void plan_ordered_solver(StepPlan* P, int solver_threads) {
  const size_t bins = 4UL * (LEVEL_CAP + 1);
  const size_t both = bins + 32UL * P->max_nb, cnt_only = bins + 8UL * P->max_nb;
  if (both <= LDS_MAX_BYTES) {
    P->solve_clds = true;
    P->solve_vlds = true;
    P->solve_lds = both;
  } else if (cnt_only <= LDS_MAX_BYTES) {
    P->solve_clds = true;
    P->solve_vlds = false;
    P->solve_lds = cnt_only;
  } else {
    P->solve_clds = false;
    P->solve_vlds = false;
    P->solve_lds = bins;
  }
  P->solve_threads = solver_threads ? solver_threads : (P->max_nb > 2048 ? 512 : 256);
}

// wsum[T / 64], progress, s_maxl
size_t ordered_static_lds(int threads) { return lds16((size_t)(threads / 64) * 4) + 2 * lds16(4); }

shp_status solver_alloc(shp_batch* b) {
  const StepPlan& P = b->plan;
  if (P.solver != SHP_SOLVER_ORDERED_F64) return colored_alloc(b);
  const void* fn = P.solve_clds ? (P.solve_vlds ? flow_kernel<true, true>(P.solve_threads)
                                                : flow_kernel<true, false>(P.solve_threads))
                                : flow_kernel<false, false>(P.solve_threads);
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)std::max<size_t>(P.solve_lds, 16)) != hipSuccess) {
    shp_set_error("hipFuncSetAttribute (ordered solver LDS) failed");
    return SHP_E_HIP;
  }
  hipFuncAttributes a{};
  if (hipFuncGetAttributes(&a, fn) != hipSuccess ||
      a.sharedSizeBytes > ordered_static_lds(P.solve_threads)) {
    shp_set_error("ordered solver: static LDS above the plan's");
    return SHP_E_HIP;
  }
  return SHP_OK;
}

template <int T>
static void launch_flow(shp_batch* b) {
  const StepPlan& P = b->plan;
  const size_t lds = std::max<size_t>(P.solve_lds, 16);
  if (P.solve_clds && P.solve_vlds)
    k_solve_flow<true, true, T><<<b->ctx.W, T, lds, b->stream>>>(b->ctx);
  else if (P.solve_clds)
    k_solve_flow<true, false, T><<<b->ctx.W, T, lds, b->stream>>>(b->ctx);
  else
    k_solve_flow<false, false, T><<<b->ctx.W, T, lds, b->stream>>>(b->ctx);
}

void launch_ordered_solver(shp_batch* b) {
  if (b->plan.solve_threads == 512)
    launch_flow<512>(b);
  else
    launch_flow<256>(b);
}
