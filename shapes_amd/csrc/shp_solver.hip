// shp_solver.hip - ordered FP64 (drop-in) solve schedule.
//
// The reference solves contacts one after another in descending key order
// (improveWorld, shapes/src/Physics/Solvers/Contact.hs:146-157), after
// applying cached lagrangians in the same order (applyCachedSlns, :74-121).
// Contact c only depends on the latest earlier contact that touches one of
// its dynamic bodies, so each world's contacts get a level
//   level(c) = 1 + max(level of earlier contacts sharing a dynamic body)
// and each level is solved in parallel.  Two contacts of one level touch
// disjoint dynamic bodies, so the order inside a level cannot change any bit
// and the result equals the sequential sweep bit for bit.  Static bodies
// (both inverse masses 0) are read-only (see store_vel).
//
// One workgroup per world:
//   1. wave 0 assigns levels, 64 contacts at a time (readlane broadcasts)
//   2. counting sort of the contacts by level -> schedule slots
//   3. every contact's constants are packed in slot order (one coalesced
//      record per slot, so a level's records are contiguous)
//   4. velocities of the world's bodies are staged in LDS when they fit
//   5. warm-start pass and sweeps, level by level; the next level's records
//      are loaded before the barrier
//   6. lagrangians and velocities are written back
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_solve_common.h"

using namespace shp;

#define LEVEL_CAP 2048
#define SOLVE_THREADS 256

// packed record fields (doubles), stride contact_cap
enum {
  PK_JN = 0,
  PK_JT = 6,
  PK_BN = 12,
  PK_RAX = 13,
  PK_RAY = 14,
  PK_RBX = 15,
  PK_RBY = 16,
  PK_NRX = 17,
  PK_NRY = 18,
  PK_MCN = 19,
  PK_MCT = 20,
  PK_MU = 21,
  PK_E = 22,
  PK_IMLA = 23,
  PK_IMRA = 24,
  PK_IMLB = 25,
  PK_IMRB = 26,
  PK_LN = 27,
  PK_LT = 28,
  PK_FIELDS = 29
};

struct Rec {
  double jn[6], jt[6];
  double bn, rax, ray, rbx, rby, nrx, nry, mcn, mct, mu, e;
  double im[6];
  double ln, lt;
  int ga, gb, flags;  // flags: 1 a static, 2 b static, 4 warm
};

__device__ __forceinline__ void load_rec(const Ctx& c, int s, Rec& r) {
  const double* P = c.pk;
  long CC = c.contact_cap;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    r.jn[q] = P[(PK_JN + q) * CC + s];
    r.jt[q] = P[(PK_JT + q) * CC + s];
  }
  r.bn = P[PK_BN * CC + s];
  r.rax = P[PK_RAX * CC + s];
  r.ray = P[PK_RAY * CC + s];
  r.rbx = P[PK_RBX * CC + s];
  r.rby = P[PK_RBY * CC + s];
  r.nrx = P[PK_NRX * CC + s];
  r.nry = P[PK_NRY * CC + s];
  r.mcn = P[PK_MCN * CC + s];
  r.mct = P[PK_MCT * CC + s];
  r.mu = P[PK_MU * CC + s];
  r.e = P[PK_E * CC + s];
  double iml_a = P[PK_IMLA * CC + s], imr_a = P[PK_IMRA * CC + s];
  double iml_b = P[PK_IMLB * CC + s], imr_b = P[PK_IMRB * CC + s];
  r.im[0] = iml_a;
  r.im[1] = iml_a;
  r.im[2] = imr_a;
  r.im[3] = iml_b;
  r.im[4] = iml_b;
  r.im[5] = imr_b;
  r.ln = P[PK_LN * CC + s];
  r.lt = P[PK_LT * CC + s];
  r.ga = c.pk_g[s];
  r.gb = c.pk_g[CC + s];
  r.flags = c.pk_f[s];
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
__device__ __forceinline__ void warm_rec(const Rec& r, const Vel<VLDS>& V) {
  double v[6];
  V.load(r.ga, v);
  V.load(r.gb, v + 3);
  apply_lag(r.ln, r.jn, r.im, v);
  apply_lag(r.lt, r.jt, r.im, v);
  if (!(r.flags & 1)) V.store(r.ga, v);
  if (!(r.flags & 2)) V.store(r.gb, v + 3);
}

// improveContactSln (Solvers/Contact.hs:124-143), operation for operation
// as improve() in shp_solve_common.h, on a packed record.
template <bool VLDS>
__device__ __forceinline__ void improve_rec(const Ctx& c, int s, const Rec& r, const Vel<VLDS>& V) {
  double v[6];
  V.load(r.ga, v);
  V.load(r.gb, v + 3);
  d2 ra = mk(r.rax, r.ray), rb = mk(r.rbx, r.rby), nr = mk(r.nrx, r.nry);
  d2 closing = add(add(add(neg(mk(v[0], v[1])), zcross(-v[2], ra)), mk(v[3], v[4])), zcross(v[5], rb));
  double bb = hs_min(0.0, r.e * dot(closing, nr));
  double bn = r.bn + bb;
  double new_n = (-(dot6(r.jn, v) + bn)) / r.mcn;
  double new_t = (-(dot6(r.jt, v) + 0.0)) / r.mct;
  double app_n = hs_max(new_n, -r.ln);
  double acc_n = r.ln + app_n;
  double thr = acc_n * r.mu;
  double acc_t = r.lt + new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - r.lt;
  apply_lag(app_n, r.jn, r.im, v);
  apply_lag(app_t, r.jt, r.im, v);
  if (!(r.flags & 1)) V.store(r.ga, v);
  if (!(r.flags & 2)) V.store(r.gb, v + 3);
  long CC = c.contact_cap;
  c.pk[PK_LN * CC + s] = acc_n;
  c.pk[PK_LT * CC + s] = accp;
}

// Wave 0 assigns levels to contacts [c0, c1), 64 at a time.  `last` holds
// per local body the level of the latest contact touching it.  Inside a
// chunk, lane k is raised above every earlier lane j that shares a dynamic
// body (lane j's level is final when broadcast); only the last lane that
// touches a body updates `last`.  Returns the maximum level.
__device__ int assign_levels(const Ctx& c, int c0, int c1, int b0, int* last) {
  int lane = threadIdx.x & 63;
  int maxl = 0;
  for (int base = c0; base < c1; base += 64) {
    int k = base + lane;
    bool valid = k < c1;
    int a = -1, b = -1;
    if (valid) {
      int ga = c.ga[k], gb = c.gb[k];
      a = c.is_static[ga] ? -1 : ga - b0;
      b = c.is_static[gb] ? -1 : gb - b0;
    }
    int la = a >= 0 ? last[a] : 0;
    int lb = b >= 0 ? last[b] : 0;
    int lvl = (la > lb ? la : lb) + 1;
    bool a_last = true, b_last = true;
    int cnt = c1 - base < 64 ? c1 - base : 64;
    for (int j = 0; j < cnt; ++j) {
      int aj = __builtin_amdgcn_readlane(a, j);
      int bj = __builtin_amdgcn_readlane(b, j);
      int lj = __builtin_amdgcn_readlane(lvl, j);
      bool sa = a >= 0 && (a == aj || a == bj);
      bool sb = b >= 0 && (b == aj || b == bj);
      if (j < lane && (sa || sb) && lj + 1 > lvl) lvl = lj + 1;
      if (j > lane) {
        if (sa) a_last = false;
        if (sb) b_last = false;
      }
    }
    if (valid) {
      c.sched_level[k] = lvl;
      if (a >= 0 && a_last) last[a] = lvl;
      if (b >= 0 && b_last) last[b] = lvl;
      if (lvl > maxl) maxl = lvl;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  }
  for (int off = 32; off > 0; off >>= 1) {
    int o = __shfl_xor(maxl, off);
    if (o > maxl) maxl = o;
  }
  return maxl;
}

// Pack contact k's constants into slot s (global slot index).
__device__ __forceinline__ void pack_contact(const Ctx& c, int k, int s) {
  double* P = c.pk;
  long CC = c.contact_cap;
  int ga = c.ga[k], gb = c.gb[k];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    P[(PK_JN + q) * CC + s] = c.jn[q * CC + k];
    P[(PK_JT + q) * CC + s] = c.jt[q * CC + k];
  }
  P[PK_BN * CC + s] = c.bn[k];
  P[PK_RAX * CC + s] = c.rax[k];
  P[PK_RAY * CC + s] = c.ray[k];
  P[PK_RBX * CC + s] = c.rbx[k];
  P[PK_RBY * CC + s] = c.rby[k];
  bool fl = c.flip[k] != 0;
  P[PK_NRX * CC + s] = fl ? -c.cnx[k] : c.cnx[k];
  P[PK_NRY * CC + s] = fl ? -c.cny[k] : c.cny[k];
  P[PK_MCN * CC + s] = c.mcn[k];
  P[PK_MCT * CC + s] = c.mct[k];
  P[PK_MU * CC + s] = (c.mu[ga] + c.mu[gb]) / 2.0;  // pairMu (Friction.hs:46-48)
  P[PK_E * CC + s] = hs_min(c.e[ga], c.e[gb]);        // min bounciness
  P[PK_IMLA * CC + s] = c.iml[ga];
  P[PK_IMRA * CC + s] = c.imr[ga];
  P[PK_IMLB * CC + s] = c.iml[gb];
  P[PK_IMRB * CC + s] = c.imr[gb];
  P[PK_LN * CC + s] = c.ln[k];
  P[PK_LT * CC + s] = c.lt[k];
  c.pk_g[s] = ga;
  c.pk_g[CC + s] = gb;
  c.pk_f[s] = (uint8_t)((c.is_static[ga] ? 1 : 0) | (c.is_static[gb] ? 2 : 0) | (c.warm[k] ? 4 : 0));
}

// Run one pass (warm-start application or a sweep) level by level.  Each
// thread owns slot s_off[l] + tid of every level (plus further slots of
// wide levels); the record of its next slot is loaded before the barrier.
template <bool VLDS, bool WARM>
__device__ void run_pass(const Ctx& c, int c0, int maxl, const int* s_off, const Vel<VLDS>& V) {
  int tid = threadIdx.x;
  Rec r, rn;
  int s = s_off[1] + tid;
  bool has = 1 <= maxl && s < s_off[2];
  if (has) load_rec(c, c0 + s, r);
  for (int l = 1; l <= maxl; ++l) {
    int sn = s_off[l + 1] + tid;
    bool hasn = l < maxl && sn < s_off[l + 2];
    if (hasn) load_rec(c, c0 + sn, rn);
    if (has) {
      if (WARM) {
        if (r.flags & 4) warm_rec<VLDS>(r, V);
      } else {
        improve_rec<VLDS>(c, c0 + s, r, V);
      }
    }
    for (int s2 = s + SOLVE_THREADS; s2 < s_off[l + 1]; s2 += SOLVE_THREADS) {
      Rec r2;
      load_rec(c, c0 + s2, r2);
      if (WARM) {
        if (r2.flags & 4) warm_rec<VLDS>(r2, V);
      } else {
        improve_rec<VLDS>(c, c0 + s2, r2, V);
      }
    }
    __syncthreads();
    r = rn;
    has = hasn;
    s = sn;
  }
}

template <bool VLDS>
__global__ __launch_bounds__(SOLVE_THREADS) void k_solve_levels(Ctx c, int pool_ints) {
  extern __shared__ int lds[];
  int* s_off = lds;                            // LEVEL_CAP + 2
  int* s_cur = lds + (LEVEL_CAP + 2);          // LEVEL_CAP + 2
  int* pool = lds + 2 * (LEVEL_CAP + 2) + 2;   // 8-byte aligned
  __shared__ int s_maxl;
  int w = blockIdx.x;
  int tid = threadIdx.x;
  int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  if (nb > pool_ints || (VLDS && 6 * nb > pool_ints)) {  // world too large: exact serial
    if (tid == 0) solve_world_serial(c, c0, c1);
    return;
  }
  int* last = pool;
  tmark(c, 0);
  for (int i = tid; i < nb; i += SOLVE_THREADS) last[i] = 0;
  __syncthreads();
  if (tid < 64) {
    int m = assign_levels(c, c0, c1, b0, last);
    if (tid == 0) {
      s_maxl = m;
      atomicMax(&c.counters[CNT_LEVELS], m);
    }
  }
  __syncthreads();
  tmark(c, 1);
  int maxl = s_maxl;
  if (maxl > LEVEL_CAP) {
    if (tid == 0) {
      atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_SCHEDULE_OVERFLOW);
      solve_world_serial(c, c0, c1);
    }
    return;
  }
  for (int i = tid; i <= maxl + 1; i += SOLVE_THREADS) s_cur[i] = 0;
  __syncthreads();
  for (int k = c0 + tid; k < c1; k += SOLVE_THREADS) atomicAdd(&s_cur[c.sched_level[k]], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int l = 0; l <= maxl + 1; ++l) {
      int h = s_cur[l];
      s_off[l] = acc;
      s_cur[l] = acc;
      acc += h;
    }
  }
  __syncthreads();
  for (int k = c0 + tid; k < c1; k += SOLVE_THREADS) {
    int pos = atomicAdd(&s_cur[c.sched_level[k]], 1);
    c.sched_order[c0 + pos] = k;
    pack_contact(c, k, c0 + pos);
  }
  Vel<VLDS> V;
  V.lds = (double*)pool;
  V.c = &c;
  V.b0 = b0;
  if (VLDS) {
    __syncthreads();  // `last` is dead; the pool now holds velocities
    for (int i = tid; i < nb; i += SOLVE_THREADS) {
      V.lds[3 * i] = c.vx[b0 + i];
      V.lds[3 * i + 1] = c.vy[b0 + i];
      V.lds[3 * i + 2] = c.w[b0 + i];
    }
  }
  __threadfence_block();
  __syncthreads();
  tmark(c, 2);
  run_pass<VLDS, true>(c, c0, maxl, s_off, V);
  tmark(c, 3);
  for (int it = 0; it < c.iterations; ++it) run_pass<VLDS, false>(c, c0, maxl, s_off, V);
  tmark(c, 4);
  long CC = c.contact_cap;
  for (int s = tid; s < c1 - c0; s += SOLVE_THREADS) {
    int k = c.sched_order[c0 + s];
    c.ln[k] = c.pk[PK_LN * CC + c0 + s];
    c.lt[k] = c.pk[PK_LT * CC + c0 + s];
  }
  if (VLDS) {
    for (int i = tid; i < nb; i += SOLVE_THREADS) {
      if (c.is_static[b0 + i]) continue;
      c.vx[b0 + i] = V.lds[3 * i];
      c.vy[b0 + i] = V.lds[3 * i + 1];
      c.w[b0 + i] = V.lds[3 * i + 2];
    }
  }
  __syncthreads();
  tmark(c, 5);
}

#define LDS_MAX_BYTES (160 * 1024 - 512)

shp_status solver_alloc(shp_batch* b) {
  Ctx& c = b->ctx;
  long CC = c.contact_cap;
  void** bufs[] = {(void**)&c.sched_level, (void**)&c.sched_order, (void**)&c.pk, (void**)&c.pk_g,
                   (void**)&c.pk_f, (void**)&c.fvx, (void**)&c.fvy, (void**)&c.fw};
  size_t sizes[] = {CC * 4UL, CC * 4UL, (size_t)PK_FIELDS * CC * 8, 2UL * CC * 4, (size_t)CC,
                    (size_t)c.NB * 4, (size_t)c.NB * 4, (size_t)c.NB * 4};
  for (int k = 0; k < 8; ++k) {
    if (hipMalloc(bufs[k], sizes[k] ? sizes[k] : 4) != hipSuccess) return SHP_E_OOM;
    b->allocs.push_back(*bufs[k]);
  }
  int max_nb = 0;
  for (int w = 0; w < c.W; ++w) max_nb = std::max(max_nb, b->h_body_off[w + 1] - b->h_body_off[w]);
  // LDS: level offsets/cursors + a pool for `last` (4 B/body) and, when it
  // fits, the velocities (24 B/body)
  size_t fixed = (2 * (LEVEL_CAP + 2) + 2) * 4;
  size_t vel = fixed + 24UL * max_nb;
  size_t lst = fixed + 4UL * max_nb;
  if (vel <= LDS_MAX_BYTES) {
    b->solve_vlds = true;
    b->solve_lds = vel;
  } else {
    b->solve_vlds = false;
    b->solve_lds = std::min<size_t>(std::max<size_t>(lst, fixed + 4 * 1024), LDS_MAX_BYTES);
  }
  auto fn = b->solve_vlds ? (const void*)k_solve_levels<true> : (const void*)k_solve_levels<false>;
  if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->solve_lds) !=
      hipSuccess)
    return SHP_E_HIP;
  return colored_alloc(b);
}

void launch_ordered_solver(shp_batch* b) {
  int pool_ints = (int)((b->solve_lds - (2 * (LEVEL_CAP + 2) + 2) * 4) / 4);
  if (b->solve_vlds)
    k_solve_levels<true><<<b->ctx.W, SOLVE_THREADS, b->solve_lds, b->stream>>>(b->ctx, pool_ints);
  else
    k_solve_levels<false><<<b->ctx.W, SOLVE_THREADS, b->solve_lds, b->stream>>>(b->ctx, pool_ints);
}
