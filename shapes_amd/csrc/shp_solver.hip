// shp_solver.hip - schedules for the contact solve.
//
// Ordered FP64 (drop-in) mode keeps the reference's sequential Gauss-Seidel
// order (improveWorld, shapes/src/Physics/Solvers/Contact.hs:146-157) bit for
// bit.  Contact c only depends on the latest earlier contact that touches
// one of its dynamic bodies, so each world's contacts are given a level
//   level(c) = 1 + max(level of earlier contacts sharing a dynamic body)
// and every level is solved in parallel: two contacts of one level touch
// disjoint dynamic bodies, so the order inside a level cannot change any
// bit.  Static bodies (both inverse masses 0) are read-only.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_solve_common.h"

using namespace shp;

#define LEVEL_CAP 2048      // levels per world handled by the parallel path
#define LAST_LDS 10240      // bodies per world whose last-level fits in LDS
#define SOLVE_THREADS 256

// Wave 0 assigns levels to the world's contacts [c0, c1), 64 at a time.
// `last` holds, per local body, the level of the latest contact touching
// it.  Within a chunk, lane k is raised above every earlier lane j that
// shares a dynamic body (lane j's level is final when it is broadcast);
// only the last lane touching a body updates `last`.  Returns the maximum
// level (valid in all lanes of wave 0).
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
      int aj = __shfl(a, j);
      int bj = __shfl(b, j);
      int lj = __shfl(lvl, j);
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

__global__ __launch_bounds__(SOLVE_THREADS) void k_solve_levels(Ctx c) {
  __shared__ int s_last[LAST_LDS];
  __shared__ int s_off[LEVEL_CAP + 2];
  __shared__ int s_cur[LEVEL_CAP + 2];
  __shared__ int s_maxl;
  int w = blockIdx.x;
  int tid = threadIdx.x;
  int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  if (nb > LAST_LDS) {  // world too large for the LDS schedule: exact serial
    if (tid == 0) solve_world_serial(c, c0, c1);
    return;
  }
  for (int i = tid; i < nb; i += SOLVE_THREADS) s_last[i] = 0;
  __syncthreads();
  if (tid < 64) {
    int m = assign_levels(c, c0, c1, b0, s_last);
    if (tid == 0) {
      s_maxl = m;
      atomicMax(&c.counters[CNT_LEVELS], m);
    }
  }
  __syncthreads();
  int maxl = s_maxl;
  if (maxl > LEVEL_CAP) {
    if (tid == 0) {
      atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_SCHEDULE_OVERFLOW);
      solve_world_serial(c, c0, c1);
    }
    return;
  }
  // bucket the contacts by level (counting sort; order inside a level is
  // irrelevant)
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
  }
  __syncthreads();
  // useCache applications (applyCachedSlns, Solvers/Contact.hs:100-114)
  for (int l = 1; l <= maxl; ++l) {
    int e = s_off[l + 1];
    for (int i = s_off[l] + tid; i < e; i += SOLVE_THREADS) {
      int k = c.sched_order[c0 + i];
      if (c.warm[k]) warm_apply(c, k);
    }
    __syncthreads();
  }
  // improveWorld sweeps
  for (int it = 0; it < c.iterations; ++it) {
    for (int l = 1; l <= maxl; ++l) {
      int e = s_off[l + 1];
      for (int i = s_off[l] + tid; i < e; i += SOLVE_THREADS) improve(c, c.sched_order[c0 + i]);
      __syncthreads();
    }
  }
}

shp_status solver_alloc(shp_batch* b) {
  Ctx& c = b->ctx;
  long CC = c.contact_cap;
  if (hipMalloc(&c.sched_level, CC * sizeof(int32_t)) != hipSuccess) return SHP_E_OOM;
  b->allocs.push_back(c.sched_level);
  if (hipMalloc(&c.sched_order, CC * sizeof(int32_t)) != hipSuccess) return SHP_E_OOM;
  b->allocs.push_back(c.sched_order);
  float** f[3] = {&c.fvx, &c.fvy, &c.fw};
  for (int k = 0; k < 3; ++k) {
    if (hipMalloc(f[k], (size_t)c.NB * sizeof(float)) != hipSuccess) return SHP_E_OOM;
    b->allocs.push_back(*f[k]);
  }
  return SHP_OK;
}

void launch_ordered_solver(shp_batch* b) {
  k_solve_levels<<<b->ctx.W, SOLVE_THREADS, 0, b->stream>>>(b->ctx);
}

