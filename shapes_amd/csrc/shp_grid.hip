// shp_grid.hip - uniform-grid broadphase for large worlds.
//
// Output contract (identical to the wavefront scan in shp_engine.hip and to
// culledKeys, shapes/src/Physics/Broadphase/Aabb.hs:168-183 / Grid.hs:67-86):
// for every body i of a world, all j < i of the same world with inclusive
// AABB overlap and not both static, segments in descending i, each segment
// in descending j.
//
// Each world has a cell size h >= the AABB extent of any of its "small"
// bodies (2 * the largest vertex radius of its dynamic bodies, plus a
// margin), fixed at creation.  Static bodies larger than that ("large",
// e.g. floors) are listed per world and tested explicitly.  A small body is
// binned by the cell of its AABB's min corner into a per-world hash table
// (counting sort, per-step), which stores a 48-byte record per body in
// bucket order (AABB, cell, index, static flag), so a bucket's candidates
// are read contiguously.  Any small j overlapping small i has its min
// corner in cells [cell(min_i) - 1, cell(max_i)] on each axis (at most 3x3
// cells), because max_j >= min_i and extent_j < h.  A large body i scans
// every j < i of its world.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "shp_engine.h"

__device__ __forceinline__ int cell_of(double x, double h) {
  double q = floor(x / h);
  if (!(q == q)) return 0;
  if (q > 1073741824.0) return 1073741824;
  if (q < -1073741824.0) return -1073741824;
  return (int)q;
}

// Bucket of cell (cx, cy): the low 16 bits of each coordinate interleaved
// (Morton order), masked to the world's table.  Nearby cells land in
// nearby buckets, so a body's 3x3 neighbourhood reads a few cache lines of
// the table; cells that alias share a bucket and are told apart by their
// stored cell coordinates.
__device__ __forceinline__ uint32_t spread_bits(uint32_t x) {
  x &= 0xFFFFu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ uint32_t cell_hash(int cx, int cy, uint32_t mask) {
  return (spread_bits((uint32_t)cx) | (spread_bits((uint32_t)cy) << 1)) & mask;
}

__device__ __forceinline__ bool aabb_hit(const Ctx& c, double aminx, double amaxx, double aminy,
                                         double amaxy, int gj) {
  double bminx = c.bminx[gj], bmaxx = c.bmaxx[gj];
  double bminy = c.bminy[gj], bmaxy = c.bmaxy[gj];
  return !(bminx > amaxx || bmaxx < aminx) && !(bminy > amaxy || bmaxy < aminy);
}

// gravity (applyExternal) + binning of small bodies; empty slots
// (World.delete) get neither
__global__ void k_grid_assign(Ctx c) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= c.NB || c.dead[g]) return;
  if (c.iml[g] != 0.0 && !c.probe) {
    c.vx[g] = c.vx[g] + c.gx * c.dt;
    c.vy[g] = c.vy[g] + c.gy * c.dt;
  }
  if (c.is_large[g]) return;
  int w = c.body_world[g];
  double h = c.cell_h[w];
  int cx = cell_of(c.bminx[g], h), cy = cell_of(c.bminy[g], h);
  c.body_cell[2 * g] = cx;
  c.body_cell[2 * g + 1] = cy;
  int t0 = c.tab_off[w];
  uint32_t mask = (uint32_t)(c.tab_off[w + 1] - t0 - 1);
  c.body_slot[g] = atomicAdd(&c.bucket_cnt[t0 + cell_hash(cx, cy, mask)], 1);
}

struct GridRec {
  double minx, maxx, miny, maxy;
  int cx, cy, g, st;
};

__device__ __forceinline__ GridRec load_rec(const Ctx& c, int q) {
  const double2* p = (const double2*)(c.grid_rec + 6L * q);
  const double2 x = p[0], y = p[1];
  const int4 m = *(const int4*)(p + 2);
  return GridRec{x.x, x.y, y.x, y.y, m.x, m.y, m.z, m.w};
}

__global__ void k_grid_fill(Ctx c) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= c.NB || c.is_large[g] || c.dead[g]) return;
  int w = c.body_world[g];
  int t0 = c.tab_off[w];
  uint32_t mask = (uint32_t)(c.tab_off[w + 1] - t0 - 1);
  const int cx = c.body_cell[2 * g], cy = c.body_cell[2 * g + 1];
  int b = t0 + cell_hash(cx, cy, mask);
  int pos = c.bucket_start[b] + c.body_slot[g];
  double2* p = (double2*)(c.grid_rec + 6L * pos);
  p[0] = make_double2(c.bminx[g], c.bmaxx[g]);
  p[1] = make_double2(c.bminy[g], c.bmaxy[g]);
  *(int4*)(p + 2) = make_int4(cx, cy, g, c.is_static[g] != 0);
}

// LANES lanes per body i (1, 4 or 16): count (fill = 0) or emit (fill = 1)
// its pairs.  With one lane (or four, which count only, splitting the
// cells), the count pass also keeps the first GRID_SLOTS partners, and the
// one-lane emit pass writes those without a second visit when they are
// all.  With 16, a small body's lanes 0..8 take one cell each of the
// <= 3x3 cells that can hold a partner's min corner and lanes 9..15 the
// world's large bodies, and a large body's lanes stride over every j < i:
// for batches too small to fill the GPU with one lane per body (the
// candidates are dependent loads).  With 1 lane per body the lane does all
// of it: for large batches, where the GPU is full anyway.  Each lane visits
// its candidates twice when emitting with 16 lanes (count, then write at
// the group's prefix sum), once with 1; lane 0 then sorts the segment into
// descending j.
template <int LANES, class F>
__device__ __forceinline__ void grid_visit(const Ctx& c, int g, int sl, F&& hit) {
  constexpr int CELL_LANES = LANES < 9 ? LANES : 9;
  if (c.dead[g]) return;  // an empty slot has no pairs; empty slots are not binned
  const int w = c.body_world[g];
  const int b0 = c.body_off[w];
  const bool si = c.is_static[g] != 0;
  const double aminx = c.bminx[g], amaxx = c.bmaxx[g], aminy = c.bminy[g], amaxy = c.bmaxy[g];
  const double h = c.cell_h[w];
  const int cx0 = cell_of(aminx, h) - 1, cx1 = cell_of(amaxx, h);
  const int cy0 = cell_of(aminy, h) - 1, cy1 = cell_of(amaxy, h);
  const long ncx = (long)cx1 - cx0 + 1, ncy = (long)cy1 - cy0 + 1;
  // A large body scans every j < i of its world.  So does a body whose AABB
  // is not finite (a small body's finite AABB spans at most 3 x 3 candidate
  // cells; a non-finite one would span up to 2^62 of them): the same
  // bounded work as a large body, by all lanes.
  const double span = (amaxx - aminx) + (amaxy - aminy);
  if (c.is_large[g] || !(span - span == 0.0) || ncx > 3 || ncy > 3) {
    for (int gj = g - 1 - sl; gj >= b0; gj -= LANES)
      if (!(si && c.is_static[gj]) && !c.dead[gj] && aabb_hit(c, aminx, amaxx, aminy, amaxy, gj))
        hit(gj);
    return;
  }
  // the world's large bodies: lanes 9.. with 16 lanes, else the last lane
  // (which also takes its share of the cells)
  constexpr bool LARGE_LANES = LANES > CELL_LANES;
  if (LARGE_LANES ? sl >= CELL_LANES : sl == LANES - 1) {
    const int l0 = LARGE_LANES ? sl - CELL_LANES : 0, ls = LARGE_LANES ? LANES - CELL_LANES : 1;
    for (int q = c.large_off[w] + l0; q < c.large_off[w + 1]; q += ls) {
      const int gj = c.large_ids[q];
      if (gj >= g) break;  // ascending
      if (!(si && c.is_static[gj]) && !c.dead[gj] && aabb_hit(c, aminx, amaxx, aminy, amaxy, gj))
        hit(gj);
    }
    if (LARGE_LANES) return;
  }
  const int t0 = c.tab_off[w];
  const uint32_t mask = (uint32_t)(c.tab_off[w + 1] - t0 - 1);
  for (long ci = sl; ci < ncx * ncy; ci += CELL_LANES) {
    const int cx = (int)(cx0 + ci % ncx), cy = (int)(cy0 + ci / ncx);
    const int b = t0 + cell_hash(cx, cy, mask);
    for (int q = c.bucket_start[b]; q < c.bucket_start[b + 1]; ++q) {
      const GridRec r = load_rec(c, q);
      if (r.g >= g || r.cx != cx || r.cy != cy) continue;
      if (!(si && r.st) && !(r.minx > amaxx || r.maxx < aminx) && !(r.miny > amaxy || r.maxy < aminy))
        hit(r.g);
    }
  }
}

template <int LANES>
__global__ void k_grid_pairs(Ctx c, int fill) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int g = (int)(t / LANES);  // the same for the LANES lanes of a group
  const int sl = (int)(t % LANES);
  if (g >= c.NB) return;
  const int w = c.body_world[g];
  const int b0 = c.body_off[w];
  const int nw = c.body_off[w + 1] - b0;
  const int rank = b0 + (nw - 1 - (g - b0));
  int base = 0, pos = 0, total = 0;
  int32_t* slots = c.grid_slots + (long)g * GRID_SLOTS;
  if (LANES == 1 && !fill) {
    int count = 0;
    grid_visit<LANES>(c, g, sl, [&](int gj) {
      if (count < GRID_SLOTS) slots[count] = gj;
      ++count;
    });
    c.cnt_rank[rank] = count;
    return;
  }
  if (LANES == 4) {  // count pass only: the cells split over 4 lanes
    int count = 0, kept[GRID_SLOTS];
    grid_visit<LANES>(c, g, sl, [&](int gj) {
#pragma unroll
      for (int k = 0; k < GRID_SLOTS; ++k)
        if (k == count) kept[k] = gj;
      ++count;
    });
    int incl = count;
    for (int o = 1; o < LANES; o <<= 1) {
      const int y = __shfl_up(incl, o, LANES);
      if (sl >= o) incl += y;
    }
    const int total = __shfl(incl, LANES - 1, LANES);
    // the first GRID_SLOTS partners of the body, lane by lane
    const int first = incl - count;
#pragma unroll
    for (int k = 0; k < GRID_SLOTS; ++k)
      if (k < count && first + k < GRID_SLOTS) slots[first + k] = kept[k];
    if (sl == 0) c.cnt_rank[rank] = total;
    return;
  }
  // Past the capacity the step reports an overflow; the slots below it are
  // still all written (a segment that straddles the capacity keeps its
  // first pairs), so no later kernel of the step reads an unwritten slot.
  if (LANES == 1) {
    // one lane: the count pass sized the segment
    base = c.off_rank[rank];
    total = c.cnt_rank[rank];
    if ((long)base + total > c.pair_cap) atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_PAIR_OVERFLOW);
    if (base >= c.pair_cap) return;
    total = min(total, c.pair_cap - base);
    if (total <= GRID_SLOTS && c.cnt_rank[rank] <= GRID_SLOTS) {
      // the kept partners, in descending order
      int v[GRID_SLOTS];
#pragma unroll
      for (int k = 0; k < GRID_SLOTS; ++k) v[k] = k < total ? slots[k] : -1;
#pragma unroll
      for (int a = 1; a < GRID_SLOTS; ++a)
#pragma unroll
        for (int q = a; q > 0; --q)
          if (v[q - 1] < v[q]) {
            const int t = v[q - 1];
            v[q - 1] = v[q];
            v[q] = t;
          }
#pragma unroll
      for (int k = 0; k < GRID_SLOTS; ++k)
        if (k < total) {
          c.pa[base + k] = g;
          c.pb[base + k] = v[k];
        }
      return;
    }
    pos = base;
    const int end = base + total;
    grid_visit<LANES>(c, g, sl, [&](int gj) {
      if (pos < end) {
        c.pa[pos] = g;
        c.pb[pos] = gj;
      }
      ++pos;
    });
  } else {
    int count = 0;
    grid_visit<LANES>(c, g, sl, [&](int) { ++count; });
    int incl = count;
    for (int o = 1; o < LANES; o <<= 1) {
      const int y = __shfl_up(incl, o, LANES);
      if (sl >= o) incl += y;
    }
    total = LANES == 1 ? incl : __shfl(incl, LANES - 1, LANES);
    if (!fill) {
      if (sl == 0) c.cnt_rank[rank] = total;
      return;
    }
    base = c.off_rank[rank];
    if ((long)base + total > c.pair_cap && sl == 0)
      atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_PAIR_OVERFLOW);
    if (base >= c.pair_cap) return;  // the same for the group's lanes
    total = min(total, c.pair_cap - base);
    const int end = base + total;
    pos = base + incl - count;
    grid_visit<LANES>(c, g, sl, [&](int gj) {
      if (pos < end) {
        c.pa[pos] = g;
        c.pb[pos] = gj;
      }
      ++pos;
    });
    // the group's stores precede lane 0's sort in this wavefront's program
    // order; the fence keeps the compiler from moving them
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  if (sl != 0) return;
  // the segment in descending j (insertion sort; segments are short)
  for (int a = base + 1; a < base + total; ++a) {
    const int v = c.pb[a];
    int q = a - 1;
    while (q >= base && c.pb[q] < v) {
      c.pb[q + 1] = c.pb[q];
      --q;
    }
    c.pb[q + 1] = v;
  }
}

// 16 lanes per body while that still fits about one wave of work per SIMD
// slot of the GPU several times over (batches up to 131072 bodies); beyond
// (the plan's grid_wide false), 4 lanes per body count and keep the first
// partners, one lane per body emits.
// The wide schedule gives each body 16 lanes in both passes.  Otherwise the
// count pass splits a body's cells over 4 lanes (their gathers overlap, and
// the pass is bound by their latency) and the emit pass takes 1 lane, which
// for most bodies only copies the partners the count pass kept.
static void launch_grid_pairs(shp_batch* b, int fill) {
  const Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  if (P.grid_wide)
    k_grid_pairs<16><<<P.g_pairs, 256, 0, b->stream>>>(c, fill);
  else if (!fill)
    k_grid_pairs<4><<<P.g_pairs_count, 256, 0, b->stream>>>(c, 0);
  else
    k_grid_pairs<1><<<P.g_pairs, 256, 0, b->stream>>>(c, 1);
}

// Grid broadphase, first half: binning and per-body pair counts into
// cnt_rank (the caller then scans cnt_rank into off_rank).
hipError_t launch_grid_count(shp_batch* b) {
  Ctx& c = b->ctx;
  hipStream_t s = b->stream;
  k_grid_assign<<<b->plan.g_bodies, 256, 0, s>>>(c);
  hipError_t e = scan_grid_buckets(b);
  if (e != hipSuccess) return e;
  k_grid_fill<<<b->plan.g_bodies, 256, 0, s>>>(c);
  launch_grid_pairs(b, 0);
  return hipGetLastError();
}

// Second half: emit the pairs at off_rank.
hipError_t launch_grid_emit(shp_batch* b) {
  launch_grid_pairs(b, 1);
  return hipGetLastError();
}
