// shp_engine.hip - gfx950 kernels and C ABI for the shapes contact step.
//
// One step (the reference's updateWorld, shapes/src/Physics/Engine/
// Main.hs:71-86) is this sequence of launches on the batch's stream:
//
//   k_reset          clear per-step counters
//   k_broadphase(0)  one wavefront per body: gravity (applyExternal) and the
//                    count of its pairs (i, j<i); or the grid's binning and
//                    count pass (shp_grid.hip)
//   scan             exclusive sum of the counts, in descending body order
//   k_broadphase(1)  same sweep, ballot/popcount compaction -> pairs in
//                    descending (i, j) order (G.culledKeys); or the grid's
//                    emit pass
//   k_sat            one lane per pair: SAT + clipping -> <= 2 contacts
//   scan             exclusive sum of contacts per pair
//   k_world_ranges   per-world pair and contact ranges
//   k_contacts       compaction, constraintGen, warm-start join; for the
//                    coloured solver, after its k_colour (on a second
//                    stream), also each manifold's record in colour order
//   solve            warm-start application + Gauss-Seidel sweeps
//                    (k_solve_flow, or k_sweep_colours + k_colour_unpack)
//   k_finish         advance + moveShapes + AABBs (k_integrate's tiles) and
//                    the solve quality maxima, in one launch
//   k_cache_store    this frame's (key, lambda) become next frame's cache,
//                    quality run sums
//
// Worlds are independent: every kernel handles all worlds of the batch at
// once, and the ordered solver runs one workgroup per world.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/shapes_gpu.h"
#include "shp_geom.h"
#include "shp_trig.h"
#include "shp_engine.h"
#include "shp_records.h"
#include "shp_solve_common.h"

using namespace shp;

// ---------------------------------------------------------------- helpers


// Key = ObjectFeatureKey ((i, j), (fi, fj)) (Constraints/Contact.hs:36-39)
// packed so that unsigned order equals the derived lexicographic order.
__device__ __forceinline__ uint64_t pack_key(int i, int j, int fi, int fj) {
  return ((uint64_t)(uint32_t)i << 40) | ((uint64_t)(uint32_t)j << 16) |
         ((uint64_t)(uint32_t)fi << 8) | (uint64_t)(uint32_t)fj;
}

__device__ __forceinline__ HullRef hull_of(const Ctx& c, int g) {
  int o = c.vert_off[g];
  HullRef h;
  h.vn = c.wv + o;
  h.sx = c.wext + o;
  h.n = c.vert_off[g + 1] - o;
  return h;
}

// ---------------------------------------------------------------- kernels

// Per-step counters and quality maxima, and the grid's bucket counts.
__global__ void k_reset(Ctx c) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < 8) c.counters[g] = 0;
  if (g < 3) c.quality[g] = 0ull;
  if (c.use_grid && g <= c.grid_buckets) c.bucket_cnt[g] = 0;
}

// listToHull in local space (ConvexHull.hs:151-167): unit edge normals and
// the per-edge (min, max) extent vertex indices, first index on ties.
__global__ void k_init_hulls(Ctx c) {
  int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= c.NB) return;
  int o = c.vert_off[g], n = c.vert_off[g + 1] - o;
  for (int e = 0; e < n; ++e) {
    int e1 = e < n - 1 ? e + 1 : 0;
    d2 dir = normalize(clockwise(sub(mk(c.lx[o + e1], c.ly[o + e1]), mk(c.lx[o + e], c.ly[o + e]))));
    double dmin = dot(mk(c.lx[o], c.ly[o]), dir), dmax = dmin;
    int imin = 0, imax = 0;
    for (int k = 1; k < n; ++k) {
      double d = dot(mk(c.lx[o + k], c.ly[o + k]), dir);
      if (d < dmin) {
        dmin = d;
        imin = k;
      }
      if (d > dmax) {
        dmax = d;
        imax = k;
      }
    }
    c.ext[o + e] = (uint16_t)(imin | (imax << 8));
  }
}

// advance (World.hs:167-169, Constraint.hs:225-229) when `advance`, then
// moveShapes (World.hs:136-140): T(pos) R(theta) applied to the local
// vertices (Transform.hs:34-38, Linear.hs:217-220), unit edge normals
// (ConvexHull.hs:184-195) and the AABB fold (Broadphase/Aabb.hs:81-110).
// With `advance`, empty slots are skipped, as E.mapM_ skips them
// (World.hs:137,168): their state and shape stay as they were when the slot
// was emptied.  Without (placement at creation or after a mutation), every
// slot is placed from its state, which for an empty slot reproduces the
// shape it was left with.
// A workgroup takes 256 consecutive bodies: one lane per body builds the
// transform, one lane per vertex transforms it into LDS, then one lane per
// vertex writes the vertex, its edge's normal and the hull's extent along
// that normal (coalesced), and one lane per body folds its AABB in vertex
// order from LDS.
constexpr int INTEGRATE_THREADS = 256;
constexpr int INTEGRATE_VERTS = INTEGRATE_THREADS * SHP_MAXV;

// The bodies [vb * INTEGRATE_THREADS, ...) of the batch (one workgroup).
__device__ __forceinline__ void integrate_tile(const Ctx& c, int advance, int vb) {
  __shared__ double s_m[6][INTEGRATE_THREADS];
  __shared__ double s_wx[INTEGRATE_VERTS], s_wy[INTEGRATE_VERTS];
  __shared__ uint8_t s_body[INTEGRATE_VERTS];
  __shared__ int s_voff[INTEGRATE_THREADS + 1];
  __shared__ uint8_t s_dead[INTEGRATE_THREADS];
  const int t = threadIdx.x;
  const int g0 = vb * INTEGRATE_THREADS;
  const int nbk = min(INTEGRATE_THREADS, c.NB - g0);
  const int g = g0 + t;
  const double one = 1.0, zer = 0.0;
  if (t < nbk) {
    const bool dead = c.dead[g] != 0;
    s_dead[t] = dead && advance;
    double px = c.px[g], py = c.py[g], th = c.th[g];
    if (advance && !dead) {
      double dt = c.dt;
      px = (dt * c.vx[g]) + px;
      py = (dt * c.vy[g]) + py;
      th = (dt * c.w[g]) + th;
      c.px[g] = px;
      c.py[g] = py;
      c.th[g] = th;
    }
    double s, co;
    shp_sincos(th, &s, &co);
    s_m[0][t] = (one * co + zer * s) + px * zer;
    s_m[1][t] = (one * (-s) + zer * co) + px * zer;
    s_m[2][t] = (one * zer + zer * zer) + px * one;
    s_m[3][t] = (zer * co + one * s) + py * zer;
    s_m[4][t] = (zer * (-s) + one * co) + py * zer;
    s_m[5][t] = (zer * zer + one * zer) + py * one;
    s_voff[t] = c.vert_off[g];
    if (t == nbk - 1) s_voff[nbk] = c.vert_off[g + 1];
    double chk = px + py + th + c.vx[g] + c.vy[g] + c.w[g];
    if (!dead && !(chk - chk == 0.0)) atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_NONFINITE);
  }
  __syncthreads();
  const int v0 = s_voff[0], nv = s_voff[nbk] - v0;
  if (t < nbk)
    for (int i = s_voff[t] - v0; i < s_voff[t + 1] - v0; ++i) s_body[i] = (uint8_t)t;
  __syncthreads();
  for (int i = t; i < nv; i += INTEGRATE_THREADS) {
    const int b = s_body[i];
    if (s_dead[b]) continue;
    const double lx = c.lx[v0 + i], ly = c.ly[v0 + i];
    s_wx[i] = (s_m[0][b] * lx + s_m[1][b] * ly) + s_m[2][b] * one;
    s_wy[i] = (s_m[3][b] * lx + s_m[4][b] * ly) + s_m[5][b] * one;
  }
  __syncthreads();
  // From the world vertices in LDS: each edge's unit normal and the hull's
  // own extent along it (extentAlong, ConvexHull.hs:81-92, at the cached
  // extent vertices), the projections the SAT reads, once per step here.
  for (int i = t; i < nv; i += INTEGRATE_THREADS) {
    const int b = s_body[i];
    if (s_dead[b]) continue;
    const int ex = c.ext[v0 + i];
    const int o = s_voff[b] - v0, last = s_voff[b + 1] - v0 - 1;
    const int i1 = i < last ? i + 1 : o;
    const double x = s_wx[i], y = s_wy[i];
    const d2 nn = normalize(clockwise(sub(mk(s_wx[i1], s_wy[i1]), mk(x, y))));
    const int lo = o + (ex & 255), hi = o + (ex >> 8);
    // vertex and edge normal in one 32-byte record, the extent in 16 (coalesced)
    c.wv[v0 + i] = make_double4(x, y, nn.x, nn.y);
    c.wext[v0 + i] = make_double2(dot(mk(s_wx[lo], s_wy[lo]), nn), dot(mk(s_wx[hi], s_wy[hi]), nn));
  }
  if (t < nbk) {
    if (s_dead[t]) return;
    const double r = c.rad[g];
    if (r > 0.0) {
      // circleToAabb (Aabb.hs:86-88) around the centre, which is the
      // body's position (setCircleTransform maps the local origin there:
      // (m00*0 + m01*0) + m02*1 = px, positions never being -0)
      const double x = c.px[g], y = c.py[g];
      c.bminx[g] = x - r;
      c.bmaxx[g] = x + r;
      c.bminy[g] = y - r;
      c.bmaxy[g] = y + r;
      return;
    }
    const int o = s_voff[t] - v0, e = s_voff[t + 1] - v0;
    if (e == o) return;  // no shape: a slot never filled since creation
    double bminx = s_wx[o], bmaxx = s_wx[o], bminy = s_wy[o], bmaxy = s_wy[o];
    for (int i = o + 1; i < e; ++i) {
      const double x = s_wx[i], y = s_wy[i];
      bminx = bminx < x ? bminx : x;
      bmaxx = bmaxx > x ? bmaxx : x;
      bminy = bminy < y ? bminy : y;
      bmaxy = bmaxy > y ? bmaxy : y;
    }
    c.bminx[g] = bminx;
    c.bmaxx[g] = bmaxx;
    c.bminy[g] = bminy;
    c.bmaxy[g] = bmaxy;
  }
}

// Per contact after the solve (FP64 state): penetration depth, normal
// complementarity residual |min(lambda_n, Jn.v + b)| and friction residual
// |lambda_t - clamp(lambda_t - Jt.v / mt, +-mu*lambda_n)|; maxima of the step
// and of the run.  Non-negative doubles order like their bit patterns, so the
// maxima are integer maxima, reduced per workgroup before one atomic each.
// Workgroup vb of nvb (grid-stride over the step's contacts).
#define QUALITY_THREADS 256

__device__ __forceinline__ unsigned long long nonneg_bits(double x) {
  if (!(x >= 0.0)) x = x != x ? 1.0e300 : 0.0;  // NaN counts as huge
  return (unsigned long long)__double_as_longlong(x);
}

__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(x, off);
    x = o > x ? o : x;
  }
  return x;
}

__device__ __forceinline__ void quality_tile(const Ctx& c, int vb, int nvb) {
  int tc = c.wcont_off[c.W];
  unsigned long long qd = 0, qn = 0, qt = 0;
  for (int k = vb * QUALITY_THREADS + threadIdx.x; k < tc; k += nvb * QUALITY_THREADS) {
    int ga = c.ga[k], gb = c.gb[k];
    double v[6], jn[6], jt[6];
    load_vel(c, ga, gb, v);
    load_rows(c, k, jn, jt);
    d2 ra = mk(c.rax[k], c.ray[k]), rb = mk(c.rbx[k], c.rby[k]);
    d2 nr = c.flip[k] ? mk(-c.cnx[k], -c.cny[k]) : mk(c.cnx[k], c.cny[k]);
    d2 closing = add(add(add(neg(mk(v[0], v[1])), zcross(-v[2], ra)), mk(v[3], v[4])), zcross(v[5], rb));
    double e = hs_min(c.e[ga], c.e[gb]);
    double bn = c.bn[k] + hs_min(0.0, e * dot(closing, nr));
    double cn = dot6(jn, v) + bn;
    double ln = c.ln[k], lt = c.lt[k];
    double rn = fabs(hs_min(ln, cn));
    double mu = (c.mu[ga] + c.mu[gb]) / 2.0;
    double thr = ln * mu;
    double t = lt - dot6(jt, v) / c.mct[k];
    double tcl = t > thr ? thr : (t < -thr ? -thr : t);
    double rt = fabs(lt - tcl);
    unsigned long long a = nonneg_bits(c.depth[k]), b = nonneg_bits(rn), d = nonneg_bits(rt);
    qd = a > qd ? a : qd;
    qn = b > qn ? b : qn;
    qt = d > qt ? d : qt;
  }
  // wavefront, then workgroup maxima: one set of atomics per workgroup
  __shared__ unsigned long long s_q[3][QUALITY_THREADS / 64];
  const int wid = threadIdx.x >> 6;
  qd = wave_max_u64(qd);
  qn = wave_max_u64(qn);
  qt = wave_max_u64(qt);
  if ((threadIdx.x & 63) == 0) {
    s_q[0][wid] = qd;
    s_q[1][wid] = qn;
    s_q[2][wid] = qt;
  }
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int q = 1; q < QUALITY_THREADS / 64; ++q) {
    qd = s_q[0][q] > qd ? s_q[0][q] : qd;
    qn = s_q[1][q] > qn ? s_q[1][q] : qn;
    qt = s_q[2][q] > qt ? s_q[2][q] : qt;
  }
  if (qd | qn | qt) {
    atomicMax(&c.quality[Q_DEPTH], qd);
    atomicMax(&c.quality[Q_NRES], qn);
    atomicMax(&c.quality[Q_FRES], qt);
    atomicMax(&c.quality[Q_RUN_DEPTH], qd);
    atomicMax(&c.quality[Q_RUN_NRES], qn);
    atomicMax(&c.quality[Q_RUN_FRES], qt);
  }
}

__global__ __launch_bounds__(INTEGRATE_THREADS) void k_integrate(Ctx c, int advance) {
  integrate_tile(c, advance, blockIdx.x);
}

// The end of a step in one launch: workgroups [0, ni) advance the bodies
// (k_integrate's tiles, first so that the longer chain starts first), the
// rest compute the solve quality.  The two read the solved velocities and
// write disjoint data (positions, transforms and AABBs; quality maxima).
static_assert(QUALITY_THREADS == INTEGRATE_THREADS, "one workgroup size");
__global__ __launch_bounds__(INTEGRATE_THREADS) void k_finish(Ctx c, int ni) {
  if ((int)blockIdx.x < ni)
    integrate_tile(c, 1, blockIdx.x);
  else
    quality_tile(c, blockIdx.x - ni, gridDim.x - ni);
}

// Broadphase: one wavefront per body i scans j = i-1 .. 0 of its world in
// chunks of 64 lanes (culledKeys, Broadphase/Aabb.hs:168-183, which is the
// pair set the grid of Grid.hs:67-86 returns).  Pass 0 counts and applies
// the constant acceleration (applyExternal, World.hs:156-158 with
// constantAccel, World/External.hs:23-28); pass 1 writes the pairs, in
// descending j within descending i, using ballot + prefix popcount.
__global__ void k_broadphase(Ctx c, int fill) {
  int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  int lane = threadIdx.x & 63;
  if (wave >= c.NB) return;
  int g = wave;
  int w = c.body_world[g];
  int b0 = c.body_off[w];
  int li = g - b0;
  int nw = c.body_off[w + 1] - b0;
  int rank = b0 + (nw - 1 - li);
  // an empty slot (World.delete) gets no acceleration and no pairs
  const bool dead = c.dead[g] != 0;
  if (!fill && lane == 0 && c.iml[g] != 0.0 && !dead && !c.probe) {
    c.vx[g] = c.vx[g] + c.gx * c.dt;
    c.vy[g] = c.vy[g] + c.gy * c.dt;
  }
  bool si = body_static(c, g);
  double aminx = c.bminx[g], amaxx = c.bmaxx[g], aminy = c.bminy[g], amaxy = c.bmaxy[g];
  int count = 0;
  int base = fill ? c.off_rank[rank] : 0;
  for (int j0 = dead ? -1 : li - 1; j0 >= 0; j0 -= 64) {
    int lj = j0 - lane;
    bool hit = false;
    if (lj >= 0) {
      int gj = b0 + lj;
      if (!(si && body_static(c, gj)) && !c.dead[gj]) {
        // aabbCheck (Aabb.hs:69-78), inclusive
        double bminx = c.bminx[gj], bmaxx = c.bmaxx[gj];
        double bminy = c.bminy[gj], bmaxy = c.bmaxy[gj];
        hit = !(bminx > amaxx || bmaxx < aminx) && !(bminy > amaxy || bmaxy < aminy);
      }
    }
    unsigned long long m = __ballot(hit);
    if (fill && hit) {
      int pos = base + count + __popcll(m & ((1ull << lane) - 1ull));
      if (pos < c.pair_cap) {
        c.pa[pos] = g;
        c.pb[pos] = b0 + lj;
      }
    }
    count += __popcll(m);
  }
  if (!fill && lane == 0) c.cnt_rank[rank] = count;
  if (fill && lane == 0 && base + count > c.pair_cap)
    atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_PAIR_OVERFLOW);
}

__device__ __forceinline__ int total_pairs(const Ctx& c) {
  int t = c.off_rank[c.NB];
  return t < c.pair_cap ? t : c.pair_cap;
}

// SAT narrowphase, one lane per pair (generateContacts, Contact.hs:22-40,
// for two hulls).  Writes up to two contacts into the pair's slots.
// Contact.generateContacts (Contact.hs:22-40) for shapes (i, j) = (pa, pb)
// when one is a circle: circle-circle keyed (0, 0) Same; circle-hull
// (0, f) Same; hull-circle (f, 0) Flip, computed circle against hull.
__device__ __forceinline__ int circle_pair(const Ctx& c, int ga, int gb, double ra, double rb,
                                           Cgeom* gm, int* flip) {
  const d2 ca = mk(c.px[ga], c.py[ga]), cb = mk(c.px[gb], c.py[gb]);
  *flip = 0;
  if (ra > 0.0 && rb > 0.0) {
    if (!circle_contact(ca, ra, cb, rb, gm)) return 0;
    gm->fa = 0;
    gm->fb = 0;
    return 1;
  }
  // one of them is a circle: the circle against the other's hull
  const bool a_circle = ra > 0.0;
  const HullRef hr = hull_of(c, a_circle ? gb : ga);
  HullV hv;
  int f = 0;
  load_hull_v(hr, hv);
  if (!circle_hull_contact(a_circle ? ca : cb, a_circle ? ra : rb, hv, hr, &f, gm)) return 0;
  gm->fa = a_circle ? 0 : f;
  gm->fb = a_circle ? f : 0;
  *flip = a_circle ? 0 : 1;
  return 1;
}

__device__ __forceinline__ void sat_pair(const Ctx& c, int p) {
  const int pga = c.pa[p], pgb = c.pb[p];
  const double ra = c.rad[pga], rb = c.rad[pgb];
  Cgeom gm[2];
  int flip = 0;
  int n;
  if (ra > 0.0 || rb > 0.0) {
    n = circle_pair(c, pga, pgb, ra, rb, gm, &flip);
  } else {
    HullRef A = hull_of(c, pga);
    HullRef B = hull_of(c, pgb);
    n = hull_contacts(A, B, gm, &flip);
  }
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k >= n) break;
    int s = 2 * p + k;
    c.s_feat[s] = (uint16_t)((gm[k].fa << 8) | gm[k].fb);
    c.s_flip[s] = (uint8_t)flip;
    c.s_nx[s] = gm[k].nx;
    c.s_ny[s] = gm[k].ny;
    c.s_px[s] = gm[k].px;
    c.s_py[s] = gm[k].py;
    c.s_depth[s] = gm[k].depth;
  }
  c.ccount[p] = n;
  if (c.solver == SHP_SOLVER_COLORED_F32) {  // the pair's word for the colouring
    const int ga = c.pa[p], gb = c.pb[p];
    uint32_t v = NO_PAIR;
    if (n > 0) {
      const int b0 = c.body_off[c.body_world[ga]];
      const uint32_t a = c.is_static[ga] ? NO_BODY : (uint32_t)(ga - b0) & 0xFFFFu;
      const uint32_t b = c.is_static[gb] ? NO_BODY : (uint32_t)(gb - b0) & 0xFFFFu;
      v = a | (b << 16);
    }
    c.pair_ab[p] = v;
  }
}

// grid-stride over this step's pairs (the launch is sized for the capacity);
// at most 128 VGPRs, so that four waves per SIMD overlap their hull loads
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_sat(Ctx c) {
  int tp = total_pairs(c);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tp; p += gridDim.x * blockDim.x) sat_pair(c, p);
}


// Per-world pair and contact ranges after the scans.
__global__ void k_world_ranges(Ctx c) {
  int w = blockIdx.x * blockDim.x + threadIdx.x;
  if (w > c.W) return;
  int ps = c.off_rank[c.body_off[w]];
  if (ps > c.pair_cap) ps = c.pair_cap;
  int cs = c.coff[ps];
  if (cs > c.contact_cap) cs = c.contact_cap;
  c.wpair_off[w] = ps;
  c.wcont_off[w] = cs;
  if (w == c.W && c.coff[ps] > c.contact_cap) atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_CONTACT_OVERFLOW);
}

// Compaction + constraintGen (Constraints/Contact.hs:60-72) + the warm
// start join of applyCachedSlns (Solvers/Contact.hs:74-121): a binary search
// of the key in last frame's cache of the same world, which is strictly
// descending, finds exactly the element descZipVector (Utils/Descending.hs:
// 47-71) would pair it with.  For the coloured solver, the pair's manifold
// record too, at the slot the colouring gave it (none: no slot).
__device__ void contacts_of_pair(const Ctx& c, int p, int n, int& hits) {
  int ga = c.pa[p], gb = c.pb[p];
  int w = c.body_world[ga];
  int b0 = c.body_off[w];
  d2 xa = mk(c.px[ga], c.py[ga]), xb = mk(c.px[gb], c.py[gb]);
  double im[6] = {c.iml[ga], c.iml[ga], c.imr[ga], c.iml[gb], c.iml[gb], c.imr[gb]};
  int cache_lo = c.cache_woff[w], cache_hi = c.cache_woff[w + 1];
  int c_lo = c.wcont_off[w], c_hi = c.wcont_off[w + 1];
  const int M = c.records_in_contacts ? c.sched_order[p] : -1;
  const int k0 = c.coff[p];
  const PackF P = packf(c);
  for (int k = 0; k < n; ++k) {
    int dst = k0 + k;
    if (dst >= c.contact_cap) continue;
    int s = 2 * p + k;
    int fa = c.s_feat[s] >> 8, fb = c.s_feat[s] & 255;
    int flip = c.s_flip[s];
    uint64_t key = pack_key(ga - b0, gb - b0, fa, fb);
    d2 nn = mk(c.s_nx[s], c.s_ny[s]);
    d2 pp = mk(c.s_px[s], c.s_py[s]);
    double depth = c.s_depth[s];
    c.key[dst] = key;
    c.ga[dst] = ga;
    c.gb[dst] = gb;
    c.flip[dst] = (uint8_t)flip;
    c.cnx[dst] = nn.x;
    c.cny[dst] = nn.y;
    c.cpx[dst] = pp.x;
    c.cpy[dst] = pp.y;
    c.depth[dst] = depth;
    double jn[6], jt[6];
    if (!flip) {
      np_jacobian(nn, pp, xa, xb, jn);
      f_jacobian(nn, pp, xa, xb, jt);
    } else {  // flipMap + flipExtract: swapped pair, halves swapped back
      double t[6];
      np_jacobian(nn, pp, xb, xa, t);
      for (int q = 0; q < 3; ++q) {
        jn[q] = t[q + 3];
        jn[q + 3] = t[q];
      }
      f_jacobian(nn, pp, xb, xa, t);
      for (int q = 0; q < 3; ++q) {
        jt[q] = t[q + 3];
        jt[q + 3] = t[q];
      }
    }
    // the other eight entries of the rows are +-n and +-cw(n), rebuilt
    // exactly from the normal and flip (row_normal)
    c.jx[dst] = make_double4(jn[2], jn[5], jt[2], jt[5]);
    // baumgarte (NonPenetration.hs:48-56)
    const double bn = depth > c.slop ? (c.beta / c.dt) * (c.slop - depth) : 0.0;
    c.bn[dst] = bn;
    // Restitution.constraintGen (Restitution.hs:21-31)
    d2 ra = sub(pp, xa), rb = sub(pp, xb);
    c.rax[dst] = ra.x;
    c.ray[dst] = ra.y;
    c.rbx[dst] = rb.x;
    c.rby[dst] = rb.y;
    const double mcn = eff_mass(jn, im), mct = eff_mass(jt, im);
    c.mcn[dst] = mcn;
    c.mct[dst] = mct;
    if (M >= 0) write_record_rows(P, M, k, im[2], im[5], jn[2], jn[5], jt[2], jt[5], bn, mcn, mct);
    // warm start lookup in the descending cache segment [cache_lo,
    // cache_hi): the first entry <= key, then equality.  This frame's
    // contacts come in the same order and mostly persist, so the search
    // starts at the contact's proportional position and gallops outwards.
    int found = -1;
    if (cache_hi > cache_lo) {
      const int nc = cache_hi - cache_lo, nw = c_hi - c_lo;
      int g = cache_lo + (int)((long)(dst - c_lo) * nc / (nw > 0 ? nw : 1));
      g = g < cache_lo ? cache_lo : (g >= cache_hi ? cache_hi - 1 : g);
      int L, R;  // the first entry <= key lies in [L, R] (R = cache_hi: none)
      if (c.cache_key[g] > key) {
        L = g + 1;
        R = cache_hi;
        for (int step = 1;; step <<= 1) {
          const int t = g + step;
          if (t >= cache_hi) break;
          if (c.cache_key[t] <= key) {
            R = t;
            break;
          }
          L = t + 1;
        }
      } else {
        L = cache_lo;
        R = g;
        for (int step = 1;; step <<= 1) {
          const int t = g - step;
          if (t < cache_lo) break;
          if (c.cache_key[t] > key) {
            L = t + 1;
            break;
          }
          R = t;
        }
      }
      while (L < R) {
        const int mid = (L + R) >> 1;
        if (c.cache_key[mid] > key) L = mid + 1;
        else R = mid;
      }
      if (R < cache_hi && c.cache_key[R] == key) found = R;
    }
    double ln = 0.0, lt = 0.0;
    if (found >= 0) {
      ln = c.cache_ln[found];
      lt = c.cache_lt[found];
    }
    if (M >= 0) write_record_lambda(P, M, k, ln, lt);
    if (found >= 0) {
      c.ln[dst] = ln;
      c.lt[dst] = lt;
      c.warm[dst] = 1;
      ++hits;
    } else {
      // newCache (Solvers/Contact.hs:91-99): zero start; the learned warm
      // start may replace it before the solve (and patch the record)
      c.ln[dst] = 0.0;
      c.lt[dst] = 0.0;
      c.warm[dst] = 0;
      if (c.collect_misses) {
        // one atomic per wavefront for the lanes that miss together
        const unsigned long long act = __ballot(1);
        const int lane = threadIdx.x & 63, leader = __ffsll((long long)act) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(&c.counters[CNT_MISS], (int)__popcll(act));
        const int slot = __shfl(base, leader) + (int)__popcll(act & ((1ull << lane) - 1ull));
        if (slot < c.miss_cap) {
          c.miss_idx[slot] = dst;
          c.miss_rec[slot] = M >= 0 ? (M << 1) | k : -1;
        }
        // the features here only for the training log (which reads them
        // after the solve); else the inference computes them from the
        // stored contact, off this kernel's divergent path
        if (slot < c.miss_cap && c.features_in_contacts) {
          double v[6] = {c.vx[ga], c.vy[ga], c.w[ga], c.vx[gb], c.vy[gb], c.w[gb]};
          double sgn = flip ? -1.0 : 1.0;
          mlp_features(sgn * nn.x, sgn * nn.y, depth, v, ra.x, ra.y, rb.x, rb.y, c.iml[ga], c.iml[gb],
                       c.imr[ga], c.imr[gb], (c.mu[ga] + c.mu[gb]) / 2.0, c.miss_feat + (size_t)slot * MLP_IN);
        }
      }
    }
  }
  if (M >= 0) {  // the manifold's words, from its first contact as stored above
    if (k0 >= c.contact_cap) {
      write_null_record(P, M, ga, gb);
    } else {
      const int nk = n < c.contact_cap - k0 ? n : c.contact_cap - k0;
      const int warm = (c.warm[k0] ? 1 : 0) | (nk > 1 && c.warm[k0 + 1] ? 2 : 0);
      write_record_head(c, P, M, ga, gb, row_normal(c, k0), im[0], im[3], nk, warm);
    }
  }
}


// 256-lane workgroups: 144 VGPRs and three waves per SIMD without spills
// (the default bound of 1024 lanes held it to 128 with 20 bytes per lane in
// scratch; a 1024-pile step 4.81 against 4.83 ms in this kernel)
__global__ __launch_bounds__(256) void k_contacts(Ctx c) {
  int tp = total_pairs(c);
  int hits = 0, mans = 0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tp; p += gridDim.x * blockDim.x) {
    int n = c.ccount[p];
    if (n > 0) {
      contacts_of_pair(c, p, n, hits);
      mans += c.coff[p] < c.contact_cap;  // a manifold with stored contacts
    }
  }
  // one atomic per workgroup for the warm-start hit and manifold counters
  __shared__ int s_cnt[2];
  if (threadIdx.x < 2) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) {
    hits += __shfl_xor(hits, off);
    mans += __shfl_xor(mans, off);
  }
  if ((threadIdx.x & 63) == 0 && hits) atomicAdd(&s_cnt[0], hits);
  if ((threadIdx.x & 63) == 0 && mans) atomicAdd(&s_cnt[1], mans);
  __syncthreads();
  if (threadIdx.x == 0 && s_cnt[0]) atomicAdd(&c.counters[CNT_WARM], s_cnt[0]);
  if (threadIdx.x == 0 && s_cnt[1]) atomicAdd(&c.counters[CNT_MANIFOLDS], s_cnt[1]);
}

// This frame's lagrangians become the next frame's EngineCache
// (Engine/Main.hs:32, wrapInitializer :61-69).
__global__ void k_cache_store(Ctx c) {
  // the step's quality maxima are final (k_finish ran before): add them to
  // the run sums, whose means over the run are less sensitive to a single
  // impact than the run maxima
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    for (int k = 0; k < 3; ++k) {
      const double m = __longlong_as_double((long long)c.quality[Q_DEPTH + k]);
      const double s = __longlong_as_double((long long)c.quality[Q_SUM_DEPTH + k]);
      c.quality[Q_SUM_DEPTH + k] = (unsigned long long)__double_as_longlong(s + m);
    }
    c.quality[Q_STEPS] += 1;
  }
  int stride = gridDim.x * blockDim.x;
  int tc = c.wcont_off[c.W];
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k <= c.W || k < tc; k += stride) {
    if (k <= c.W) c.cache_woff[k] = c.wcont_off[k];
    if (k < tc) {
      c.cache_key[k] = c.key[k];
      c.cache_ln[k] = c.ln[k];
      c.cache_lt[k] = c.lt[k];
    }
  }
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// ---------------------------------------------------------------- host side

static thread_local std::string g_err;

void shp_set_error(const std::string& msg) { g_err = msg; }

static shp_status fail(shp_status s, const std::string& msg) {
  g_err = msg;
  return s;
}

#define HIPCHK(expr)                                                            \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess)                                                       \
      return fail(SHP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

shp_status dalloc_bytes(shp_batch* b, void** p, size_t bytes) {
  void* q = nullptr;
  if (bytes < 16) bytes = 16;
  if (hipMalloc(&q, bytes) != hipSuccess)
    return fail(SHP_E_OOM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
  b->allocs.push_back(q);
  b->device_bytes += (long)bytes;
  *p = q;
  return SHP_OK;
}

template <typename T>
static shp_status dalloc(shp_batch* b, T** p, size_t n) {
  return dalloc_bytes(b, (void**)p, n * sizeof(T));
}

#define TRY(expr)                  \
  do {                             \
    shp_status _s = (expr);        \
    if (_s != SHP_OK) return _s;   \
  } while (0)

static inline double nz(double x) { return x == 0.0 ? 0.0 : x; }

static long blocks(long n, int bs) { return (n + bs - 1) / bs; }
// grid of a grid-stride kernel over at most n items
static long gs_blocks(long n) { return std::max(1L, std::min(blocks(n, 256), 4096L)); }

// With timing enabled, an event is recorded on the batch's stream before
// every launch and after the last one, so phase k's duration is
// event[k+1] - event[k] (kernel time plus the launch gap).
static inline void mark(shp_batch* b, int k) {
  if (b->timing) (void)hipEventRecord(b->events[k], b->stream);
}


// ---- exclusive sums ---------------------------------------------------------
// Counts in[0, n): item i is in[i] for i < valid and 0 beyond, where valid =
// min(*limit, cap) when a limit is given (the contact counts of this step's
// pairs) and n otherwise.  out[i] = the sum of the items before i, written
// for i <= valid (i < n): what the consumers read.  Sums are formed in 64
// bits and stored saturated at SCAN_SAT (above every capacity, so a
// saturated offset is an overflow the consumers skip, never a wrapped
// negative index).  Up to 20480 counts take one workgroup (k_scan_small),
// more take three launches: k_scan_up sums each workgroup's SCAN_ITEMS
// counts, k_scan_mid scans those sums in one workgroup, k_scan_down scans
// each workgroup's counts from its offset.  Workgroups whose counts all lie
// beyond `valid` skip their work, so a scan sized for the capacity costs
// what the step's pairs need.
constexpr int SCAN_T = 256;                       // k_scan_up / k_scan_down lanes
constexpr int SCAN_V = 8;                         // int4 loads per lane
constexpr int SCAN_ITEMS = SCAN_T * 4 * SCAN_V;  // counts per workgroup
constexpr long long SCAN_SAT = 1LL << 30;

__device__ __forceinline__ int sat_int(long long x) { return (int)(x < SCAN_SAT ? x : SCAN_SAT); }

__device__ __forceinline__ long scan_valid(long n, const int* limit, long cap) {
  if (!limit) return n;
  const long t = *limit;
  return t < cap ? t : cap;
}

// four counts at i (16-byte aligned), zero at and beyond `valid`
__device__ __forceinline__ int4 load_counts(const int* in, long i, long valid) {
  if (i + 3 < valid) return *(const int4*)(in + i);
  int4 x;
  x.x = i < valid ? in[i] : 0;
  x.y = i + 1 < valid ? in[i + 1] : 0;
  x.z = i + 2 < valid ? in[i + 2] : 0;
  x.w = i + 3 < valid ? in[i + 3] : 0;
  return x;
}

__device__ __forceinline__ long long sum4(int4 x) {
  return ((long long)x.x + x.y) + ((long long)x.z + x.w);
}

// Workgroup of T lanes scans the V * 4 * T counts from `base` (lane t:
// counts base + u*4T + 4t .. +3 of chunk u): wavefront scans of the lane
// sums, wavefront 0 scans the (chunk, wavefront) totals in order, and each
// lane writes its counts' exclusive sums plus `offset`, for indices below
// `wend`.
template <int T, int V>
__device__ __forceinline__ void scan_block(const int* __restrict__ in, int* out, long base,
                                           long valid, long wend, long long offset) {
  constexpr int NW = T / 64, P = V * NW, E = (P + 63) / 64;
  __shared__ long long s_part[P];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int4 x[V];
  long long sum[V], incl[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    x[u] = load_counts(in, base + (long)u * 4 * T + 4 * tid, valid);
    sum[u] = sum4(x[u]);
    incl[u] = sum[u];
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const long long y = __shfl_up(incl[u], o);
      if (lane >= o) incl[u] += y;
    }
  }
  if (lane == 63) {
#pragma unroll
    for (int u = 0; u < V; ++u) s_part[u * NW + wv] = incl[u];
  }
  __syncthreads();
  if (wv == 0) {
    long long v[E], t = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane * E + e;
      v[e] = k < P ? s_part[k] : 0;
      t += v[e];
    }
    long long sc = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(sc, o);
      if (lane >= o) sc += y;
    }
    long long ex = sc - t;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane * E + e;
      if (k < P) s_part[k] = ex;
      ex += v[e];
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < V; ++u) {
    const long i = base + (long)u * 4 * T + 4 * tid;
    const long long e0 = offset + s_part[u * NW + wv] + incl[u] - sum[u];
    const long long e1 = e0 + x[u].x, e2 = e1 + x[u].y, e3 = e2 + x[u].z;
    int4 e;
    e.x = sat_int(e0);
    e.y = sat_int(e1);
    e.z = sat_int(e2);
    e.w = sat_int(e3);
    if (i + 3 < wend) {
      *(int4*)(out + i) = e;
    } else {
      if (i < wend) out[i] = e.x;
      if (i + 1 < wend) out[i + 1] = e.y;
      if (i + 2 < wend) out[i + 2] = e.z;
    }
  }
}

// up to V * 4096 counts in one workgroup of 1024 lanes
template <int V>
__global__ __launch_bounds__(1024) void k_scan_small(const int* __restrict__ in, int* out, long n,
                                                     const int* limit, long cap) {
  const long valid = scan_valid(n, limit, cap);
  scan_block<1024, V>(in, out, 0, valid, valid + 1 < n ? valid + 1 : n, 0);
}

// per-workgroup sums, saturated (a saturated sum saturates every later offset)
__global__ __launch_bounds__(SCAN_T) void k_scan_up(const int* __restrict__ in, long n,
                                                    const int* limit, long cap, int* part) {
  __shared__ long long s_w[SCAN_T / 64];
  const long valid = scan_valid(n, limit, cap);
  const long base = (long)blockIdx.x * SCAN_ITEMS;
  long long s = 0;
  if (base < valid) {
#pragma unroll
    for (int u = 0; u < SCAN_V; ++u)
      s += sum4(load_counts(in, base + (long)u * 4 * SCAN_T + 4 * threadIdx.x, valid));
  }
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int k = 0; k < SCAN_T / 64; ++k) t += s_w[k];
    part[blockIdx.x] = sat_int(t);
  }
}

// exclusive sums of part[0, parts) in place (saturated), 1024 at a time
// with a carry
__global__ __launch_bounds__(1024) void k_scan_mid(int* part, int parts) {
  __shared__ long long s_w[16];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  long long carry = 0;
  for (int c0 = 0; c0 < parts; c0 += 1024) {
    const int i = c0 + tid;
    const long long x = i < parts ? part[i] : 0;
    long long incl = x;
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(incl, o);
      if (lane >= o) incl += y;
    }
    if (lane == 63) s_w[wv] = incl;
    __syncthreads();
    long long before = 0, total = 0;
    for (int k = 0; k < 16; ++k) {
      const long long v = s_w[k];
      before += k < wv ? v : 0;
      total += v;
    }
    if (i < parts) part[i] = sat_int(carry + before + incl - x);
    carry += total;
    __syncthreads();  // s_w is rewritten by the next chunk
  }
}

__global__ __launch_bounds__(SCAN_T) void k_scan_down(const int* __restrict__ in, int* out, long n,
                                                      const int* limit, long cap,
                                                      const int* part) {
  const long valid = scan_valid(n, limit, cap);
  const long base = (long)blockIdx.x * SCAN_ITEMS;
  if (base > valid) return;  // no output of this workgroup is read
  scan_block<SCAN_T, SCAN_V>(in, out, base, valid, valid + 1 < n ? valid + 1 : n,
                             part[blockIdx.x]);
}

// The plan of an exclusive sum of n counts.
static ScanPlan plan_scan(long n) {
  ScanPlan s;
  s.n = n;
  if (n <= 20480) {
    s.small_v = n <= 4096 ? 1 : n <= 8192 ? 2 : n <= 16384 ? 4 : 5;
  } else {
    s.parts = (int)((n + SCAN_ITEMS - 1) / SCAN_ITEMS);
  }
  return s;
}

static void scan_launches(const ScanPlan& s, const char* name, std::vector<LaunchRec>* out) {
  if (s.parts == 0) {
    out->push_back({name, 1, 1024, 0, lds16((size_t)s.small_v * 16 * 8)});
  } else {
    out->push_back({name, s.parts, SCAN_T, 0, lds16((SCAN_T / 64) * 8)});
    out->push_back({name, 1, 1024, 0, lds16(16 * 8)});
    out->push_back({name, s.parts, SCAN_T, 0, lds16((size_t)SCAN_V * (SCAN_T / 64) * 8)});
  }
}

// out[i] = in[0] + ... + in[i-1] on stream st, as planned
static hipError_t exclusive_sum(shp_batch* b, const ScanPlan& sp, const int* in, int* out,
                                const int* limit, long cap, hipStream_t st) {
  const long n = sp.n;
  switch (sp.small_v) {
    case 1: k_scan_small<1><<<1, 1024, 0, st>>>(in, out, n, limit, cap); break;
    case 2: k_scan_small<2><<<1, 1024, 0, st>>>(in, out, n, limit, cap); break;
    case 4: k_scan_small<4><<<1, 1024, 0, st>>>(in, out, n, limit, cap); break;
    case 5: k_scan_small<5><<<1, 1024, 0, st>>>(in, out, n, limit, cap); break;
    default:
      k_scan_up<<<sp.parts, SCAN_T, 0, st>>>(in, n, limit, cap, b->scan_part);
      k_scan_mid<<<1, 1024, 0, st>>>(b->scan_part, sp.parts);
      k_scan_down<<<sp.parts, SCAN_T, 0, st>>>(in, out, n, limit, cap, b->scan_part);
  }
  return hipGetLastError();
}

hipError_t scan_grid_buckets(shp_batch* b) {
  return exclusive_sum(b, b->plan.scan_buckets, b->ctx.bucket_cnt, b->ctx.bucket_start, nullptr,
                       0, b->stream);
}

// The step's front: broadphase, SAT, contact generation and the warm-start
// join; with `fork` the colouring starts on the second stream after k_sat.
// The coloured solver's colouring on the second stream, from this point of
// the batch's stream; join_event marks its end.
static shp_status fork_colouring(shp_batch* b) {
  HIPCHK(hipEventRecord(b->fork_event, b->stream));
  HIPCHK(hipStreamWaitEvent(b->side_stream, b->fork_event, 0));
  if (b->timing) HIPCHK(hipEventRecord(b->side_events[0], b->side_stream));
  launch_colouring(b, b->side_stream);
  if (b->timing) HIPCHK(hipEventRecord(b->side_events[1], b->side_stream));
  HIPCHK(hipEventRecord(b->join_event, b->side_stream));
  return SHP_OK;
}

static shp_status enqueue_front(shp_batch* b, bool fork) {
  Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  hipStream_t s = b->stream;
  mark(b, 0);
  k_reset<<<P.g_reset, 256, 0, s>>>(c);
  mark(b, 1);
  if (c.use_grid)
    HIPCHK(launch_grid_count(b));
  else
    k_broadphase<<<P.g_broadphase, 256, 0, s>>>(c, 0);
  mark(b, 2);
  HIPCHK(exclusive_sum(b, P.scan_bodies, c.cnt_rank, c.off_rank, nullptr, 0, s));
  mark(b, 3);
  if (c.use_grid)
    HIPCHK(launch_grid_emit(b));
  else
    k_broadphase<<<P.g_broadphase, 256, 0, s>>>(c, 1);
  mark(b, 4);
  k_sat<<<P.g_sat, 256, 0, s>>>(c);
  mark(b, 5);
  // The colouring needs only the pairs and which of them have contacts.
  // When its workgroups would take every CU anyway (records_in_contacts),
  // it starts after the contact scan, whose short passes would otherwise
  // wait for CUs behind it, and ends before the contact kernel, which then
  // writes each manifold's record at its colour slot (phase "world_ranges"
  // holds the wait) so that no packing pass reads the contacts back.
  // Otherwise it runs beside the scan, the contact kernel and the learned
  // inference, and joins before the packing pass.
  if (fork && !P.records_in_contacts) TRY(fork_colouring(b));
  // contacts per pair: this step's pairs only (the pair count is on the
  // device)
  HIPCHK(exclusive_sum(b, P.scan_pairs, c.ccount, c.coff, c.off_rank + c.NB, c.pair_cap, s));
  if (fork && P.records_in_contacts) TRY(fork_colouring(b));
  mark(b, 6);
  k_world_ranges<<<P.g_ranges, 256, 0, s>>>(c);
  if (fork && P.records_in_contacts) HIPCHK(hipStreamWaitEvent(s, b->join_event, 0));
  mark(b, 7);
  k_contacts<<<P.g_contacts, 256, 0, s>>>(c);
  mark(b, 8);
  HIPCHK(hipGetLastError());
  return SHP_OK;
}

static shp_status enqueue_step(shp_batch* b) {
  Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  hipStream_t s = b->stream;
  const bool fork = c.solver == SHP_SOLVER_COLORED_F32;
  TRY(enqueue_front(b, fork));
  // the learned warm start of the cache misses (which also patches their
  // manifold records when the contact kernel wrote them)
  if (c.warm_mode == SHP_WARM_LEARNED && c.mlp_ready) mlp_launch_infer(b);
  mark(b, 9);
  if (fork) {
    if (!P.records_in_contacts) {  // phase "solve_setup" holds the colouring's end
      HIPCHK(hipStreamWaitEvent(s, b->join_event, 0));
      launch_colored_pack(b);
    }
    mark(b, 10);
    launch_colored_sweep(b);
    mark(b, 11);
    launch_colored_unpack(b);
  } else {
    mark(b, 10);
    launch_ordered_solver(b);
    mark(b, 11);
  }
  if (b->logging) mlp_launch_log(b);
  mark(b, 12);
  // advance + moveShapes beside the solve quality, one launch
  k_finish<<<P.g_finish_ni + P.g_finish_nq, INTEGRATE_THREADS, 0, s>>>(c, (int)P.g_finish_ni);
  mark(b, 13);
  // the run sums read the quality maxima
  k_cache_store<<<P.g_cache, 256, 0, s>>>(c);
  mark(b, SHP_MAIN_PHASES);
  HIPCHK(hipGetLastError());
  return SHP_OK;
}

static shp_status accumulate_timing(shp_batch* b) {
  HIPCHK(hipEventSynchronize(b->events[SHP_MAIN_PHASES]));
  unsigned long long ts[8];
  HIPCHK(hipMemcpy(ts, b->ctx.tstamp, sizeof(ts), hipMemcpyDeviceToHost));
  // wall_clock64 runs at 100 MHz on CDNA3/4
  for (int k = 0; k < 5; ++k)
    if (ts[k + 1] >= ts[k] && ts[k] != 0) b->solve_us[k] += (double)(ts[k + 1] - ts[k]) * 0.01;
  unsigned long long zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(hipMemcpy(b->ctx.tstamp, zero, sizeof(zero), hipMemcpyHostToDevice));
  for (int k = 0; k < SHP_MAIN_PHASES; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->events[k], b->events[k + 1]));
    b->phase_ms[k] += ms;
  }
  if (b->ctx.solver == SHP_SOLVER_COLORED_F32) {  // the colouring, on the second stream
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->side_events[0], b->side_events[1]));
    b->phase_ms[SHP_MAIN_PHASES] += ms;
  }
  b->timed_steps += 1;
  return SHP_OK;
}


// ---- batch creation: validation, host staging, plan, audit, upload --------

// static LDS of the engine's kernels (the __shared__ arrays they declare)
constexpr size_t FINISH_STATIC_LDS =
    lds16(6 * INTEGRATE_THREADS * 8) + 2 * lds16(INTEGRATE_VERTS * 8) + lds16(INTEGRATE_VERTS) +
    lds16((INTEGRATE_THREADS + 1) * 4) + lds16(INTEGRATE_THREADS) +
    lds16(3 * (QUALITY_THREADS / 64) * 8);
// 16 lanes per body in the grid's pair passes up to this many bodies
#define GRID_WIDE_MAX_BODIES 131072

// Every device buffer of a batch and its size in bytes, in allocation
// order; both the footprint of the plan and the allocation walk this list.
struct Buf {
  void** p;
  size_t bytes;
};

static void batch_buffers(const StepPlan& P, Ctx& c, int32_t** scan_part, std::vector<Buf>* out) {
  const size_t NB = P.NB, NV = P.NV, W = P.W, PC = P.pair_cap, CC = P.contact_cap;
  auto add = [&](auto** ptr, size_t bytes) { out->push_back({(void**)ptr, bytes}); };
  double** body[10] = {&c.px, &c.py, &c.th, &c.vx, &c.vy, &c.w, &c.iml, &c.imr, &c.mu, &c.e};
  for (auto* q : body) add(q, NB * 8);
  add(&c.rad, NB * 8);
  add(&c.is_static, NB);
  add(&c.dead, NB);
  add(&c.body_world, NB * 4);
  add(&c.vert_off, (NB + 1) * 4);
  add(&c.body_off, (W + 1) * 4);
  add(&c.world_order, W * 4);
  add(&c.lx, NV * 8);
  add(&c.ly, NV * 8);
  add(&c.ext, NV * 2);
  add(&c.wv, NV * 32);
  add(&c.wext, NV * 16);
  for (double** q : {&c.bminx, &c.bmaxx, &c.bminy, &c.bmaxy}) add(q, NB * 8);
  add(&c.cnt_rank, (NB + 1) * 4);
  add(&c.off_rank, (NB + 1) * 4);
  add(&c.pa, PC * 4);
  add(&c.pb, PC * 4);
  add(&c.ccount, (PC + 1) * 4);
  add(&c.coff, (PC + 1) * 4);
  add(&c.s_feat, 2 * PC * 2);
  add(&c.s_flip, 2 * PC);
  for (double** q : {&c.s_nx, &c.s_ny, &c.s_px, &c.s_py, &c.s_depth}) add(q, 2 * PC * 8);
  add(&c.wpair_off, (W + 1) * 4);
  add(&c.wcont_off, (W + 1) * 4);
  add(&c.key, CC * 8);
  add(&c.ga, CC * 4);
  add(&c.gb, CC * 4);
  add(&c.flip, CC);
  for (double** q : {&c.cnx, &c.cny, &c.cpx, &c.cpy, &c.depth}) add(q, CC * 8);
  add(&c.jx, CC * 32);
  for (double** q : {&c.bn, &c.rax, &c.ray, &c.rbx, &c.rby, &c.mcn, &c.mct, &c.ln, &c.lt})
    add(q, CC * 8);
  add(&c.warm, CC);
  add(&c.cache_key, CC * 8);
  add(&c.cache_ln, CC * 8);
  add(&c.cache_lt, CC * 8);
  add(&c.cache_woff, (W + 1) * 4);
  if (P.miss_cap > 0) {
    add(&c.miss_idx, (size_t)P.miss_cap * 4);
    add(&c.miss_rec, (size_t)P.miss_cap * 4);
    add(&c.miss_feat, (size_t)P.miss_cap * MLP_IN * 4);
  }
  add(&c.log_count, 8);
  add(&c.counters, 8 * 4);
  add(&c.quality, Q_ENTRIES * 8);
  if (P.scan_parts_max > 0) add(scan_part, (size_t)P.scan_parts_max * 4);
  if (P.use_grid) {
    const size_t TB = P.grid_buckets;
    add(&c.cell_h, W * 8);
    add(&c.is_large, NB);
    add(&c.body_cell, 2 * NB * 4);
    add(&c.tab_off, (W + 1) * 4);
    add(&c.bucket_cnt, (TB + 1) * 4);
    add(&c.bucket_start, (TB + 1) * 4);
    add(&c.body_slot, NB * 4);
    add(&c.grid_rec, 6 * NB * 8);
    add(&c.grid_slots, GRID_SLOTS * NB * 4);
    add(&c.large_off, (W + 1) * 4);
    add(&c.large_ids, std::max<size_t>(P.n_large, 1) * 4);
  }
  if (P.solver == SHP_SOLVER_ORDERED_F64) {
    add(&c.pk, FLOW_RECORD_DOUBLES * CC * 8);
    add(&c.pk_g, 2 * CC * 4);
    add(&c.pk_f, CC);
    add(&c.flow_list, 2 * CC * 4);
    add(&c.flow_rk, CC * 16);
    add(&c.flow_off, NB * 4);
    add(&c.flow_cur, NB * 4);
    add(&c.flow_lvl, CC * 4);
    add(&c.flow_order, CC * 4);
  } else {
    add(&c.pk, COLOR_PACK_FLOATS(PC) * 4);
    add(&c.sched_level, PC * 4);
    add(&c.sched_order, PC * 4);
    add(&c.fv, NB * 16);
    add(&c.col_bu, NB * 16);
    add(&c.col_list, 2 * PC * 8);
    add(&c.col_off, W * (COLOR_CAP + 3) * 4);
    add(&c.col_info, W * 4);
    add(&c.pair_ab, PC * 4);
  }
}

static long footprint(const StepPlan& P) {
  Ctx c{};
  int32_t* part = nullptr;
  std::vector<Buf> bufs;
  batch_buffers(P, c, &part, &bufs);
  long t = 0;
  for (const Buf& x : bufs) t += (long)std::max<size_t>(x.bytes, 16);
  return t;
}

shp_status prepare_batch(const shp_world_desc* worlds, int n_worlds, const shp_config& cfg,
                         HostBatch* hb) {
  if (!worlds || n_worlds <= 0) return fail(SHP_E_INVAL, "null argument");
  if (!(cfg.dt > 0.0)) return fail(SHP_E_INVAL, "dt must be positive");
  if (cfg.iterations < 0) return fail(SHP_E_INVAL, "iterations must be >= 0");
  if (cfg.solver != SHP_SOLVER_ORDERED_F64 && cfg.solver != SHP_SOLVER_COLORED_F32)
    return fail(SHP_E_INVAL, "unknown solver mode");
  if (cfg.warm_start != SHP_WARM_ZERO && cfg.warm_start != SHP_WARM_LEARNED)
    return fail(SHP_E_INVAL, "unknown warm-start mode");
  if (cfg.broadphase < SHP_BROADPHASE_AUTO || cfg.broadphase > SHP_BROADPHASE_GRID_16)
    return fail(SHP_E_INVAL, "unknown broadphase mode");
  if (cfg.solver_threads != 0 && cfg.solver_threads != 256 && cfg.solver_threads != 512 &&
      !((cfg.solver_threads == 1024 || cfg.solver_threads == 64) &&
        cfg.solver == SHP_SOLVER_COLORED_F32))
    return fail(SHP_E_INVAL,
                "solver_threads must be 0, 256 or 512 (or 64 or 1024, coloured solver)");
  if (cfg.max_pairs_per_body < 0) return fail(SHP_E_INVAL, "max_pairs_per_body must be >= 0");
  long NB = 0, NV = 0;
  for (int w = 0; w < n_worlds; ++w) {
    const shp_world_desc& d = worlds[w];
    if (d.n_bodies <= 0) return fail(SHP_E_INVAL, "world with no bodies");
    if (d.capacity != 0 && d.capacity < d.n_bodies)
      return fail(SHP_E_INVAL, "capacity below the number of bodies");
    const long slots = std::max(d.capacity, d.n_bodies);
    if (slots >= (1 << 24)) return fail(SHP_E_UNSUPPORTED, "world exceeds 2^24 slots");
    if (cfg.solver == SHP_SOLVER_COLORED_F32 && slots >= (long)NO_BODY)
      return fail(SHP_E_UNSUPPORTED, "world exceeds 65534 slots (coloured solver)");
    if (!d.px || !d.py || !d.theta || !d.vx || !d.vy || !d.omega || !d.mass || !d.inertia ||
        !d.mu || !d.bounce || !d.hull_off || !d.hull_x || !d.hull_y)
      return fail(SHP_E_INVAL, "null array in world description");
    if (d.hull_off[0] != 0) return fail(SHP_E_INVAL, "hull_off[0] must be 0");
    for (int i = 0; i < d.n_bodies; ++i) {
      int nv = d.hull_off[i + 1] - d.hull_off[i];
      const double r = d.radius ? d.radius[i] : 0.0;
      if (!(r >= 0.0) || r - r != 0.0) return fail(SHP_E_INVAL, "radius must be finite and >= 0");
      if (r > 0.0) {  // a circle
        if (nv != 0) return fail(SHP_E_INVAL, "a circle's hull range must be empty");
        continue;
      }
      if (nv < 3) return fail(SHP_E_INVAL, "hull with fewer than 3 vertices");
      if (nv > SHP_MAXV) return fail(SHP_E_UNSUPPORTED, "hull with more than 8 vertices");
    }
    NB += slots;
    NV += d.hull_off[d.n_bodies];
    if (NB >= (1L << 30) || NV >= (1L << 30)) return fail(SHP_E_UNSUPPORTED, "too many bodies");
  }

  // host staging in SoA, -0 normalised to +0 (static velocities stay as
  // loaded); slots [n_bodies, capacity) start empty, with no vertices
  for (auto& v : hb->h) v.assign(NB, 0.0);
  hb->rad.assign(NB, 0.0);
  hb->body_world.assign(NB, 0);
  hb->vert_off.assign(NB + 1, 0);
  hb->body_off.assign(n_worlds + 1, 0);
  hb->is_static.assign(NB, 0);
  hb->dead.assign(NB, 0);
  hb->lx.assign(NV, 0.0);
  hb->ly.assign(NV, 0.0);
  long g = 0, vo = 0;
  for (int w = 0; w < n_worlds; ++w) {
    const shp_world_desc& d = worlds[w];
    hb->body_off[w] = (int32_t)g;
    const int slots = std::max(d.capacity, d.n_bodies);
    for (int i = 0; i < slots; ++i, ++g) {
      hb->body_world[g] = w;
      hb->vert_off[g] = (int32_t)vo;
      if (i >= d.n_bodies) {
        hb->dead[g] = 1;
        hb->is_static[g] = 1;
        continue;
      }
      hb->h[0][g] = nz(d.px[i]);
      hb->h[1][g] = nz(d.py[i]);
      hb->h[2][g] = nz(d.theta[i]);
      hb->h[3][g] = nz(d.vx[i]);
      hb->h[4][g] = nz(d.vy[i]);
      hb->h[5][g] = nz(d.omega[i]);
      // toInvMass2 (Constraint.hs:79-83)
      hb->h[6][g] = d.mass[i] == 0.0 ? 0.0 : 1.0 / d.mass[i];
      hb->h[7][g] = d.inertia[i] == 0.0 ? 0.0 : 1.0 / d.inertia[i];
      hb->h[8][g] = d.mu[i];
      hb->h[9][g] = d.bounce[i];
      hb->rad[g] = d.radius ? d.radius[i] : 0.0;
      hb->is_static[g] = (hb->h[6][g] == 0.0 && hb->h[7][g] == 0.0) ? 1 : 0;
      for (int k = d.hull_off[i]; k < d.hull_off[i + 1]; ++k, ++vo) {
        hb->lx[vo] = d.hull_x[k];
        hb->ly[vo] = d.hull_y[k];
      }
    }
  }
  hb->body_off[n_worlds] = (int32_t)g;
  hb->vert_off[NB] = (int32_t)vo;
  return plan_staged(cfg, hb);
}

// The plan of a staged batch: capacities, world order, grid set-up, scans,
// launch sizes and their audit.
shp_status plan_staged(const shp_config& cfg, HostBatch* hb) {
  const int n_worlds = (int)hb->body_off.size() - 1;
  const long NB = hb->body_off[n_worlds], NV = hb->vert_off[NB];
  long max_nb = 0;
  for (int w = 0; w < n_worlds; ++w)
    max_nb = std::max<long>(max_nb, hb->body_off[w + 1] - hb->body_off[w]);
  StepPlan& P = hb->P;
  P = StepPlan{};
  P.NB = NB;
  P.NV = NV;
  P.W = n_worlds;
  P.max_nb = max_nb;
  P.solver = cfg.solver;
  P.warm = cfg.warm_start;
  P.ppb = cfg.max_pairs_per_body > 0 ? cfg.max_pairs_per_body : 8;
  P.pair_cap = NB * (long)P.ppb + 1024;
  // pair and contact offsets stay below the scans' saturation value (2^30)
  if (P.pair_cap >= (1L << 29)) return fail(SHP_E_UNSUPPORTED, "pair capacity too large (2^29)");
  // the coloured sweeps address a world's records with 32-bit byte offsets
  if (cfg.solver == SHP_SOLVER_COLORED_F32 && P.pair_cap >= (1L << 28))
    return fail(SHP_E_UNSUPPORTED, "pair capacity too large for the coloured solver (2^28)");
  P.contact_cap = 2 * P.pair_cap;
  P.miss_cap = cfg.warm_start == SHP_WARM_LEARNED ? std::min(2 * P.pair_cap, 4 * NB + 4096) : 0;
  const std::vector<int32_t>& body_off = hb->body_off;
  // per-world solver workgroups take the largest worlds first, so the
  // longest solves start at once instead of trailing the batch
  hb->order.resize(n_worlds);
  for (int w = 0; w < n_worlds; ++w) hb->order[w] = w;
  std::stable_sort(hb->order.begin(), hb->order.end(), [&](int32_t x, int32_t y) {
    return body_off[x + 1] - body_off[x] > body_off[y + 1] - body_off[y];
  });

  // grid broadphase set-up: cell size per world from the largest vertex
  // radius of its dynamic bodies; static bodies larger than a cell are
  // "large" and tested explicitly
  const int mode = cfg.broadphase;
  P.use_grid = mode >= SHP_BROADPHASE_GRID || (mode == SHP_BROADPHASE_AUTO && max_nb > 2048);
  if (P.use_grid) {
    std::vector<double> rmax(hb->rad);  // a circle's extent is its radius
    for (long gg = 0; gg < NB; ++gg)
      for (int k = hb->vert_off[gg]; k < hb->vert_off[gg + 1]; ++k)
        rmax[gg] = std::max(rmax[gg], std::sqrt(hb->lx[k] * hb->lx[k] + hb->ly[k] * hb->ly[k]));
    hb->cell.assign(n_worlds, 1.0);
    hb->tab_off.assign(n_worlds + 1, 0);
    hb->large_off.assign(n_worlds + 1, 0);
    hb->large_ids.clear();
    hb->is_large.assign(NB, 0);
    long tb_total = 0;
    for (int w = 0; w < n_worlds; ++w) {
      double hmax = 0.0;
      for (int gg = body_off[w]; gg < body_off[w + 1]; ++gg)
        if (!hb->is_static[gg] && !hb->dead[gg]) hmax = std::max(hmax, 2.0 * rmax[gg]);
      hb->cell[w] = hmax > 0.0 ? hmax * (1.0 + 1e-6) + 1e-9 : 1.0;
      hb->large_off[w] = (int32_t)hb->large_ids.size();
      for (int gg = body_off[w]; gg < body_off[w + 1]; ++gg)
        if (hb->is_static[gg] && !hb->dead[gg] && 2.0 * rmax[gg] >= hb->cell[w]) {
          hb->is_large[gg] = 1;
          hb->large_ids.push_back(gg);
        }
      const int nw = body_off[w + 1] - body_off[w];
      long t = 64;
      while (t < nw) t <<= 1;
      hb->tab_off[w] = (int32_t)tb_total;
      tb_total += t;
    }
    if (tb_total >= (1L << 30)) return fail(SHP_E_UNSUPPORTED, "grid table too large");
    hb->tab_off[n_worlds] = (int32_t)tb_total;
    hb->large_off[n_worlds] = (int32_t)hb->large_ids.size();
    P.grid_buckets = tb_total;
    P.n_large = (long)hb->large_ids.size();
    P.grid_wide = mode == SHP_BROADPHASE_GRID_16 ||
                  (mode != SHP_BROADPHASE_GRID_1 && NB <= GRID_WIDE_MAX_BODIES);
    P.scan_buckets = plan_scan(tb_total + 1);
  }
  P.scan_bodies = plan_scan(NB + 1);
  P.scan_pairs = plan_scan(P.pair_cap + 1);
  P.scan_parts_max =
      std::max({(long)P.scan_buckets.parts, (long)P.scan_bodies.parts, (long)P.scan_pairs.parts});
  if (cfg.solver == SHP_SOLVER_ORDERED_F64)
    plan_ordered_solver(&P, cfg.solver_threads);
  else
    plan_colored_solver(&P, cfg.solver_threads);
  P.sweep_big = n_worlds;
  if (cfg.solver == SHP_SOLVER_COLORED_F32 && cfg.solver_threads == 64) {
    // worlds by descending size: the small ones form the tail of the order
    long big = 0, small_max = 0;
    for (int k = 0; k < n_worlds; ++k) {
      const long nw = body_off[hb->order[k] + 1] - body_off[hb->order[k]];
      if (nw > SWEEP_WAVE_MAX_BODIES)
        ++big;
      else
        small_max = std::max(small_max, nw);
    }
    P.sweep_big = big;
    P.sweep_slice = (int)((3 * small_max + 3) & ~3L);
    P.g_sweep_waves = (n_worlds - big + SWEEP_WAVES - 1) / SWEEP_WAVES;
  }

  const long PC = P.pair_cap, CC = P.contact_cap;
  P.g_reset = blocks(std::max<long>(64, P.use_grid ? P.grid_buckets + 1 : 0), 256);
  P.g_bodies = blocks(NB, 256);
  P.g_pairs = P.grid_wide ? blocks(NB * 16, 256) : blocks(NB, 256);
  P.g_pairs_count = P.grid_wide ? P.g_pairs : blocks(NB * 4, 256);
  P.g_broadphase = blocks(NB * 64, 256);
  P.g_sat = gs_blocks(PC);
  P.g_ranges = blocks(n_worlds + 1, 256);
  P.g_contacts = gs_blocks(PC);
  P.g_pack = gs_blocks(PC);
  P.g_infer = std::min<long>(std::max<long>(1, blocks(P.miss_cap, 64)), 256);
  P.g_unpack = std::max(1L, std::min(blocks(std::max(PC, NB), 256), 4096L));
  P.g_log = std::max<long>(1, blocks(P.miss_cap, 256));
  P.g_finish_ni = blocks(NB, INTEGRATE_THREADS);
  P.g_finish_nq = std::max(1L, std::min(blocks(CC, 256), 2048L));
  P.g_cache = gs_blocks(std::max<long>(CC, n_worlds + 1));
  P.device_bytes = footprint(P);

  std::vector<LaunchRec> L;
  plan_launches(P, cfg.warm_start == SHP_WARM_LEARNED, false, &L);
  TRY(audit_plan(P, L));
  return SHP_OK;
}

void plan_launches(const StepPlan& P, bool infer, bool logging, std::vector<LaunchRec>* out) {
  out->clear();
  auto add = [&](const char* n, long g, int b, size_t dyn, size_t st) {
    out->push_back({n, g, b, dyn, st});
  };
  add("k_reset", P.g_reset, 256, 0, 0);
  if (P.use_grid) {
    add("k_grid_assign", P.g_bodies, 256, 0, 0);
    scan_launches(P.scan_buckets, "scan_buckets", out);
    add("k_grid_fill", P.g_bodies, 256, 0, 0);
    add("k_grid_pairs", P.g_pairs_count, 256, 0, 0);
  } else {
    add("k_broadphase", P.g_broadphase, 256, 0, 0);
  }
  scan_launches(P.scan_bodies, "scan_bodies", out);
  if (P.use_grid)
    add("k_grid_pairs", P.g_pairs, 256, 0, 0);
  else
    add("k_broadphase", P.g_broadphase, 256, 0, 0);
  add("k_sat", P.g_sat, 256, 0, 0);
  const bool col = P.solver == SHP_SOLVER_COLORED_F32;
  if (col) add("k_colour", P.W, P.color_threads >= 512 ? 1024 : 256, P.color_lds,
               colored_static_lds(false));
  scan_launches(P.scan_pairs, "scan_pairs", out);
  add("k_world_ranges", P.g_ranges, 256, 0, 0);
  add("k_contacts", P.g_contacts, 256, 0, lds16(4));
  if (infer) add("k_mlp_infer", P.g_infer, 256, 0, mlp_infer_static_lds());
  if (col) {
    if (!P.records_in_contacts) add("k_pack_manifolds", P.g_pack, 256, 0, 0);
    if (P.sweep_big > 0)
      add("k_sweep_colours", P.sweep_big, P.color_threads, P.color_pool, colored_static_lds(true));
    if (P.g_sweep_waves > 0)
      add("k_sweep_waves", P.g_sweep_waves, 64 * SWEEP_WAVES,
          (size_t)SWEEP_WAVES * P.sweep_slice * 4, 0);
    add("k_colour_unpack", P.g_unpack, 256, 0, 0);
  } else {
    add("k_solve_flow", P.W, P.solve_threads, std::max<size_t>(P.solve_lds, 16),
        ordered_static_lds(P.solve_threads));
  }
  if (logging) add("k_mlp_log", P.g_log, 256, 0, 0);
  add("k_finish", P.g_finish_ni + P.g_finish_nq, INTEGRATE_THREADS, 0, FINISH_STATIC_LDS);
  add("k_cache_store", P.g_cache, 256, 0, 0);
}

shp_status audit_plan(const StepPlan& P, const std::vector<LaunchRec>& L) {
  for (const LaunchRec& r : L) {
    const std::string n = r.name;
    if (r.grid < 1 || r.grid > 0x7FFFFFFFL)
      return fail(SHP_E_UNSUPPORTED, n + ": " + std::to_string(r.grid) + " workgroups");
    if (r.block < 64 || r.block > 1024)
      return fail(SHP_E_UNSUPPORTED, n + ": workgroup of " + std::to_string(r.block));
    if (r.grid * (long)r.block > 0xFFFFFFFFL)
      return fail(SHP_E_UNSUPPORTED, n + ": more than 2^32 work-items");
    if (r.dyn_lds + r.static_lds > (size_t)LDS_BYTES_PER_CU)
      return fail(SHP_E_UNSUPPORTED, n + ": " + std::to_string(r.dyn_lds + r.static_lds) +
                                         " bytes of LDS");
  }
  // 32-bit index arithmetic: contact slots 2p + k, contact offsets, the
  // grid table, vertex offsets, the coloured solver's 16-byte record offsets
  const long I32 = 0x7FFFFFFFL;
  if (2 * P.pair_cap + 1 > I32 || P.contact_cap + 1 > I32 || P.grid_buckets + 1 > I32 ||
      P.NV + 1 > I32 || P.NB * 16 > 0xFFFFFFFFL)
    return fail(SHP_E_UNSUPPORTED, "an index exceeds 32 bits");
  if (P.solver == SHP_SOLVER_COLORED_F32 && 16 * P.pair_cap > 0xFFFFFFFFL)
    return fail(SHP_E_UNSUPPORTED, "coloured record offsets exceed 32 bits");
  if (P.miss_cap > 0 && P.miss_cap > P.contact_cap)
    return fail(SHP_E_UNSUPPORTED, "miss capacity above the contact capacity");
  return SHP_OK;
}

static void fill_plan_out(const StepPlan& P, const std::vector<LaunchRec>& L, shp_plan* o,
                          shp_launch* launches, int32_t cap) {
  memset(o, 0, sizeof(*o));
  o->bodies = P.NB;
  o->vertices = P.NV;
  o->worlds = P.W;
  o->max_world_bodies = P.max_nb;
  o->pair_capacity = P.pair_cap;
  o->contact_capacity = P.contact_cap;
  o->miss_capacity = P.miss_cap;
  o->grid_buckets = P.grid_buckets;
  o->device_bytes = P.device_bytes;
  o->launches = (int32_t)L.size();
  o->broadphase = P.use_grid ? SHP_BROADPHASE_GRID : SHP_BROADPHASE_SCAN;
  o->solver_threads = P.solver == SHP_SOLVER_COLORED_F32 ? P.color_threads : P.solve_threads;
  for (size_t k = 0; k < L.size(); ++k) {
    o->max_workitems = std::max<int64_t>(o->max_workitems, L[k].grid * (int64_t)L[k].block);
    o->max_lds_bytes = std::max<int64_t>(o->max_lds_bytes, (int64_t)(L[k].dyn_lds + L[k].static_lds));
    if (launches && (int32_t)k < cap) {
      shp_launch& x = launches[k];
      memset(&x, 0, sizeof(x));
      strncpy(x.name, L[k].name, sizeof(x.name) - 1);
      x.workgroups = L[k].grid;
      x.workgroup_size = L[k].block;
      x.lds_bytes = (int32_t)(L[k].dyn_lds + L[k].static_lds);
    }
  }
}

extern "C" {

const char* shp_last_error(void) { return g_err.c_str(); }
const char* shp_version(void) { return "shapes-mi355x 0.2 (gfx950)"; }

shp_status shp_batch_plan(const shp_world_desc* worlds, int32_t n_worlds, const shp_config* cfg,
                          shp_plan* out, shp_launch* launches, int32_t launch_cap) {
  if (!worlds || n_worlds <= 0 || !cfg || !out || launch_cap < 0)
    return fail(SHP_E_INVAL, "bad argument");
  HostBatch hb;
  shp_status st = prepare_batch(worlds, n_worlds, *cfg, &hb);
  if (hb.P.NB > 0) {  // validated: the plan is filled in even when refused
    std::vector<LaunchRec> L;
    plan_launches(hb.P, cfg->warm_start == SHP_WARM_LEARNED, false, &L);
    fill_plan_out(hb.P, L, out, launches, launch_cap);
  }
  return st;
}

shp_status shp_batch_create(const shp_world_desc* worlds, int32_t n_worlds, const shp_config* cfg,
                            shp_batch** out) {
  if (!worlds || n_worlds <= 0 || !cfg || !out) return fail(SHP_E_INVAL, "null argument");
  HostBatch hb;
  TRY(prepare_batch(worlds, n_worlds, *cfg, &hb));
  shp_batch* b = new shp_batch();
  b->cfg = *cfg;
  b->plan = hb.P;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    delete b;
    return fail(SHP_E_HIP, "no HIP device");
  }
  b->device = dev;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail(SHP_E_HIP, "hipStreamCreate failed");
  }
  if (hipStreamCreateWithFlags(&b->side_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork_event, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join_event, hipEventDisableTiming) != hipSuccess) {
    shp_batch_destroy(b);
    return fail(SHP_E_HIP, "stream or event creation failed");
  }
  shp_status st = batch_alloc(b, hb);
  if (st != SHP_OK) {
    std::string keep = g_err;
    shp_batch_destroy(b);
    g_err = keep;
    return st;
  }
  b->table_allocs = b->allocs.size();
  b->table_bytes = b->device_bytes;
  // EmptiesVector.new capacity, then fromList's appends fill [0, n)
  // (EmptiesVector.hs:27-42, World.hs:118-124)
  b->ev_next.resize(n_worlds);
  b->ev_head.resize(n_worlds);
  b->ev_filled.resize(n_worlds);
  for (int w = 0; w < n_worlds; ++w) {
    const int n = worlds[w].n_bodies, cap = b->h_body_off[w + 1] - b->h_body_off[w];
    std::vector<int32_t>& nx = b->ev_next[w];
    nx.assign(cap, -1);
    for (int s = n; s < cap; ++s) nx[s] = s + 1;
    b->ev_head[w] = n;
    b->ev_filled[w] = n;
  }
  *out = b;
  return SHP_OK;
}

}  // extern "C"

// The plan's static LDS of a kernel must cover what the compiler gave it.
static shp_status check_static_lds(const void* fn, size_t planned, const char* name) {
  hipFuncAttributes a{};
  HIPCHK(hipFuncGetAttributes(&a, fn));
  if (a.sharedSizeBytes > planned)
    return fail(SHP_E_HIP, std::string(name) + " uses " + std::to_string(a.sharedSizeBytes) +
                               " bytes of static LDS, the plan " + std::to_string(planned));
  return SHP_OK;
}

shp_status batch_alloc(shp_batch* b, const HostBatch& hb) {
  const StepPlan& P = b->plan;
  Ctx& c = b->ctx;
  memset(&c, 0, sizeof(c));
  const long NB = P.NB, NV = P.NV;
  const int n_worlds = (int)P.W;
  c.W = n_worlds;
  c.NB = (int)NB;
  c.NV = (int)NV;
  c.pair_cap = (int)P.pair_cap;
  c.contact_cap = (int)P.contact_cap;
  c.miss_cap = (int)P.miss_cap;
  c.dt = b->cfg.dt;
  c.beta = b->cfg.baumgarte;
  c.slop = b->cfg.slop;
  c.gx = b->cfg.gravity_x;
  c.gy = b->cfg.gravity_y;
  c.iterations = b->cfg.iterations;
  c.solver = b->cfg.solver;
  c.warm_mode = b->cfg.warm_start;
  c.use_grid = P.use_grid ? 1 : 0;
  c.records_in_contacts = P.records_in_contacts ? 1 : 0;
  c.grid_buckets = (int)P.grid_buckets;
  b->h_body_off = hb.body_off;
  b->h_vert_off = hb.vert_off;
  b->h_cell = hb.cell;
  b->h_is_large = hb.is_large;

  std::vector<Buf> bufs;
  batch_buffers(P, c, &b->scan_part, &bufs);
  for (const Buf& x : bufs) TRY(dalloc_bytes(b, x.p, x.bytes));
  if (b->device_bytes != P.device_bytes)
    return fail(SHP_E_HIP, "allocated " + std::to_string(b->device_bytes) + " bytes, planned " +
                               std::to_string(P.device_bytes));

  double** dptr[10] = {&c.px, &c.py, &c.th, &c.vx, &c.vy, &c.w, &c.iml, &c.imr, &c.mu, &c.e};
  for (int k = 0; k < 10; ++k)
    HIPCHK(hipMemcpy(*dptr[k], hb.h[k].data(), NB * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.rad, hb.rad.data(), NB * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.is_static, hb.is_static.data(), NB, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.dead, hb.dead.data(), NB, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.body_world, hb.body_world.data(), NB * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.vert_off, hb.vert_off.data(), (NB + 1) * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.body_off, hb.body_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.world_order, hb.order.data(), n_worlds * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.lx, hb.lx.data(), NV * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.ly, hb.ly.data(), NV * 8, hipMemcpyHostToDevice));
  if (P.use_grid) {
    HIPCHK(hipMemcpy(c.cell_h, hb.cell.data(), n_worlds * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.is_large, hb.is_large.data(), NB, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.tab_off, hb.tab_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.large_off, hb.large_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
    if (!hb.large_ids.empty())
      HIPCHK(hipMemcpy(c.large_ids, hb.large_ids.data(), hb.large_ids.size() * 4,
                       hipMemcpyHostToDevice));
  }
  TRY(solver_alloc(b));
  TRY(check_static_lds((const void*)k_finish, FINISH_STATIC_LDS, "k_finish"));

  hipStream_t s = b->stream;
  k_fill_i32<<<1, 1, 0, s>>>(c.cnt_rank + NB, 0, 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.cache_woff, 0, n_worlds + 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.wcont_off, 0, n_worlds + 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.wpair_off, 0, n_worlds + 1);
  k_fill_i32<<<1, 64, 0, s>>>(c.counters, 0, 8);
  k_fill_i32<<<1, 64, 0, s>>>((int32_t*)c.quality, 0, 2 * Q_ENTRIES);
  k_fill_i32<<<1, 64, 0, s>>>((int32_t*)c.log_count, 0, 2);
  // makeWorldObj places each shape with moveShape (World.hs:142-149)
  k_init_hulls<<<blocks(NB, 256), 256, 0, s>>>(c);
  k_integrate<<<blocks(NB, INTEGRATE_THREADS), INTEGRATE_THREADS, 0, s>>>(c, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return SHP_OK;
}

static shp_status check_flags(shp_batch* b) {
  int32_t cnt[8];
  HIPCHK(hipMemcpyAsync(cnt, b->ctx.counters, sizeof(cnt), hipMemcpyDeviceToHost, b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  b->last_flags |= cnt[CNT_FLAGS];
  if (b->last_flags & (SHP_FLAG_PAIR_OVERFLOW | SHP_FLAG_CONTACT_OVERFLOW))
    return fail(SHP_E_OVERFLOW, "pair/contact capacity exceeded; raise max_pairs_per_body");
  if (b->last_flags & SHP_FLAG_NONFINITE) return fail(SHP_E_NONFINITE, "non-finite body state");
  return SHP_OK;
}

static shp_status capture_graph(shp_batch* b) {
  if (b->graph_ready) return SHP_OK;
  if (b->graph_made) {  // kernel arguments changed since the last capture
    HIPCHK(hipStreamSynchronize(b->stream));
    HIPCHK(hipGraphExecDestroy(b->graph_exec));
    b->graph_made = false;
  }
  hipGraph_t graph;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  bool timing = b->timing;
  b->timing = false;  // events are never captured
  shp_status st = enqueue_step(b);
  b->timing = timing;
  hipError_t e = hipStreamEndCapture(b->stream, &graph);
  if (st != SHP_OK) return st;
  if (e != hipSuccess) return fail(SHP_E_HIP, "graph capture failed");
  hipError_t ie = hipGraphInstantiate(&b->graph_exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ie != hipSuccess) return fail(SHP_E_HIP, "graph instantiation failed");
  b->graph_made = true;
  b->graph_ready = true;
  return SHP_OK;
}

extern "C" {

shp_status shp_batch_destroy(shp_batch* b) {
  if (!b) return SHP_OK;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph_made) (void)hipGraphExecDestroy(b->graph_exec);
  if (b->events_made)
    for (int k = 0; k <= SHP_MAIN_PHASES; ++k) (void)hipEventDestroy(b->events[k]);
  if (b->events_made)
    for (int k = 0; k < 2; ++k) (void)hipEventDestroy(b->side_events[k]);
  mlp_params_free(&b->mlp);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->side_stream) {
    (void)hipStreamSynchronize(b->side_stream);
    (void)hipStreamDestroy(b->side_stream);
  }
  if (b->fork_event) (void)hipEventDestroy(b->fork_event);
  if (b->join_event) (void)hipEventDestroy(b->join_event);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return SHP_OK;
}

shp_status shp_batch_step_async(shp_batch* b, int32_t n_steps, int32_t use_graph) {
  if (!b || n_steps < 0) return fail(SHP_E_INVAL, "bad argument");
  if (use_graph) TRY(capture_graph(b));
  for (int s = 0; s < n_steps; ++s) {
    if (use_graph) {
      HIPCHK(hipGraphLaunch(b->graph_exec, b->stream));
    } else {
      TRY(enqueue_step(b));
      if (b->timing) TRY(accumulate_timing(b));
    }
  }
  return SHP_OK;
}

shp_status shp_batch_set_timing(shp_batch* b, int32_t enable) {
  if (!b) return fail(SHP_E_INVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  if (enable && !b->events_made) {
    for (int k = 0; k <= SHP_MAIN_PHASES; ++k) HIPCHK(hipEventCreate(&b->events[k]));
    for (int k = 0; k < 2; ++k) HIPCHK(hipEventCreate(&b->side_events[k]));
    b->events_made = true;
    TRY(dalloc(b, &b->ctx.tstamp_store, 8));
    unsigned long long zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    HIPCHK(hipMemcpy(b->ctx.tstamp_store, zero, sizeof(zero), hipMemcpyHostToDevice));
  }
  b->timing = enable != 0;
  b->ctx.tstamp = b->timing ? b->ctx.tstamp_store : nullptr;
  b->graph_ready = false;  // kernel arguments changed
  for (int k = 0; k < 8; ++k) b->solve_us[k] = 0.0;
  for (int k = 0; k < SHP_NUM_PHASES; ++k) b->phase_ms[k] = 0.0;
  b->timed_steps = 0;
  return SHP_OK;
}

shp_status shp_batch_read_timing(shp_batch* b, double* ms_per_phase, int32_t* steps) {
  if (!b || !ms_per_phase || !steps) return fail(SHP_E_INVAL, "bad argument");
  for (int k = 0; k < SHP_NUM_PHASES; ++k) ms_per_phase[k] = b->phase_ms[k];
  *steps = b->timed_steps;
  return SHP_OK;
}

shp_status shp_batch_read_solve_profile(shp_batch* b, double* us) {
  if (!b || !us) return fail(SHP_E_INVAL, "bad argument");
  int n = b->timed_steps > 0 ? b->timed_steps : 1;
  for (int k = 0; k < 5; ++k) us[k] = b->solve_us[k] / n;
  return SHP_OK;
}

shp_status shp_batch_sync(shp_batch* b) {
  if (!b) return fail(SHP_E_INVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  return check_flags(b);
}

shp_status shp_batch_step(shp_batch* b, int32_t n_steps) {
  TRY(shp_batch_step_async(b, n_steps, 0));
  return shp_batch_sync(b);
}

shp_status shp_batch_num_bodies(const shp_batch* b, int32_t world, int32_t* n) {
  if (!b || !n || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  *n = b->h_body_off[world + 1] - b->h_body_off[world];
  return SHP_OK;
}

shp_status shp_batch_read_bodies(shp_batch* b, int32_t world, double* px, double* py, double* theta,
                                 double* vx, double* vy, double* omega) {
  if (!b || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  int o = b->h_body_off[world], n = b->h_body_off[world + 1] - o;
  double* outs[6] = {px, py, theta, vx, vy, omega};
  double* srcs[6] = {b->ctx.px, b->ctx.py, b->ctx.th, b->ctx.vx, b->ctx.vy, b->ctx.w};
  for (int k = 0; k < 6; ++k)
    if (outs[k]) HIPCHK(hipMemcpy(outs[k], srcs[k] + o, n * sizeof(double), hipMemcpyDeviceToHost));
  return SHP_OK;
}

shp_status shp_batch_read_hulls(shp_batch* b, int32_t world, double* out) {
  if (!b || !out || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  int v0 = b->h_vert_off[b->h_body_off[world]], v1 = b->h_vert_off[b->h_body_off[world + 1]];
  int n = v1 - v0;
  // the device keeps (x, y, nx, ny) per vertex, the layout of `out`
  if (n) HIPCHK(hipMemcpy(out, b->ctx.wv + v0, (size_t)n * 32, hipMemcpyDeviceToHost));
  return SHP_OK;
}

shp_status shp_batch_read_pairs(shp_batch* b, int32_t world, int32_t* count, int32_t* ij) {
  if (!b || !count || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  int32_t r[2];
  HIPCHK(hipMemcpy(r, b->ctx.wpair_off + world, 8, hipMemcpyDeviceToHost));
  int n = r[1] - r[0];
  *count = n;
  if (!ij || n == 0) return SHP_OK;
  std::vector<int32_t> a(n), bb(n);
  HIPCHK(hipMemcpy(a.data(), b->ctx.pa + r[0], n * 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(bb.data(), b->ctx.pb + r[0], n * 4, hipMemcpyDeviceToHost));
  int o = b->h_body_off[world];
  for (int k = 0; k < n; ++k) {
    ij[2 * k] = a[k] - o;
    ij[2 * k + 1] = bb[k] - o;
  }
  return SHP_OK;
}

shp_status shp_batch_read_contacts(shp_batch* b, int32_t world, int32_t* count, int32_t* keys,
                                   int32_t* flip, double* geom, double* lam, int32_t* warm) {
  if (!b || !count || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  const Ctx& c = b->ctx;
  int32_t r[2];
  HIPCHK(hipMemcpy(r, c.wcont_off + world, 8, hipMemcpyDeviceToHost));
  int n = r[1] - r[0];
  *count = n;
  if (n == 0) return SHP_OK;
  if (keys) {
    std::vector<uint64_t> k(n);
    HIPCHK(hipMemcpy(k.data(), c.key + r[0], n * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) {
      keys[4 * q + 0] = (int32_t)(k[q] >> 40);
      keys[4 * q + 1] = (int32_t)((k[q] >> 16) & 0xFFFFFF);
      keys[4 * q + 2] = (int32_t)((k[q] >> 8) & 0xFF);
      keys[4 * q + 3] = (int32_t)(k[q] & 0xFF);
    }
  }
  if (flip || warm) {
    std::vector<uint8_t> f(n), wm(n);
    HIPCHK(hipMemcpy(f.data(), c.flip + r[0], n, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(wm.data(), c.warm + r[0], n, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) {
      if (flip) flip[q] = f[q];
      if (warm) warm[q] = wm[q];
    }
  }
  if (geom) {
    double* src[5] = {c.cnx, c.cny, c.cpx, c.cpy, c.depth};
    std::vector<double> t(n);
    for (int k = 0; k < 5; ++k) {
      HIPCHK(hipMemcpy(t.data(), src[k] + r[0], n * 8, hipMemcpyDeviceToHost));
      for (int q = 0; q < n; ++q) geom[5 * q + k] = t[q];
    }
  }
  if (lam) {
    std::vector<double> t(n);
    HIPCHK(hipMemcpy(t.data(), c.ln + r[0], n * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) lam[2 * q] = t[q];
    HIPCHK(hipMemcpy(t.data(), c.lt + r[0], n * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) lam[2 * q + 1] = t[q];
  }
  return SHP_OK;
}

shp_status shp_batch_last_stats(shp_batch* b, shp_step_stats* out) {
  if (!b || !out) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  const Ctx& c = b->ctx;
  int32_t cnt[8];
  HIPCHK(hipMemcpy(cnt, c.counters, sizeof(cnt), hipMemcpyDeviceToHost));
  int32_t pr[1], cr[1];
  HIPCHK(hipMemcpy(pr, c.wpair_off + c.W, 4, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(cr, c.wcont_off + c.W, 4, hipMemcpyDeviceToHost));
  out->pairs = pr[0];
  out->contacts = cr[0];
  out->warm_hits = cnt[CNT_WARM];
  out->learned_starts = cnt[CNT_LEARNED];
  out->misses = cnt[CNT_MISS];
  out->max_levels = cnt[CNT_LEVELS];
  out->manifolds = cnt[CNT_MANIFOLDS];
  out->flags = cnt[CNT_FLAGS] | b->last_flags;
  unsigned long long q[Q_ENTRIES];
  HIPCHK(hipMemcpy(q, c.quality, sizeof(q), hipMemcpyDeviceToHost));
  double* dst[6] = {&out->max_depth, &out->normal_residual, &out->friction_residual,
                    &out->run_max_depth, &out->run_normal_residual, &out->run_friction_residual};
  for (int k = 0; k < 6; ++k) memcpy(dst[k], &q[k], sizeof(double));
  double* mean[3] = {&out->mean_max_depth, &out->mean_normal_residual,
                     &out->mean_friction_residual};
  for (int k = 0; k < 3; ++k) {
    double sum;
    memcpy(&sum, &q[Q_SUM_DEPTH + k], sizeof(double));
    *mean[k] = q[Q_STEPS] ? sum / (double)q[Q_STEPS] : 0.0;
  }
  return SHP_OK;
}

shp_status shp_batch_set_mlp(shp_batch* b, const float* params) {
  if (!b || !params) return fail(SHP_E_INVAL, "null argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  if (!b->mlp.w) TRY(mlp_params_alloc(&b->mlp));
  TRY(mlp_params_upload(&b->mlp, params, b->stream));
  b->ctx.mlp_w = b->mlp.w;
  b->ctx.mlp_wbf = b->mlp.wbf;
  b->ctx.mlp_ready = 1;
  b->ctx.collect_misses = (b->cfg.warm_start == SHP_WARM_LEARNED) || b->logging;
  b->ctx.features_in_contacts = b->logging;
  b->graph_ready = false;  // kernel arguments changed
  return SHP_OK;
}

shp_status shp_batch_set_logging(shp_batch* b, int64_t capacity) {
  if (!b || capacity < 0) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  Ctx& c = b->ctx;
  if (capacity > 0 && c.miss_cap == 0) {  // a zero-start batch: the miss buffers now
    StepPlan& P = b->plan;
    P.miss_cap = std::min(2 * P.pair_cap, 4 * P.NB + 4096);
    P.g_log = std::max<long>(1, blocks(P.miss_cap, 256));
    TRY(dalloc(b, &c.miss_idx, P.miss_cap));
    TRY(dalloc(b, &c.miss_rec, P.miss_cap));
    TRY(dalloc(b, &c.miss_feat, (size_t)P.miss_cap * MLP_IN));
    c.miss_cap = (int)P.miss_cap;
  }
  if (capacity > c.log_cap) {
    TRY(dalloc(b, &c.log_feat, capacity * MLP_IN));
    TRY(dalloc(b, &c.log_tgt, capacity * 2));
    c.log_cap = capacity;
  }
  unsigned long long zero = 0;
  HIPCHK(hipMemcpy(c.log_count, &zero, 8, hipMemcpyHostToDevice));
  b->logging = capacity > 0;
  c.collect_misses = (b->cfg.warm_start == SHP_WARM_LEARNED && c.mlp_ready) || b->logging;
  c.features_in_contacts = b->logging;
  b->graph_ready = false;
  return SHP_OK;
}

shp_status shp_batch_read_log(shp_batch* b, float* feats, float* targets, int64_t* count) {
  if (!b || !count) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  unsigned long long n = 0;
  HIPCHK(hipMemcpy(&n, b->ctx.log_count, 8, hipMemcpyDeviceToHost));
  if ((long)n > b->ctx.log_cap) n = b->ctx.log_cap;
  if (feats && targets && *count > 0) {
    long m = (long)n < *count ? (long)n : *count;
    if (m > 0) {
      HIPCHK(hipMemcpy(feats, b->ctx.log_feat, m * MLP_IN * 4, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(targets, b->ctx.log_tgt, m * 2 * 4, hipMemcpyDeviceToHost));
    }
    n = m;
  }
  *count = (int64_t)n;
  return SHP_OK;
}

shp_status shp_batch_set_cache(shp_batch* b, int32_t world, int32_t count, const int32_t* keys,
                               const double* lam) {
  if (!b || world < 0 || world >= b->ctx.W || count < 0 || (count > 0 && (!keys || !lam)))
    return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  Ctx& c = b->ctx;
  int W = c.W;
  std::vector<int32_t> off(W + 1);
  HIPCHK(hipMemcpy(off.data(), c.cache_woff, (W + 1) * 4, hipMemcpyDeviceToHost));
  int total = off[W];
  std::vector<uint64_t> k(total);
  std::vector<double> ln(total), lt(total);
  if (total) {
    HIPCHK(hipMemcpy(k.data(), c.cache_key, total * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ln.data(), c.cache_ln, total * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lt.data(), c.cache_lt, total * 8, hipMemcpyDeviceToHost));
  }
  std::vector<uint64_t> nk;
  std::vector<double> nln, nlt;
  std::vector<int32_t> noff(W + 1);
  const int nbw = b->h_body_off[world + 1] - b->h_body_off[world];
  for (int w = 0; w < W; ++w) {
    noff[w] = (int32_t)nk.size();
    if (w == world) {
      for (int q = 0; q < count; ++q) {
        const int32_t* kk = keys + 4 * q;
        // bodies of this world (below 2^24, the key's body fields), feature
        // indices below 256
        if (kk[0] < 0 || kk[1] < 0 || kk[0] >= nbw || kk[1] >= nbw || kk[2] < 0 || kk[3] < 0 ||
            kk[2] > 255 || kk[3] > 255)
          return fail(SHP_E_INVAL, "bad cache key (bodies must be below the world's body count, "
                                   "features below 256)");
        uint64_t pk = ((uint64_t)kk[0] << 40) | ((uint64_t)kk[1] << 16) | ((uint64_t)kk[2] << 8) |
                      (uint64_t)kk[3];
        if (q > 0 && !(pk < nk.back())) return fail(SHP_E_INVAL, "cache keys must be strictly descending");
        nk.push_back(pk);
        nln.push_back(lam[2 * q]);
        nlt.push_back(lam[2 * q + 1]);
      }
    } else {
      for (int q = off[w]; q < off[w + 1]; ++q) {
        nk.push_back(k[q]);
        nln.push_back(ln[q]);
        nlt.push_back(lt[q]);
      }
    }
  }
  noff[W] = (int32_t)nk.size();
  if ((long)nk.size() > c.contact_cap) return fail(SHP_E_OVERFLOW, "cache larger than contact capacity");
  if (!nk.empty()) {
    HIPCHK(hipMemcpy(c.cache_key, nk.data(), nk.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.cache_ln, nln.data(), nln.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.cache_lt, nlt.data(), nlt.size() * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c.cache_woff, noff.data(), (W + 1) * 4, hipMemcpyHostToDevice));
  return SHP_OK;
}

// ---- world mutation (World.append / World.delete, World.hs:77-87) --------

// A body description staged like prepare_batch stages one; validated as
// shp_batch_create validates a body.
struct StagedBody {
  double h[10];
  double rad;
  uint8_t is_static;
  std::vector<double> x, y;
};

static shp_status stage_body(const shp_body_desc* d, StagedBody* s) {
  const double r = d->radius;
  if (!(r >= 0.0) || r - r != 0.0) return fail(SHP_E_INVAL, "radius must be finite and >= 0");
  if (r > 0.0 && d->n_vertices != 0) return fail(SHP_E_INVAL, "a circle has no vertices");
  if (r == 0.0 && d->n_vertices < 3) return fail(SHP_E_INVAL, "hull with fewer than 3 vertices");
  if (d->n_vertices > SHP_MAXV) return fail(SHP_E_UNSUPPORTED, "hull with more than 8 vertices");
  if (d->n_vertices > 0 && (!d->hull_x || !d->hull_y)) return fail(SHP_E_INVAL, "null hull");
  const double in[10] = {d->px, d->py, d->theta, d->vx, d->vy, d->omega,
                         d->mass, d->inertia, d->mu, d->bounce};
  for (int k = 0; k < 6; ++k) s->h[k] = nz(in[k]);
  s->h[6] = d->mass == 0.0 ? 0.0 : 1.0 / d->mass;  // toInvMass2 (Constraint.hs:79-83)
  s->h[7] = d->inertia == 0.0 ? 0.0 : 1.0 / d->inertia;
  s->h[8] = d->mu;
  s->h[9] = d->bounce;
  s->rad = r;
  s->is_static = (s->h[6] == 0.0 && s->h[7] == 0.0) ? 1 : 0;
  s->x.assign(d->hull_x, d->hull_x + d->n_vertices);
  s->y.assign(d->hull_y, d->hull_y + d->n_vertices);
  return SHP_OK;
}

// The body's state, materials and shape into slot g, which keeps its
// vertex range (same vertex count); then its hull's normals and extents
// and every body's placement (moveShape, World.hs:142-149: the other
// bodies' placements are recomputed from unchanged state, bit for bit).
static shp_status write_slot_in_place(shp_batch* b, long g, const StagedBody& s) {
  Ctx& c = b->ctx;
  double* dst[10] = {c.px, c.py, c.th, c.vx, c.vy, c.w, c.iml, c.imr, c.mu, c.e};
  for (int k = 0; k < 10; ++k)
    HIPCHK(hipMemcpy(dst[k] + g, &s.h[k], 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.rad + g, &s.rad, 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.is_static + g, &s.is_static, 1, hipMemcpyHostToDevice));
  const long v0 = b->h_vert_off[g];
  if (!s.x.empty()) {
    HIPCHK(hipMemcpy(c.lx + v0, s.x.data(), s.x.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.ly + v0, s.y.data(), s.y.size() * 8, hipMemcpyHostToDevice));
  }
  const uint8_t zero = 0;
  HIPCHK(hipMemcpy(c.dead + g, &zero, 1, hipMemcpyHostToDevice));
  k_init_hulls<<<blocks(c.NB, 256), 256, 0, b->stream>>>(c);
  k_integrate<<<blocks(c.NB, INTEGRATE_THREADS), INTEGRATE_THREADS, 0, b->stream>>>(c, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(b->stream));
  return SHP_OK;
}

// The batch re-laid with slot g holding the staged body: the current state
// is read back, staged again with the new vertex ranges and grid set-up,
// and uploaded into new buffers; the warm-start caches, the run's quality
// sums, the training log and the learned model carry over.
static shp_status rebuild_with(shp_batch* b, long g, const StagedBody& s) {
  Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  const long NB = P.NB, NV = P.NV;
  const int W = (int)P.W;
  HostBatch hb;
  double* src[10] = {c.px, c.py, c.th, c.vx, c.vy, c.w, c.iml, c.imr, c.mu, c.e};
  for (int k = 0; k < 10; ++k) {
    hb.h[k].resize(NB);
    HIPCHK(hipMemcpy(hb.h[k].data(), src[k], NB * 8, hipMemcpyDeviceToHost));
  }
  hb.rad.resize(NB);
  hb.is_static.resize(NB);
  hb.dead.resize(NB);
  HIPCHK(hipMemcpy(hb.rad.data(), c.rad, NB * 8, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hb.is_static.data(), c.is_static, NB, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(hb.dead.data(), c.dead, NB, hipMemcpyDeviceToHost));
  std::vector<double> lx(NV), ly(NV);
  if (NV) {
    HIPCHK(hipMemcpy(lx.data(), c.lx, NV * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ly.data(), c.ly, NV * 8, hipMemcpyDeviceToHost));
  }
  for (int k = 0; k < 10; ++k) hb.h[k][g] = s.h[k];
  hb.rad[g] = s.rad;
  hb.is_static[g] = s.is_static;
  hb.dead[g] = 0;
  hb.body_off = b->h_body_off;
  hb.body_world.resize(NB);
  for (int w = 0; w < W; ++w)
    for (int q = hb.body_off[w]; q < hb.body_off[w + 1]; ++q) hb.body_world[q] = w;
  hb.vert_off.resize(NB + 1);
  long vo = 0;
  for (long q = 0; q < NB; ++q) {
    hb.vert_off[q] = (int32_t)vo;
    if (q == g) {
      for (size_t k = 0; k < s.x.size(); ++k, ++vo) {
        hb.lx.push_back(s.x[k]);
        hb.ly.push_back(s.y[k]);
      }
    } else {
      for (int k = b->h_vert_off[q]; k < b->h_vert_off[q + 1]; ++k, ++vo) {
        hb.lx.push_back(lx[k]);
        hb.ly.push_back(ly[k]);
      }
    }
  }
  hb.vert_off[NB] = (int32_t)vo;
  if (vo >= (1L << 30)) return fail(SHP_E_UNSUPPORTED, "too many vertices");
  TRY(plan_staged(b->cfg, &hb));

  // what one frame hands to the next
  std::vector<int32_t> woff(W + 1);
  HIPCHK(hipMemcpy(woff.data(), c.cache_woff, (W + 1) * 4, hipMemcpyDeviceToHost));
  const long total = woff[W];
  std::vector<uint64_t> ck(total);
  std::vector<double> cln(total), clt(total);
  if (total) {
    HIPCHK(hipMemcpy(ck.data(), c.cache_key, total * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cln.data(), c.cache_ln, total * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(clt.data(), c.cache_lt, total * 8, hipMemcpyDeviceToHost));
  }
  unsigned long long q[Q_ENTRIES], logc = 0;
  HIPCHK(hipMemcpy(q, c.quality, sizeof(q), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(&logc, c.log_count, 8, hipMemcpyDeviceToHost));
  const Ctx keep = c;
  const StepPlan old_plan = P;

  // the plan's buffers are replaced; buffers added later (timing stamps,
  // logging) are kept
  if (b->graph_made) {
    HIPCHK(hipGraphExecDestroy(b->graph_exec));
    b->graph_made = false;
  }
  b->graph_ready = false;
  for (size_t k = 0; k < b->table_allocs; ++k) (void)hipFree(b->allocs[k]);
  b->allocs.erase(b->allocs.begin(), b->allocs.begin() + b->table_allocs);
  const long extra = b->device_bytes - b->table_bytes;
  b->device_bytes = 0;
  b->plan = hb.P;
  const size_t a0 = b->allocs.size();
  const shp_status st = batch_alloc(b, hb);
  std::rotate(b->allocs.begin(), b->allocs.begin() + a0, b->allocs.end());
  b->table_allocs = b->allocs.size() - a0;
  b->table_bytes = b->device_bytes;
  b->device_bytes += extra;
  if (st != SHP_OK) return st;

  c.tstamp_store = keep.tstamp_store;
  c.tstamp = keep.tstamp;
  c.mlp_w = keep.mlp_w;
  c.mlp_wbf = keep.mlp_wbf;
  c.mlp_ready = keep.mlp_ready;
  c.collect_misses = keep.collect_misses;
  c.features_in_contacts = keep.features_in_contacts;
  c.log_feat = keep.log_feat;
  c.log_tgt = keep.log_tgt;
  c.log_cap = keep.log_cap;
  if (b->plan.miss_cap == 0 && keep.miss_cap > 0) {  // miss buffers added for logging
    c.miss_idx = keep.miss_idx;
    c.miss_rec = keep.miss_rec;
    c.miss_feat = keep.miss_feat;
    c.miss_cap = keep.miss_cap;
    b->plan.miss_cap = old_plan.miss_cap;
    b->plan.g_log = old_plan.g_log;
  }
  if (total) {
    HIPCHK(hipMemcpy(c.cache_key, ck.data(), total * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.cache_ln, cln.data(), total * 8, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c.cache_lt, clt.data(), total * 8, hipMemcpyHostToDevice));
  }
  HIPCHK(hipMemcpy(c.cache_woff, woff.data(), (W + 1) * 4, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.quality, q, sizeof(q), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.log_count, &logc, 8, hipMemcpyHostToDevice));
  return SHP_OK;
}

static double body_extent(const StagedBody& s) {
  double r = s.rad;
  for (size_t k = 0; k < s.x.size(); ++k) r = std::max(r, std::sqrt(s.x[k] * s.x[k] + s.y[k] * s.y[k]));
  return 2.0 * r;
}

shp_status shp_batch_append(shp_batch* b, int32_t world, const shp_body_desc* body,
                            int32_t* out_slot) {
  if (!b || !body || !out_slot || world < 0 || world >= b->ctx.W)
    return fail(SHP_E_INVAL, "bad argument");
  StagedBody s;
  TRY(stage_body(body, &s));
  std::vector<int32_t>& nx = b->ev_next[world];
  const int cap = (int)nx.size();
  const int slot = b->ev_head[world];
  // EmptiesVector.append (EmptiesVector.hs:70-85): the head slot is taken
  // and its successor becomes the head; a full world has no head slot
  if (slot >= cap) return fail(SHP_E_INVAL, "world is full: no empty slot to append into");
  const int next = nx[slot];
  if (next == -1) return fail(SHP_E_INVAL, "the empty-slot list points at a filled slot");
  HIPCHK(hipStreamSynchronize(b->stream));
  const long g = (long)b->h_body_off[world] + slot;
  bool in_place = (long)s.x.size() == (long)(b->h_vert_off[g + 1] - b->h_vert_off[g]);
  if (in_place && b->plan.use_grid) {
    // the grid's cells must stay at least as wide as every small body, and
    // the large static bodies are listed at set-up
    const double ext = body_extent(s), cell = b->h_cell[world];
    if (b->h_is_large[g]) in_place = false;
    else if (s.is_static) in_place = ext < cell;
    else in_place = ext * (1.0 + 1e-6) + 1e-9 <= cell;
  }
  if (in_place)
    TRY(write_slot_in_place(b, g, s));
  else
    TRY(rebuild_with(b, g, s));
  nx[slot] = -1;
  b->ev_head[world] = next;
  b->ev_filled[world] += 1;
  *out_slot = slot;
  return SHP_OK;
}

shp_status shp_batch_delete(shp_batch* b, int32_t world, int32_t slot) {
  if (!b || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  std::vector<int32_t>& nx = b->ev_next[world];
  if (slot < 0 || slot >= (int)nx.size()) return fail(SHP_E_INVAL, "slot out of range");
  // safeDelete (EmptiesVector.hs:57-68): the slot must be filled; it becomes
  // the head of the empty-slot list
  if (nx[slot] != -1) return fail(SHP_E_INVAL, "slot " + std::to_string(slot) + " is empty");
  HIPCHK(hipStreamSynchronize(b->stream));
  const long g = (long)b->h_body_off[world] + slot;
  const uint8_t one = 1;
  HIPCHK(hipMemcpy(b->ctx.dead + g, &one, 1, hipMemcpyHostToDevice));
  nx[slot] = b->ev_head[world];
  b->ev_head[world] = slot;
  b->ev_filled[world] -= 1;
  return SHP_OK;
}

shp_status shp_batch_num_vertices(const shp_batch* b, int32_t world, int32_t* n) {
  if (!b || !n || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  *n = b->h_vert_off[b->h_body_off[world + 1]] - b->h_vert_off[b->h_body_off[world]];
  return SHP_OK;
}

shp_status shp_batch_filled(const shp_batch* b, int32_t world, int32_t* out_filled) {
  if (!b || !out_filled || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  *out_filled = b->ev_filled[world];
  return SHP_OK;
}

shp_status shp_batch_read_slots(const shp_batch* b, int32_t world, uint8_t* out_filled) {
  if (!b || !out_filled || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  const std::vector<int32_t>& nx = b->ev_next[world];
  for (size_t s = 0; s < nx.size(); ++s) out_filled[s] = nx[s] == -1 ? 1 : 0;
  return SHP_OK;
}

// ---- debug overlays and the dump format (shapes-demo IOWorld.hs:74-90) ---

shp_status shp_batch_probe_contacts(shp_batch* b) {
  if (!b) return fail(SHP_E_INVAL, "null batch");
  HIPCHK(hipStreamSynchronize(b->stream));
  const bool timing = b->timing;
  b->timing = false;
  // no colouring runs: the contact kernel writes no manifold records
  const int records = b->ctx.records_in_contacts;
  b->ctx.probe = 1;
  b->ctx.records_in_contacts = 0;
  const shp_status st = enqueue_front(b, false);
  b->ctx.probe = 0;
  b->ctx.records_in_contacts = records;
  b->timing = timing;
  TRY(st);
  return shp_batch_sync(b);
}

shp_status shp_batch_read_aabbs(shp_batch* b, int32_t world, int32_t* count, int32_t* slots,
                                double* boxes) {
  if (!b || !count || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  const std::vector<int32_t>& nx = b->ev_next[world];
  const int o = b->h_body_off[world], n = (int)nx.size();
  int m = 0;
  for (int s = 0; s < n; ++s) m += nx[s] == -1;
  *count = m;
  if (!slots && !boxes) return SHP_OK;
  std::vector<double> t[4];
  double* src[4] = {b->ctx.bminx, b->ctx.bmaxx, b->ctx.bminy, b->ctx.bmaxy};
  for (int k = 0; k < 4; ++k) {
    t[k].resize(n);
    if (boxes && n) HIPCHK(hipMemcpy(t[k].data(), src[k] + o, n * 8, hipMemcpyDeviceToHost));
  }
  int q = 0;
  for (int s = 0; s < n; ++s) {
    if (nx[s] != -1) continue;
    if (slots) slots[q] = s;
    if (boxes)
      for (int k = 0; k < 4; ++k) boxes[4 * q + k] = t[k][s];
    ++q;
  }
  return SHP_OK;
}

shp_status shp_batch_read_cache(shp_batch* b, int32_t world, int32_t* count, int32_t* keys,
                                double* lam) {
  if (!b || !count || world < 0 || world >= b->ctx.W) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  const Ctx& c = b->ctx;
  int32_t r[2];
  HIPCHK(hipMemcpy(r, c.cache_woff + world, 8, hipMemcpyDeviceToHost));
  const int n = r[1] - r[0];
  *count = n;
  if (n == 0) return SHP_OK;
  if (keys) {
    std::vector<uint64_t> k(n);
    HIPCHK(hipMemcpy(k.data(), c.cache_key + r[0], n * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) {
      keys[4 * q + 0] = (int32_t)(k[q] >> 40);
      keys[4 * q + 1] = (int32_t)((k[q] >> 16) & 0xFFFFFF);
      keys[4 * q + 2] = (int32_t)((k[q] >> 8) & 0xFF);
      keys[4 * q + 3] = (int32_t)(k[q] & 0xFF);
    }
  }
  if (lam) {
    std::vector<double> ln(n), lt(n);
    HIPCHK(hipMemcpy(ln.data(), c.cache_ln + r[0], n * 8, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(lt.data(), c.cache_lt + r[0], n * 8, hipMemcpyDeviceToHost));
    for (int q = 0; q < n; ++q) {
      lam[2 * q] = ln[q];
      lam[2 * q + 1] = lt[q];
    }
  }
  return SHP_OK;
}

}  // extern "C"

// The dump's byte stream: little-endian, every record a multiple of 8 bytes.
struct DumpWriter {
  std::vector<uint8_t> out;
  template <class T>
  void put(T v) {
    const uint8_t* p = (const uint8_t*)&v;
    out.insert(out.end(), p, p + sizeof(T));
  }
  void tag(const char* t, uint32_t record_bytes, uint64_t count) {
    out.insert(out.end(), t, t + 4);
    put<uint32_t>(record_bytes);
    put<uint64_t>(count);
  }
};

extern "C" {

shp_status shp_batch_dump(shp_batch* b, int32_t world, uint32_t sections, void* buf,
                          int64_t capacity, int64_t* size) {
  if (!b || !size || world < 0 || world >= b->ctx.W || capacity < 0 ||
      (sections & ~(uint32_t)SHP_DUMP_ALL) != 0)
    return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  unsigned long long q[Q_ENTRIES];
  HIPCHK(hipMemcpy(q, b->ctx.quality, sizeof(q), hipMemcpyDeviceToHost));
  DumpWriter w;
  int n_sec = 0;
  for (int k = 0; k < 5; ++k) n_sec += (sections >> k) & 1;
  w.out.insert(w.out.end(), {'S', 'H', 'P', 'D'});
  w.put<uint16_t>(SHP_DUMP_VERSION);
  w.put<uint16_t>((uint16_t)n_sec);
  w.put<int32_t>(world);
  w.put<int32_t>(0);
  w.put<int64_t>((int64_t)q[Q_STEPS]);
  const std::vector<int32_t>& nx = b->ev_next[world];
  const int n = (int)nx.size();
  if (sections & SHP_DUMP_BODIES) {
    std::vector<double> v[6];
    for (auto& x : v) x.resize(n);
    TRY(shp_batch_read_bodies(b, world, v[0].data(), v[1].data(), v[2].data(), v[3].data(),
                              v[4].data(), v[5].data()));
    int m = 0;
    for (int s = 0; s < n; ++s) m += nx[s] == -1;
    w.tag("BODY", 56, (uint64_t)m);
    for (int s = 0; s < n; ++s) {
      if (nx[s] != -1) continue;
      w.put<int32_t>(s);
      w.put<int32_t>(0);
      for (int k = 0; k < 6; ++k) w.put<double>(v[k][s]);
    }
  }
  if (sections & SHP_DUMP_AABBS) {
    int32_t m = 0;
    TRY(shp_batch_read_aabbs(b, world, &m, nullptr, nullptr));
    std::vector<int32_t> sl(m);
    std::vector<double> bx(4 * (size_t)m);
    TRY(shp_batch_read_aabbs(b, world, &m, sl.data(), bx.data()));
    w.tag("AABB", 40, (uint64_t)m);
    for (int k = 0; k < m; ++k) {
      w.put<int32_t>(sl[k]);
      w.put<int32_t>(0);
      for (int e = 0; e < 4; ++e) w.put<double>(bx[4 * k + e]);
    }
  }
  if (sections & SHP_DUMP_PAIRS) {
    int32_t m = 0;
    TRY(shp_batch_read_pairs(b, world, &m, nullptr));
    std::vector<int32_t> ij(2 * (size_t)m + 2);
    TRY(shp_batch_read_pairs(b, world, &m, ij.data()));
    w.tag("PAIR", 8, (uint64_t)m);
    for (int k = 0; k < 2 * m; ++k) w.put<int32_t>(ij[k]);
  }
  if (sections & SHP_DUMP_CONTACTS) {
    int32_t m = 0;
    TRY(shp_batch_read_contacts(b, world, &m, nullptr, nullptr, nullptr, nullptr, nullptr));
    std::vector<int32_t> keys(4 * (size_t)m + 4), flip(m + 1), warm(m + 1);
    std::vector<double> geom(5 * (size_t)m + 5), lam(2 * (size_t)m + 2);
    TRY(shp_batch_read_contacts(b, world, &m, keys.data(), flip.data(), geom.data(), lam.data(),
                                warm.data()));
    w.tag("CONT", 72, (uint64_t)m);
    for (int k = 0; k < m; ++k) {
      w.put<int32_t>(keys[4 * k]);
      w.put<int32_t>(keys[4 * k + 1]);
      w.put<uint8_t>((uint8_t)keys[4 * k + 2]);
      w.put<uint8_t>((uint8_t)keys[4 * k + 3]);
      w.put<uint8_t>((uint8_t)flip[k]);
      w.put<uint8_t>((uint8_t)warm[k]);
      w.put<uint32_t>(0);
      for (int e = 0; e < 5; ++e) w.put<double>(geom[5 * k + e]);
      w.put<double>(lam[2 * k]);
      w.put<double>(lam[2 * k + 1]);
    }
  }
  if (sections & SHP_DUMP_CACHE) {
    int32_t m = 0;
    TRY(shp_batch_read_cache(b, world, &m, nullptr, nullptr));
    std::vector<int32_t> keys(4 * (size_t)m + 4);
    std::vector<double> lam(2 * (size_t)m + 2);
    TRY(shp_batch_read_cache(b, world, &m, keys.data(), lam.data()));
    w.tag("CACH", 32, (uint64_t)m);
    for (int k = 0; k < m; ++k) {
      for (int e = 0; e < 4; ++e) w.put<int32_t>(keys[4 * k + e]);
      w.put<double>(lam[2 * k]);
      w.put<double>(lam[2 * k + 1]);
    }
  }
  *size = (int64_t)w.out.size();
  if (!buf) return SHP_OK;
  if (capacity < *size) return fail(SHP_E_INVAL, "dump buffer too small");
  memcpy(buf, w.out.data(), w.out.size());
  return SHP_OK;
}

}  // extern "C"
