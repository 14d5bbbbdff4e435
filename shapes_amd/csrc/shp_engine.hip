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
//   k_contacts       compaction, constraintGen, warm-start join (beside it,
//                    on a second stream, the coloured solver's k_colour)
//   solve            warm-start application + Gauss-Seidel sweeps
//                    (k_solve_flow, or k_pack_manifolds + k_sweep_colours +
//                    k_colour_unpack)
//   k_finish         advance + moveShapes + AABBs (k_integrate's tiles) and
//                    the solve quality maxima, in one launch
//   k_cache_store    this frame's (key, lambda) become next frame's cache,
//                    quality run sums
//
// Worlds are independent: every kernel handles all worlds of the batch at
// once, and the ordered solver runs one workgroup per world.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

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
#include "shp_solve_common.h"

using namespace shp;

// ---------------------------------------------------------------- helpers


// Key = ObjectFeatureKey ((i, j), (fi, fj)) (Constraints/Contact.hs:194-197)
// packed so that unsigned order equals the derived lexicographic order.
__device__ __forceinline__ uint64_t pack_key(int i, int j, int fi, int fj) {
  return ((uint64_t)(uint32_t)i << 40) | ((uint64_t)(uint32_t)j << 16) |
         ((uint64_t)(uint32_t)fi << 8) | (uint64_t)(uint32_t)fj;
}

__device__ __forceinline__ HullRef hull_of(const Ctx& c, int g) {
  int o = c.vert_off[g];
  HullRef h;
  h.wx = c.wx + o;
  h.wy = c.wy + o;
  h.nx = c.wnx + o;
  h.ny = c.wny + o;
  h.emin = c.ext_min + o;
  h.emax = c.ext_max + o;
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
    c.ext_min[o + e] = (uint8_t)imin;
    c.ext_max[o + e] = (uint8_t)imax;
  }
}

// advance (World.hs:194-196, Constraint.hs:225-229) when `advance`, then
// moveShapes (World.hs:163-167): T(pos) R(theta) applied to the local
// vertices (Transform.hs:34-38, Linear.hs:217-220), unit edge normals
// (ConvexHull.hs:184-195) and the AABB fold (Broadphase/Aabb.hs:81-110).
// A workgroup takes 256 consecutive bodies: one lane per body builds the
// transform, one lane per vertex transforms it and its successor and writes
// the vertex and edge normal (coalesced), and one lane per body folds its
// AABB in vertex order from LDS.
constexpr int INTEGRATE_THREADS = 256;
constexpr int INTEGRATE_VERTS = INTEGRATE_THREADS * SHP_MAXV;

// The bodies [vb * INTEGRATE_THREADS, ...) of the batch (one workgroup).
__device__ __forceinline__ void integrate_tile(const Ctx& c, int advance, int vb) {
  __shared__ double s_m[6][INTEGRATE_THREADS];
  __shared__ double s_wx[INTEGRATE_VERTS], s_wy[INTEGRATE_VERTS];
  __shared__ uint8_t s_body[INTEGRATE_VERTS];
  __shared__ int s_voff[INTEGRATE_THREADS + 1];
  const int t = threadIdx.x;
  const int g0 = vb * INTEGRATE_THREADS;
  const int nbk = min(INTEGRATE_THREADS, c.NB - g0);
  const int g = g0 + t;
  const double one = 1.0, zer = 0.0;
  if (t < nbk) {
    double px = c.px[g], py = c.py[g], th = c.th[g];
    if (advance) {
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
    if (!(chk - chk == 0.0)) atomicOr(&c.counters[CNT_FLAGS], SHP_FLAG_NONFINITE);
  }
  __syncthreads();
  const int v0 = s_voff[0], nv = s_voff[nbk] - v0;
  if (t < nbk)
    for (int i = s_voff[t] - v0; i < s_voff[t + 1] - v0; ++i) s_body[i] = (uint8_t)t;
  __syncthreads();
  for (int i = t; i < nv; i += INTEGRATE_THREADS) {
    const int b = s_body[i];
    const int last = s_voff[b + 1] - v0 - 1;
    const int i1 = i < last ? i + 1 : s_voff[b] - v0;
    const double m00 = s_m[0][b], m01 = s_m[1][b], m02 = s_m[2][b];
    const double m10 = s_m[3][b], m11 = s_m[4][b], m12 = s_m[5][b];
    const double lx = c.lx[v0 + i], ly = c.ly[v0 + i];
    const double lx1 = c.lx[v0 + i1], ly1 = c.ly[v0 + i1];
    const double x = (m00 * lx + m01 * ly) + m02 * one;
    const double y = (m10 * lx + m11 * ly) + m12 * one;
    const double x1 = (m00 * lx1 + m01 * ly1) + m02 * one;
    const double y1 = (m10 * lx1 + m11 * ly1) + m12 * one;
    c.wx[v0 + i] = x;
    c.wy[v0 + i] = y;
    s_wx[i] = x;
    s_wy[i] = y;
    d2 nn = normalize(clockwise(sub(mk(x1, y1), mk(x, y))));
    c.wnx[v0 + i] = nn.x;
    c.wny[v0 + i] = nn.y;
  }
  __syncthreads();
  if (t < nbk) {
    const int o = s_voff[t] - v0, e = s_voff[t + 1] - v0;
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
// the constant acceleration (applyExternal, World.hs:183-185 with
// constantAccel, World/External.hs:249-254); pass 1 writes the pairs, in
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
  if (!fill && lane == 0 && c.iml[g] != 0.0) {
    c.vx[g] = c.vx[g] + c.gx * c.dt;
    c.vy[g] = c.vy[g] + c.gy * c.dt;
  }
  bool si = body_static(c, g);
  double aminx = c.bminx[g], amaxx = c.bmaxx[g], aminy = c.bminy[g], amaxy = c.bmaxy[g];
  int count = 0;
  int base = fill ? c.off_rank[rank] : 0;
  for (int j0 = li - 1; j0 >= 0; j0 -= 64) {
    int lj = j0 - lane;
    bool hit = false;
    if (lj >= 0) {
      int gj = b0 + lj;
      if (!(si && body_static(c, gj))) {
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
__device__ __forceinline__ void sat_pair(const Ctx& c, int p) {
  HullRef A = hull_of(c, c.pa[p]);
  HullRef B = hull_of(c, c.pb[p]);
  Cgeom gm[2];
  int flip = 0;
  int n = hull_contacts(A, B, gm, &flip);
  for (int k = 0; k < n; ++k) {
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

// grid-stride over this step's pairs (the launch is sized for the capacity)
__global__ void k_sat(Ctx c) {
  int tp = total_pairs(c);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tp; p += gridDim.x * blockDim.x) sat_pair(c, p);
}

// Contact count of pair p for the contact scan: 0 beyond this step's pairs.
struct PairCount {
  const int32_t* ccount;
  const int32_t* total;  // off_rank[NB]
  int cap;
  __host__ __device__ int operator()(int p) const {
    int t = *total;
    t = t < cap ? t : cap;
    return p < t ? ccount[p] : 0;
  }
};

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

// Compaction + constraintGen (Constraints/Contact.hs:218-229) + the warm
// start join of applyCachedSlns (Solvers/Contact.hs:74-121): a binary search
// of the key in last frame's cache of the same world, which is strictly
// descending, finds exactly the element descZipVector (Utils/Descending.hs:
// 47-71) would pair it with.
__device__ void contacts_of_pair(const Ctx& c, int p, int n, int& hits) {
  int ga = c.pa[p], gb = c.pb[p];
  int w = c.body_world[ga];
  int b0 = c.body_off[w];
  d2 xa = mk(c.px[ga], c.py[ga]), xb = mk(c.px[gb], c.py[gb]);
  double im[6] = {c.iml[ga], c.iml[ga], c.imr[ga], c.iml[gb], c.iml[gb], c.imr[gb]};
  int cache_lo = c.cache_woff[w], cache_hi = c.cache_woff[w + 1];
  int c_lo = c.wcont_off[w], c_hi = c.wcont_off[w + 1];
  for (int k = 0; k < n; ++k) {
    int dst = c.coff[p] + k;
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
    for (int q = 0; q < 6; ++q) {
      c.jn[q * c.contact_cap + dst] = jn[q];
      c.jt[q * c.contact_cap + dst] = jt[q];
    }
    // baumgarte (NonPenetration.hs:48-56)
    c.bn[dst] = depth > c.slop ? (c.beta / c.dt) * (c.slop - depth) : 0.0;
    // Restitution.constraintGen (Restitution.hs:21-31)
    d2 ra = sub(pp, xa), rb = sub(pp, xb);
    c.rax[dst] = ra.x;
    c.ray[dst] = ra.y;
    c.rbx[dst] = rb.x;
    c.rby[dst] = rb.y;
    c.mcn[dst] = eff_mass(jn, im);
    c.mct[dst] = eff_mass(jt, im);
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
    if (found >= 0) {
      c.ln[dst] = c.cache_ln[found];
      c.lt[dst] = c.cache_lt[found];
      c.warm[dst] = 1;
      ++hits;
    } else {
      // newCache (Solvers/Contact.hs:91-99): zero start; the learned warm
      // start may replace it before the solve
      c.ln[dst] = 0.0;
      c.lt[dst] = 0.0;
      c.warm[dst] = 0;
      if (c.collect_misses) {
        int slot = atomicAdd(&c.counters[CNT_MISS], 1);
        if (slot < c.miss_cap) {
          c.miss_idx[slot] = dst;
          double v[6] = {c.vx[ga], c.vy[ga], c.w[ga], c.vx[gb], c.vy[gb], c.w[gb]};
          double sgn = flip ? -1.0 : 1.0;
          mlp_features(sgn * nn.x, sgn * nn.y, depth, v, ra.x, ra.y, rb.x, rb.y, c.iml[ga], c.iml[gb],
                       c.imr[ga], c.imr[gb], (c.mu[ga] + c.mu[gb]) / 2.0, c.miss_feat + (size_t)slot * MLP_IN);
        }
      }
    }
  }
}


__global__ void k_contacts(Ctx c) {
  int tp = total_pairs(c);
  int hits = 0;
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tp; p += gridDim.x * blockDim.x) {
    int n = c.ccount[p];
    if (n > 0) contacts_of_pair(c, p, n, hits);
  }
  // one atomic per workgroup for the warm-start hit counter
  __shared__ int s_hits;
  if (threadIdx.x == 0) s_hits = 0;
  __syncthreads();
  for (int off = 32; off > 0; off >>= 1) hits += __shfl_xor(hits, off);
  if ((threadIdx.x & 63) == 0 && hits) atomicAdd(&s_hits, hits);
  __syncthreads();
  if (threadIdx.x == 0 && s_hits) atomicAdd(&c.counters[CNT_WARM], s_hits);
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

template <typename T>
static shp_status dalloc(shp_batch* b, T** p, size_t n) {
  void* q = nullptr;
  size_t bytes = n * sizeof(T);
  if (bytes == 0) bytes = sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(SHP_E_OOM, "hipMalloc failed");
  b->allocs.push_back(q);
  *p = (T*)q;
  return SHP_OK;
}

#define TRY(expr)                  \
  do {                             \
    shp_status _s = (expr);        \
    if (_s != SHP_OK) return _s;   \
  } while (0)

static inline double nz(double x) { return x == 0.0 ? 0.0 : x; }

static int blocks(long n, int bs) { return (int)((n + bs - 1) / bs); }
// grid of a grid-stride kernel over at most n items
static int gs_blocks(long n) { return std::max(1, std::min(blocks(n, 256), 4096)); }

// With timing enabled, an event is recorded on the batch's stream before
// every launch and after the last one, so phase k's duration is
// event[k+1] - event[k] (kernel time plus the launch gap).
static inline void mark(shp_batch* b, int k) {
  if (b->timing) (void)hipEventRecord(b->events[k], b->stream);
}

static hipError_t exclusive_sum(shp_batch* b, const int* in, int* out, long n, hipStream_t s);

static shp_status enqueue_step(shp_batch* b) {
  Ctx& c = b->ctx;
  hipStream_t s = b->stream;
  mark(b, 0);
  k_reset<<<blocks(std::max<long>(64, c.use_grid ? (long)c.grid_buckets + 1 : 0), 256), 256, 0,
            s>>>(c);
  mark(b, 1);
  if (c.use_grid)
    HIPCHK(launch_grid_count(b));
  else
    k_broadphase<<<blocks((long)c.NB * 64, 256), 256, 0, s>>>(c, 0);
  mark(b, 2);
  HIPCHK(exclusive_sum(b, c.cnt_rank, c.off_rank, (long)c.NB + 1, s));
  mark(b, 3);
  if (c.use_grid)
    HIPCHK(launch_grid_emit(b));
  else
    k_broadphase<<<blocks((long)c.NB * 64, 256), 256, 0, s>>>(c, 1);
  mark(b, 4);
  k_sat<<<gs_blocks(c.pair_cap), 256, 0, s>>>(c);
  // the colouring needs only the pairs and which of them have contacts: it
  // runs beside the contact scan, contact generation and learned warm start
  const bool fork = c.solver == SHP_SOLVER_COLORED_F32;
  if (fork) {
    HIPCHK(hipEventRecord(b->fork_event, s));
    HIPCHK(hipStreamWaitEvent(b->side_stream, b->fork_event, 0));
    launch_colouring(b, b->side_stream);
    HIPCHK(hipEventRecord(b->join_event, b->side_stream));
  }
  mark(b, 5);
  size_t tb = b->scan_tmp_bytes;
  {
    hipcub::CountingInputIterator<int> idx(0);
    hipcub::TransformInputIterator<int, PairCount, hipcub::CountingInputIterator<int>> cnt(
        idx, PairCount{c.ccount, c.off_rank + c.NB, c.pair_cap});
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, tb, cnt, c.coff, c.pair_cap + 1, s));
  }
  mark(b, 6);
  k_world_ranges<<<blocks(c.W + 1, 256), 256, 0, s>>>(c);
  mark(b, 7);
  k_contacts<<<gs_blocks(c.pair_cap), 256, 0, s>>>(c);
  if (c.warm_mode == SHP_WARM_LEARNED && c.mlp_ready) mlp_launch_infer(b);
  mark(b, 8);
  if (fork) {
    HIPCHK(hipStreamWaitEvent(s, b->join_event, 0));
    launch_colored_sweeps(b);
  } else {
    launch_ordered_solver(b);
  }
  mark(b, 9);
  if (b->logging) mlp_launch_log(b);
  {  // advance + moveShapes beside the solve quality, one launch
    const int ni = blocks(c.NB, INTEGRATE_THREADS);
    const int nq = (int)std::max<long>(1, std::min<long>((c.contact_cap + 255) / 256, 2048));
    k_finish<<<ni + nq, INTEGRATE_THREADS, 0, s>>>(c, ni);
  }
  mark(b, 10);
  // the run sums read the quality maxima
  k_cache_store<<<gs_blocks(std::max(c.contact_cap, c.W + 1)), 256, 0, s>>>(c);
  mark(b, 11);
  mark(b, 12);
  HIPCHK(hipGetLastError());
  return SHP_OK;
}

// ---- exclusive sums of up to 20480 counts in one workgroup ---------------
// Chunk u = entries [4096u, 4096u + 4096): thread t loads entries 4t..4t+3
// of every chunk as int4 (coalesced, all loads in flight at once), the
// wavefronts scan their thread sums, wavefront 0 scans the (chunk,
// wavefront) totals in chunk-major order, and every thread writes its
// entries.  One launch against the library scan's two (state set-up and
// scan), for the per-body pair counts and grid bucket counts of batches up
// to 20479 bodies / buckets.
template <int V>
__global__ __launch_bounds__(1024) void k_scan_small(const int* __restrict__ in, int* out, int n) {
  constexpr int P = V * 16, E = (P + 63) / 64;
  __shared__ int s_part[P];  // (chunk, wavefront) totals, then their exclusive sums
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  int4 x[V];
  int sum[V], incl[V];
#pragma unroll
  for (int u = 0; u < V; ++u) {
    const int i = u * 4096 + 4 * tid;
    if (i + 3 < n) {
      x[u] = *(const int4*)(in + i);
    } else {
      x[u].x = i < n ? in[i] : 0;
      x[u].y = i + 1 < n ? in[i + 1] : 0;
      x[u].z = i + 2 < n ? in[i + 2] : 0;
      x[u].w = 0;
    }
    sum[u] = (x[u].x + x[u].y) + (x[u].z + x[u].w);
    incl[u] = sum[u];
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
#pragma unroll
    for (int u = 0; u < V; ++u) {
      const int y = __shfl_up(incl[u], o);
      if (lane >= o) incl[u] += y;
    }
  }
  if (lane == 63) {
#pragma unroll
    for (int u = 0; u < V; ++u) s_part[u * 16 + wv] = incl[u];
  }
  __syncthreads();
  if (wv == 0) {
    int v[E], t = 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
      const int k = lane * E + e;
      v[e] = k < P ? s_part[k] : 0;
      t += v[e];
    }
    int sc = t;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int y = __shfl_up(sc, o);
      if (lane >= o) sc += y;
    }
    int ex = sc - t;
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
    const int i = u * 4096 + 4 * tid;
    int4 e;
    e.x = s_part[u * 16 + wv] + incl[u] - sum[u];
    e.y = e.x + x[u].x;
    e.z = e.y + x[u].y;
    e.w = e.z + x[u].z;
    if (i + 3 < n) {
      *(int4*)(out + i) = e;
    } else {
      if (i < n) out[i] = e.x;
      if (i + 1 < n) out[i + 1] = e.y;
      if (i + 2 < n) out[i + 2] = e.z;
    }
  }
}

// out[i] = in[0] + ... + in[i-1] for i < n, on stream s: one workgroup for
// n <= 20480 (int4-aligned arrays), the library scan beyond.
static hipError_t exclusive_sum(shp_batch* b, const int* in, int* out, long n, hipStream_t s) {
  if (n <= 4096) {
    k_scan_small<1><<<1, 1024, 0, s>>>(in, out, (int)n);
  } else if (n <= 8192) {
    k_scan_small<2><<<1, 1024, 0, s>>>(in, out, (int)n);
  } else if (n <= 16384) {
    k_scan_small<4><<<1, 1024, 0, s>>>(in, out, (int)n);
  } else if (n <= 20480) {  // e.g. a power-of-two bucket table + 1
    k_scan_small<5><<<1, 1024, 0, s>>>(in, out, (int)n);
  } else {  // (measured: from 8 chunks on, the wavefront scans cost more than the library's launch)
    size_t tb = b->scan_tmp_bytes;
    return hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, tb, in, out, (int)n, s);
  }
  return hipGetLastError();
}

hipError_t scan_grid_buckets(shp_batch* b) {
  return exclusive_sum(b, b->ctx.bucket_cnt, b->ctx.bucket_start, (long)b->ctx.grid_buckets + 1,
                       b->stream);
}

static shp_status accumulate_timing(shp_batch* b) {
  HIPCHK(hipEventSynchronize(b->events[SHP_NUM_PHASES]));
  unsigned long long ts[8];
  HIPCHK(hipMemcpy(ts, b->ctx.tstamp, sizeof(ts), hipMemcpyDeviceToHost));
  // wall_clock64 runs at 100 MHz on CDNA3/4
  for (int k = 0; k < 5; ++k)
    if (ts[k + 1] >= ts[k] && ts[k] != 0) b->solve_us[k] += (double)(ts[k + 1] - ts[k]) * 0.01;
  unsigned long long zero[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(hipMemcpy(b->ctx.tstamp, zero, sizeof(zero), hipMemcpyHostToDevice));
  for (int k = 0; k < SHP_NUM_PHASES; ++k) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, b->events[k], b->events[k + 1]));
    b->phase_ms[k] += ms;
  }
  b->timed_steps += 1;
  return SHP_OK;
}


extern "C" {

const char* shp_last_error(void) { return g_err.c_str(); }
const char* shp_version(void) { return "shapes-mi355x 0.1 (gfx950)"; }

shp_status shp_batch_create(const shp_world_desc* worlds, int32_t n_worlds, const shp_config* cfg,
                            shp_batch** out) {
  if (!worlds || n_worlds <= 0 || !cfg || !out) return fail(SHP_E_INVAL, "null argument");
  if (!(cfg->dt > 0.0)) return fail(SHP_E_INVAL, "dt must be positive");
  if (cfg->iterations < 0) return fail(SHP_E_INVAL, "iterations must be >= 0");
  if (cfg->solver != SHP_SOLVER_ORDERED_F64 && cfg->solver != SHP_SOLVER_COLORED_F32)
    return fail(SHP_E_INVAL, "unknown solver mode");
  if (cfg->warm_start != SHP_WARM_ZERO && cfg->warm_start != SHP_WARM_LEARNED)
    return fail(SHP_E_INVAL, "unknown warm-start mode");
  if (cfg->broadphase < SHP_BROADPHASE_AUTO || cfg->broadphase > SHP_BROADPHASE_GRID)
    return fail(SHP_E_INVAL, "unknown broadphase mode");
  if (cfg->solver_threads != 0 && cfg->solver_threads != 256 && cfg->solver_threads != 512 &&
      !(cfg->solver_threads == 1024 && cfg->solver == SHP_SOLVER_COLORED_F32))
    return fail(SHP_E_INVAL, "solver_threads must be 0, 256 or 512 (or 1024, coloured solver)");
  long NB = 0, NV = 0;
  for (int w = 0; w < n_worlds; ++w) {
    const shp_world_desc& d = worlds[w];
    if (d.n_bodies <= 0) return fail(SHP_E_INVAL, "world with no bodies");
    if (d.n_bodies >= (1 << 24)) return fail(SHP_E_UNSUPPORTED, "world exceeds 2^24 bodies");
    if (!d.px || !d.py || !d.theta || !d.vx || !d.vy || !d.omega || !d.mass || !d.inertia ||
        !d.mu || !d.bounce || !d.hull_off || !d.hull_x || !d.hull_y)
      return fail(SHP_E_INVAL, "null array in world description");
    if (d.hull_off[0] != 0) return fail(SHP_E_INVAL, "hull_off[0] must be 0");
    for (int i = 0; i < d.n_bodies; ++i) {
      int nv = d.hull_off[i + 1] - d.hull_off[i];
      if (nv < 3) return fail(SHP_E_INVAL, "hull with fewer than 3 vertices");
      if (nv > SHP_MAXV) return fail(SHP_E_UNSUPPORTED, "hull with more than 8 vertices");
    }
    NB += d.n_bodies;
    NV += d.hull_off[d.n_bodies];
  }
  if (NB >= (1L << 30)) return fail(SHP_E_UNSUPPORTED, "too many bodies");
  int ppb = cfg->max_pairs_per_body > 0 ? cfg->max_pairs_per_body : 8;
  long pair_cap = NB * (long)ppb + 1024;
  if (pair_cap >= (1L << 30)) return fail(SHP_E_UNSUPPORTED, "pair capacity too large");
  // the coloured sweeps address a world's records with 32-bit byte offsets
  if (cfg->solver == SHP_SOLVER_COLORED_F32 && pair_cap >= (1L << 28))
    return fail(SHP_E_UNSUPPORTED, "pair capacity too large for the coloured solver (2^28)");

  shp_batch* b = new shp_batch();
  b->cfg = *cfg;
  b->ppb = ppb;
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
  shp_status st = batch_alloc(b, worlds, n_worlds, NB, NV, pair_cap);
  if (st != SHP_OK) {
    std::string keep = g_err;
    shp_batch_destroy(b);
    g_err = keep;
    return st;
  }
  *out = b;
  return SHP_OK;
}

}  // extern "C"

shp_status batch_alloc(shp_batch* b, const shp_world_desc* worlds, int n_worlds, long NB, long NV,
                       long pair_cap) {
  Ctx& c = b->ctx;
  memset(&c, 0, sizeof(c));
  c.W = n_worlds;
  c.NB = (int)NB;
  c.NV = (int)NV;
  c.pair_cap = (int)pair_cap;
  c.contact_cap = (int)(2 * pair_cap);
  c.dt = b->cfg.dt;
  c.beta = b->cfg.baumgarte;
  c.slop = b->cfg.slop;
  c.gx = b->cfg.gravity_x;
  c.gy = b->cfg.gravity_y;
  c.iterations = b->cfg.iterations;
  c.solver = b->cfg.solver;
  c.warm_mode = b->cfg.warm_start;

  // host staging in SoA, -0 normalised to +0 (static velocities stay as loaded)
  std::vector<double> h[10];
  for (auto& v : h) v.resize(NB);
  std::vector<int32_t> body_world(NB), vert_off(NB + 1), body_off(n_worlds + 1);
  std::vector<uint8_t> is_static(NB);
  std::vector<double> lx(NV), ly(NV);
  long g = 0, vo = 0;
  for (int w = 0; w < n_worlds; ++w) {
    const shp_world_desc& d = worlds[w];
    body_off[w] = (int32_t)g;
    for (int i = 0; i < d.n_bodies; ++i, ++g) {
      h[0][g] = nz(d.px[i]);
      h[1][g] = nz(d.py[i]);
      h[2][g] = nz(d.theta[i]);
      h[3][g] = nz(d.vx[i]);
      h[4][g] = nz(d.vy[i]);
      h[5][g] = nz(d.omega[i]);
      // toInvMass2 (Constraint.hs:79-83)
      h[6][g] = d.mass[i] == 0.0 ? 0.0 : 1.0 / d.mass[i];
      h[7][g] = d.inertia[i] == 0.0 ? 0.0 : 1.0 / d.inertia[i];
      h[8][g] = d.mu[i];
      h[9][g] = d.bounce[i];
      is_static[g] = (h[6][g] == 0.0 && h[7][g] == 0.0) ? 1 : 0;
      body_world[g] = w;
      vert_off[g] = (int32_t)vo;
      for (int k = d.hull_off[i]; k < d.hull_off[i + 1]; ++k, ++vo) {
        lx[vo] = d.hull_x[k];
        ly[vo] = d.hull_y[k];
      }
    }
  }
  body_off[n_worlds] = (int32_t)g;
  vert_off[NB] = (int32_t)vo;
  b->h_body_off = body_off;
  b->h_vert_off = vert_off;

  double** dptr[10] = {&c.px, &c.py, &c.th, &c.vx, &c.vy, &c.w, &c.iml, &c.imr, &c.mu, &c.e};
  for (int k = 0; k < 10; ++k) {
    TRY(dalloc(b, dptr[k], NB));
    HIPCHK(hipMemcpy(*dptr[k], h[k].data(), NB * sizeof(double), hipMemcpyHostToDevice));
  }
  TRY(dalloc(b, &c.is_static, NB));
  HIPCHK(hipMemcpy(c.is_static, is_static.data(), NB, hipMemcpyHostToDevice));
  TRY(dalloc(b, &c.body_world, NB));
  HIPCHK(hipMemcpy(c.body_world, body_world.data(), NB * 4, hipMemcpyHostToDevice));
  TRY(dalloc(b, &c.vert_off, NB + 1));
  HIPCHK(hipMemcpy(c.vert_off, vert_off.data(), (NB + 1) * 4, hipMemcpyHostToDevice));
  TRY(dalloc(b, &c.body_off, n_worlds + 1));
  HIPCHK(hipMemcpy(c.body_off, body_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
  // per-world solver workgroups take the largest worlds first, so the
  // longest solves start at once instead of trailing the batch
  std::vector<int32_t> order(n_worlds);
  for (int w = 0; w < n_worlds; ++w) order[w] = w;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
    return body_off[x + 1] - body_off[x] > body_off[y + 1] - body_off[y];
  });
  TRY(dalloc(b, &c.world_order, n_worlds));
  HIPCHK(hipMemcpy(c.world_order, order.data(), n_worlds * 4, hipMemcpyHostToDevice));
  TRY(dalloc(b, &c.lx, NV));
  TRY(dalloc(b, &c.ly, NV));
  HIPCHK(hipMemcpy(c.lx, lx.data(), NV * 8, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c.ly, ly.data(), NV * 8, hipMemcpyHostToDevice));
  TRY(dalloc(b, &c.ext_min, NV));
  TRY(dalloc(b, &c.ext_max, NV));
  TRY(dalloc(b, &c.wx, NV));
  TRY(dalloc(b, &c.wy, NV));
  TRY(dalloc(b, &c.wnx, NV));
  TRY(dalloc(b, &c.wny, NV));
  TRY(dalloc(b, &c.bminx, NB));
  TRY(dalloc(b, &c.bmaxx, NB));
  TRY(dalloc(b, &c.bminy, NB));
  TRY(dalloc(b, &c.bmaxy, NB));

  long PC = pair_cap, CC = 2 * pair_cap;
  TRY(dalloc(b, &c.cnt_rank, NB + 1));
  TRY(dalloc(b, &c.off_rank, NB + 1));
  TRY(dalloc(b, &c.pa, PC));
  TRY(dalloc(b, &c.pb, PC));
  TRY(dalloc(b, &c.ccount, PC + 1));
  TRY(dalloc(b, &c.coff, PC + 1));
  TRY(dalloc(b, &c.s_feat, 2 * PC));
  TRY(dalloc(b, &c.s_flip, 2 * PC));
  TRY(dalloc(b, &c.s_nx, 2 * PC));
  TRY(dalloc(b, &c.s_ny, 2 * PC));
  TRY(dalloc(b, &c.s_px, 2 * PC));
  TRY(dalloc(b, &c.s_py, 2 * PC));
  TRY(dalloc(b, &c.s_depth, 2 * PC));
  TRY(dalloc(b, &c.wpair_off, n_worlds + 1));
  TRY(dalloc(b, &c.wcont_off, n_worlds + 1));
  TRY(dalloc(b, &c.key, CC));
  TRY(dalloc(b, &c.ga, CC));
  TRY(dalloc(b, &c.gb, CC));
  TRY(dalloc(b, &c.flip, CC));
  TRY(dalloc(b, &c.cnx, CC));
  TRY(dalloc(b, &c.cny, CC));
  TRY(dalloc(b, &c.cpx, CC));
  TRY(dalloc(b, &c.cpy, CC));
  TRY(dalloc(b, &c.depth, CC));
  TRY(dalloc(b, &c.jn, 6 * CC));
  TRY(dalloc(b, &c.jt, 6 * CC));
  TRY(dalloc(b, &c.bn, CC));
  TRY(dalloc(b, &c.rax, CC));
  TRY(dalloc(b, &c.ray, CC));
  TRY(dalloc(b, &c.rbx, CC));
  TRY(dalloc(b, &c.rby, CC));
  TRY(dalloc(b, &c.mcn, CC));
  TRY(dalloc(b, &c.mct, CC));
  TRY(dalloc(b, &c.ln, CC));
  TRY(dalloc(b, &c.lt, CC));
  TRY(dalloc(b, &c.warm, CC));
  TRY(dalloc(b, &c.cache_key, CC));
  TRY(dalloc(b, &c.cache_ln, CC));
  TRY(dalloc(b, &c.cache_lt, CC));
  TRY(dalloc(b, &c.cache_woff, n_worlds + 1));
  c.miss_cap = (int)std::min<long>(2 * pair_cap, 4 * NB + 4096);
  TRY(dalloc(b, &c.miss_idx, c.miss_cap));
  TRY(dalloc(b, &c.miss_feat, (long)c.miss_cap * MLP_IN));
  TRY(dalloc(b, &c.log_count, 1));
  TRY(dalloc(b, &c.counters, 8));
  TRY(dalloc(b, &c.quality, Q_ENTRIES));
  TRY(solver_alloc(b));

  // grid broadphase set-up: cell size per world from the largest vertex
  // radius of its dynamic bodies; static bodies larger than a cell are
  // "large" and tested explicitly
  {
    int max_n = 0;
    for (int w = 0; w < n_worlds; ++w) max_n = std::max(max_n, body_off[w + 1] - body_off[w]);
    int mode = b->cfg.broadphase;
    c.use_grid = mode == SHP_BROADPHASE_GRID || (mode == SHP_BROADPHASE_AUTO && max_n > 2048);
    std::vector<double> rmax(NB, 0.0), cell(n_worlds, 1.0);
    for (long gg = 0; gg < NB; ++gg)
      for (int k = vert_off[gg]; k < vert_off[gg + 1]; ++k)
        rmax[gg] = std::max(rmax[gg], std::sqrt(lx[k] * lx[k] + ly[k] * ly[k]));
    std::vector<int32_t> tab_off(n_worlds + 1), large_off(n_worlds + 1), large_ids;
    std::vector<uint8_t> large(NB, 0);
    long tb_total = 0;
    for (int w = 0; w < n_worlds; ++w) {
      double hmax = 0.0;
      for (int gg = body_off[w]; gg < body_off[w + 1]; ++gg)
        if (!is_static[gg]) hmax = std::max(hmax, 2.0 * rmax[gg]);
      cell[w] = hmax > 0.0 ? hmax * (1.0 + 1e-6) + 1e-9 : 1.0;
      large_off[w] = (int32_t)large_ids.size();
      for (int gg = body_off[w]; gg < body_off[w + 1]; ++gg)
        if (is_static[gg] && 2.0 * rmax[gg] >= cell[w]) {
          large[gg] = 1;
          large_ids.push_back(gg);
        }
      int nw = body_off[w + 1] - body_off[w];
      long t = 64;
      while (t < nw) t <<= 1;
      tab_off[w] = (int32_t)tb_total;
      tb_total += t;
    }
    tab_off[n_worlds] = (int32_t)tb_total;
    large_off[n_worlds] = (int32_t)large_ids.size();
    c.grid_buckets = (int)tb_total;
    TRY(dalloc(b, &c.cell_h, n_worlds));
    HIPCHK(hipMemcpy(c.cell_h, cell.data(), n_worlds * 8, hipMemcpyHostToDevice));
    TRY(dalloc(b, &c.is_large, NB));
    HIPCHK(hipMemcpy(c.is_large, large.data(), NB, hipMemcpyHostToDevice));
    TRY(dalloc(b, &c.body_cell, 2 * NB));
    TRY(dalloc(b, &c.tab_off, n_worlds + 1));
    HIPCHK(hipMemcpy(c.tab_off, tab_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
    TRY(dalloc(b, &c.bucket_cnt, tb_total + 1));
    TRY(dalloc(b, &c.bucket_start, tb_total + 1));
    TRY(dalloc(b, &c.body_slot, NB));
    TRY(dalloc(b, &c.grid_rec, 6L * NB));
    TRY(dalloc(b, &c.grid_slots, (size_t)GRID_SLOTS * NB));
    TRY(dalloc(b, &c.large_off, n_worlds + 1));
    HIPCHK(hipMemcpy(c.large_off, large_off.data(), (n_worlds + 1) * 4, hipMemcpyHostToDevice));
    TRY(dalloc(b, &c.large_ids, std::max<size_t>(large_ids.size(), 1)));
    if (!large_ids.empty())
      HIPCHK(hipMemcpy(c.large_ids, large_ids.data(), large_ids.size() * 4, hipMemcpyHostToDevice));
  }

  size_t t1 = 0, t2 = 0;
  HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t1, c.cnt_rank, c.off_rank, c.NB + 1, b->stream));
  {
    hipcub::CountingInputIterator<int> idx(0);
    hipcub::TransformInputIterator<int, PairCount, hipcub::CountingInputIterator<int>> cnt(
        idx, PairCount{c.ccount, c.off_rank + c.NB, c.pair_cap});
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, t2, cnt, c.coff, c.pair_cap + 1, b->stream));
  }
  b->scan_tmp_bytes = std::max(std::max(t1, t2), grid_scan_bytes(b));
  TRY(dalloc(b, (uint8_t**)&b->scan_tmp, b->scan_tmp_bytes));

  hipStream_t s = b->stream;
  k_fill_i32<<<1, 1, 0, s>>>(c.cnt_rank + NB, 0, 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.cache_woff, 0, n_worlds + 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.wcont_off, 0, n_worlds + 1);
  k_fill_i32<<<blocks(n_worlds + 1, 256), 256, 0, s>>>(c.wpair_off, 0, n_worlds + 1);
  k_fill_i32<<<1, 64, 0, s>>>(c.counters, 0, 8);
  k_fill_i32<<<1, 64, 0, s>>>((int32_t*)c.quality, 0, 2 * Q_ENTRIES);
  k_fill_i32<<<1, 64, 0, s>>>((int32_t*)c.log_count, 0, 2);
  // makeWorldObj places each shape with moveShape (World.hs:169-176)
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
    for (int k = 0; k <= SHP_NUM_PHASES; ++k) (void)hipEventDestroy(b->events[k]);
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
    for (int k = 0; k <= SHP_NUM_PHASES; ++k) HIPCHK(hipEventCreate(&b->events[k]));
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
  std::vector<double> t(4 * (size_t)n);
  double* src[4] = {b->ctx.wx, b->ctx.wy, b->ctx.wnx, b->ctx.wny};
  for (int k = 0; k < 4; ++k)
    HIPCHK(hipMemcpy(t.data() + (size_t)k * n, src[k] + v0, n * 8, hipMemcpyDeviceToHost));
  for (int q = 0; q < n; ++q)
    for (int k = 0; k < 4; ++k) out[4 * q + k] = t[(size_t)k * n + q];
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
  b->graph_ready = false;  // kernel arguments changed
  return SHP_OK;
}

shp_status shp_batch_set_logging(shp_batch* b, int64_t capacity) {
  if (!b || capacity < 0) return fail(SHP_E_INVAL, "bad argument");
  HIPCHK(hipStreamSynchronize(b->stream));
  Ctx& c = b->ctx;
  if (capacity > c.log_cap) {
    TRY(dalloc(b, &c.log_feat, capacity * MLP_IN));
    TRY(dalloc(b, &c.log_tgt, capacity * 2));
    c.log_cap = capacity;
  }
  unsigned long long zero = 0;
  HIPCHK(hipMemcpy(c.log_count, &zero, 8, hipMemcpyHostToDevice));
  b->logging = capacity > 0;
  c.collect_misses = (b->cfg.warm_start == SHP_WARM_LEARNED && c.mlp_ready) || b->logging;
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
  for (int w = 0; w < W; ++w) {
    noff[w] = (int32_t)nk.size();
    if (w == world) {
      for (int q = 0; q < count; ++q) {
        const int32_t* kk = keys + 4 * q;
        if (kk[0] < 0 || kk[1] < 0 || kk[2] < 0 || kk[3] < 0 || kk[2] > 255 || kk[3] > 255)
          return fail(SHP_E_INVAL, "bad cache key");
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

}  // extern "C"
