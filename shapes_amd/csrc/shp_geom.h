// shp_geom.h - FP64 geometry and constraint arithmetic for gfx950 kernels.
//
// Every expression keeps the operation order of the reference's Haskell
// definitions (cited per function) so that, compiled with
// -ffp-contract=off, the device produces the same IEEE-754 results as the
// reference's SSE2 doubles.  Dot products are left folds
// (shapes-math/src/Shapes/Linear/Template.hs:108-110).
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SHP_MAXV 8  // vertices per hull supported by the kernels

namespace shp {

struct d2 {
  double x, y;
};

__device__ __forceinline__ d2 mk(double x, double y) { return d2{x, y}; }
__device__ __forceinline__ d2 sub(d2 a, d2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ d2 add(d2 a, d2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ d2 neg(d2 a) { return mk(-a.x, -a.y); }
// dotV2 (Template.hs:108-110)
__device__ __forceinline__ double dot(d2 a, d2 b) { return a.x * b.x + a.y * b.y; }
// crossV2 (Physics/Linear.hs:118-120)
__device__ __forceinline__ double cross(d2 a, d2 b) { return a.x * b.y - a.y * b.x; }
// clockwiseV2 (Linear.hs:161-163)
__device__ __forceinline__ d2 clockwise(d2 a) { return mk(a.y, -a.x); }
// zcrossV2 (Linear.hs:127-130)
__device__ __forceinline__ d2 zcross(double w, d2 b) { return mk(-(w * b.y), w * b.x); }
// normalizeV2 (Linear.hs:165-168); sqrt and / are IEEE correctly rounded
__device__ __forceinline__ d2 normalize(d2 a) {
  double n = __builtin_sqrt(a.x * a.x + a.y * a.y);
  return mk(a.x / n, a.y / n);
}
// Haskell's default Ord methods for Double
__device__ __forceinline__ double hs_min(double x, double y) { return x <= y ? x : y; }
__device__ __forceinline__ double hs_max(double x, double y) { return x <= y ? y : x; }

// dotV6 left fold
__device__ __forceinline__ double dot6(const double* a, const double* b) {
  return ((((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3]) + a[4] * b[4]) +
         a[5] * b[5];
}

// intersect2 (Linear.hs:244-251): invM2x2 (:194-199) then mul2x2c
__device__ __forceinline__ d2 intersect2(d2 p, d2 n, d2 q, d2 m) {
  double b0 = dot(p, n);
  double b1 = dot(q, m);
  double det = n.x * m.y - n.y * m.x;
  double inv = 1.0 / det;
  double i00 = m.y * inv, i01 = (-n.y) * inv, i10 = (-m.x) * inv, i11 = n.x * inv;
  return mk(i00 * b0 + i01 * b1, i10 * b0 + i11 * b1);
}

enum { CLIP_LEFT = 0, CLIP_RIGHT = 1, CLIP_BOTH = 2, CLIP_NONE = 3 };

// clipSegment (Linear.hs:327-343): boundary (bp, bn), incident line (ip, in)
__device__ __forceinline__ int clip_segment(d2 bp, d2 bn, d2 ip, d2 inn, d2 a, d2 b,
                                            d2* c_out) {
  d2 c = intersect2(bp, bn, ip, inn);
  double a_ = dot(a, bn), b_ = dot(b, bn), c_ = dot(c, bn);
  *c_out = c;
  if (a_ < c_) return b_ < c_ ? CLIP_BOTH : CLIP_LEFT;
  if (b_ < c_) return CLIP_RIGHT;
  return CLIP_NONE;
}

// A hull as the kernels see it: CSR offset, vertex count and device arrays.
// A hull in world space: per vertex (x, y, the unit normal of the edge to
// the next vertex) in 32 contiguous bytes, and per edge the extent vertex
// indices (min | max << 8), so a hull of n vertices spans 32n + 2n bytes.
struct HullRef {
  const double4* vn;
  const double2* sx;  // per edge: the hull's own extent along the edge normal
  int n;
  __device__ __forceinline__ d2 v(int k) const {
    const double2 p = *(const double2*)(vn + k);
    return mk(p.x, p.y);
  }
  __device__ __forceinline__ d2 nrm(int k) const {
    const double2 p = *((const double2*)(vn + k) + 1);
    return mk(p.x, p.y);
  }
};

struct Overlap {
  int edge;
  int pen;
  double depth;
};

// A hull's world vertices held in registers (the projections below read
// every vertex of the other hull once per edge).
struct HullV {
  double x[SHP_MAXV], y[SHP_MAXV];
  int n;
};

__device__ __forceinline__ void load_hull_v(const HullRef& h, HullV& r) {
  r.n = h.n;
#pragma unroll
  for (int k = 0; k < SHP_MAXV; ++k)
    if (k < h.n) {
      const d2 p = h.v(k);
      r.x[k] = p.x;
      r.y[k] = p.y;
    }
}

// overlap + minOverlap fold (Contact/SAT.hs:103-138): first separating
// edge wins; otherwise the first strictly smallest depth.  Returns false if
// separated.  se's own extent along each edge normal (extentAlong,
// ConvexHull.hs:81-92) comes precomputed with the step's transform; every
// edge's axis and extent are loaded before the first projection, so that a
// pair waits on one round of loads per hull rather than on two per edge.
__device__ __forceinline__ bool min_overlap(const HullRef& se, const HullV& sp, Overlap* out) {
  double ax[SHP_MAXV], ay[SHP_MAXV], lo[SHP_MAXV], hi[SHP_MAXV];
#pragma unroll
  for (int e = 0; e < SHP_MAXV; ++e)
    if (e < se.n) {
      const d2 d = se.nrm(e);
      const double2 s = se.sx[e];
      ax[e] = d.x;
      ay[e] = d.y;
      lo[e] = s.x;
      hi[e] = s.y;
    }
  Overlap best{0, 0, 0.0};
#pragma unroll
  for (int e = 0; e < SHP_MAXV; ++e) {
    if (e >= se.n) break;
    const d2 dir = mk(ax[e], ay[e]);
    const double smin = lo[e], smax = hi[e];
    double pmin = dot(mk(sp.x[0], sp.y[0]), dir), pmax = pmin;
    int imin = 0;
#pragma unroll
    for (int k = 1; k < SHP_MAXV; ++k) {
      if (k >= sp.n) break;
      double d = dot(mk(sp.x[k], sp.y[k]), dir);
      if (d < pmin) {
        pmin = d;
        imin = k;
      }
      if (d > pmax) pmax = d;
    }
    if (pmin > smax || pmax < smin) return false;
    double depth = smax - pmin;
    if (e == 0 || depth < best.depth) best = Overlap{e, imin, depth};
  }
  *out = best;
  return true;
}

struct Cgeom {
  int fa, fb;
  double nx, ny, px, py, depth;
};

__device__ __forceinline__ int next_i(int mx, int i) { return i < mx ? i + 1 : 0; }
__device__ __forceinline__ int prev_i(int mx, int i) { return i > 0 ? i - 1 : mx; }

// contact_ -> penetratingEdge -> clipEdge -> flatten (Contact/SAT.hs:152-267,
// Contact/HullVsHull.hs:24-76).  se is penetrated by sp.
__device__ __forceinline__ int clip_contacts(const HullRef& se, const HullRef& sp,
                                             const Overlap& ov, bool flip, Cgeom* out) {
  int mxp = sp.n - 1;
  d2 n = se.nrm(ov.edge);
  int bi = ov.pen;
  int ci = next_i(mxp, bi), ai = prev_i(mxp, bi);
  d2 aa = sp.v(ai), bb = sp.v(bi), cc = sp.v(ci);
  double abn = fabs(dot(sub(bb, aa), n));
  double bcn = fabs(dot(sub(cc, bb), n));
  d2 p0, p1;
  int i0, i1;
  if (bcn < abn) {
    p0 = bb; i0 = bi; p1 = cc; i1 = ci;
  } else {
    p0 = aa; i0 = ai; p1 = bb; i1 = bi;
  }
  d2 a = se.v(ov.edge);
  d2 b = se.v(next_i(se.n - 1, ov.edge));
  d2 cdp = p0;
  d2 cdn = clockwise(sub(p1, p0));  // toLine2 of the original incident edge
  d2 c;
  int r = clip_segment(a, sub(b, a), cdp, cdn, p0, p1, &c);  // perpLine2 a b
  if (r == CLIP_BOTH) return 0;
  if (r == CLIP_LEFT) p0 = c;
  else if (r == CLIP_RIGHT) p1 = c;
  r = clip_segment(b, sub(a, b), cdp, cdn, p0, p1, &c);  // perpLine2 b a
  if (r == CLIP_BOTH) return 0;
  if (r == CLIP_LEFT) p0 = c;
  else if (r == CLIP_RIGHT) p1 = c;
  r = clip_segment(a, neg(n), cdp, cdn, p0, p1, &c);  // Line2 a (negateV2 n)
  if (r == CLIP_BOTH) return 0;
  d2 q[2];
  int qi[2];
  int np;
  if (r == CLIP_LEFT) {
    q[0] = p1; qi[0] = i1; np = 1;
  } else if (r == CLIP_RIGHT) {
    q[0] = p0; qi[0] = i0; np = 1;
  } else {
    if (i0 > i1) {
      q[0] = p0; qi[0] = i0; q[1] = p1; qi[1] = i1;
    } else {
      q[0] = p1; qi[0] = i1; q[1] = p0; qi[1] = i0;
    }
    np = 2;
  }
  double da = dot(a, n);
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    if (k >= np) break;
    out[k].depth = da - dot(q[k], n);  // contactDepth (HullVsHull.hs:24-37)
    out[k].nx = n.x;
    out[k].ny = n.y;
    out[k].px = q[k].x;
    out[k].py = q[k].y;
    if (!flip) {
      out[k].fa = ov.edge;
      out[k].fb = qi[k];
    } else {
      out[k].fa = qi[k];
      out[k].fb = ov.edge;
    }
  }
  return np;
}

// HH.generateContacts via contactDebug + eitherBranchBoth (SAT.hs:238-249,
// Utils/Utils.hs:230-235).  *flip_out: 0 Same (A penetrated), 1 Flip.
__device__ __forceinline__ int hull_contacts(const HullRef& A, const HullRef& B, Cgeom* out,
                                             int* flip_out) {
  // one hull's vertices in registers at a time: B's against A's edges, then
  // A's against B's edges (the same reads of the same values)
  Overlap oab, oba;
  HullV hv;
  load_hull_v(B, hv);
  if (!min_overlap(A, hv, &oab)) return 0;
  load_hull_v(A, hv);
  if (!min_overlap(B, hv, &oba)) return 0;
  if (oab.depth < oba.depth) {
    *flip_out = 0;
    return clip_contacts(A, B, oab, false, out);
  }
  *flip_out = 1;
  return clip_contacts(B, A, oba, true, out);
}

// ---- circles and GJK ---------------------------------------------------------

// Circle.contact (Contact/Circle.hs:30-57): A penetrated, B penetrating, the
// normal out of A; a contact when (ra + rb)^2 >= |b - a|^2.
__device__ __forceinline__ bool circle_contact(d2 a, double ra, d2 b, double rb, Cgeom* g) {
  const d2 ab = sub(b, a);
  const double rab = ra + rb;
  const double ab_sq = ab.x * ab.x + ab.y * ab.y;  // sqLengthV2
  if (!(rab * rab >= ab_sq)) return false;
  const double len = __builtin_sqrt(ab_sq);
  const d2 abn = mk(ab.x / len, ab.y / len);                      // sdivV2
  const d2 ea = mk(abn.x * ra + a.x, abn.y * ra + a.y);           // edge of A
  const d2 eb = mk(abn.x * (-rb) + b.x, abn.y * (-rb) + b.y);     // edge of B
  g->px = (ea.x + eb.x) / 2.0;                                    // midpointP2
  g->py = (ea.y + eb.y) / 2.0;
  g->depth = (ra + rb) - len;
  g->nx = abn.x;
  g->ny = abn.y;
  return true;
}

// support (Contact/ConvexHull.hs:124-128): the first vertex of maximal
// dot(v, dir)
__device__ __forceinline__ int gjk_support(const HullV& h, d2 dir) {
  int best = 0;
  double bd = dot(mk(h.x[0], h.y[0]), dir);
#pragma unroll
  for (int k = 1; k < SHP_MAXV; ++k) {
    if (k >= h.n) break;
    const double d = dot(mk(h.x[k], h.y[k]), dir);
    if (d > bd) {
      bd = d;
      best = k;
    }
  }
  return best;
}

__device__ __forceinline__ bool same_direction(d2 a, d2 b) { return dot(a, b) > 0.0; }
// crossV2V2 (Linear.hs:135-139) and crosszV2 (:122-125)
__device__ __forceinline__ d2 cross_v2v2(d2 a, d2 b, d2 c) {
  const double abz = a.x * b.y - a.y * b.x;
  return mk(-(abz * c.y), abz * c.x);
}
__device__ __forceinline__ d2 crossz(d2 a, double z) { return mk(a.y * z, -(a.x * z)); }

#define GJK_MAX_ITERATIONS 32

// GJK.closestSimplex (Contact/GJK.hs:52-146) of a hull to point o: the
// simplex's vertex indices, most recently added first; 3 = encloses o.  The
// support search reads the vertices held in registers (h), the simplex's
// vertices by index come from the hull in memory (hr; the same values).  The
// reference's loop has no bound; this one keeps the current simplex after
// GJK_MAX_ITERATIONS (a hull of <= 8 vertices needs a few).
__device__ __forceinline__ int closest_simplex(const HullV& h, const HullRef& hr, d2 o, int* sv) {
  int n = 1;
  sv[0] = 0;
  d2 d = sub(o, mk(h.x[0], h.y[0]));
  for (int it = 0; it < GJK_MAX_ITERATIONS; ++it) {
    const int ai = gjk_support(h, d);
    // extendSimplex (GJK.hs:70-86): a repeat ends the search
    if (sv[0] == ai || (n == 2 && sv[1] == ai)) return n;
    const d2 a = hr.v(ai);
    const d2 ao = sub(o, a);
    if (n == 1) {  // shiftSimplex2 (GJK.hs:95-109)
      const int bi = sv[0];
      const d2 ab = sub(hr.v(bi), a);
      if (same_direction(ab, ao)) {
        n = 2;
        sv[0] = ai;
        sv[1] = bi;
        d = cross_v2v2(ab, ao, ab);
      } else {
        sv[0] = ai;
        d = ao;
      }
      continue;
    }
    // shiftSimplex3 (GJK.hs:111-137) on (a, b, c)
    const int bi = sv[0], ci = sv[1];
    const d2 ab = sub(hr.v(bi), a), ac = sub(hr.v(ci), a);
    const double abc = cross(ab, ac);
    const d2 abcac = zcross(abc, ac);
    const d2 ababc = crossz(ab, abc);
    bool star;
    if (same_direction(abcac, ao)) {
      if (same_direction(ac, ao)) {
        sv[0] = ai;
        sv[1] = ci;
        d = cross_v2v2(ac, ao, ac);
        continue;
      }
      star = true;
    } else {
      star = same_direction(ababc, ao);
    }
    if (!star) {  // the triangle encloses o
      sv[0] = ai;
      sv[1] = bi;
      sv[2] = ci;
      return 3;
    }
    if (same_direction(ab, ao)) {
      sv[0] = ai;
      sv[1] = bi;
      d = cross_v2v2(ab, ao, ab);
    } else {
      n = 1;
      sv[0] = ai;
      d = ao;
    }
  }
  return n;
}

// CircleVsHull.generateContacts (Contact/CircleVsHull.hs:18-69): the hull
// feature (the simplex's most recent vertex) and one contact, the circle
// the penetrator; none for a centre inside the hull or farther than the
// radius.
__device__ __forceinline__ bool circle_hull_contact(d2 center, double r, const HullV& h,
                                                    const HullRef& hr, int* feature, Cgeom* g) {
  int sv[3];
  const int n = closest_simplex(h, hr, center, sv);
  if (n == 3) return false;
  d2 a = hr.v(sv[0]);
  if (n == 2) {  // closestAlong (:135-144)
    const d2 p = a, q = hr.v(sv[1]);
    const d2 ao = sub(center, p);
    const d2 abn = normalize(sub(q, p));
    const double t = dot(ao, abn);
    a = mk(abn.x * t + p.x, abn.y * t + p.y);
  }
  // processSimplex_ (:117-133)
  const d2 ab = sub(center, a);
  const double ab_sq = ab.x * ab.x + ab.y * ab.y;
  if (r * r < ab_sq) return false;
  const double len = __builtin_sqrt(ab_sq);
  g->nx = -(ab.x / len);
  g->ny = -(ab.y / len);
  g->px = a.x;
  g->py = a.y;
  g->depth = r - len;
  *feature = sv[0];
  return true;
}

// NonPenetration.jacobian (Constraints/Contact/NonPenetration.hs:34-44)
__device__ __forceinline__ void np_jacobian(d2 n, d2 p, d2 xa, d2 xb, double* j) {
  d2 na = neg(n);
  j[0] = na.x;
  j[1] = na.y;
  j[2] = cross(sub(xa, p), n);
  j[3] = n.x;
  j[4] = n.y;
  j[5] = cross(sub(p, xb), n);
}

// Friction.jacobian (Constraints/Contact/Friction.hs:31-44)
__device__ __forceinline__ void f_jacobian(d2 n, d2 p, d2 xa, d2 xb, double* j) {
  d2 tb = clockwise(n);
  d2 ta = neg(tb);
  j[0] = ta.x;
  j[1] = ta.y;
  j[2] = cross(sub(p, xa), ta);
  j[3] = tb.x;
  j[4] = tb.y;
  j[5] = cross(sub(p, xb), tb);
}

// effMassM2 (Constraint.hs:173-179)
__device__ __forceinline__ double eff_mass(const double* j, const double* im) {
  double ji[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) ji[k] = j[k] * im[k];
  return dot6(ji, j);
}

}  // namespace shp
