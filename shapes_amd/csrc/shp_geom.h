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
struct HullRef {
  const double* wx;
  const double* wy;
  const double* nx;
  const double* ny;
  const uint8_t* emin;
  const uint8_t* emax;
  int n;
  __device__ __forceinline__ d2 v(int k) const { return mk(wx[k], wy[k]); }
  __device__ __forceinline__ d2 nrm(int k) const { return mk(nx[k], ny[k]); }
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
      r.x[k] = h.wx[k];
      r.y[k] = h.wy[k];
    }
}

// overlap + minOverlap fold (Contact/SAT.hs:103-138; extentAlong',
// ConvexHull.hs:81-92): first separating edge wins; otherwise the first
// strictly smallest depth.  Returns false if separated.
__device__ __forceinline__ bool min_overlap(const HullRef& se, const HullV& sp, Overlap* out) {
  Overlap best{0, 0, 0.0};
  for (int e = 0; e < se.n; ++e) {
    d2 dir = se.nrm(e);
    double smin = dot(se.v(se.emin[e]), dir);
    double smax = dot(se.v(se.emax[e]), dir);
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
  for (int k = 0; k < np; ++k) {
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
  Overlap oab, oba;
  HullV va, vb;
  load_hull_v(A, va);
  load_hull_v(B, vb);
  if (!min_overlap(A, vb, &oab)) return 0;
  if (!min_overlap(B, va, &oba)) return 0;
  if (oab.depth < oba.depth) {
    *flip_out = 0;
    return clip_contacts(A, B, oab, false, out);
  }
  *flip_out = 1;
  return clip_contacts(B, A, oba, true, out);
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
