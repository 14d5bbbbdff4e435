// shapes_oracle.cpp - CPU FP64 restatement of the shapes per-step path.
//
// TEST INFRASTRUCTURE ONLY (see shapes_oracle.h).  Each function names the
// reference definition it restates.  Arithmetic follows the reference's
// expression trees operation by operation (left-fold dot products from
// shapes-math/src/Shapes/Linear/Template.hs:108-110, matrix products from
// MatrixTemplate.hs:47-87), compiled with -ffp-contract=off so no fused
// multiply-add changes a rounding.
//
// Haskell's default Ord methods are used by the reference for Double
// (min x y = if x <= y then x else y; max x y = if x <= y then y else x);
// hs_min/hs_max below restate them, NaN behaviour included.

#include "shapes_oracle.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <vector>

#include "../shapes_amd/csrc/shp_trig.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct V2 {
  double x, y;
};

inline V2 v2(double x, double y) { return V2{x, y}; }
inline V2 sub(V2 a, V2 b) { return v2(a.x - b.x, a.y - b.y); }   // minusV2
inline V2 add(V2 a, V2 b) { return v2(a.x + b.x, a.y + b.y); }   // plusV2
inline V2 neg(V2 a) { return v2(-a.x, -a.y); }                    // negateV2
inline double dot(V2 a, V2 b) { return a.x * b.x + a.y * b.y; }   // dotV2
inline double cross(V2 a, V2 b) { return a.x * b.y - a.y * b.x; } // crossV2
inline V2 clockwise(V2 a) { return v2(a.y, -a.x); }              // clockwiseV2
// zcrossV2 (Linear.hs:127-130)
inline V2 zcross(double w, V2 b) { return v2(-(w * b.y), w * b.x); }
// normalizeV2 (Linear.hs:165-168)
inline V2 normalize(V2 a) {
  double n = sqrt(a.x * a.x + a.y * a.y);
  return v2(a.x / n, a.y / n);
}
inline double hs_min(double x, double y) { return x <= y ? x : y; }
inline double hs_max(double x, double y) { return x <= y ? y : x; }
inline double nz(double x) { return x == 0.0 ? 0.0 : x; }  // -0 -> +0

// dotV6 left fold (Template.hs:108-110)
inline double dot6(const double* a, const double* b) {
  return ((((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3]) +
          a[4] * b[4]) +
         a[5] * b[5];
}

// ---- bodies ---------------------------------------------------------------

struct Phys {  // PhysicalObj (Constraint.hs:52-57)
  V2 vel;
  double w;
  V2 pos;
  double th;
  double iml, imr;  // InvMass2
};

// toInvMass2 (Constraint.hs:79-83)
inline double invert_mass(double m) { return m == 0.0 ? 0.0 : 1.0 / m; }
// isStatic (Constraint.hs:123-125), isStaticLin (:128-130)
inline bool is_static(const Phys& p) { return p.iml == 0.0 && p.imr == 0.0; }
inline bool is_static_lin(const Phys& p) { return 0.0 == p.iml; }

struct Material {  // World.hs:35-38
  double mu, bounce;
};

// ConvexHull (ConvexHull.hs:61-68): vertex count, world vertices, world unit
// edge normals, per-edge cached extent indices, local vertices.
// A circle (Contact/Circle.hs:19-22) is a Hull with radius > 0, no
// vertices and its world centre.
struct Hull {
  int n;
  double radius = 0.0;
  V2 center{0.0, 0.0};
  std::vector<V2> verts;
  std::vector<V2> normals;
  std::vector<int> ext_min, ext_max;
  std::vector<V2> local;
};

inline int next_index(int max_i, int i) { return i < max_i ? i + 1 : 0; }
inline int prev_index(int max_i, int i) { return i > 0 ? i - 1 : max_i; }

// unitEdgeNormal (ConvexHull.hs:218-226)
inline V2 unit_edge_normal(const std::vector<V2>& vs, int i) {
  int mx = (int)vs.size() - 1;
  return normalize(clockwise(sub(vs[next_index(mx, i)], vs[i])));
}

// extentAlong' index fold (ConvexHull.hs:81-92): first index wins ties.
inline void extent_along(const std::vector<V2>& vs, V2 dir, int* mn, int* mx) {
  int imin = 0, imax = 0;
  double dmin = dot(vs[0], dir), dmax = dmin;
  for (int k = 1; k < (int)vs.size(); ++k) {
    double d = dot(vs[k], dir);
    if (d < dmin) { dmin = d; imin = k; }
    if (d > dmax) { dmax = d; imax = k; }
  }
  *mn = imin;
  *mx = imax;
}

// listToHull (ConvexHull.hs:151-167): normals and extents from the given
// vertices, which also become the local vertices.
Hull list_to_hull(const std::vector<V2>& vs) {
  Hull h;
  h.n = (int)vs.size();
  h.verts = vs;
  h.local = vs;
  h.normals.resize(h.n);
  h.ext_min.resize(h.n);
  h.ext_max.resize(h.n);
  for (int i = 0; i < h.n; ++i) h.normals[i] = unit_edge_normal(vs, i);
  for (int i = 0; i < h.n; ++i)
    extent_along(vs, h.normals[i], &h.ext_min[i], &h.ext_max[i]);
  return h;
}

// The forward matrix of toTransform pos ori = T(pos) * R(ori)
// (Transform.hs:34-38,73-77; Linear.hs:349-377), each entry a 3-term
// left-fold dot product (MatrixTemplate.hs:47-67).  The translation column
// reduces to pos exactly (GHC folds 0*... + p*1 with literal identities),
// which the -0 -> +0 normalisation at load makes equal to the IEEE sum.
struct Affine {
  double m00, m01, m02, m10, m11, m12;
};

Affine to_transform(V2 pos, double th, int trig) {
  double c, s;
  if (trig == ORC_TRIG_LIBM) {
    c = cos(th);
    s = sin(th);
  } else {
    shp_sincos(th, &s, &c);
  }
  const double one = 1.0, zer = 0.0;
  Affine m;
  // T = [[1,0,x],[0,1,y],[0,0,1]], R = [[c,-s,0],[s,c,0],[0,0,1]]
  m.m00 = (one * c + zer * s) + pos.x * zer;
  m.m01 = (one * (-s) + zer * c) + pos.x * zer;
  m.m02 = (one * zer + zer * zer) + pos.x * one;
  m.m10 = (zer * c + one * s) + pos.y * zer;
  m.m11 = (zer * (-s) + one * c) + pos.y * zer;
  m.m12 = (zer * zer + one * zer) + pos.y * one;
  return m;
}

// afmul (Linear.hs:217-220) = mul3x3c t (V3 a b 1)
inline V2 afmul(const Affine& t, V2 v) {
  return v2((t.m00 * v.x + t.m01 * v.y) + t.m02 * 1.0,
            (t.m10 * v.x + t.m11 * v.y) + t.m12 * 1.0);
}

// setHullTransform (ConvexHull.hs:184-195) via moveShape (World.hs:132-134);
// for a circle setCircleTransform (Circle.hs:60-64): the image of the local
// origin.
void set_hull_transform(Hull& h, const Phys& p, int trig) {
  Affine t = to_transform(p.pos, p.th, trig);
  if (h.radius > 0.0) {
    h.center = afmul(t, v2(0.0, 0.0));
    return;
  }
  for (int i = 0; i < h.n; ++i) h.verts[i] = afmul(t, h.local[i]);
  for (int i = 0; i < h.n; ++i) h.normals[i] = unit_edge_normal(h.verts, i);
}

struct Aabb {
  double minx, maxx, miny, maxy;
};

// toAabb (Aabb.hs:81-93): hullToAabb = foldl1 mergeAabb over the
// vertices, or circleToAabb = (x - r, x + r) x (y - r, y + r).
Aabb hull_aabb(const Hull& h) {
  if (h.radius > 0.0) {
    const double x = h.center.x, y = h.center.y, r = h.radius;
    return Aabb{x - r, x + r, y - r, y + r};
  }
  Aabb a{h.verts[0].x, h.verts[0].x, h.verts[0].y, h.verts[0].y};
  for (int i = 1; i < h.n; ++i) {
    double x = h.verts[i].x, y = h.verts[i].y;
    a.minx = a.minx < x ? a.minx : x;  // mergeRange: if a < c then a else c
    a.maxx = a.maxx > x ? a.maxx : x;
    a.miny = a.miny < y ? a.miny : y;
    a.maxy = a.maxy > y ? a.maxy : y;
  }
  return a;
}

// boundsOverlap / aabbCheck (Aabb.hs:69-78): inclusive, NaN overlaps.
inline bool bounds_overlap(double a, double b, double c, double d) {
  return !(c > b || d < a);
}
inline bool aabb_check(const Aabb& p, const Aabb& q) {
  return bounds_overlap(p.minx, p.maxx, q.minx, q.maxx) &&
         bounds_overlap(p.miny, p.maxy, q.miny, q.maxy);
}

// ---- SAT and clipping -------------------------------------------------------

struct Overlap {  // SAT.hs:32-37
  int edge;        // index of the penetrated edge in the edge hull
  double depth;
  int penetrator;  // index of the penetrating vertex in the other hull
};

// overlap (SAT.hs:103-118) + minOverlap fold (SAT.hs:121-138).
// Returns false when a separating axis exists.
bool min_overlap(const Hull& se, const Hull& sp, Overlap* out) {
  bool have = false;
  Overlap best{0, 0.0, 0};
  for (int e = 0; e < se.n; ++e) {
    V2 dir = se.normals[e];
    // extentAlongSelf: cached extent indices, projected in world space
    double smin = dot(se.verts[se.ext_min[e]], dir);
    double smax = dot(se.verts[se.ext_max[e]], dir);
    int pmi, pma;
    extent_along(sp.verts, dir, &pmi, &pma);
    double pmin = dot(sp.verts[pmi], dir);
    double pmax = dot(sp.verts[pma], dir);
    // overlapTest (SAT.hs:74-83)
    if (pmin > smax || pmax < smin) return false;  // first Separated wins
    double depth = smax - pmin;                     // overlapAmount
    if (!have || depth < best.depth) {
      best = Overlap{e, depth, pmi};
      have = true;
    }
  }
  *out = best;
  return true;
}

struct Line2 {
  V2 p, n;
};
inline Line2 to_line2(V2 a, V2 b) { return Line2{a, clockwise(sub(b, a))}; }
inline Line2 perp_line2(V2 a, V2 b) { return Line2{a, sub(b, a)}; }

// intersect2 (Linear.hs:244-251) with invM2x2 (:194-199) and mul2x2c.
V2 intersect2(const Line2& l1, const Line2& l2) {
  double a = l1.n.x, b = l1.n.y, c = l2.n.x, d = l2.n.y;
  double b0 = dot(l1.p, l1.n);
  double b1 = dot(l2.p, l2.n);
  double det = a * d - b * c;
  double inv = 1.0 / det;
  double i00 = d * inv, i01 = (-b) * inv, i10 = (-c) * inv, i11 = a * inv;
  return v2(i00 * b0 + i01 * b1, i10 * b0 + i11 * b1);
}

enum ClipKind { CLIP_LEFT, CLIP_RIGHT, CLIP_BOTH, CLIP_NONE };

// clipSegment (Linear.hs:327-343)
ClipKind clip_segment(const Line2& boundary, const Line2& incident, V2 a, V2 b,
                      V2* cpt) {
  V2 c = intersect2(boundary, incident);
  V2 n = boundary.n;
  double a_ = dot(a, n), b_ = dot(b, n), c_ = dot(c, n);
  *cpt = c;
  if (a_ < c_) return b_ < c_ ? CLIP_BOTH : CLIP_LEFT;
  if (b_ < c_) return CLIP_RIGHT;
  return CLIP_NONE;
}

struct Pt {  // a clipped incident-edge endpoint keeps its vertex index
  V2 p;
  int idx;
};

struct ContactGeom {
  int fa, fb;  // feature key (first shape's feature, second shape's)
  bool flip;
  V2 n, p;
  double depth;
};

// contact_ (SAT.hs:261-267) -> clipEdge (:190-218) -> flatten
// (HullVsHull.hs:54-76).  `se` is penetrated, `sp` penetrates.
int clip_contacts(const Hull& se, const Hull& sp, const Overlap& ov,
                  bool flip, ContactGeom* out) {
  int mxp = sp.n - 1;
  V2 n = se.normals[ov.edge];
  // penetratingEdge (SAT.hs:152-166)
  int bi = ov.penetrator;
  int ci = next_index(mxp, bi), ai = prev_index(mxp, bi);
  V2 aa = sp.verts[ai], bb = sp.verts[bi], cc = sp.verts[ci];
  double abn = fabs(dot(sub(bb, aa), n));
  double bcn = fabs(dot(sub(cc, bb), n));
  Pt i0, i1;
  if (bcn < abn) {
    i0 = Pt{bb, bi};
    i1 = Pt{cc, ci};
  } else {
    i0 = Pt{aa, ai};
    i1 = Pt{bb, bi};
  }
  // penetratedEdge (SAT.hs:169-171)
  V2 a = se.verts[ov.edge];
  V2 b = se.verts[next_index(se.n - 1, ov.edge)];
  Line2 cd = to_line2(i0.p, i1.p);  // plane of the original incident edge
  Line2 a_bound = perp_line2(a, b);
  Line2 b_bound = perp_line2(b, a);
  Line2 ab_bound = Line2{a, neg(n)};
  V2 c;
  switch (clip_segment(a_bound, cd, i0.p, i1.p, &c)) {  // lApplyClip'
    case CLIP_BOTH: return 0;
    case CLIP_LEFT: i0.p = c; break;
    case CLIP_RIGHT: i1.p = c; break;
    case CLIP_NONE: break;
  }
  switch (clip_segment(b_bound, cd, i0.p, i1.p, &c)) {
    case CLIP_BOTH: return 0;
    case CLIP_LEFT: i0.p = c; break;
    case CLIP_RIGHT: i1.p = c; break;
    case CLIP_NONE: break;
  }
  Pt pts[2];
  int np;
  switch (clip_segment(ab_bound, cd, i0.p, i1.p, &c)) {  // applyClip''
    case CLIP_BOTH: return 0;
    case CLIP_LEFT: pts[0] = i1; np = 1; break;
    case CLIP_RIGHT: pts[0] = i0; np = 1; break;
    default:
      // flattenContactPoints (SAT.hs:181-187): descending vertex index
      if (i0.idx > i1.idx) { pts[0] = i0; pts[1] = i1; }
      else { pts[0] = i1; pts[1] = i0; }
      np = 2;
      break;
  }
  for (int k = 0; k < np; ++k) {
    ContactGeom& g = out[k];
    // contactDepth (HullVsHull.hs:24-37)
    g.depth = dot(a, n) - dot(pts[k].p, n);
    g.n = n;
    g.p = pts[k].p;
    g.flip = flip;
    if (!flip) { g.fa = ov.edge; g.fb = pts[k].idx; }
    else { g.fa = pts[k].idx; g.fb = ov.edge; }  // flipExtractPair swaps
  }
  return np;
}

// HH.generateContacts (HullVsHull.hs:85-90) via contactDebug
// (SAT.hs:238-249) and eitherBranchBoth (Utils.hs:230-235).
int hull_contacts(const Hull& A, const Hull& B, ContactGeom* out, bool* sep) {
  Overlap oab, oba;
  bool ab = min_overlap(A, B, &oab);
  *sep = true;
  if (!ab) return 0;
  bool ba = min_overlap(B, A, &oba);
  if (!ba) return 0;
  *sep = false;
  if (oab.depth < oba.depth) return clip_contacts(A, B, oab, false, out);
  return clip_contacts(B, A, oba, true, out);
}

// ---- circles and GJK --------------------------------------------------------

// Circle.contact (Circle.hs:30-57): A penetrated, B penetrating; the normal
// points out of A.  Contact when (ra + rb)^2 >= |b - a|^2.
bool circle_contact(V2 a, double ra, V2 b, double rb, ContactGeom* c) {
  const V2 ab = sub(b, a);                          // diffP2 b a
  const double rab = ra + rb;
  const double ab_sq = ab.x * ab.x + ab.y * ab.y;   // sqLengthV2
  if (!(rab * rab >= ab_sq)) return false;
  const double len = std::sqrt(ab_sq);
  const V2 abn = v2(ab.x / len, ab.y / len);        // sdivV2 abLength ab
  const V2 ea = v2(abn.x * ra + a.x, abn.y * ra + a.y);        // vplusP2 (smulV2 ra abN) a
  const V2 eb = v2(abn.x * (-rb) + b.x, abn.y * (-rb) + b.y);  // vplusP2 (smulV2 (-rb) abN) b
  c->p = v2((ea.x + eb.x) / 2.0, (ea.y + eb.y) / 2.0);         // midpointP2
  c->depth = (ra + rb) - len;
  c->n = abn;
  return true;
}

// GJK.closestSimplex (GJK.hs:52-146) of a hull to a point: vertex indices of
// the simplex, most recently added first (a 1-, 2- or 3-simplex).
struct Simplex {
  int n;
  int v[3];
};

// support (ConvexHull.hs:124-128): the first vertex of maximal dot(v, dir)
inline int support(const Hull& h, V2 dir) {
  int best = 0;
  double bd = dot(h.verts[0], dir);
  for (int k = 1; k < h.n; ++k) {
    const double d = dot(h.verts[k], dir);
    if (d > bd) {
      bd = d;
      best = k;
    }
  }
  return best;
}

inline bool same_direction(V2 a, V2 b) { return dot(a, b) > 0.0; }  // GJK.hs:139-140
// crossV2V2 (Linear.hs:135-139): (a x b) x c
inline V2 cross_v2v2(V2 a, V2 b, V2 c) {
  const double abz = a.x * b.y - a.y * b.x;
  return v2(-(abz * c.y), abz * c.x);
}
// crosszV2 (Linear.hs:122-125)
inline V2 crossz(V2 a, double z) { return v2(a.y * z, -(a.x * z)); }

// The loop stops when the support point repeats (a found simplex) or the
// 2-simplex encloses the point.  The reference has no iteration bound; the
// loop here stops after GJK_MAX_ITERATIONS, keeping the current simplex
// (no hull of <= 8 vertices takes more than a few).
#define GJK_MAX_ITERATIONS 32
Simplex closest_simplex(const Hull& h, V2 o) {
  Simplex s{1, {0, 0, 0}};
  V2 d = sub(o, h.verts[0]);
  for (int it = 0; it < GJK_MAX_ITERATIONS; ++it) {
    const int ai = support(h, d);
    // extendSimplex (GJK.hs:70-86): a repeat ends the search
    for (int k = 0; k < s.n; ++k)
      if (s.v[k] == ai) return s;
    if (s.n == 1) {
      // shiftSimplex2 (GJK.hs:95-109) on (a, b)
      const int bi = s.v[0];
      const V2 a = h.verts[ai], b = h.verts[bi];
      const V2 ab = sub(b, a), ao = sub(o, a);
      if (same_direction(ab, ao)) {
        s = Simplex{2, {ai, bi, 0}};
        d = cross_v2v2(ab, ao, ab);
      } else {
        s = Simplex{1, {ai, 0, 0}};
        d = ao;
      }
    } else {
      // shiftSimplex3 (GJK.hs:111-137) on (a, b, c)
      const int bi = s.v[0], ci = s.v[1];
      const V2 a = h.verts[ai], b = h.verts[bi], c = h.verts[ci];
      const V2 ab = sub(b, a), ac = sub(c, a), ao = sub(o, a);
      const double abc = cross(ab, ac);
      const V2 abcac = zcross(abc, ac);
      const V2 ababc = crossz(ab, abc);
      bool star;
      if (same_direction(abcac, ao)) {
        if (same_direction(ac, ao)) {
          s = Simplex{2, {ai, ci, 0}};
          d = cross_v2v2(ac, ao, ac);
          continue;
        }
        star = true;
      } else {
        star = same_direction(ababc, ao);
      }
      if (!star) return Simplex{3, {ai, bi, ci}};  // encloses the point
      if (same_direction(ab, ao)) {
        s = Simplex{2, {ai, bi, 0}};
        d = cross_v2v2(ab, ao, ab);
      } else {
        s = Simplex{1, {ai, 0, 0}};
        d = ao;
      }
    }
  }
  return s;
}

// CircleVsHull.generateContacts (CircleVsHull.hs:18-69): the hull feature
// (the simplex's most recent vertex) and the contact, the circle always the
// penetrator; none when the centre is inside the hull (3-simplex) or
// farther than the radius from the closest point.
bool circle_hull_contact(V2 center, double r, const Hull& h, int* feature, ContactGeom* c) {
  const Simplex s = closest_simplex(h, center);
  if (s.n == 3) return false;
  V2 a = h.verts[s.v[0]];
  if (s.n == 2) {  // closestAlong (:135-144)
    const V2 p = h.verts[s.v[0]], q = h.verts[s.v[1]];
    const V2 ao = sub(center, p);
    const V2 abn = normalize(sub(q, p));
    const double t = dot(ao, abn);
    a = v2(abn.x * t + p.x, abn.y * t + p.y);
  }
  // processSimplex_ (:117-133)
  const V2 ab = sub(center, a);
  const double ab_sq = ab.x * ab.x + ab.y * ab.y;
  if (r * r < ab_sq) return false;
  const double len = std::sqrt(ab_sq);
  const V2 nrm = v2(ab.x / len, ab.y / len);
  c->n = neg(nrm);
  c->p = a;
  c->depth = r - len;
  *feature = s.v[0];
  return true;
}

// Contact.generateContacts (Contact.hs:22-40) for a pair (A, B) = (shape i,
// shape j): circle-circle (0, 0) Same; circle-hull (0, f) Same; hull-circle
// (f, 0) Flip (the contact computed circle against hull); hull-hull as SAT.
int shape_contacts(const Hull& A, const Hull& B, ContactGeom* out, bool* sep) {
  const bool ca = A.radius > 0.0, cb = B.radius > 0.0;
  if (!ca && !cb) return hull_contacts(A, B, out, sep);
  *sep = false;
  ContactGeom g;
  int f = 0;
  if (ca && cb) {
    if (!circle_contact(A.center, A.radius, B.center, B.radius, &g)) return 0;
    g.fa = 0;
    g.fb = 0;
    g.flip = false;
  } else if (ca) {
    if (!circle_hull_contact(A.center, A.radius, B, &f, &g)) return 0;
    g.fa = 0;
    g.fb = f;
    g.flip = false;
  } else {
    if (!circle_hull_contact(B.center, B.radius, A, &f, &g)) return 0;
    g.fa = f;
    g.fb = 0;
    g.flip = true;
  }
  out[0] = g;
  return 1;
}

// ---- constraints -------------------------------------------------------------

struct Key {  // ObjectFeatureKey (Constraints/Contact.hs:36-39)
  int i, j, fi, fj;
};
inline bool key_less(const Key& a, const Key& b) {
  if (a.i != b.i) return a.i < b.i;
  if (a.j != b.j) return a.j < b.j;
  if (a.fi != b.fi) return a.fi < b.fi;
  return a.fj < b.fj;
}
inline bool key_eq(const Key& a, const Key& b) {
  return a.i == b.i && a.j == b.j && a.fi == b.fi && a.fj == b.fj;
}

struct Constraint {  // ContactConstraint (Constraints/Types.hs:47-51)
  double jn[6], bn;  // non-penetration row and Baumgarte bias
  double jt[6];      // friction row (bias 0)
  V2 ra, rb, nr;     // RestitutionConstraint
};

struct Contact {
  Key key;
  ContactGeom g;
  Constraint cc;
  double ln, lt;  // cached/accumulated lagrangians
  bool warm;
};

// NP.jacobian (NonPenetration.hs:34-44) on (a,b) = (penetrated, penetrator)
void np_jacobian(const ContactGeom& c, V2 xa, V2 xb, double* j) {
  V2 n = c.n;
  V2 na = neg(n);
  j[0] = na.x; j[1] = na.y; j[2] = cross(sub(xa, c.p), n);
  j[3] = n.x;  j[4] = n.y;  j[5] = cross(sub(c.p, xb), n);
}

// F.jacobian (Friction.hs:31-44)
void f_jacobian(const ContactGeom& c, V2 xa, V2 xb, double* j) {
  V2 tb = clockwise(c.n);
  V2 ta = neg(tb);
  j[0] = ta.x; j[1] = ta.y; j[2] = cross(sub(c.p, xa), ta);
  j[3] = tb.x; j[4] = tb.y; j[5] = cross(sub(c.p, xb), tb);
}

inline void flip3v3(double* j) {  // Linear.hs:149-151
  for (int k = 0; k < 3; ++k) std::swap(j[k], j[k + 3]);
}

// constraintGen (Constraints/Contact.hs:60-72): NP and F are generated
// for the swapped pair and their halves swapped back when the contact is
// Flip (flipMap + flipExtract, Utils.hs:168-209); the restitution radii
// always use (a,b) = (obj_i, obj_j) (Restitution.hs:21-31).
Constraint constraint_gen(double beta, double slop, double dt,
                          const ContactGeom& c, const Phys& a, const Phys& b) {
  Constraint k;
  if (!c.flip) {
    np_jacobian(c, a.pos, b.pos, k.jn);
    f_jacobian(c, a.pos, b.pos, k.jt);
  } else {
    np_jacobian(c, b.pos, a.pos, k.jn);
    flip3v3(k.jn);
    f_jacobian(c, b.pos, a.pos, k.jt);
    flip3v3(k.jt);
  }
  // baumgarte (NonPenetration.hs:48-56)
  double d = c.depth;
  k.bn = d > slop ? (beta / dt) * (slop - d) : 0.0;
  k.ra = sub(c.p, a.pos);
  k.rb = sub(c.p, b.pos);
  k.nr = c.flip ? neg(c.n) : c.n;
  return k;
}

inline void vel6(const Phys& a, const Phys& b, double* v) {  // velocity2
  v[0] = a.vel.x; v[1] = a.vel.y; v[2] = a.w;
  v[3] = b.vel.x; v[4] = b.vel.y; v[5] = b.w;
}
inline void im6(const Phys& a, const Phys& b, double* m) {  // invMassM2
  m[0] = a.iml; m[1] = a.iml; m[2] = a.imr;
  m[3] = b.iml; m[4] = b.iml; m[5] = b.imr;
}

// effMassM2 (Constraint.hs:173-179): (J (*) im) . J
inline double eff_mass(const double* j, const double* im) {
  double ji[6];
  for (int k = 0; k < 6; ++k) ji[k] = j[k] * im[k];
  return dot6(ji, j);
}

// lagrangian2 (Constraint.hs:164-169)
inline double lagrangian2(const double* j, double b, const double* v,
                          double mc) {
  return (-(dot6(j, v) + b)) / mc;
}

// applyLagrangian (Constraint.hs:182-221): v_k + ((j_k*l)*im_k)
inline void apply_lagrangian(double l, const double* j, Phys& a, Phys& b) {
  double v[6], im[6];
  vel6(a, b, v);
  im6(a, b, im);
  for (int k = 0; k < 6; ++k) v[k] = v[k] + ((j[k] * l) * im[k]);
  a.vel = v2(v[0], v[1]); a.w = v[2];
  b.vel = v2(v[3], v[4]); b.w = v[5];
}

// applySln (Solvers/Contact.hs:55-65): normal first, then friction
inline void apply_sln(double ln, double lt, const Constraint& c, Phys& a,
                      Phys& b) {
  apply_lagrangian(ln, c.jn, a, b);
  apply_lagrangian(lt, c.jt, a, b);
}

// bounceB (Restitution.hs:34-47)
inline double bounce_b(double ea, double eb, const Constraint& c,
                       const Phys& a, const Phys& b) {
  V2 nva = neg(a.vel);
  V2 nwa = zcross(-a.w, c.ra);
  V2 wb = zcross(b.w, c.rb);
  V2 closing = add(add(add(nva, nwa), b.vel), wb);
  double e = hs_min(ea, eb);
  return hs_min(0.0, e * dot(closing, c.nr));
}

}  // namespace

// ---- engine ----------------------------------------------------------------

struct orc_engine {
  orc_params prm;
  std::vector<Phys> objs;
  std::vector<Material> mats;
  std::vector<Hull> hulls;
  // EmptiesVector (Utils/EmptiesVector.hs:17-22): ev[s] == -1 for a filled
  // slot, else the next empty slot (the capacity ends the list)
  std::vector<int> ev;
  int ev_head = 0, ev_filled = 0;
  std::vector<std::pair<int, int>> pairs;  // last step
  std::vector<int> pair_first;             // pairs + 1: first contact of each pair
  std::vector<Contact> contacts;           // last step == the cache
  std::vector<Contact> probe;              // contacts of the last probe
  bool probed = false;                     // readbacks show the probe
  orc_stats stats;
};

namespace {

// culledKeys (Aabb.hs:168-183): {(i,j) | i > j, not both static, inclusive
// AABB overlap} in descending lexicographic order, by testing every pair.
void culled_keys_brute(const std::vector<Aabb>& boxes, const std::vector<char>& st,
                       std::vector<std::pair<int, int>>& out) {
  const int n = (int)boxes.size();
  out.clear();
  for (int i = n - 1; i >= 1; --i)
    for (int j = i - 1; j >= 0; --j)
      if (!(st[i] && st[j]) && aabb_check(boxes[i], boxes[j]))
        out.push_back(std::make_pair(i, j));
}

// axialIndex (Grid.hs:127-129) with GridAxis 20 1 (-10) (Engine/Main.hs:
// 40-41): floor((v - origin) / unit).  Clamped to +-2^30 so that huge or
// non-finite values stay defined (the reference's floor :: Double -> Int
// has no defined value for them; finite worlds never reach the clamp).
inline int axial_index(double v) {
  double q = std::floor((v - -10.0) / 1.0);
  if (!(q == q)) return 0;
  if (q > 1073741824.0) return 1073741824;
  if (q < -1073741824.0) return -1073741824;
  return (int)q;
}

// toGrid + culledKeys (Grid.hs:67-100,102-141): every AABB is inserted into
// each flattened cell x + 20 y its index ranges cover (boxIndices, flattened
// indices alias across rows); each cell's shapes in descending key order
// give allPairs, pairs of two static shapes are skipped and the rest kept
// on aabbCheck; the concatenation is sorted descending and deduplicated
// (uniq).  Cells are gathered here by sorting (cell, key) entries instead
// of IntMap inserts: the same cells with the same members.
void culled_keys_grid(const std::vector<Aabb>& boxes, const std::vector<char>& st,
                      std::vector<std::pair<int, int>>& out) {
  const int n = (int)boxes.size();
  std::vector<std::pair<long long, int>> ent;  // (cell, -key): keys descending in a cell
  for (int k = 0; k < n; ++k) {
    const Aabb& b = boxes[k];
    const int x0 = axial_index(b.minx), x1 = axial_index(b.maxx);
    const int y0 = axial_index(b.miny), y1 = axial_index(b.maxy);
    for (long long x = x0; x <= x1; ++x)
      for (long long y = y0; y <= y1; ++y) ent.push_back(std::make_pair(x + y * 20, -k));
  }
  std::sort(ent.begin(), ent.end());
  out.clear();
  size_t a = 0;
  while (a < ent.size()) {
    size_t e = a;
    while (e < ent.size() && ent[e].first == ent[a].first) ++e;
    // a shape inserted twice into one flattened cell (aliasing) is one
    // IntMap entry
    std::vector<int> keys;
    for (size_t q = a; q < e; ++q)
      if (keys.empty() || keys.back() != -ent[q].second) keys.push_back(-ent[q].second);
    for (size_t p = 0; p < keys.size(); ++p)
      for (size_t q = p + 1; q < keys.size(); ++q) {
        const int i = keys[p], j = keys[q];  // i > j
        if (!(st[i] && st[j]) && aabb_check(boxes[i], boxes[j])) out.push_back(std::make_pair(i, j));
      }
    a = e;
  }
  std::sort(out.begin(), out.end(), [](const std::pair<int, int>& x, const std::pair<int, int>& y) {
    return x > y;
  });
  out.erase(std::unique(out.begin(), out.end()), out.end());
}

inline bool filled(const orc_engine* e, int i) { return e->ev[i] == -1; }

// toTaggedAabbs (Aabb.hs:140-148) lists the filled slots only (E.mapM),
// keyed by slot: the pairs are found among them and named by their slots.
void culled_keys(orc_engine* e, std::vector<std::pair<int, int>>& out) {
  int n = (int)e->objs.size();
  std::vector<int> slot;
  for (int i = 0; i < n; ++i)
    if (filled(e, i)) slot.push_back(i);
  const int m = (int)slot.size();
  std::vector<Aabb> boxes(m);
  std::vector<char> st(m);
  for (int k = 0; k < m; ++k) {
    boxes[k] = hull_aabb(e->hulls[slot[k]]);
    st[k] = is_static(e->objs[slot[k]]);
  }
  if (e->prm.broadphase == ORC_BROADPHASE_BRUTE)
    culled_keys_brute(boxes, st, out);
  else
    culled_keys_grid(boxes, st, out);
  for (auto& p : out) p = std::make_pair(slot[p.first], slot[p.second]);
}

// prepareFrame (Solvers/Contact.hs:40-52) + keyedContacts
// (Constraints/Contact.hs:49-57)
void prepare_frame(orc_engine* e, std::vector<Contact>& out) {
  out.clear();
  e->pair_first.clear();
  for (auto& pr : e->pairs) {
    e->pair_first.push_back((int)out.size());
    ContactGeom g[2];
    bool sep;
    int k = shape_contacts(e->hulls[pr.first], e->hulls[pr.second], g, &sep);
    for (int q = 0; q < k; ++q) {
      Contact c;
      c.key = Key{pr.first, pr.second, g[q].fa, g[q].fb};
      c.g = g[q];
      c.ln = c.lt = 0.0;
      c.warm = false;
      out.push_back(c);
    }
  }
  e->pair_first.push_back((int)out.size());
}

// applyCachedSlns (Solvers/Contact.hs:74-121) with descZipVector
// (Utils/Descending.hs:47-71): merge-join against last frame's cache,
// applying reused lagrangians immediately in contact order; new contacts
// start at zero (:95).
void apply_cached_slns(orc_engine* e, std::vector<Contact>& cs,
                       const std::vector<Contact>& old, bool apply) {
  size_t that_i = 0;
  for (auto& c : cs) {
    Phys& a = e->objs[c.key.i];
    Phys& b = e->objs[c.key.j];
    c.cc = constraint_gen(e->prm.baumgarte, e->prm.slop, e->prm.dt, c.g, a, b);
    bool hit = false;
    while (that_i < old.size()) {
      const Key& tk = old[that_i].key;
      if (key_less(c.key, tk)) { ++that_i; continue; }
      if (key_eq(c.key, tk)) hit = true;
      break;
    }
    if (hit) {
      c.ln = old[that_i].ln;
      c.lt = old[that_i].lt;
      c.warm = true;
      if (apply) apply_sln(c.ln, c.lt, c.cc, a, b);
      ++that_i;
    } else {
      c.ln = 0.0;
      c.lt = 0.0;
      c.warm = false;
    }
  }
}

// improveContactSln (Solvers/Contact.hs:124-143) with contactLagrangian
// (Constraints/Contact.hs:86-95) and solutionProcessor (:101-110):
// positive (SolutionProcessors.hs:33-35) then clampAbs (:42-53).
void improve_contact(orc_engine* e, Contact& c) {
  Phys& a = e->objs[c.key.i];
  Phys& b = e->objs[c.key.j];
  const Material& ma = e->mats[c.key.i];
  const Material& mb = e->mats[c.key.j];
  double v[6], im[6];
  vel6(a, b, v);
  im6(a, b, im);
  double bn = c.cc.bn + bounce_b(ma.bounce, mb.bounce, c.cc, a, b);
  double new_n = lagrangian2(c.cc.jn, bn, v, eff_mass(c.cc.jn, im));
  double new_t = lagrangian2(c.cc.jt, 0.0, v, eff_mass(c.cc.jt, im));
  // positive: apply = max new (-cached)
  double app_n = hs_max(new_n, -c.ln);
  double acc_n = c.ln + app_n;
  // clampAbs with threshold acc_n * pairMu (Friction.hs:46-48,50-55)
  double mu = (ma.mu + mb.mu) / 2.0;
  double thr = acc_n * mu;
  double acc_t = c.lt + new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - c.lt;
  apply_sln(app_n, app_t, c.cc, a, b);
  c.ln = acc_n;
  c.lt = accp;
}

// ---- graph-coloured FP32 Gauss-Seidel (the GPU's SHP_SOLVER_COLORED_F32) ----
//
// Not a reference function: the same per-contact update as improve_contact
// in FP32 (FP64 accumulators), in colour order.  Manifolds (pairs with
// contacts) get colours from Jones-Plassmann rounds over hashed priorities;
// a colour's manifolds share no dynamic body, so their order inside a
// colour does not matter; the colours run in order, the warm-start
// application (applyCachedSlns) too.

uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// Priority of manifold q (pair index in the world) of world w: 15 hash bits,
// then the index (distinct within a world).
uint64_t colour_prio(int q, int w) {
  return ((uint64_t)(mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) & 0x7FFFu) << 32) |
         (uint64_t)((uint32_t)q + 1u);
}

const int COLOUR_CAP = 256;
const int GREEDY_COLOURS = 16;                          // colours from the body masks
const int JP_ROUNDS = COLOUR_CAP - GREEDY_COLOURS - 3;  // rounds before the serial tail

// colour[q] >= 0 for pairs with contacts.  Jones-Plassmann rounds select the
// manifolds whose priority is the largest among the uncoloured ones on each
// of their dynamic bodies; a selected manifold takes the smallest colour
// that no manifold on its dynamic bodies has (colours below 16, tracked per
// body), else 16 + the round.  The last colour is the serial tail (pair
// order) when JP_ROUNDS rounds were not enough.  Returns the number of
// colours.
int colour_pairs(orc_engine* e, std::vector<int>& colour, bool* tail) {
  const int np = (int)e->pairs.size();
  const int n = (int)e->objs.size();
  colour.assign(np, -2);
  int left = 0;
  for (int q = 0; q < np; ++q)
    if (e->pair_first[q + 1] > e->pair_first[q]) {
      colour[q] = -1;
      ++left;
    }
  std::vector<uint64_t> best(n);
  std::vector<uint32_t> used(n, 0u);
  int rounds = 0, top = 0;
  while (left > 0 && rounds < JP_ROUNDS) {
    std::fill(best.begin(), best.end(), 0ull);
    for (int q = 0; q < np; ++q) {
      if (colour[q] != -1) continue;
      uint64_t pr = colour_prio(q, e->prm.world);
      int a = e->pairs[q].first, b = e->pairs[q].second;
      if (!is_static(e->objs[a]) && pr > best[a]) best[a] = pr;
      if (!is_static(e->objs[b]) && pr > best[b]) best[b] = pr;
    }
    std::vector<int> take;
    for (int q = 0; q < np; ++q) {
      if (colour[q] != -1) continue;
      uint64_t pr = colour_prio(q, e->prm.world);
      int a = e->pairs[q].first, b = e->pairs[q].second;
      if ((is_static(e->objs[a]) || best[a] == pr) && (is_static(e->objs[b]) || best[b] == pr))
        take.push_back(q);
    }
    for (int q : take) {
      int a = e->pairs[q].first, b = e->pairs[q].second;
      const bool sa = is_static(e->objs[a]), sb = is_static(e->objs[b]);
      const uint32_t u = (sa ? 0u : used[a]) | (sb ? 0u : used[b]);
      int l = GREEDY_COLOURS + rounds;
      for (int k = 0; k < GREEDY_COLOURS; ++k)
        if (!((u >> k) & 1u)) {
          l = k;
          break;
        }
      colour[q] = l;
      if (l < GREEDY_COLOURS) {
        if (!sa) used[a] |= 1u << l;
        if (!sb) used[b] |= 1u << l;
      }
      top = std::max(top, l + 1);
    }
    left -= (int)take.size();
    ++rounds;
  }
  *tail = left > 0;
  for (int q = 0; q < np; ++q)
    if (colour[q] == -1) colour[q] = top;
  return top + (*tail ? 1 : 0);
}

// The coloured FP32 solver's arithmetic (shp_colored.hip, solve_man):
// a . b as a product and five fused multiply-adds in index order; v += l*w
// fused per entry.  fmaf rounds once, on the CPU as on the GPU.
inline float dot6f(const float* a, const float* b) {
  float d = a[0] * b[0];
  for (int q = 1; q < 6; ++q) d = fmaf(a[q], b[q], d);
  return d;
}

inline void applyw(float l, const float* w, float* v) {
  for (int q = 0; q < 6; ++q) v[q] = fmaf(l, w[q], v[q]);
}

// Pass `pass` (0: warm start) over the contacts [k0, k1) of one manifold
// between bodies a and b, FP32 velocities vf (3 per body).
void solve_manifold_f32(orc_engine* e, std::vector<Contact>& cs, int k0, int k1, int pass,
                        std::vector<float>& vf) {
  const int ia = cs[k0].key.i, ib = cs[k0].key.j;
  const Phys& A = e->objs[ia];
  const Phys& B = e->objs[ib];
  const Material& ma = e->mats[ia];
  const Material& mb = e->mats[ib];
  // the row normal of the manifold's first contact times the inverse
  // masses, formed in FP64 and rounded once
  const double ux = cs[k0].cc.jn[3], uy = cs[k0].cc.jn[4];
  const float ax = (float)(A.iml * ux), ay = (float)(A.iml * uy);
  const float bx = (float)(B.iml * ux), by = (float)(B.iml * uy);
  const float mu = (float)((ma.mu + mb.mu) / 2.0);
  const float eb = (float)hs_min(ma.bounce, mb.bounce);
  float v[6] = {vf[3 * ia], vf[3 * ia + 1], vf[3 * ia + 2], vf[3 * ib], vf[3 * ib + 1], vf[3 * ib + 2]};
  bool changed = false;
  for (int k = k0; k < k1; ++k) {
    Contact& c = cs[k];
    float jn[6], jt[6];
    for (int q = 0; q < 6; ++q) {
      jn[q] = (float)c.cc.jn[q];
      jt[q] = (float)c.cc.jt[q];
    }
    // the rows times the inverse masses: linear parts from the manifold's
    // products (the friction row's direction is clockwise of the normal),
    // angular parts rounded once from FP64
    const float wn[6] = {-ax, -ay, (float)(A.imr * c.cc.jn[2]), bx, by, (float)(B.imr * c.cc.jn[5])};
    const float wt[6] = {-ay, ax, (float)(A.imr * c.cc.jt[2]), by, -bx, (float)(B.imr * c.cc.jt[5])};
    if (pass == 0) {
      if (!c.warm) continue;
      applyw((float)c.ln, wn, v);
      applyw((float)c.lt, wt, v);
      changed = true;
      continue;
    }
    double imd[6];
    im6(A, B, imd);
    // reciprocal effective masses, rounded once from FP64
    const float icn = (float)(1.0 / eff_mass(c.cc.jn, imd));
    const float ict = (float)(1.0 / eff_mass(c.cc.jt, imd));
    // the bounce term's closing velocity (bounceB) is the normal row times
    // the velocities
    const float dn = dot6f(jn, v), dt = dot6f(jt, v);
    float bb = eb * dn;
    bb = 0.0f <= bb ? 0.0f : bb;
    float new_n = -(dn + ((float)c.cc.bn + bb)) * icn;
    float new_t = -dt * ict;
    double app_n = hs_max((double)new_n, -c.ln);
    double acc_n = c.ln + app_n;
    double thr = acc_n * (double)mu;
    double acc_t = c.lt + (double)new_t;
    double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
    double app_t = accp - c.lt;
    applyw((float)app_n, wn, v);
    applyw((float)app_t, wt, v);
    c.ln = acc_n;
    c.lt = accp;
    changed = true;
  }
  if (!changed) return;
  if (!is_static(A)) {
    vf[3 * ia] = v[0];
    vf[3 * ia + 1] = v[1];
    vf[3 * ia + 2] = v[2];
  }
  if (!is_static(B)) {
    vf[3 * ib] = v[3];
    vf[3 * ib + 1] = v[4];
    vf[3 * ib + 2] = v[5];
  }
}

void solve_coloured(orc_engine* e, std::vector<Contact>& cs) {
  std::vector<int> colour;
  bool tail = false;
  const int ncol = colour_pairs(e, colour, &tail);
  const int n = (int)e->objs.size();
  std::vector<float> vf(3 * n);
  for (int i = 0; i < n; ++i) {
    vf[3 * i] = (float)e->objs[i].vel.x;
    vf[3 * i + 1] = (float)e->objs[i].vel.y;
    vf[3 * i + 2] = (float)e->objs[i].w;
  }
  std::vector<std::vector<int>> members(ncol);
  for (int q = 0; q < (int)colour.size(); ++q)
    if (colour[q] >= 0) members[colour[q]].push_back(q);
  for (int pass = 0; pass <= e->prm.iterations; ++pass)
    for (int l = 0; l < ncol; ++l)
      for (int q : members[l])
        solve_manifold_f32(e, cs, e->pair_first[q], e->pair_first[q + 1], pass, vf);
  for (int i = 0; i < n; ++i) {
    Phys& o = e->objs[i];
    if (is_static(o) || !filled(e, i)) continue;
    o.vel = v2(o.vel.x + (double)(vf[3 * i] - (float)o.vel.x),
               o.vel.y + (double)(vf[3 * i + 1] - (float)o.vel.y));
    o.w = o.w + (double)(vf[3 * i + 2] - (float)o.w);
  }
}

void compute_stats(orc_engine* e) {
  orc_stats& s = e->stats;
  s.pairs = (int32_t)e->pairs.size();
  s.contacts = (int32_t)e->contacts.size();
  s.warm_hits = 0;
  s.max_depth = 0.0;
  s.normal_residual = 0.0;
  s.friction_residual = 0.0;
  int n = (int)e->objs.size();
  std::vector<int> last(n, 0);
  int levels = 0;
  for (auto& c : e->contacts) {
    if (c.warm) ++s.warm_hits;
    if (c.g.depth > s.max_depth) s.max_depth = c.g.depth;
    int la = is_static(e->objs[c.key.i]) ? 0 : last[c.key.i];
    int lb = is_static(e->objs[c.key.j]) ? 0 : last[c.key.j];
    int l = 1 + (la > lb ? la : lb);
    last[c.key.i] = l;
    last[c.key.j] = l;
    if (l > levels) levels = l;
  }
  s.dag_levels = levels;
}

}  // namespace

// residuals are measured after advance would move bodies, so they are
// computed before advance in step_once (see below).
static void residuals(orc_engine* e) {
  orc_stats& s = e->stats;
  for (auto& c : e->contacts) {
    const Phys& a = e->objs[c.key.i];
    const Phys& b = e->objs[c.key.j];
    const Material& ma = e->mats[c.key.i];
    const Material& mb = e->mats[c.key.j];
    double v[6], im[6];
    vel6(a, b, v);
    im6(a, b, im);
    double bn = c.cc.bn + bounce_b(ma.bounce, mb.bounce, c.cc, a, b);
    double cn = dot6(c.cc.jn, v) + bn;
    double rn = fabs(hs_min(c.ln, cn));
    if (rn > s.normal_residual) s.normal_residual = rn;
    double mt = eff_mass(c.cc.jt, im);
    double mu = (ma.mu + mb.mu) / 2.0;
    double thr = c.ln * mu;
    double t = c.lt - dot6(c.cc.jt, v) / mt;
    double tc = t > thr ? thr : (t < -thr ? -thr : t);
    double rt = fabs(c.lt - tc);
    if (rt > s.friction_residual) s.friction_residual = rt;
  }
}

// updateWorld (Engine/Main.hs:71-86)
static void step_once(orc_engine* e) {
  const orc_params& p = e->prm;
  int n = (int)e->objs.size();
  e->probed = false;
  // 1. broadphase on the shapes placed by the previous moveShapes
  culled_keys(e, e->pairs);
  // 2. applyExternal constantAccel (World.hs:156-158, External.hs:23-28)
  for (int i = 0; i < n; ++i) {
    Phys& o = e->objs[i];
    if (filled(e, i) && !is_static_lin(o)) o.vel = add(o.vel, v2(p.gx * p.dt, p.gy * p.dt));
  }
  // 3. prepareFrame
  std::vector<Contact> cs;
  prepare_frame(e, cs);
  // 4. applyCachedSlns against the previous frame's cache
  const bool coloured = p.solver == ORC_SOLVER_COLORED;
  apply_cached_slns(e, cs, e->contacts, !coloured);
  e->contacts.swap(cs);
  // 5. improveWorld sweeps in descending key order (or the coloured
  //    schedule, warm start included)
  if (coloured) {
    solve_coloured(e, e->contacts);
  } else {
    for (int it = 0; it < p.iterations; ++it)
      for (auto& c : e->contacts) improve_contact(e, c);
  }
  compute_stats(e);
  residuals(e);
  // 6. advance (World.hs:167-169, Constraint.hs:225-229), filled slots
  for (int i = 0; i < n; ++i) {
    if (!filled(e, i)) continue;
    Phys& o = e->objs[i];
    o.pos = add(v2(p.dt * o.vel.x, p.dt * o.vel.y), o.pos);
    o.th = (p.dt * o.w) + o.th;
  }
  // 7. moveShapes (World.hs:136-140), filled slots
  for (int i = 0; i < n; ++i)
    if (filled(e, i)) set_hull_transform(e->hulls[i], e->objs[i], p.trig);
}

extern "C" {

orc_engine* orc_engine_create(int32_t n, const double* px, const double* py,
                              const double* theta, const double* vx,
                              const double* vy, const double* omega,
                              const double* mass, const double* inertia,
                              const double* mu, const double* bounce,
                              const int32_t* hull_off, const double* hull_x,
                              const double* hull_y, const double* radius,
                              const orc_params* prm, int32_t capacity) {
  orc_engine* e = new orc_engine();
  e->prm = *prm;
  e->objs.resize(n);
  e->mats.resize(n);
  e->hulls.resize(n);
  memset(&e->stats, 0, sizeof(e->stats));
  for (int i = 0; i < n; ++i) {
    Phys& o = e->objs[i];
    // makePhysicalObj (Engine.hs:27-39); -0 normalised to +0 at load
    o.vel = v2(nz(vx[i]), nz(vy[i]));
    o.w = nz(omega[i]);
    o.pos = v2(nz(px[i]), nz(py[i]));
    o.th = nz(theta[i]);
    o.iml = invert_mass(mass[i]);
    o.imr = invert_mass(inertia[i]);
    e->mats[i] = Material{mu[i], bounce[i]};
    std::vector<V2> vs;
    for (int k = hull_off[i]; k < hull_off[i + 1]; ++k)
      vs.push_back(v2(hull_x[k], hull_y[k]));
    if (radius && radius[i] > 0.0) {  // makeCircle (Engine.hs:53-54)
      e->hulls[i] = Hull();
      e->hulls[i].n = 0;
      e->hulls[i].radius = radius[i];
    } else {
      e->hulls[i] = list_to_hull(vs);
    }
    // makeWorldObj places the shape with moveShape (World.hs:142-149)
    set_hull_transform(e->hulls[i], o, prm->trig);
  }
  // World.new capacity + fromList's appends (World.hs:61-75,118-124,
  // EmptiesVector.hs:27-42,70-85): slots [0, n) filled, [n, capacity) empty
  const int cap = capacity > n ? capacity : n;
  e->objs.resize(cap, Phys{});
  e->mats.resize(cap, Material{0.0, 0.0});
  e->hulls.resize(cap);
  e->ev.assign(cap, -1);
  for (int s = n; s < cap; ++s) {
    e->ev[s] = s + 1;
    e->hulls[s].n = 0;
    e->hulls[s].radius = 0.0;
  }
  e->ev_head = n;
  e->ev_filled = n;
  return e;
}

// World.append (World.hs:77-84) over EmptiesVector.append
// (EmptiesVector.hs:70-85): the head empty slot takes the object, placed by
// moveShape (makeWorldObj, World.hs:142-149); its successor becomes the
// head.  Returns the slot, or -1 when the world is full (the reference's
// out-of-bounds read of the vector).
int32_t orc_engine_append(orc_engine* e, const double* b10, double radius, int32_t nv,
                          const double* hx, const double* hy) {
  const int cap = (int)e->ev.size();
  const int slot = e->ev_head;
  if (slot >= cap || e->ev[slot] == -1) return -1;
  Phys o;
  o.vel = v2(nz(b10[3]), nz(b10[4]));
  o.w = nz(b10[5]);
  o.pos = v2(nz(b10[0]), nz(b10[1]));
  o.th = nz(b10[2]);
  o.iml = invert_mass(b10[6]);
  o.imr = invert_mass(b10[7]);
  e->objs[slot] = o;
  e->mats[slot] = Material{b10[8], b10[9]};
  if (radius > 0.0) {
    e->hulls[slot] = Hull();
    e->hulls[slot].n = 0;
    e->hulls[slot].radius = radius;
  } else {
    std::vector<V2> vs;
    for (int k = 0; k < nv; ++k) vs.push_back(v2(hx[k], hy[k]));
    e->hulls[slot] = list_to_hull(vs);
  }
  set_hull_transform(e->hulls[slot], o, e->prm.trig);
  e->ev_head = e->ev[slot];
  e->ev[slot] = -1;
  e->ev_filled += 1;
  return slot;
}

// World.delete = EmptiesVector.safeDelete (World.hs:86-87,
// EmptiesVector.hs:57-68): 0, or -1 when the slot is already empty.
int32_t orc_engine_delete(orc_engine* e, int32_t slot) {
  if (slot < 0 || slot >= (int)e->ev.size() || e->ev[slot] != -1) return -1;
  e->ev[slot] = e->ev_head;
  e->ev_head = slot;
  e->ev_filled -= 1;
  return 0;
}

int32_t orc_engine_filled(const orc_engine* e) { return e->ev_filled; }

// EmptiesVector.read (EmptiesVector.hs:87-92) of every slot
void orc_engine_read_slots(const orc_engine* e, uint8_t* out) {
  for (size_t s = 0; s < e->ev.size(); ++s) out[s] = e->ev[s] == -1 ? 1 : 0;
}

void orc_engine_destroy(orc_engine* e) { delete e; }

void orc_engine_step(orc_engine* e, int32_t steps) {
  for (int s = 0; s < steps; ++s) step_once(e);
}

void orc_engine_read_bodies(const orc_engine* e, double* out) {
  for (size_t i = 0; i < e->objs.size(); ++i) {
    const Phys& o = e->objs[i];
    out[6 * i + 0] = o.pos.x;
    out[6 * i + 1] = o.pos.y;
    out[6 * i + 2] = o.th;
    out[6 * i + 3] = o.vel.x;
    out[6 * i + 4] = o.vel.y;
    out[6 * i + 5] = o.w;
  }
}

int32_t orc_engine_num_vertices(const orc_engine* e) {
  int32_t s = 0;
  for (auto& h : e->hulls) s += h.n;
  return s;
}

void orc_engine_read_hulls(const orc_engine* e, double* out) {
  size_t q = 0;
  for (auto& h : e->hulls)
    for (int k = 0; k < h.n; ++k) {
      out[q++] = h.verts[k].x;
      out[q++] = h.verts[k].y;
      out[q++] = h.normals[k].x;
      out[q++] = h.normals[k].y;
    }
}

void orc_engine_read_aabbs(const orc_engine* e, double* out) {
  for (size_t i = 0; i < e->hulls.size(); ++i) {
    if (e->hulls[i].n == 0 && !(e->hulls[i].radius > 0.0)) {  // never filled
      for (int k = 0; k < 4; ++k) out[4 * i + k] = NAN;
      continue;
    }
    Aabb a = hull_aabb(e->hulls[i]);
    out[4 * i + 0] = a.minx;
    out[4 * i + 1] = a.maxx;
    out[4 * i + 2] = a.miny;
    out[4 * i + 3] = a.maxy;
  }
}

int32_t orc_engine_num_pairs(const orc_engine* e) {
  return (int32_t)e->pairs.size();
}

void orc_engine_read_pairs(const orc_engine* e, int32_t* ij) {
  for (size_t k = 0; k < e->pairs.size(); ++k) {
    ij[2 * k] = e->pairs[k].first;
    ij[2 * k + 1] = e->pairs[k].second;
  }
}

// worldContacts (shapes-demo IOWorld.hs:74-81): culledKeys and prepareFrame
// of the world as it stands, joined against the cache without applying it
// (the lagrangians a step would start from); the cache is left as it is.
void orc_engine_probe(orc_engine* e) {
  culled_keys(e, e->pairs);
  prepare_frame(e, e->probe);
  apply_cached_slns(e, e->probe, e->contacts, false);
  e->probed = true;
}

int32_t orc_engine_num_cache(const orc_engine* e) { return (int32_t)e->contacts.size(); }

void orc_engine_read_cache(const orc_engine* e, int32_t* keys, double* lam) {
  for (size_t k = 0; k < e->contacts.size(); ++k) {
    const Contact& c = e->contacts[k];
    keys[4 * k + 0] = c.key.i;
    keys[4 * k + 1] = c.key.j;
    keys[4 * k + 2] = c.key.fi;
    keys[4 * k + 3] = c.key.fj;
    lam[2 * k] = c.ln;
    lam[2 * k + 1] = c.lt;
  }
}

int32_t orc_engine_num_contacts(const orc_engine* e) {
  return (int32_t)(e->probed ? e->probe : e->contacts).size();
}

void orc_engine_read_contacts(const orc_engine* e, int32_t* keys,
                              int32_t* flip, double* geom, double* lam,
                              int32_t* warm) {
  const std::vector<Contact>& shown = e->probed ? e->probe : e->contacts;
  for (size_t k = 0; k < shown.size(); ++k) {
    const Contact& c = shown[k];
    if (keys) {
      keys[4 * k + 0] = c.key.i;
      keys[4 * k + 1] = c.key.j;
      keys[4 * k + 2] = c.key.fi;
      keys[4 * k + 3] = c.key.fj;
    }
    if (flip) flip[k] = c.g.flip ? 1 : 0;
    if (geom) {
      geom[5 * k + 0] = c.g.n.x;
      geom[5 * k + 1] = c.g.n.y;
      geom[5 * k + 2] = c.g.p.x;
      geom[5 * k + 3] = c.g.p.y;
      geom[5 * k + 4] = c.g.depth;
    }
    if (lam) {
      lam[2 * k] = c.ln;
      lam[2 * k + 1] = c.lt;
    }
    if (warm) warm[k] = c.warm ? 1 : 0;
  }
}

void orc_engine_stats(const orc_engine* e, orc_stats* out) { *out = e->stats; }

void orc_engine_set_cache(orc_engine* e, int32_t n, const int32_t* keys,
                          const double* lam) {
  e->contacts.clear();
  for (int k = 0; k < n; ++k) {
    Contact c;
    memset(&c, 0, sizeof(c));
    c.key = Key{keys[4 * k], keys[4 * k + 1], keys[4 * k + 2], keys[4 * k + 3]};
    c.ln = lam[2 * k];
    c.lt = lam[2 * k + 1];
    e->contacts.push_back(c);
  }
}

int32_t orc_hull_contacts(int32_t na, const double* ax, const double* ay,
                          int32_t nb, const double* bx, const double* by,
                          int32_t* out_keys, int32_t* out_flip,
                          double* out_geom, int32_t* out_sep) {
  std::vector<V2> va, vb;
  for (int k = 0; k < na; ++k) va.push_back(v2(ax[k], ay[k]));
  for (int k = 0; k < nb; ++k) vb.push_back(v2(bx[k], by[k]));
  Hull A = list_to_hull(va), B = list_to_hull(vb);
  ContactGeom g[2];
  bool sep;
  int n = hull_contacts(A, B, g, &sep);
  *out_sep = sep ? 1 : 0;
  for (int k = 0; k < n; ++k) {
    out_keys[2 * k] = g[k].fa;
    out_keys[2 * k + 1] = g[k].fb;
    out_flip[k] = g[k].flip ? 1 : 0;
    out_geom[5 * k + 0] = g[k].n.x;
    out_geom[5 * k + 1] = g[k].n.y;
    out_geom[5 * k + 2] = g[k].p.x;
    out_geom[5 * k + 3] = g[k].p.y;
    out_geom[5 * k + 4] = g[k].depth;
  }
  return n;
}

// A shape for the test entry points: n > 0 world-space hull vertices, or a
// circle (n == 0) of radius r at (cx, cy).
static Hull test_shape(int32_t n, const double* x, const double* y, double r, double cx, double cy) {
  if (n == 0) {
    Hull h;
    h.n = 0;
    h.radius = r;
    h.center = v2(cx, cy);
    return h;
  }
  std::vector<V2> v;
  for (int k = 0; k < n; ++k) v.push_back(v2(x[k], y[k]));
  return list_to_hull(v);
}

int32_t orc_shape_contacts(int32_t na, const double* ax, const double* ay, const double* ac,
                           int32_t nb, const double* bx, const double* by, const double* bc,
                           int32_t* out_keys, int32_t* out_flip, double* out_geom) {
  const Hull A = test_shape(na, ax, ay, ac[0], ac[1], ac[2]);
  const Hull B = test_shape(nb, bx, by, bc[0], bc[1], bc[2]);
  ContactGeom g[2];
  bool sep;
  const int n = shape_contacts(A, B, g, &sep);
  for (int k = 0; k < n; ++k) {
    out_keys[2 * k] = g[k].fa;
    out_keys[2 * k + 1] = g[k].fb;
    out_flip[k] = g[k].flip ? 1 : 0;
    out_geom[5 * k + 0] = g[k].n.x;
    out_geom[5 * k + 1] = g[k].n.y;
    out_geom[5 * k + 2] = g[k].p.x;
    out_geom[5 * k + 3] = g[k].p.y;
    out_geom[5 * k + 4] = g[k].depth;
  }
  return n;
}

int32_t orc_closest_simplex(int32_t n, const double* x, const double* y, double ox, double oy,
                            int32_t* idx) {
  const Hull h = test_shape(n, x, y, 0.0, 0.0, 0.0);
  const Simplex s = closest_simplex(h, v2(ox, oy));
  for (int k = 0; k < s.n; ++k) idx[k] = s.v[k];
  return s.n;
}

void orc_solve_constraint(const double* j, double b, double* vel,
                          const double* im4) {
  Phys a, c;
  memset(&a, 0, sizeof(a));
  memset(&c, 0, sizeof(c));
  a.vel = v2(vel[0], vel[1]); a.w = vel[2];
  c.vel = v2(vel[3], vel[4]); c.w = vel[5];
  a.iml = im4[0]; a.imr = im4[1];
  c.iml = im4[2]; c.imr = im4[3];
  double v[6], im[6];
  vel6(a, c, v);
  im6(a, c, im);
  double l = lagrangian2(j, b, v, eff_mass(j, im));
  apply_lagrangian(l, j, a, c);
  vel[0] = a.vel.x; vel[1] = a.vel.y; vel[2] = a.w;
  vel[3] = c.vel.x; vel[4] = c.vel.y; vel[5] = c.w;
}

int32_t orc_unordered_pairs(int32_t n, int32_t* out) {
  if (n < 2) return 0;
  int32_t q = 0;
  for (int x = n - 1; x >= 1; --x)
    for (int y = x - 1; y >= 0; --y) {
      out[2 * q] = x;
      out[2 * q + 1] = y;
      ++q;
    }
  return q;
}

void orc_sincos(double x, int32_t trig, double* s, double* c) {
  if (trig == ORC_TRIG_LIBM) {
    *s = sin(x);
    *c = cos(x);
  } else {
    shp_sincos(x, s, c);
  }
}

double orc_engines_run(orc_engine** engines, int32_t n_engines, int32_t steps,
                       int32_t threads) {
  auto t0 = std::chrono::steady_clock::now();
#ifdef _OPENMP
  if (threads > 1) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int w = 0; w < n_engines; ++w) orc_engine_step(engines[w], steps);
  } else
#endif
  {
    (void)threads;
    for (int w = 0; w < n_engines; ++w) orc_engine_step(engines[w], steps);
  }
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
