// shp_records.h - the coloured solver's manifold records: their layout,
// and how one is written from a pair's contacts (by the contact kernel,
// which has every value in registers) and patched by the learned warm start.
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "shp_engine.h"
#include "shp_geom.h"

namespace shp {

// Manifold M's record: 7 words of 16 bytes, word q at
// rec[(M / RB) * 7 * RB + q * RB + M % RB] (blocks of RB manifolds)
//   0  ux, uy, mu, e       the row normal (flip applied), materials
//   1  ax, ay, bx, by      the normal times the inverse masses of a / b
//   2  ga, gb, flags, 0    bodies (global ids), flags
//   3  jn2, jn5, jt2, jt5  the first contact: angular entries of the
//                          normal / friction rows
//   4  bn, icn, ict, wn2   bias, reciprocal effective masses, angular
//   5  wn5, wt2, wt5, 0      entries times the inverse inertias
//   6  ln, lt              its lagrangians (FP64)
// and the second contact's 4 words (3-6 above) at
// rec2[(M / RB) * 4 * RB + q * RB + M % RB].  Products and reciprocals are
// formed in FP64 and rounded once.
enum { REC_WORDS = 7, REC2_WORDS = 4 };
// manifolds per block, a power of two dividing 64: 4 against 64 and 8, a
// step of 1024 piles 21.23 against 21.36 ms and of 4096 small worlds 1.599
// against 1.613 (2: 21.68 / 1.647)
constexpr int RB = 4;
static_assert(16 * 64 * (REC_WORDS + REC2_WORDS) == 4 * COLOR_PACK_FLOATS(64),
              "record words differ from the buffer size");
// manifold flags: 1 a static, 2 b static, 4 first contact warm, 8 second
// contact present, 16 second contact warm
enum { MFL_A_STATIC = 1, MFL_B_STATIC = 2, MFL_WARM0 = 4, MFL_TWO = 8, MFL_WARM1 = 16 };

struct PackF {
  float4* rec;   // REC_WORDS x PC (PC rounded up to blocks)
  float4* rec2;  // REC2_WORDS x PC
};

__device__ __forceinline__ PackF packf(const Ctx& c) {
  PackF p;
  p.rec = (float4*)c.pk;
  p.rec2 = p.rec + (long)REC_WORDS * ((c.pair_cap + 63) / 64 * 64);
  return p;
}

// Word 0 of manifold M's record (words q at + RB q) and of its second
// contact's.
__device__ __forceinline__ float4* rec_of(float4* rec, int M) {
  return rec + (long)((unsigned)M / RB) * (REC_WORDS * RB) + ((unsigned)M % RB);
}
__device__ __forceinline__ float4* rec2_of(float4* rec2, int M) {
  return rec2 + (long)((unsigned)M / RB) * (REC2_WORDS * RB) + ((unsigned)M % RB);
}

// Word 3 of manifold M's record (its first contact) or of its second
// contact's (words at + RB q).
__device__ __forceinline__ float4* rec_contact(const PackF& P, int M, int j) {
  return j ? rec2_of(P.rec2, M) : rec_of(P.rec, M) + 3 * RB;
}

// Contact j (0 or 1) of manifold M as the contact generation computes it:
// the rows' angular entries, bias and effective masses, with the inverse
// inertias ira / irb of the two bodies ...
__device__ __forceinline__ void write_record_rows(const PackF& P, int M, int j, double ira,
                                                  double irb, double jn2, double jn5, double jt2,
                                                  double jt5, double bn, double mcn, double mct) {
  float4* w = rec_contact(P, M, j);
  w[0] = make_float4((float)jn2, (float)jn5, (float)jt2, (float)jt5);
  w[RB] = make_float4((float)bn, (float)(1.0 / mcn), (float)(1.0 / mct), (float)(ira * jn2));
  w[2 * RB] = make_float4((float)(irb * jn5), (float)(ira * jt2), (float)(irb * jt5), 0.0f);
}

// ... and its warm-start lagrangians.
__device__ __forceinline__ void write_record_lambda(const PackF& P, int M, int j, double ln,
                                                    double lt) {
  *(double2*)(rec_contact(P, M, j) + 3 * RB) = make_double2(ln, lt);
}

// The rest of manifold M's record: row normal u (flip applied), materials,
// the normal times the inverse masses ima / imb, bodies and flags for its
// nk (1 or 2) stored contacts, bit j of `warm` set when contact j is warm.
__device__ __forceinline__ void write_record_head(const Ctx& c, const PackF& P, int M, int ga,
                                                  int gb, d2 u, double ima, double imb, int nk,
                                                  int warm) {
  const int flags = (c.is_static[ga] ? MFL_A_STATIC : 0) | (c.is_static[gb] ? MFL_B_STATIC : 0) |
                    ((warm & 1) ? MFL_WARM0 : 0) | (nk > 1 ? MFL_TWO : 0) |
                    (nk > 1 && (warm & 2) ? MFL_WARM1 : 0);
  float4* r = rec_of(P.rec, M);
  r[0] = make_float4((float)u.x, (float)u.y, (float)((c.mu[ga] + c.mu[gb]) / 2.0),
                     (float)hs_min(c.e[ga], c.e[gb]));
  r[RB] = make_float4((float)(ima * u.x), (float)(ima * u.y), (float)(imb * u.x),
                      (float)(imb * u.y));
  *(int4*)(r + 2 * RB) = make_int4(ga, gb, flags, 0);
}

// A manifold whose contacts lie beyond the contact capacity (the step then
// reports the overflow): a record that touches no velocity and no contact.
__device__ __forceinline__ void write_null_record(const PackF& P, int M, int ga, int gb) {
  float4* r = rec_of(P.rec, M);
  const float4 z = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  r[0] = z;
  r[RB] = z;
  *(int4*)(r + 2 * RB) = make_int4(ga, gb, MFL_A_STATIC | MFL_B_STATIC, 0);
  r[3 * RB] = z;
  r[4 * RB] = z;
  r[5 * RB] = z;
  r[6 * RB] = z;
}

// Contact j of manifold M given a warm start after its record was written
// (the learned one): its lagrangians and warm flag.
__device__ __forceinline__ void patch_record_warm(const PackF& P, int M, int j, double ln, double lt) {
  write_record_lambda(P, M, j, ln, lt);
  atomicOr((int*)(rec_of(P.rec, M) + 2 * RB) + 2, j ? MFL_WARM1 : MFL_WARM0);
}

}  // namespace shp
