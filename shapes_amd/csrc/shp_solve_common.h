// shp_solve_common.h - small FP64 solve helpers shared by the solvers.
// Static bodies (both inverse masses 0) receive v + (+-0) in the
// reference, which leaves a finite velocity unchanged (-0 is normalised at
// load), so the solvers never write their velocities.
#pragma once

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include "shp_engine.h"
#include "shp_geom.h"

namespace shp {

__device__ __forceinline__ bool body_static(const Ctx& c, int g) { return c.is_static[g] != 0; }

// Solver sub-phase timestamp k of world 0 (timing mode only).
__device__ __forceinline__ void tmark(const Ctx& c, int k) {
  if (c.tstamp && blockIdx.x == 0 && threadIdx.x == 0) c.tstamp[k] = wall_clock64();
}


__device__ __forceinline__ void load_vel(const Ctx& c, int ga, int gb, double* v) {
  v[0] = c.vx[ga];
  v[1] = c.vy[ga];
  v[2] = c.w[ga];
  v[3] = c.vx[gb];
  v[4] = c.vy[gb];
  v[5] = c.w[gb];
}

// applyLagrangian (Constraint.hs:182-221): v_k + ((j_k * l) * im_k)
__device__ __forceinline__ void apply_lag(double l, const double* j, const double* im, double* v) {
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = v[k] + ((j[k] * l) * im[k]);
}

// The row normal u of contact k: jn[3..4] as constraint generation computes
// them, n for a pair in order and -n for a flipped one (whose halves are
// swapped back).  Then jn = (-u, jn2, u, jn5) and jt = (-cw(u), jt2, cw(u),
// jt5) in both cases, and negation is exact, so the rows are rebuilt
// bit for bit from u and the stored cross terms.
__device__ __forceinline__ d2 row_normal(const Ctx& c, int k) {
  const double nx = c.cnx[k], ny = c.cny[k];
  return c.flip[k] ? mk(-nx, -ny) : mk(nx, ny);
}

__device__ __forceinline__ void load_rows(const Ctx& c, int k, double* jn, double* jt) {
  const d2 u = row_normal(c, k), t = clockwise(u);
  const double4 x = c.jx[k];
  jn[0] = -u.x;
  jn[1] = -u.y;
  jn[2] = x.x;
  jn[3] = u.x;
  jn[4] = u.y;
  jn[5] = x.y;
  jt[0] = -t.x;
  jt[1] = -t.y;
  jt[2] = x.z;
  jt[3] = t.x;
  jt[4] = t.y;
  jt[5] = x.w;
}

}  // namespace shp
