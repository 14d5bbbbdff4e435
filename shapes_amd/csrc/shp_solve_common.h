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

__device__ __forceinline__ void load_rows(const Ctx& c, int k, double* jn, double* jt) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    jn[q] = c.jn[(size_t)q * c.contact_cap + k];
    jt[q] = c.jt[(size_t)q * c.contact_cap + k];
  }
}

}  // namespace shp
