// shp_solve_common.h - per-contact FP64 solve steps shared by the ordered
// solver schedules (serial and level-parallel).
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


struct Vel6 {
  double v[6];
};

__device__ __forceinline__ void load_vel(const Ctx& c, int ga, int gb, double* v) {
  v[0] = c.vx[ga];
  v[1] = c.vy[ga];
  v[2] = c.w[ga];
  v[3] = c.vx[gb];
  v[4] = c.vy[gb];
  v[5] = c.w[gb];
}

// Static bodies (both inverse masses 0) receive v + (+-0) in the reference,
// which leaves a finite velocity unchanged (-0 is normalised at load), so
// their velocities are never written.
__device__ __forceinline__ void store_vel(const Ctx& c, int ga, int gb, const double* v) {
  if (!body_static(c, ga)) {
    c.vx[ga] = v[0];
    c.vy[ga] = v[1];
    c.w[ga] = v[2];
  }
  if (!body_static(c, gb)) {
    c.vx[gb] = v[3];
    c.vy[gb] = v[4];
    c.w[gb] = v[5];
  }
}

// applyLagrangian (Constraint.hs:182-221): v_k + ((j_k * l) * im_k)
__device__ __forceinline__ void apply_lag(double l, const double* j, const double* im, double* v) {
#pragma unroll
  for (int k = 0; k < 6; ++k) v[k] = v[k] + ((j[k] * l) * im[k]);
}

__device__ __forceinline__ void load_rows(const Ctx& c, int k, double* jn, double* jt) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    jn[q] = c.jn[q * c.contact_cap + k];
    jt[q] = c.jt[q * c.contact_cap + k];
  }
}

// useCache branch of applyCachedSlns (Solvers/Contact.hs:100-114):
// applySln = applyFriction . applyNonPen (:55-65)
__device__ __forceinline__ void warm_apply(const Ctx& c, int k) {
  int ga = c.ga[k], gb = c.gb[k];
  double v[6], jn[6], jt[6];
  double im[6] = {c.iml[ga], c.iml[ga], c.imr[ga], c.iml[gb], c.iml[gb], c.imr[gb]};
  load_vel(c, ga, gb, v);
  load_rows(c, k, jn, jt);
  apply_lag(c.ln[k], jn, im, v);
  apply_lag(c.lt[k], jt, im, v);
  store_vel(c, ga, gb, v);
}

// improveContactSln (Solvers/Contact.hs:124-143): contactLagrangian with
// bounceB (Constraints/Contact.hs:232-253, Restitution.hs:34-47), then
// solutionProcessor = positive / clampAbs (Constraints/Contact.hs:259-268,
// SolutionProcessors.hs:427-447), then applySln.
__device__ __forceinline__ void improve(const Ctx& c, int k) {
  int ga = c.ga[k], gb = c.gb[k];
  double v[6], jn[6], jt[6];
  double im[6] = {c.iml[ga], c.iml[ga], c.imr[ga], c.iml[gb], c.iml[gb], c.imr[gb]};
  load_vel(c, ga, gb, v);
  load_rows(c, k, jn, jt);
  d2 ra = mk(c.rax[k], c.ray[k]), rb = mk(c.rbx[k], c.rby[k]);
  d2 nr = c.flip[k] ? mk(-c.cnx[k], -c.cny[k]) : mk(c.cnx[k], c.cny[k]);
  d2 closing = add(add(add(neg(mk(v[0], v[1])), zcross(-v[2], ra)), mk(v[3], v[4])), zcross(v[5], rb));
  double e = hs_min(c.e[ga], c.e[gb]);
  double bb = hs_min(0.0, e * dot(closing, nr));
  double bn = c.bn[k] + bb;
  double new_n = (-(dot6(jn, v) + bn)) / c.mcn[k];
  double new_t = (-(dot6(jt, v) + 0.0)) / c.mct[k];
  double ln = c.ln[k], lt = c.lt[k];
  double app_n = hs_max(new_n, -ln);
  double acc_n = ln + app_n;
  double mu = (c.mu[ga] + c.mu[gb]) / 2.0;
  double thr = acc_n * mu;
  double acc_t = lt + new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - lt;
  apply_lag(app_n, jn, im, v);
  apply_lag(app_t, jt, im, v);
  store_vel(c, ga, gb, v);
  c.ln[k] = acc_n;
  c.lt[k] = accp;
}

// Reference order for one world by one lane: the warm-start pass then
// `iterations` improveWorld sweeps (Solvers/Contact.hs:146-157), each over
// the world's contacts in descending key order.
__device__ __forceinline__ void solve_world_serial(const Ctx& c, int c0, int c1) {
  for (int k = c0; k < c1; ++k)
    if (c.warm[k]) warm_apply(c, k);
  for (int it = 0; it < c.iterations; ++it)
    for (int k = c0; k < c1; ++k) improve(c, k);
}

}  // namespace shp
