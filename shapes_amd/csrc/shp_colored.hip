// shp_colored.hip - graph-coloured Gauss-Seidel solver (FP32 math, FP64
// impulse accumulators) and the per-step solve-quality measurements.
//
// The reference solves contacts strictly one after another
// (improveWorld, shapes/src/Physics/Solvers/Contact.hs:146-157).  This mode
// keeps the same per-contact update (lagrangian2, positive, clampAbs,
// applySln) but reorders the sweep: a world's manifolds (pairs with
// contacts; one lane solves a pair's <= 2 points in key order) are coloured
// so that no two manifolds of a colour share a dynamic body, then each
// colour is solved in parallel.  Colours come from deterministic
// Jones-Plassmann rounds: in round r every uncoloured manifold whose hashed
// priority is the largest among the uncoloured manifolds on each of its
// dynamic bodies takes colour r.  Velocities and Jacobian rows are FP32;
// the accumulated lagrangians stay FP64; each body's FP64 velocity receives
// the FP32 change of the solve, so bodies without contacts keep their FP64
// velocity exactly.
#include <hip/hip_runtime.h>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_solve_common.h"

using namespace shp;

#define COLOR_CAP 512
#define COLOR_BODIES 10240  // bodies per world for the LDS colouring
#define COLOR_PAIRS 32767   // pairs per world (15-bit tie-break in priorities)
#define COLOR_THREADS 256

__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

__device__ __forceinline__ float dot6f(const float* a, const float* b) {
  return ((((a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]) + a[3] * b[3]) + a[4] * b[4]) + a[5] * b[5];
}

struct F32Ctx {
  float *vx, *vy, *w;
};

__device__ __forceinline__ void load_f(const Ctx& c, int ga, int gb, float* v, float* im,
                                       float* jn, float* jt, int k) {
  v[0] = c.fvx[ga];
  v[1] = c.fvy[ga];
  v[2] = c.fw[ga];
  v[3] = c.fvx[gb];
  v[4] = c.fvy[gb];
  v[5] = c.fw[gb];
  im[0] = im[1] = (float)c.iml[ga];
  im[2] = (float)c.imr[ga];
  im[3] = im[4] = (float)c.iml[gb];
  im[5] = (float)c.imr[gb];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    jn[q] = (float)c.jn[q * c.contact_cap + k];
    jt[q] = (float)c.jt[q * c.contact_cap + k];
  }
}

__device__ __forceinline__ void store_f(const Ctx& c, int ga, int gb, const float* v) {
  if (!body_static(c, ga)) {
    c.fvx[ga] = v[0];
    c.fvy[ga] = v[1];
    c.fw[ga] = v[2];
  }
  if (!body_static(c, gb)) {
    c.fvx[gb] = v[3];
    c.fvy[gb] = v[4];
    c.fw[gb] = v[5];
  }
}

__device__ __forceinline__ void apply_f(float l, const float* j, const float* im, float* v) {
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = v[q] + ((j[q] * l) * im[q]);
}

__device__ void warm_apply_f32(const Ctx& c, int k) {
  int ga = c.ga[k], gb = c.gb[k];
  float v[6], im[6], jn[6], jt[6];
  load_f(c, ga, gb, v, im, jn, jt, k);
  apply_f((float)c.ln[k], jn, im, v);
  apply_f((float)c.lt[k], jt, im, v);
  store_f(c, ga, gb, v);
}

__device__ void improve_f32(const Ctx& c, int k) {
  int ga = c.ga[k], gb = c.gb[k];
  float v[6], im[6], jn[6], jt[6];
  load_f(c, ga, gb, v, im, jn, jt, k);
  float rax = (float)c.rax[k], ray = (float)c.ray[k], rbx = (float)c.rbx[k], rby = (float)c.rby[k];
  float nx = (float)c.cnx[k], ny = (float)c.cny[k];
  if (c.flip[k]) {
    nx = -nx;
    ny = -ny;
  }
  // closing velocity of bounceB (Restitution.hs:34-47)
  float cx = ((-v[0] + v[2] * ray) + v[3]) - v[5] * rby;
  float cy = ((-v[1] - v[2] * rax) + v[4]) + v[5] * rbx;
  float ea = (float)c.e[ga], eb = (float)c.e[gb];
  float e = ea <= eb ? ea : eb;
  float bb = e * (cx * nx + cy * ny);
  bb = 0.0f <= bb ? 0.0f : bb;
  float bn = (float)c.bn[k] + bb;
  float new_n = -(dot6f(jn, v) + bn) / (float)c.mcn[k];
  float new_t = -dot6f(jt, v) / (float)c.mct[k];
  double ln = c.ln[k], lt = c.lt[k];
  double app_n = hs_max((double)new_n, -ln);
  double acc_n = ln + app_n;
  double mu = (c.mu[ga] + c.mu[gb]) / 2.0;
  double thr = acc_n * mu;
  double acc_t = lt + (double)new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - lt;
  apply_f((float)app_n, jn, im, v);
  apply_f((float)app_t, jt, im, v);
  store_f(c, ga, gb, v);
  c.ln[k] = acc_n;
  c.lt[k] = accp;
}

// One workgroup per world.
__global__ __launch_bounds__(COLOR_THREADS) void k_solve_colored(Ctx c) {
  __shared__ uint32_t s_best[COLOR_BODIES];
  __shared__ int s_off[COLOR_CAP + 3];  // colours 0..COLOR_CAP (serial tail) + end
  __shared__ int s_cur[COLOR_CAP + 3];
  __shared__ int s_left, s_taken;
  int w = blockIdx.x, tid = threadIdx.x;
  int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  int p0 = c.wpair_off[w], p1 = c.wpair_off[w + 1];
  int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  if (nb > COLOR_BODIES || p1 - p0 > COLOR_PAIRS) {
    if (tid == 0) solve_world_serial(c, c0, c1);  // exact FP64 fallback
    return;
  }
  int32_t* col = c.sched_level;  // colour per pair (indexed by pair)
  for (int g = b0 + tid; g < b0 + nb; g += COLOR_THREADS) {
    c.fvx[g] = (float)c.vx[g];
    c.fvy[g] = (float)c.vy[g];
    c.fw[g] = (float)c.w[g];
  }
  if (tid == 0) s_left = 0;
  __syncthreads();
  int mine = 0;
  for (int p = p0 + tid; p < p1; p += COLOR_THREADS) {
    bool has = c.ccount[p] > 0;
    col[p] = has ? -1 : -2;
    mine += has;
  }
  atomicAdd(&s_left, mine);
  __syncthreads();
  int rounds = 0;
  while (s_left > 0 && rounds < COLOR_CAP) {
    for (int i = tid; i < nb; i += COLOR_THREADS) s_best[i] = 0;
    if (tid == 0) s_taken = 0;
    __syncthreads();
    for (int p = p0 + tid; p < p1; p += COLOR_THREADS) {
      if (col[p] != -1) continue;
      uint32_t pl = (uint32_t)(p - p0);
      uint32_t prio = ((mix32(pl * 0x9E3779B9u + (uint32_t)w) & 0x1FFFFu) << 15) | pl;
      prio = prio + 1u;  // never 0
      int ga = c.pa[p], gb = c.pb[p];
      if (!c.is_static[ga]) atomicMax(&s_best[ga - b0], prio);
      if (!c.is_static[gb]) atomicMax(&s_best[gb - b0], prio);
    }
    __syncthreads();
    int taken = 0;
    for (int p = p0 + tid; p < p1; p += COLOR_THREADS) {
      if (col[p] != -1) continue;
      uint32_t pl = (uint32_t)(p - p0);
      uint32_t prio = (((mix32(pl * 0x9E3779B9u + (uint32_t)w) & 0x1FFFFu) << 15) | pl) + 1u;
      int ga = c.pa[p], gb = c.pb[p];
      bool ok = (c.is_static[ga] || s_best[ga - b0] == prio) &&
                (c.is_static[gb] || s_best[gb - b0] == prio);
      if (ok) {
        col[p] = rounds;
        ++taken;
      }
    }
    atomicAdd(&s_taken, taken);
    __syncthreads();
    if (tid == 0) s_left -= s_taken;
    ++rounds;
    __syncthreads();
  }
  int ncol = rounds + (s_left > 0 ? 1 : 0);  // leftovers: one serial colour
  if (tid == 0) atomicMax(&c.counters[CNT_LEVELS], ncol);
  for (int i = tid; i <= ncol + 1; i += COLOR_THREADS) s_cur[i] = 0;
  __syncthreads();
  for (int p = p0 + tid; p < p1; p += COLOR_THREADS) {
    int cl = col[p];
    if (cl == -1) col[p] = cl = rounds;
    if (cl >= 0) atomicAdd(&s_cur[cl], 1);
  }
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int l = 0; l <= ncol; ++l) {
      int h = s_cur[l];
      s_off[l] = acc;
      s_cur[l] = acc;
      acc += h;
    }
    s_off[ncol + 1] = acc;
  }
  __syncthreads();
  for (int p = p0 + tid; p < p1; p += COLOR_THREADS) {
    int cl = col[p];
    if (cl >= 0) c.sched_order[p0 + atomicAdd(&s_cur[cl], 1)] = p;
  }
  __syncthreads();
  bool serial_tail = s_left > 0;
  // warm start, then the sweeps, colour by colour
  for (int pass = 0; pass <= c.iterations; ++pass) {
    for (int l = 0; l < ncol; ++l) {
      int s = s_off[l], e = s_off[l + 1];
      bool tail = serial_tail && l == ncol - 1;
      int step = tail ? 1 : COLOR_THREADS;
      if (!tail || tid == 0) {
        for (int i = s + (tail ? 0 : tid); i < e; i += step) {
          int p = c.sched_order[p0 + i];
          int k0 = c.coff[p], k1 = k0 + c.ccount[p];
          if (k1 > c1) k1 = c1;  // contacts beyond the capacity were not stored
          for (int k = k0; k < k1; ++k) {
            if (pass == 0) {
              if (c.warm[k]) warm_apply_f32(c, k);
            } else {
              improve_f32(c, k);
            }
          }
        }
      }
      __syncthreads();
    }
  }
  for (int g = b0 + tid; g < b0 + nb; g += COLOR_THREADS) {
    if (c.is_static[g]) continue;
    c.vx[g] = c.vx[g] + (double)(c.fvx[g] - (float)c.vx[g]);
    c.vy[g] = c.vy[g] + (double)(c.fvy[g] - (float)c.vy[g]);
    c.w[g] = c.w[g] + (double)(c.fw[g] - (float)c.w[g]);
  }
}

void launch_colored_solver(shp_batch* b) {
  k_solve_colored<<<b->ctx.W, COLOR_THREADS, 0, b->stream>>>(b->ctx);
}

// ---- solve quality -----------------------------------------------------------

// Non-negative doubles order like their bit patterns, so atomicMax on the
// 64-bit pattern is a max of the values.
__device__ __forceinline__ void max_nonneg(unsigned long long* p, double x) {
  if (!(x >= 0.0)) x = x != x ? 1.0e300 : 0.0;  // NaN counts as huge
  atomicMax(p, (unsigned long long)__double_as_longlong(x));
}

// Per contact after the last sweep (FP64 state): penetration depth, normal
// complementarity residual |min(lambda_n, Jn.v + b)| and friction residual
// |lambda_t - clamp(lambda_t - Jt.v / mt, +-mu*lambda_n)|; maxima of the step
// and of the run.
__global__ void k_quality(Ctx c) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  int tc = c.wcont_off[c.W];
  if (k >= tc) return;
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
  double d = c.depth[k];
  max_nonneg(&c.quality[Q_DEPTH], d);
  max_nonneg(&c.quality[Q_NRES], rn);
  max_nonneg(&c.quality[Q_FRES], rt);
  max_nonneg(&c.quality[Q_RUN_DEPTH], d);
  max_nonneg(&c.quality[Q_RUN_NRES], rn);
  max_nonneg(&c.quality[Q_RUN_FRES], rt);
}

void launch_quality(shp_batch* b) {
  int n = b->ctx.contact_cap;
  k_quality<<<(n + 255) / 256, 256, 0, b->stream>>>(b->ctx);
}
