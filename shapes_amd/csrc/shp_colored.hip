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
// dynamic bodies takes colour r.  Velocities and constraint rows are FP32
// (velocities staged in LDS); the accumulated lagrangians stay FP64; each
// body's FP64 velocity receives the FP32 change of the solve, so bodies
// without contacts keep their FP64 velocity exactly.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_solve_common.h"

using namespace shp;

#define COLOR_CAP 256
#define COLOR_THREADS 256
#define NO_BODY 0xFFFFu

// packed FP32 record fields, stride contact_cap; lambdas are FP64 and live
// after the FP32 fields in the same buffer
enum {
  PF_JN = 0,
  PF_JT = 6,
  PF_BN = 12,
  PF_RAX = 13,
  PF_RAY = 14,
  PF_RBX = 15,
  PF_RBY = 16,
  PF_NRX = 17,
  PF_NRY = 18,
  PF_MCN = 19,
  PF_MCT = 20,
  PF_MU = 21,
  PF_E = 22,
  PF_IMLA = 23,
  PF_IMRA = 24,
  PF_IMLB = 25,
  PF_IMRB = 26,
  PF_FIELDS = 28  // even, so the FP64 lambdas that follow stay 8-byte aligned
};

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

struct RecF {
  float jn[6], jt[6];
  float bn, rax, ray, rbx, rby, nrx, nry, mcn, mct, mu, e;
  float im[6];
  double ln, lt;
  int ga, gb, flags;  // 1 a static, 2 b static, 4 warm, 8 second contact follows
};

struct PackF {
  float* f;        // PF_FIELDS x CC
  double* lam;     // 2 x CC (normal, friction)
  long CC;
};

__device__ __forceinline__ PackF packf(const Ctx& c) {
  PackF p;
  p.CC = c.contact_cap;
  p.f = (float*)c.pk;
  p.lam = (double*)(p.f + (long)PF_FIELDS * p.CC);
  return p;
}

__device__ __forceinline__ void load_recf(const Ctx& c, const PackF& P, int s, RecF& r) {
  long CC = P.CC;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    r.jn[q] = P.f[(PF_JN + q) * CC + s];
    r.jt[q] = P.f[(PF_JT + q) * CC + s];
  }
  r.bn = P.f[PF_BN * CC + s];
  r.rax = P.f[PF_RAX * CC + s];
  r.ray = P.f[PF_RAY * CC + s];
  r.rbx = P.f[PF_RBX * CC + s];
  r.rby = P.f[PF_RBY * CC + s];
  r.nrx = P.f[PF_NRX * CC + s];
  r.nry = P.f[PF_NRY * CC + s];
  r.mcn = P.f[PF_MCN * CC + s];
  r.mct = P.f[PF_MCT * CC + s];
  r.mu = P.f[PF_MU * CC + s];
  r.e = P.f[PF_E * CC + s];
  float ia = P.f[PF_IMLA * CC + s], ra = P.f[PF_IMRA * CC + s];
  float ib = P.f[PF_IMLB * CC + s], rb = P.f[PF_IMRB * CC + s];
  r.im[0] = ia;
  r.im[1] = ia;
  r.im[2] = ra;
  r.im[3] = ib;
  r.im[4] = ib;
  r.im[5] = rb;
  r.ln = P.lam[s];
  r.lt = P.lam[CC + s];
  r.ga = c.pk_g[s];
  r.gb = c.pk_g[CC + s];
  r.flags = c.pk_f[s];
}

__device__ __forceinline__ void pack_contact_f(const Ctx& c, const PackF& P, int k, int s, int extra) {
  long CC = P.CC;
  int ga = c.ga[k], gb = c.gb[k];
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    P.f[(PF_JN + q) * CC + s] = (float)c.jn[q * CC + k];
    P.f[(PF_JT + q) * CC + s] = (float)c.jt[q * CC + k];
  }
  P.f[PF_BN * CC + s] = (float)c.bn[k];
  P.f[PF_RAX * CC + s] = (float)c.rax[k];
  P.f[PF_RAY * CC + s] = (float)c.ray[k];
  P.f[PF_RBX * CC + s] = (float)c.rbx[k];
  P.f[PF_RBY * CC + s] = (float)c.rby[k];
  bool fl = c.flip[k] != 0;
  P.f[PF_NRX * CC + s] = (float)(fl ? -c.cnx[k] : c.cnx[k]);
  P.f[PF_NRY * CC + s] = (float)(fl ? -c.cny[k] : c.cny[k]);
  P.f[PF_MCN * CC + s] = (float)c.mcn[k];
  P.f[PF_MCT * CC + s] = (float)c.mct[k];
  P.f[PF_MU * CC + s] = (float)((c.mu[ga] + c.mu[gb]) / 2.0);
  P.f[PF_E * CC + s] = (float)hs_min(c.e[ga], c.e[gb]);
  P.f[PF_IMLA * CC + s] = (float)c.iml[ga];
  P.f[PF_IMRA * CC + s] = (float)c.imr[ga];
  P.f[PF_IMLB * CC + s] = (float)c.iml[gb];
  P.f[PF_IMRB * CC + s] = (float)c.imr[gb];
  P.lam[s] = c.ln[k];
  P.lam[CC + s] = c.lt[k];
  c.pk_g[s] = ga;
  c.pk_g[CC + s] = gb;
  c.pk_k[s] = k;
  c.pk_f[s] = (uint8_t)((c.is_static[ga] ? 1 : 0) | (c.is_static[gb] ? 2 : 0) | (c.warm[k] ? 4 : 0) |
                        (extra ? 8 : 0));
}

__device__ __forceinline__ void applyf(float l, const float* j, const float* im, float* v) {
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = v[q] + ((j[q] * l) * im[q]);
}

__device__ __forceinline__ void vload(const float* V, int la, int lb, float* v) {
  v[0] = V[3 * la];
  v[1] = V[3 * la + 1];
  v[2] = V[3 * la + 2];
  v[3] = V[3 * lb];
  v[4] = V[3 * lb + 1];
  v[5] = V[3 * lb + 2];
}

__device__ __forceinline__ void vstore(float* V, int la, int lb, int flags, const float* v) {
  if (!(flags & 1)) {
    V[3 * la] = v[0];
    V[3 * la + 1] = v[1];
    V[3 * la + 2] = v[2];
  }
  if (!(flags & 2)) {
    V[3 * lb] = v[3];
    V[3 * lb + 1] = v[4];
    V[3 * lb + 2] = v[5];
  }
}

// useCache's applySln in FP32
__device__ __forceinline__ void warm_f(const RecF& r, float* V, int b0) {
  if (!(r.flags & 4)) return;
  float v[6];
  int la = r.ga - b0, lb = r.gb - b0;
  vload(V, la, lb, v);
  applyf((float)r.ln, r.jn, r.im, v);
  applyf((float)r.lt, r.jt, r.im, v);
  vstore(V, la, lb, r.flags, v);
}

// improveContactSln with FP32 velocities/rows and FP64 accumulators
__device__ __forceinline__ void improve_f(const PackF& P, int s, const RecF& r, float* V, int b0) {
  float v[6];
  int la = r.ga - b0, lb = r.gb - b0;
  vload(V, la, lb, v);
  float cx = ((-v[0] + v[2] * r.ray) + v[3]) - v[5] * r.rby;  // bounceB closing velocity
  float cy = ((-v[1] - v[2] * r.rax) + v[4]) + v[5] * r.rbx;
  float bb = r.e * (cx * r.nrx + cy * r.nry);
  bb = 0.0f <= bb ? 0.0f : bb;
  float bn = r.bn + bb;
  float new_n = -(dot6f(r.jn, v) + bn) / r.mcn;
  float new_t = -dot6f(r.jt, v) / r.mct;
  double app_n = hs_max((double)new_n, -r.ln);
  double acc_n = r.ln + app_n;
  double thr = acc_n * (double)r.mu;
  double acc_t = r.lt + (double)new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - r.lt;
  applyf((float)app_n, r.jn, r.im, v);
  applyf((float)app_t, r.jt, r.im, v);
  vstore(V, la, lb, r.flags, v);
  P.lam[s] = acc_n;
  P.lam[P.CC + s] = accp;
}

// Solve one manifold (1 or 2 packed contacts starting at slot s).
template <bool WARM>
__device__ __forceinline__ void manifold(const Ctx& c, const PackF& P, int s, const RecF& r0,
                                         float* V, int b0) {
  if (WARM) warm_f(r0, V, b0);
  else improve_f(P, s, r0, V, b0);
  if (r0.flags & 8) {
    RecF r1;
    load_recf(c, P, s + 1, r1);
    if (WARM) warm_f(r1, V, b0);
    else improve_f(P, s + 1, r1, V, b0);
  }
}

// One workgroup per world.  LDS pool: during colouring s_best (4 B/body)
// and per-manifold bodies + colour (6 B/pair); during the solve FP32
// velocities (12 B/body).
__global__ __launch_bounds__(COLOR_THREADS) void k_solve_colored(Ctx c, int pool_bytes) {
  extern __shared__ int lds[];
  int* s_off = lds;                      // COLOR_CAP + 3
  int* s_cur = lds + (COLOR_CAP + 3);    // COLOR_CAP + 3
  unsigned char* pool = (unsigned char*)(lds + 2 * (COLOR_CAP + 3) + 2);
  __shared__ int s_left, s_taken;
  int w = blockIdx.x, tid = threadIdx.x;
  int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  int p0 = c.wpair_off[w], p1 = c.wpair_off[w + 1];
  int np = p1 - p0;
  int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  long need_col = 4L * nb + 6L * np, need_vel = 12L * nb;
  if (nb >= (int)NO_BODY || np > 32767 || need_col > pool_bytes || need_vel > pool_bytes)
    return;  // k_solve_colored_global takes this world
  tmark(c, 0);
  uint32_t* s_best = (uint32_t*)pool;
  uint32_t* m_ab = s_best + nb;                       // np: local a | local b << 16
  uint16_t* m_col = (uint16_t*)(m_ab + np);           // np: colour, 0xFFFF uncoloured
  if (tid == 0) s_left = 0;
  __syncthreads();
  int mine = 0;
  for (int q = tid; q < np; q += COLOR_THREADS) {
    int p = p0 + q;
    int ga = c.pa[p], gb = c.pb[p];
    uint32_t a = c.is_static[ga] ? NO_BODY : (uint32_t)(ga - b0);
    uint32_t b = c.is_static[gb] ? NO_BODY : (uint32_t)(gb - b0);
    m_ab[q] = a | (b << 16);
    bool has = c.ccount[p] > 0 && c.coff[p] < c1;  // contacts stored within capacity
    m_col[q] = has ? 0xFFFF : 0xFFFE;
    mine += has;
  }
  atomicAdd(&s_left, mine);
  __syncthreads();
  int rounds = 0;
  while (s_left > 0 && rounds < COLOR_CAP) {
    for (int i = tid; i < nb; i += COLOR_THREADS) s_best[i] = 0;
    if (tid == 0) s_taken = 0;
    __syncthreads();
    for (int q = tid; q < np; q += COLOR_THREADS) {
      if (m_col[q] != 0xFFFF) continue;
      uint32_t prio = (((mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) & 0x1FFFFu) << 15) | (uint32_t)q) + 1u;
      uint32_t ab = m_ab[q];
      uint32_t a = ab & 0xFFFF, b = ab >> 16;
      if (a != NO_BODY) atomicMax(&s_best[a], prio);
      if (b != NO_BODY) atomicMax(&s_best[b], prio);
    }
    __syncthreads();
    int taken = 0;
    for (int q = tid; q < np; q += COLOR_THREADS) {
      if (m_col[q] != 0xFFFF) continue;
      uint32_t prio = (((mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) & 0x1FFFFu) << 15) | (uint32_t)q) + 1u;
      uint32_t ab = m_ab[q];
      uint32_t a = ab & 0xFFFF, b = ab >> 16;
      if ((a == NO_BODY || s_best[a] == prio) && (b == NO_BODY || s_best[b] == prio)) {
        m_col[q] = (uint16_t)rounds;
        ++taken;
      }
    }
    atomicAdd(&s_taken, taken);
    __syncthreads();
    if (tid == 0) s_left -= s_taken;
    ++rounds;
    __syncthreads();
  }
  bool serial_tail = s_left > 0;
  int ncol = rounds + (serial_tail ? 1 : 0);
  tmark(c, 1);
  if (tid == 0) atomicMax(&c.counters[CNT_LEVELS], ncol);
  for (int i = tid; i <= ncol + 1; i += COLOR_THREADS) s_cur[i] = 0;
  __syncthreads();
  for (int q = tid; q < np; q += COLOR_THREADS) {
    int cl = m_col[q];
    if (cl == 0xFFFF) m_col[q] = (uint16_t)(cl = rounds);
    if (cl != 0xFFFE) atomicAdd(&s_cur[cl], 1);
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
  // manifold m of the colour order owns slots 2 (p0 + m) and + 1
  PackF P = packf(c);
  for (int q = tid; q < np; q += COLOR_THREADS) {
    int cl = m_col[q];
    if (cl == 0xFFFE) continue;
    int m = atomicAdd(&s_cur[cl], 1);
    int p = p0 + q;
    int k0 = c.coff[p], k1 = k0 + c.ccount[p];
    if (k1 > c1) k1 = c1;  // contacts beyond the capacity were not stored
    int s = 2 * (p0 + m);
    c.sched_order[p0 + m] = s;
    if (k0 < k1) pack_contact_f(c, P, k0, s, k1 - k0 > 1);
    if (k0 + 1 < k1) pack_contact_f(c, P, k0 + 1, s + 1, 0);
  }
  __syncthreads();  // colouring data is dead: the pool now holds velocities
  float* V = (float*)pool;
  for (int i = tid; i < nb; i += COLOR_THREADS) {
    V[3 * i] = (float)c.vx[b0 + i];
    V[3 * i + 1] = (float)c.vy[b0 + i];
    V[3 * i + 2] = (float)c.w[b0 + i];
  }
  __threadfence_block();
  __syncthreads();
  tmark(c, 2);
  for (int pass = 0; pass <= c.iterations; ++pass) {
    if (pass == 1) tmark(c, 3);
    for (int l = 0; l < ncol; ++l) {
      int e = s_off[l + 1];
      bool tail = serial_tail && l == ncol - 1;
      if (!tail || tid == 0) {
        int step = tail ? 1 : COLOR_THREADS;
        for (int i = s_off[l] + (tail ? 0 : tid); i < e; i += step) {
          int s = c.sched_order[p0 + i];
          RecF r;
          load_recf(c, P, s, r);
          if (pass == 0) manifold<true>(c, P, s, r, V, b0);
          else manifold<false>(c, P, s, r, V, b0);
        }
      }
      __syncthreads();
    }
  }
  tmark(c, 4);
  // lagrangians back to contact order; FP32 velocity change into FP64
  for (int i = tid; i < s_off[ncol]; i += COLOR_THREADS) {
    int s = c.sched_order[p0 + i];
    int k = c.pk_k[s];
    c.ln[k] = P.lam[s];
    c.lt[k] = P.lam[P.CC + s];
    if (c.pk_f[s] & 8) {
      int k1 = c.pk_k[s + 1];
      c.ln[k1] = P.lam[s + 1];
      c.lt[k1] = P.lam[P.CC + s + 1];
    }
  }
  for (int i = tid; i < nb; i += COLOR_THREADS) {
    int g = b0 + i;
    if (c.is_static[g]) continue;
    c.vx[g] = c.vx[g] + (double)(V[3 * i] - (float)c.vx[g]);
    c.vy[g] = c.vy[g] + (double)(V[3 * i + 1] - (float)c.vy[g]);
    c.w[g] = c.w[g] + (double)(V[3 * i + 2] - (float)c.w[g]);
  }
  __syncthreads();
  tmark(c, 5);
}


// FP32 velocities in the global arrays (worlds too large for LDS)
__device__ __forceinline__ void gload(const Ctx& c, int ga, int gb, float* v) {
  v[0] = c.fvx[ga];
  v[1] = c.fvy[ga];
  v[2] = c.fw[ga];
  v[3] = c.fvx[gb];
  v[4] = c.fvy[gb];
  v[5] = c.fw[gb];
}

__device__ __forceinline__ void gstore(const Ctx& c, int ga, int gb, int flags, const float* v) {
  if (!(flags & 1)) {
    c.fvx[ga] = v[0];
    c.fvy[ga] = v[1];
    c.fw[ga] = v[2];
  }
  if (!(flags & 2)) {
    c.fvx[gb] = v[3];
    c.fvy[gb] = v[4];
    c.fw[gb] = v[5];
  }
}

__device__ __forceinline__ void solve_one_g(const Ctx& c, const PackF& P, int s, const RecF& r, bool warm) {
  float v[6];
  gload(c, r.ga, r.gb, v);
  if (warm) {
    if (!(r.flags & 4)) return;
    applyf((float)r.ln, r.jn, r.im, v);
    applyf((float)r.lt, r.jt, r.im, v);
    gstore(c, r.ga, r.gb, r.flags, v);
    return;
  }
  float cx = ((-v[0] + v[2] * r.ray) + v[3]) - v[5] * r.rby;
  float cy = ((-v[1] - v[2] * r.rax) + v[4]) + v[5] * r.rbx;
  float bb = r.e * (cx * r.nrx + cy * r.nry);
  bb = 0.0f <= bb ? 0.0f : bb;
  float bn = r.bn + bb;
  float new_n = -(dot6f(r.jn, v) + bn) / r.mcn;
  float new_t = -dot6f(r.jt, v) / r.mct;
  double app_n = hs_max((double)new_n, -r.ln);
  double acc_n = r.ln + app_n;
  double thr = acc_n * (double)r.mu;
  double acc_t = r.lt + (double)new_t;
  double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
  double app_t = accp - r.lt;
  applyf((float)app_n, r.jn, r.im, v);
  applyf((float)app_t, r.jt, r.im, v);
  gstore(c, r.ga, r.gb, r.flags, v);
  P.lam[s] = acc_n;
  P.lam[P.CC + s] = accp;
}

template <bool WARM>
__device__ __forceinline__ void manifold_g(const Ctx& c, const PackF& P, int s, const RecF& r0) {
  solve_one_g(c, P, s, r0, WARM);
  if (r0.flags & 8) {
    RecF r1;
    load_recf(c, P, s + 1, r1);
    solve_one_g(c, P, s + 1, r1, WARM);
  }
}

// The same algorithm for worlds whose colouring or velocities do not fit in
// LDS: 64-bit priorities per body, colours per pair and FP32 velocities in
// global memory (L2-resident), 1024 lanes per world.  Worlds that fit are
// left to k_solve_colored (the fit test is the same).
#define COLOR_THREADS_G 1024

__global__ __launch_bounds__(COLOR_THREADS_G) void k_solve_colored_global(Ctx c, int pool_bytes) {
  __shared__ int s_off[COLOR_CAP + 3];
  __shared__ int s_cur[COLOR_CAP + 3];
  __shared__ int s_left, s_taken;
  int w = blockIdx.x, tid = threadIdx.x;
  int c0 = c.wcont_off[w], c1 = c.wcont_off[w + 1];
  if (c1 <= c0) return;
  int p0 = c.wpair_off[w], p1 = c.wpair_off[w + 1];
  int np = p1 - p0;
  int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  long need_col = 4L * nb + 6L * np, need_vel = 12L * nb;
  if (!(nb >= (int)NO_BODY || np > 32767 || need_col > pool_bytes || need_vel > pool_bytes)) return;
  unsigned long long* best = c.col_best + b0;
  int32_t* col = c.sched_level;  // per pair: colour, -1 uncoloured, -2 none
  if (tid == 0) s_left = 0;
  __syncthreads();
  int mine = 0;
  for (int q = tid; q < np; q += COLOR_THREADS_G) {
    int p = p0 + q;
    bool has = c.ccount[p] > 0 && c.coff[p] < c1;
    col[p] = has ? -1 : -2;
    mine += has;
  }
  atomicAdd(&s_left, mine);
  __syncthreads();
  int rounds = 0;
  while (s_left > 0 && rounds < COLOR_CAP) {
    for (int i = tid; i < nb; i += COLOR_THREADS_G) best[i] = 0ull;
    if (tid == 0) s_taken = 0;
    __syncthreads();
    for (int q = tid; q < np; q += COLOR_THREADS_G) {
      int p = p0 + q;
      if (col[p] != -1) continue;
      unsigned long long prio =
          ((unsigned long long)mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) << 32) | (uint32_t)(q + 1);
      int ga = c.pa[p], gb = c.pb[p];
      if (!c.is_static[ga]) atomicMax(&best[ga - b0], prio);
      if (!c.is_static[gb]) atomicMax(&best[gb - b0], prio);
    }
    __syncthreads();
    int taken = 0;
    for (int q = tid; q < np; q += COLOR_THREADS_G) {
      int p = p0 + q;
      if (col[p] != -1) continue;
      unsigned long long prio =
          ((unsigned long long)mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) << 32) | (uint32_t)(q + 1);
      int ga = c.pa[p], gb = c.pb[p];
      if ((c.is_static[ga] || best[ga - b0] == prio) && (c.is_static[gb] || best[gb - b0] == prio)) {
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
  bool serial_tail = s_left > 0;
  int ncol = rounds + (serial_tail ? 1 : 0);
  if (tid == 0) atomicMax(&c.counters[CNT_LEVELS], ncol);
  for (int i = tid; i <= ncol + 1; i += COLOR_THREADS_G) s_cur[i] = 0;
  __syncthreads();
  for (int q = tid; q < np; q += COLOR_THREADS_G) {
    int p = p0 + q;
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
  PackF P = packf(c);
  for (int q = tid; q < np; q += COLOR_THREADS_G) {
    int p = p0 + q;
    int cl = col[p];
    if (cl < 0) continue;
    int m = atomicAdd(&s_cur[cl], 1);
    int k0 = c.coff[p], k1 = k0 + c.ccount[p];
    if (k1 > c1) k1 = c1;
    int s = 2 * (p0 + m);
    c.sched_order[p0 + m] = s;
    if (k0 < k1) pack_contact_f(c, P, k0, s, k1 - k0 > 1);
    if (k0 + 1 < k1) pack_contact_f(c, P, k0 + 1, s + 1, 0);
  }
  for (int i = tid; i < nb; i += COLOR_THREADS_G) {
    c.fvx[b0 + i] = (float)c.vx[b0 + i];
    c.fvy[b0 + i] = (float)c.vy[b0 + i];
    c.fw[b0 + i] = (float)c.w[b0 + i];
  }
  __threadfence_block();
  __syncthreads();
  // velocities as a flat FP32 array of this world: V[3*i + k] is not the
  // global layout, so use per-component accessors through a local view
  for (int pass = 0; pass <= c.iterations; ++pass) {
    for (int l = 0; l < ncol; ++l) {
      int e = s_off[l + 1];
      bool tail = serial_tail && l == ncol - 1;
      if (!tail || tid == 0) {
        int step = tail ? 1 : COLOR_THREADS_G;
        for (int i = s_off[l] + (tail ? 0 : tid); i < e; i += step) {
          int s = c.sched_order[p0 + i];
          RecF r;
          load_recf(c, P, s, r);
          if (pass == 0) manifold_g<true>(c, P, s, r);
          else manifold_g<false>(c, P, s, r);
        }
      }
      __syncthreads();
    }
  }
  for (int i = tid; i < s_off[ncol]; i += COLOR_THREADS_G) {
    int s = c.sched_order[p0 + i];
    int k = c.pk_k[s];
    c.ln[k] = P.lam[s];
    c.lt[k] = P.lam[P.CC + s];
    if (c.pk_f[s] & 8) {
      int k1 = c.pk_k[s + 1];
      c.ln[k1] = P.lam[s + 1];
      c.lt[k1] = P.lam[P.CC + s + 1];
    }
  }
  for (int i = tid; i < nb; i += COLOR_THREADS_G) {
    int g = b0 + i;
    if (c.is_static[g]) continue;
    c.vx[g] = c.vx[g] + (double)(c.fvx[g] - (float)c.vx[g]);
    c.vy[g] = c.vy[g] + (double)(c.fvy[g] - (float)c.vy[g]);
    c.w[g] = c.w[g] + (double)(c.fw[g] - (float)c.w[g]);
  }
}

#define LDS_MAX_BYTES (160 * 1024 - 512)

shp_status colored_alloc(shp_batch* b) {
  Ctx& c = b->ctx;
  if (hipMalloc(&c.pk_k, (size_t)c.contact_cap * 4) != hipSuccess) return SHP_E_OOM;
  b->allocs.push_back(c.pk_k);
  long need = 0;
  for (int w = 0; w < c.W; ++w) {
    long nb = b->h_body_off[w + 1] - b->h_body_off[w];
    // pairs per world are unknown before a step: budget 8 per body (a world
    // that needs more falls back to the exact serial solve)
    long np = std::min<long>(32767, 8 * nb + 64);
    need = std::max(need, std::max(4 * nb + 6 * np, 12 * nb));
  }
  size_t fixed = (2 * (COLOR_CAP + 3) + 2) * 4;
  b->color_lds = std::min<size_t>(fixed + need, LDS_MAX_BYTES);
  if (hipMalloc(&c.col_best, (size_t)c.NB * 8) != hipSuccess) return SHP_E_OOM;
  b->allocs.push_back(c.col_best);
  b->color_global = false;
  for (int w = 0; w < c.W; ++w)
    if (12L * (b->h_body_off[w + 1] - b->h_body_off[w]) > (long)(b->color_lds - fixed)) b->color_global = true;
  if (hipFuncSetAttribute((const void*)k_solve_colored, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)b->color_lds) != hipSuccess)
    return SHP_E_HIP;
  return SHP_OK;
}

void launch_colored_solver(shp_batch* b) {
  int pool = (int)(b->color_lds - (2 * (COLOR_CAP + 3) + 2) * 4);
  k_solve_colored<<<b->ctx.W, COLOR_THREADS, b->color_lds, b->stream>>>(b->ctx, pool);
  // worlds that do not fit in LDS (by body count at creation, or by their
  // pair count this step) are solved by the global-memory variant; it exits
  // at once for the others
  k_solve_colored_global<<<b->ctx.W, COLOR_THREADS_G, 0, b->stream>>>(b->ctx, pool);
}

// ---- solve quality -----------------------------------------------------------

__device__ __forceinline__ unsigned long long nonneg_bits(double x) {
  if (!(x >= 0.0)) x = x != x ? 1.0e300 : 0.0;  // NaN counts as huge
  return (unsigned long long)__double_as_longlong(x);
}

__device__ __forceinline__ unsigned long long wave_max(unsigned long long x) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(x, off);
    x = o > x ? o : x;
  }
  return x;
}

// Per contact after the last sweep (FP64 state): penetration depth, normal
// complementarity residual |min(lambda_n, Jn.v + b)| and friction residual
// |lambda_t - clamp(lambda_t - Jt.v / mt, +-mu*lambda_n)|; maxima of the step
// and of the run.  Non-negative doubles order like their bit patterns, so the
// maxima are integer maxima, reduced per wavefront before one atomic each.
__global__ void k_quality(Ctx c) {
  int tc = c.wcont_off[c.W];
  unsigned long long qd = 0, qn = 0, qt = 0;
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < tc; k += gridDim.x * blockDim.x) {
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
  qd = wave_max(qd);
  qn = wave_max(qn);
  qt = wave_max(qt);
  if ((threadIdx.x & 63) == 0 && (qd | qn | qt)) {
    atomicMax(&c.quality[Q_DEPTH], qd);
    atomicMax(&c.quality[Q_NRES], qn);
    atomicMax(&c.quality[Q_FRES], qt);
    atomicMax(&c.quality[Q_RUN_DEPTH], qd);
    atomicMax(&c.quality[Q_RUN_NRES], qn);
    atomicMax(&c.quality[Q_RUN_FRES], qt);
  }
}

void launch_quality(shp_batch* b) {
  long n = b->ctx.contact_cap;
  int grid = (int)std::max<long>(1, std::min<long>((n + 255) / 256, 2048));
  k_quality<<<grid, 256, 0, b->stream>>>(b->ctx);
}
