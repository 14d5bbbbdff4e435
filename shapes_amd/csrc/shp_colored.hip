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
// Jones-Plassmann rounds: in each round every uncoloured manifold whose
// hashed priority is the largest among the uncoloured manifolds on each of
// its dynamic bodies is selected (the selected manifolds share no dynamic
// body) and takes the smallest colour that no manifold on its dynamic
// bodies has (a 16-bit mask per body; past 16 colours, colour 16 + round).  Velocities and constraint rows are FP32
// (velocities staged in LDS when they fit); the accumulated lagrangians stay
// FP64; each body's FP64 velocity receives the FP32 change of the solve, so
// bodies without contacts keep their FP64 velocity exactly.
//
// Records are per manifold: the normal, materials, inverse masses and
// bodies once, then per contact the two cross terms of each row, bias,
// lever arms, effective masses and the FP64 lagrangians; a second contact
// has a record of its own (a manifold with one contact reads nothing of
// it).  Records are interleaved word by word in blocks of four manifolds
// (shp_records.h): a wavefront reading 64 consecutive manifolds reads each
// word as 16 runs of 64 bytes, the rest of each 128-byte line being the
// next word, which its next load reads, with 16-byte loads from one
// address per lane; one manifold's record spans about four lines, which
// is what its writer stores to (blocks of 64: seven).  The rows are
// rebuilt from them exactly (jn = (-u, jn2, u, jn5), jt = (-cw(u), jt2,
// cw(u), jt5) with u the row normal; negation commutes with rounding to
// FP32).  A lane keeps its slot of a colour in every pass, so the record
// of its manifold of the next colour (lagrangians included) is loaded
// before the barrier that ends the current colour.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "shp_engine.h"
#include "shp_geom.h"
#include "shp_records.h"
#include "shp_solve_common.h"

using namespace shp;

// colours 0..15 are assigned smallest-free from the per-body masks; a
// manifold whose bodies already use all 16 takes GREEDY_COLOURS + round,
// so the Jones-Plassmann rounds stop at JP_ROUNDS to keep every colour and
// the serial tail below COLOR_CAP (and below the uint8 sentinel 0xFF)
#define GREEDY_COLOURS 16
#define JP_ROUNDS (COLOR_CAP - GREEDY_COLOURS - 3)


__device__ __forceinline__ uint32_t mix32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// a . b as a product and five fused multiply-adds, in index order
__device__ __forceinline__ float dot6f(const float* a, const float* b) {
  float d = a[0] * b[0];
#pragma unroll
  for (int q = 1; q < 6; ++q) d = fmaf(a[q], b[q], d);
  return d;
}

struct ConF {
  float jn2, jn5, jt2, jt5, bn, icn, ict, wn2, wn5, wt2, wt5;
  double ln, lt;
};

struct ManF {
  float ux, uy, mu, e, ax, ay, bx, by;
  int ga, gb, flags;
  ConF k[2];
};

// The records and the first manifold of one world.
struct RecW {
  float4* rec;
  float4* rec2;
  int p0;
};

__device__ __forceinline__ RecW recw(const Ctx& c, const PackF& P, int p0) {
  RecW r;
  r.rec = P.rec;
  r.rec2 = P.rec2;
  r.p0 = p0;
  return r;
}

// Words 3-6 of a record: one contact.
__device__ __forceinline__ void load_contact(const float4* w, ConF& k) {
  const float4 a = w[0], b = w[RB], d = w[2 * RB];
  const double2 l = *(const double2*)(w + 3 * RB);
  k.jn2 = a.x;
  k.jn5 = a.y;
  k.jt2 = a.z;
  k.jt5 = a.w;
  k.bn = b.x;
  k.icn = b.y;
  k.ict = b.z;
  k.wn2 = b.w;
  k.wn5 = d.x;
  k.wt2 = d.y;
  k.wt5 = d.z;
  k.ln = l.x;
  k.lt = l.y;
}

// Manifold m of the world (local index) and its contacts; the second
// contact's record loads only when the manifold has one.
__device__ __forceinline__ void load_man(const RecW& R, int m_loc, ManF& m) {
  const float4* r = rec_of(R.rec, R.p0 + m_loc);
  const float4 a = r[0], b = r[RB];
  const int4 g = *(const int4*)(r + 2 * RB);
  m.ux = a.x;
  m.uy = a.y;
  m.mu = a.z;
  m.e = a.w;
  m.ax = b.x;
  m.ay = b.y;
  m.bx = b.z;
  m.by = b.w;
  m.ga = g.x;
  m.gb = g.y;
  m.flags = g.z;
  load_contact(r + 3 * RB, m.k[0]);
  if (m.flags & MFL_TWO) load_contact(rec2_of(R.rec2, R.p0 + m_loc), m.k[1]);
}

// v += l * w, fused per entry (w: a row times the inverse masses)
__device__ __forceinline__ void applyw(float l, const float* w, float* v) {
#pragma unroll
  for (int q = 0; q < 6; ++q) v[q] = fmaf(l, w[q], v[q]);
}

// FP32 velocities: LDS (3 floats per local body) or fv (global memory).
template <bool VLDS>
struct VelF {
  float* lds;
  const Ctx* c;
  int b0;
  __device__ __forceinline__ void load(int g, float* v) const {
    if (VLDS) {
      const float* p = lds + 3 * (g - b0);
      v[0] = p[0];
      v[1] = p[1];
      v[2] = p[2];
    } else {
      const float4 x = c->fv[g];
      v[0] = x.x;
      v[1] = x.y;
      v[2] = x.z;
    }
  }
  __device__ __forceinline__ void store(int g, const float* v) const {
    if (VLDS) {
      float* p = lds + 3 * (g - b0);
      p[0] = v[0];
      p[1] = v[1];
      p[2] = v[2];
    } else {
      c->fv[g] = make_float4(v[0], v[1], v[2], 0.0f);
    }
  }
};

// One manifold: warm start (pass 0: useCache's applySln) or one sweep
// (improveContactSln with FP32 velocities/rows and FP64 accumulators) of
// its contacts in key order.  m_loc: the world-local manifold index for the
// lagrangian stores.  The FP32 arithmetic is this solver's own: the bounce
// term's closing velocity is the normal row times the velocities (the same
// quantity as bounceB's, Solvers/Contact.hs), the effective masses enter as
// reciprocals, and products with velocities are fused.
template <bool VLDS>
__device__ __forceinline__ void solve_man(const RecW& R, int m_loc, ManF& m, bool warm,
                                          const VelF<VLDS>& V) {
  float v[6];
  V.load(m.ga, v);
  V.load(m.gb, v + 3);
  const int nk = (m.flags & MFL_TWO) ? 2 : 1;
  bool changed = false;
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    if (j >= nk) break;
    ConF& k = m.k[j];
    // rows (the friction row's direction is clockwise(u)) and the rows
    // times the inverse masses
    const float jn[6] = {-m.ux, -m.uy, k.jn2, m.ux, m.uy, k.jn5};
    const float jt[6] = {-m.uy, m.ux, k.jt2, m.uy, -m.ux, k.jt5};
    const float wn[6] = {-m.ax, -m.ay, k.wn2, m.bx, m.by, k.wn5};
    const float wt[6] = {-m.ay, m.ax, k.wt2, m.by, -m.bx, k.wt5};
    if (warm) {
      if (!(m.flags & (j == 0 ? MFL_WARM0 : MFL_WARM1))) continue;
      applyw((float)k.ln, wn, v);
      applyw((float)k.lt, wt, v);
      changed = true;
      continue;
    }
    const float dn = dot6f(jn, v), dt = dot6f(jt, v);
    float bb = m.e * dn;  // bounce term
    bb = 0.0f <= bb ? 0.0f : bb;
    const float new_n = -(dn + (k.bn + bb)) * k.icn;
    const float new_t = -dt * k.ict;
    double app_n = hs_max((double)new_n, -k.ln);
    double acc_n = k.ln + app_n;
    double thr = acc_n * (double)m.mu;
    double acc_t = k.lt + (double)new_t;
    double accp = acc_t > thr ? thr : (acc_t < -thr ? -thr : acc_t);
    double app_t = accp - k.lt;
    applyw((float)app_n, wn, v);
    applyw((float)app_t, wt, v);
    k.ln = acc_n;
    k.lt = accp;
    changed = true;
  }
  if (!changed) return;
  if (!warm) {  // the solved contacts' lagrangians
    *(double2*)(rec_of(R.rec, R.p0 + m_loc) + 6 * RB) = make_double2(m.k[0].ln, m.k[0].lt);
    if (nk > 1)
      *(double2*)(rec2_of(R.rec2, R.p0 + m_loc) + 3 * RB) = make_double2(m.k[1].ln, m.k[1].lt);
  }
  if (!(m.flags & MFL_A_STATIC)) V.store(m.ga, v);
  if (!(m.flags & MFL_B_STATIC)) V.store(m.gb, v + 3);
}

// Colours per pair, in global memory: -1 uncoloured, -2 no contacts.
__device__ __forceinline__ int colour_of(int x) { return x; }

// Colour of a manifold selected in Jones-Plassmann round r whose dynamic
// bodies use the colours in `used`.
__device__ __forceinline__ int greedy_colour(uint32_t used, int r) {
  return used != 0xFFFFu ? __builtin_ctz(~used) : GREEDY_COLOURS + r;
}

__device__ __forceinline__ int wave_sum(int x) {
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
  return x;
}
__device__ __forceinline__ int wave_max(int x) {
  for (int o = 32; o > 0; o >>= 1) x = max(x, __shfl_xor(x, o));
  return x;
}

// atomicAdd(&ctr[key], 1) for every lane with key >= 0, one atomic per
// distinct key of the wavefront (lanes of equal key take consecutive
// values); returns the lane's old value, or -1.  Every lane of the
// wavefront calls it.
__device__ __forceinline__ int wave_count(int* ctr, int key) {
  const int lane = threadIdx.x & 63;
  unsigned long long pending = __ballot(key >= 0);
  int res = -1;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const int k = __shfl(key, leader);
    const unsigned long long grp = __ballot(key == k) & pending;
    int base = 0;
    if (lane == leader) base = atomicAdd(&ctr[k], (int)__popcll(grp));
    base = __shfl(base, leader);
    if ((grp >> lane) & 1ull) res = base + (int)__popcll(grp & ((1ull << lane) - 1ull));
    pending &= ~grp;
  }
  return res;
}

// Slots of the coloured manifolds: colour l owns manifold indices
// [p0 + s_off[l], p0 + s_off[l + 1]); scur holds the colour cursors;
// slot[q] (pair q of the world) gets its manifold index, or -1.  The order
// inside a colour is free, except in the serial tail colour (tail >= 0),
// whose manifolds may share bodies: thread 0 places those in pair order.
template <int T, class CT>
__device__ void assign_slots(int p0, int np, const CT* colour, int* scur, int tail,
                             int32_t* slot) {
  for (int q0 = 0; q0 < np; q0 += T) {
    const int q = q0 + threadIdx.x;
    const int cl = q < np ? colour_of(colour[q]) : -1;
    const bool in_tail = cl == tail && tail >= 0;
    const int m = wave_count(scur, in_tail ? -1 : cl);
    if (q < np && !in_tail) slot[q] = cl < 0 ? -1 : p0 + m;
  }
  if (tail >= 0 && threadIdx.x == 0) {
    int m = scur[tail];
    for (int q = 0; q < np; ++q)
      if (colour_of(colour[q]) == tail) slot[q] = p0 + m++;
  }
}

// Warm-start pass and sweeps, colour by colour, with prefetching.  Phase
// h = pass * ncol + l; lane t of STEP takes manifolds off[l] + t, + STEP,
// ... of colour l (the last colour serially by lane 0 when colouring
// stopped after JP_ROUNDS), and `sync` ends each phase.  Two records per
// lane: the lane's next manifold of a colour loads into the other record
// while the current one solves, and its first manifold of the next colour
// loads before the phase ends (not with one colour: that is the manifold
// just solved, whose lagrangians the pass changes).  The colour bounds of
// this phase (cs, ce) and the next (ns, ne) stay in registers; those of
// the phase after next are read early in the phase.
template <bool VLDS, int STEP, class Sync>
__device__ __forceinline__ void sweep_prefetched(const Ctx& c, const RecW& R, int ncol,
                                                 bool serial_tail, const int* off, int lane,
                                                 const VelF<VLDS>& V, Sync sync) {
  const int phases = (c.iterations + 1) * ncol;
  auto first_of = [&](int l, int b0_, int b1_) {
    const bool tail = serial_tail && l == ncol - 1;
    return tail ? (lane == 0 ? b0_ : b1_) : b0_ + lane;
  };
  int cs = off[0], ce = off[1];
  int ns = off[1 % ncol], ne = off[1 % ncol + 1];
  ManF m, n;
  bool loaded = false;
  for (int h = 0; h < phases; ++h) {
    const int l = h % ncol, l1 = (h + 1) % ncol, l2 = (h + 2) % ncol;
    const int s2 = off[l2], e2 = off[l2 + 1];
    const bool warm = h < ncol;
    const int di = serial_tail && l == ncol - 1 ? 1 : STEP;
    const int end = ce;
    int i = first_of(l, cs, ce);
    if (i < end && !loaded) load_man(R, i, m);
    while (i < end) {
      int inext = i + di;
      bool more = inext < end;
      if (more) load_man(R, inext, n);
      solve_man<VLDS>(R, i, m, warm, V);
      if (!more) break;
      i = inext;
      inext = i + di;
      more = inext < end;
      if (more) load_man(R, inext, m);
      solve_man<VLDS>(R, i, n, warm, V);
      if (!more) break;
      i = inext;
    }
    loaded = false;
    if (ncol > 1 && h + 1 < phases) {
      const int j = first_of(l1, ns, ne);
      if (j < ne) {
        load_man(R, j, m);
        loaded = true;
      }
    }
    sync();
    cs = ns;
    ce = ne;
    ns = s2;
    ne = e2;
  }
}

// Warm-start pass and sweeps of one world by one workgroup.
template <bool VLDS, int T>
__device__ void sweep_colours(const Ctx& c, const RecW& R, int ncol, bool serial_tail,
                              const int* s_off, const VelF<VLDS>& V) {
  sweep_prefetched<VLDS, T>(c, R, ncol, serial_tail, s_off, (int)threadIdx.x, V,
                            [] { __syncthreads(); });
}

// Colour-order bookkeeping shared by both kernels: per-colour counts of the
// coloured manifolds -> s_off (exclusive) and s_cur (cursors).
template <int T, class CT>
__device__ void colour_offsets(const CT* colour, int np, int ncol, int* s_off, int* s_cur) {
  const int tid = threadIdx.x;
  for (int i = tid; i <= ncol + 1; i += T) s_cur[i] = 0;
  __syncthreads();
  for (int q0 = 0; q0 < np; q0 += T) {
    const int q = q0 + tid;
    (void)wave_count(s_cur, q < np ? colour_of(colour[q]) : -1);
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
}

// Colouring in LDS: local body ids and pair indices fit 16 and 15 bits and
// the colouring (6 B per body, 7 B per manifold) fits the pool.
__device__ __forceinline__ bool colour_fits_lds(int nb, int np, int nm, int pool_bytes) {
  return nb < (int)NO_BODY && np <= 32767 && 4L * nb + 2L * ((nb + 1) & ~1) + 7L * nm <= pool_bytes;
}
// FP32 velocities in LDS: 12 B per body.
__device__ __forceinline__ bool vel_fits_lds(int nb, int pool_bytes) { return 12L * nb <= pool_bytes; }

// Pair indices below this fit prio32 (17 bits, prio32 - 1 below 2^32).
#define PRIO32_PAIRS 131071
// Colouring with the per-body priorities (4 B per body) and, when they fit
// too, colour masks (2 B) in LDS and the manifold lists in global memory.
__device__ __forceinline__ bool colour_fits_hybrid(int nb, int np, int pool_bytes) {
  return nb < (int)NO_BODY && np <= PRIO32_PAIRS && 4L * nb <= pool_bytes;
}
__device__ __forceinline__ bool colour_masks_fit(int nb, int pool_bytes) {
  return 4L * nb + 2L * ((nb + 1) & ~1) <= pool_bytes;
}

// Hashed priority of manifold q of world w: 15 hash bits, then the index
// (distinct within a world).  prio32 packs it for q < PRIO32_PAIRS (LDS
// colouring), prio64 for any q (global colouring); both order manifolds
// alike, so the colours do not depend on which path colours a world.
__device__ __forceinline__ uint32_t prio32(int q, int w) {
  return (((mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) & 0x7FFFu) << 17) | (uint32_t)q) + 1u;
}
__device__ __forceinline__ unsigned long long prio64(int q, int w) {
  return ((unsigned long long)(mix32((uint32_t)q * 0x9E3779B9u + (uint32_t)w) & 0x7FFFu) << 32) |
         (unsigned long long)((uint32_t)q + 1u);
}

// After colouring world w (colours 0..top-1, then the serial tail colour
// top if any): per-colour offsets into col_off[w], the colour count into
// col_info[w], and each pair's manifold slot into sched_order (pair index).
template <int T, class CT>
__device__ void colour_finish(const Ctx& c, int w, int p0, int np, const CT* colour, int top,
                              bool serial_tail, int* s_off, int* s_cur) {
  const int ncol = top + (serial_tail ? 1 : 0);
  if (threadIdx.x == 0) atomicMax(&c.counters[CNT_LEVELS], ncol);
  colour_offsets<T>(colour, np, ncol, s_off, s_cur);
  assign_slots<T>(p0, np, colour, s_cur, serial_tail ? ncol - 1 : -1, c.sched_order + p0);
  int* off = c.col_off + (long)w * (COLOR_CAP + 3);
  for (int l = threadIdx.x; l <= ncol; l += T) off[l] = s_off[l];
  if (threadIdx.x == 0) c.col_info[w] = ncol | (serial_tail ? 1 << 16 : 0);
}

// Number of manifolds (pairs with stored contacts) of a world, summed over
// the workgroup; every thread gets it.
template <int T>
__device__ int count_manifolds(const Ctx& c, int p0, int np, int* s_cnt) {
  if (threadIdx.x == 0) *s_cnt = 0;
  __syncthreads();
  int mine = 0;
  for (int q = threadIdx.x; q < np; q += T) mine += c.pair_ab[p0 + q] != NO_PAIR;
  mine = wave_sum(mine);
  if ((threadIdx.x & 63) == 0) atomicAdd(s_cnt, mine);
  __syncthreads();
  return *s_cnt;
}

// The rounds of colour_world_lds with each lane's manifolds (m = lane,
// lane + T, ...) held in registers with their bodies, priorities and
// colours: a round reads only the bodies' priorities and colour masks, a
// lane's coloured manifolds drop out of its later rounds, and the
// per-colour counts and manifold slots come from the registers.  The taken count is double-buffered, so a round needs three
// barriers.  Returns false (colours written to `colour` per pair instead)
// when the rounds stopped at JP_ROUNDS: the caller then places the serial
// tail in pair order.
#define COLOUR_REG_ITEMS 16
template <int T>
__device__ bool colour_regs(const Ctx& c, int w, int p0, int nm, int nb, uint32_t* s_best,
                            uint16_t* s_used, const uint32_t* m_ab, const uint16_t* m_q,
                            int32_t* colour, int* s_off, int* s_cur, int* sm) {
  int* s_tk = sm;  // [2]: taken this round, next round
  int& s_top = sm[2];
  const int tid = threadIdx.x;
  // pair index q = (prio32 - 1) & 0x1FFFF
  uint32_t pr[COLOUR_REG_ITEMS], ab[COLOUR_REG_ITEMS];
  uint32_t col[COLOUR_REG_ITEMS / 4];  // colour of manifold k in byte k % 4
  uint32_t alive = 0;
#pragma unroll
  for (int k = 0; k < COLOUR_REG_ITEMS; ++k) {
    const int m = tid + k * T;
    pr[k] = 0;
    ab[k] = 0;
    if (k % 4 == 0) col[k / 4] = 0;
    if (m < nm) {
      pr[k] = prio32(m_q[m], w);
      ab[k] = m_ab[m];
      alive |= 1u << k;
    }
  }
  if (tid == 0) {
    s_tk[0] = 0;
    s_tk[1] = 0;
  }
  int left = nm, rounds = 0;
  while (left > 0 && rounds < JP_ROUNDS) {
    for (int i = tid; i < nb; i += T) s_best[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < COLOUR_REG_ITEMS; ++k) {
      if (!((alive >> k) & 1u)) continue;
      const uint32_t a = ab[k] & 0xFFFF, b = ab[k] >> 16;
      if (a != NO_BODY) atomicMax(&s_best[a], pr[k]);
      if (b != NO_BODY) atomicMax(&s_best[b], pr[k]);
    }
    __syncthreads();
    // the selected manifolds share no dynamic body: each updates the masks
    // of its own bodies
    int taken = 0;
#pragma unroll
    for (int k = 0; k < COLOUR_REG_ITEMS; ++k) {
      if (!((alive >> k) & 1u)) continue;
      const uint32_t a = ab[k] & 0xFFFF, b = ab[k] >> 16;
      if ((a == NO_BODY || s_best[a] == pr[k]) && (b == NO_BODY || s_best[b] == pr[k])) {
        const uint32_t used = (a != NO_BODY ? s_used[a] : 0u) | (b != NO_BODY ? s_used[b] : 0u);
        const int l = greedy_colour(used, rounds);
        col[k / 4] |= (uint32_t)l << (8 * (k % 4));
        if (l < GREEDY_COLOURS) {
          if (a != NO_BODY) s_used[a] = (uint16_t)(s_used[a] | (1u << l));
          if (b != NO_BODY) s_used[b] = (uint16_t)(s_used[b] | (1u << l));
        }
        alive &= ~(1u << k);
        ++taken;
      }
    }
    taken = wave_sum(taken);
    if ((tid & 63) == 0) atomicAdd(&s_tk[rounds & 1], taken);
    // every lane read last round's count before this round's first barrier
    if (tid == 0) s_tk[(rounds + 1) & 1] = 0;
    __syncthreads();
    left -= s_tk[rounds & 1];
    ++rounds;
  }
  auto colour_at = [&](int k) { return (int)((col[k / 4] >> (8 * (k % 4))) & 0xFFu); };
  {  // the colour count: the highest colour of a coloured manifold, plus one
    int top = 0;
#pragma unroll
    for (int k = 0; k < COLOUR_REG_ITEMS; ++k)
      if (tid + k * T < nm && !((alive >> k) & 1u)) top = max(top, colour_at(k) + 1);
    top = wave_max(top);
    if ((tid & 63) == 0) atomicMax(&s_top, top);
    __syncthreads();
  }
  const int ncol = s_top;
  if (left > 0) {
    // serial tail: colours per pair, the tail (colour ncol) for the rest
#pragma unroll
    for (int k = 0; k < COLOUR_REG_ITEMS; ++k)
      if (tid + k * T < nm) colour[(pr[k] - 1u) & 0x1FFFFu] = (alive >> k) & 1u ? ncol : colour_at(k);
    __syncthreads();
    return false;
  }
  // per-colour counts -> offsets, then each manifold's slot in its colour
  // (LDS atomics; the order inside a colour is free, since its manifolds
  // share no dynamic body)
  for (int i = tid; i <= ncol + 1; i += T) s_cur[i] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < COLOUR_REG_ITEMS; ++k)
    if (tid + k * T < nm) atomicAdd(&s_cur[colour_at(k)], 1);
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int l = 0; l <= ncol; ++l) {
      const int h = s_cur[l];
      s_off[l] = acc;
      s_cur[l] = acc;
      acc += h;
    }
    s_off[ncol + 1] = acc;
    atomicMax(&c.counters[CNT_LEVELS], ncol);
    c.col_info[w] = ncol;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < COLOUR_REG_ITEMS; ++k)
    if (tid + k * T < nm) {
      const int m = atomicAdd(&s_cur[colour_at(k)], 1);
      c.sched_order[p0 + ((pr[k] - 1u) & 0x1FFFFu)] = p0 + m;
    }
  int* off = c.col_off + (long)w * (COLOR_CAP + 3);
  for (int l = tid; l <= ncol; l += T) off[l] = s_off[l];
  return true;
}

// Jones-Plassmann colouring of world w's manifolds in LDS.  The world's
// manifolds are first gathered into a list (any order: the colours depend
// only on the priorities).  LDS pool: s_best (4 B/body), s_used (2 B/body),
// per manifold its local bodies (4 B), pair index (2 B) and colour (1 B:
// 0xFF uncoloured), room for `cap` manifolds.  Colours go to sched_level per
// pair (-2 for pairs without contacts).
template <int T>
__device__ void colour_world_lds(const Ctx& c, int w, int p0, int np, int b0, int nb,
                                 int cap, unsigned char* pool, int* s_off, int* s_cur, int* sm) {
  int &s_left = sm[0], &s_taken = sm[1], &s_top = sm[2], &s_fill = sm[3];
  const int tid = threadIdx.x;
  uint32_t* s_best = (uint32_t*)pool;
  uint16_t* s_used = (uint16_t*)(s_best + nb);                  // nb: colours taken at the body
  uint32_t* m_ab = (uint32_t*)(s_used + ((nb + 1) & ~1));       // cap: local a | local b << 16
  uint16_t* m_q = (uint16_t*)(m_ab + cap);                      // cap: pair index
  uint8_t* m_col = (uint8_t*)(m_q + cap);                       // cap
  int32_t* colour = c.sched_level + p0;                         // per pair
  if (tid == 0) {
    s_top = 0;
    s_fill = 0;
  }
  for (int i = tid; i < nb; i += T) s_used[i] = 0;
  __syncthreads();
  // the manifold list: four pairs per lane in flight, one atomic per
  // wavefront per batch
  constexpr int LB = 4;
  const unsigned long long below = (1ull << (tid & 63)) - 1ull;
  for (int q0 = 0; q0 < np; q0 += LB * T) {
    uint32_t abv[LB];
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int q = q0 + u * T + tid;
      abv[u] = q < np ? c.pair_ab[p0 + q] : NO_PAIR;
    }
    unsigned long long bal[LB];
    int tot = 0;
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      bal[u] = __ballot(abv[u] != NO_PAIR);
      tot += (int)__popcll(bal[u]);
    }
    int base = 0;
    if ((tid & 63) == 0 && tot) base = atomicAdd(&s_fill, tot);
    base = __shfl(base, 0);
#pragma unroll
    for (int u = 0; u < LB; ++u) {
      const int q = q0 + u * T + tid;
      if (abv[u] != NO_PAIR) {
        const int m = base + (int)__popcll(bal[u] & below);
        m_ab[m] = abv[u];
        m_q[m] = (uint16_t)q;
        m_col[m] = 0xFFu;
      } else if (q < np) {
        colour[q] = -2;
        c.sched_order[p0 + q] = -1;
      }
      base += (int)__popcll(bal[u]);
    }
  }
  __syncthreads();
  const int nm = s_fill;
  if (tid == 0) s_left = nm;
  __syncthreads();
  if constexpr (T >= 1024) {
    if (nm <= COLOUR_REG_ITEMS * T) {
      if (!colour_regs<T>(c, w, p0, nm, nb, s_best, s_used, m_ab, m_q, colour, s_off, s_cur,
                          sm))
        colour_finish<T>(c, w, p0, np, colour, s_top, true, s_off, s_cur);
      return;
    }
  }
  int rounds = 0;
  while (s_left > 0 && rounds < JP_ROUNDS) {
    for (int i = tid; i < nb; i += T) s_best[i] = 0;
    if (tid == 0) s_taken = 0;
    __syncthreads();
    for (int m = tid; m < nm; m += T) {
      if (m_col[m] != 0xFFu) continue;
      const uint32_t prio = prio32(m_q[m], w);
      const uint32_t ab = m_ab[m], a = ab & 0xFFFF, b = ab >> 16;
      if (a != NO_BODY) atomicMax(&s_best[a], prio);
      if (b != NO_BODY) atomicMax(&s_best[b], prio);
    }
    __syncthreads();
    // the selected manifolds share no dynamic body: each updates the masks
    // of its own bodies
    int taken = 0, top = 0;
    for (int m = tid; m < nm; m += T) {
      if (m_col[m] != 0xFFu) continue;
      const uint32_t prio = prio32(m_q[m], w);
      const uint32_t ab = m_ab[m], a = ab & 0xFFFF, b = ab >> 16;
      if ((a == NO_BODY || s_best[a] == prio) && (b == NO_BODY || s_best[b] == prio)) {
        const uint32_t used = (a != NO_BODY ? s_used[a] : 0u) | (b != NO_BODY ? s_used[b] : 0u);
        const int l = greedy_colour(used, rounds);
        m_col[m] = (uint8_t)l;
        if (l < GREEDY_COLOURS) {
          if (a != NO_BODY) s_used[a] = (uint16_t)(s_used[a] | (1u << l));
          if (b != NO_BODY) s_used[b] = (uint16_t)(s_used[b] | (1u << l));
        }
        top = max(top, l + 1);
        ++taken;
      }
    }
    taken = wave_sum(taken);
    top = wave_max(top);
    if ((tid & 63) == 0) {
      atomicAdd(&s_taken, taken);
      atomicMax(&s_top, top);
    }
    __syncthreads();
    if (tid == 0) s_left -= s_taken;
    ++rounds;
    __syncthreads();
  }
  const bool serial_tail = s_left > 0;
  const int top = s_top;
  for (int m = tid; m < nm; m += T) colour[m_q[m]] = m_col[m] == 0xFFu ? top : (int)m_col[m];
  __syncthreads();
  colour_finish<T>(c, w, p0, np, colour, top, serial_tail, s_off, s_cur);
}

// The same colouring for worlds whose colouring does not fit in LDS:
// 64-bit priorities and colour masks per body (col_bu) and colours per
// pair (sched_level) in global memory (L2-resident).  The uncoloured
// manifolds of a round are kept as a list of (pair index, body word)
// entries (col_list, two buffers of the world's pair range), so a round
// reads no pair or body arrays beside it: a round visits only them, clears
// only their bodies' priorities, and passes the manifolds it does not
// colour on to the next round's list.  The selection of a round depends
// only on which manifolds are uncoloured and on their priorities, so the
// colours are those of a round over every pair.
template <int T>
__device__ void colour_world_global(const Ctx& c, int w, int p0, int np, int b0, int nb,
                                    int* s_off, int* s_cur, int* sm) {
  int &s_left = sm[0], &s_taken = sm[1], &s_top = sm[2], &s_fill = sm[3];
  const int tid = threadIdx.x;
  // per body: the round's largest priority (x) and the colours taken (y),
  // one 16-byte word, so a body's test reads one cache line
  ulonglong2* bu = c.col_bu + b0;
  int32_t* colour = c.sched_level + p0;  // per pair: colour, -1 uncoloured, -2 none
  // list entries: the pair index and its pair_ab word (world-local bodies)
  uint2* list[2] = {c.col_list + p0, c.col_list + c.pair_cap + p0};
  const unsigned long long below = (1ull << (tid & 63)) - 1ull;
  if (tid == 0) {
    s_top = 0;
    s_fill = 0;
  }
  for (int i = tid; i < nb; i += T) bu[i].y = 0ull;
  __syncthreads();
  // the first list: every pair with contacts
  for (int q0 = 0; q0 < np; q0 += T) {
    const int q = q0 + tid;
    const uint32_t ab = q < np ? c.pair_ab[p0 + q] : NO_PAIR;
    const bool has = ab != NO_PAIR;
    if (q < np) colour[q] = has ? -1 : -2;
    const unsigned long long bal = __ballot(has);
    int base = 0;
    if ((tid & 63) == 0 && bal) base = atomicAdd(&s_fill, (int)__popcll(bal));
    base = __shfl(base, 0);
    if (has) list[0][base + (int)__popcll(bal & below)] = make_uint2((uint32_t)q, ab);
  }
  __syncthreads();
  if (tid == 0) s_left = s_fill;
  __syncthreads();
  int rounds = 0, cur = 0;
  while (s_left > 0 && rounds < JP_ROUNDS) {
    const int n = s_left;
    const uint2* in = list[cur];
    uint2* out = list[cur ^ 1];
    for (int k = tid; k < n; k += T) {
      const uint32_t ab = in[k].y, a = ab & 0xFFFFu, b = ab >> 16;
      if (a != NO_BODY) bu[a].x = 0ull;
      if (b != NO_BODY) bu[b].x = 0ull;
    }
    if (tid == 0) {
      s_taken = 0;
      s_fill = 0;
    }
    __syncthreads();
    for (int k = tid; k < n; k += T) {
      const uint2 e = in[k];
      const unsigned long long prio = prio64((int)e.x, w);
      const uint32_t a = e.y & 0xFFFFu, b = e.y >> 16;
      if (a != NO_BODY) atomicMax(&bu[a].x, prio);
      if (b != NO_BODY) atomicMax(&bu[b].x, prio);
    }
    __syncthreads();
    int taken = 0, top = 0;
    for (int k0 = 0; k0 < n; k0 += T) {
      const int k = k0 + tid;
      bool keep = false;
      uint2 e = make_uint2(0u, 0u);
      if (k < n) {
        e = in[k];
        const unsigned long long prio = prio64((int)e.x, w);
        const uint32_t a = e.y & 0xFFFFu, b = e.y >> 16;
        const bool sa = a == NO_BODY, sb = b == NO_BODY;
        const ulonglong2 ra = sa ? make_ulonglong2(0ull, 0ull) : bu[a];
        const ulonglong2 rb = sb ? make_ulonglong2(0ull, 0ull) : bu[b];
        if ((sa || ra.x == prio) && (sb || rb.x == prio)) {
          const uint32_t used = (uint32_t)ra.y | (uint32_t)rb.y;
          const int l = greedy_colour(used, rounds);
          colour[e.x] = l;
          if (l < GREEDY_COLOURS) {
            if (!sa) bu[a].y = ra.y | (1ull << l);
            if (!sb) bu[b].y = rb.y | (1ull << l);
          }
          top = max(top, l + 1);
          ++taken;
        } else {
          keep = true;
        }
      }
      const unsigned long long bal = __ballot(keep);
      int base = 0;
      if ((tid & 63) == 0 && bal) base = atomicAdd(&s_fill, (int)__popcll(bal));
      base = __shfl(base, 0);
      if (keep) out[base + (int)__popcll(bal & below)] = e;
    }
    taken = wave_sum(taken);
    top = wave_max(top);
    if ((tid & 63) == 0) {
      atomicAdd(&s_taken, taken);
      atomicMax(&s_top, top);
    }
    __syncthreads();
    if (tid == 0) s_left = s_fill;
    cur ^= 1;
    ++rounds;
    __syncthreads();
  }
  const bool serial_tail = s_left > 0;
  const int top = s_top;
  // the manifolds the rounds did not reach form the serial tail colour
  for (int k = tid; k < s_left; k += T) colour[list[cur][k].x] = top;
  __syncthreads();
  colour_finish<T>(c, w, p0, np, colour, top, serial_tail, s_off, s_cur);
}

// The same colouring with the per-body priorities (prio32) in LDS, the
// colour masks in LDS (MASKS_LDS) or in col_bu, and the uncoloured
// manifolds' lists in global memory (as in colour_world_global): for
// worlds whose bodies fit LDS but whose manifolds do not.  A mask is read
// and written only by the manifold a round selects on the body.
template <int T, bool MASKS_LDS>
__device__ void colour_world_hybrid(const Ctx& c, int w, int p0, int np, int b0, int nb,
                                    unsigned char* pool, int* s_off, int* s_cur, int* sm) {
  int &s_left = sm[0], &s_taken = sm[1], &s_top = sm[2], &s_fill = sm[3];
  const int tid = threadIdx.x;
  uint32_t* s_best = (uint32_t*)pool;
  uint16_t* s_used = (uint16_t*)(s_best + nb);  // MASKS_LDS
  unsigned long long* g_used = &c.col_bu[b0].y;  // else: body i's mask at g_used[2 i]
  auto used_of = [&](uint32_t i) -> uint32_t {
    return MASKS_LDS ? (uint32_t)s_used[i] : (uint32_t)g_used[2 * i];
  };
  auto set_used = [&](uint32_t i, uint32_t m) {
    if (MASKS_LDS)
      s_used[i] = (uint16_t)m;
    else
      g_used[2 * i] = m;
  };
  int32_t* colour = c.sched_level + p0;  // per pair: colour, -1 uncoloured, -2 none
  uint2* list[2] = {c.col_list + p0, c.col_list + c.pair_cap + p0};
  const unsigned long long below = (1ull << (tid & 63)) - 1ull;
  if (tid == 0) {
    s_top = 0;
    s_fill = 0;
  }
  for (int i = tid; i < nb; i += T) set_used(i, 0u);
  __syncthreads();
  for (int q0 = 0; q0 < np; q0 += T) {
    const int q = q0 + tid;
    const uint32_t ab = q < np ? c.pair_ab[p0 + q] : NO_PAIR;
    const bool has = ab != NO_PAIR;
    if (q < np) colour[q] = has ? -1 : -2;
    const unsigned long long bal = __ballot(has);
    int base = 0;
    if ((tid & 63) == 0 && bal) base = atomicAdd(&s_fill, (int)__popcll(bal));
    base = __shfl(base, 0);
    if (has) list[0][base + (int)__popcll(bal & below)] = make_uint2((uint32_t)q, ab);
  }
  __syncthreads();
  if (tid == 0) s_left = s_fill;
  __syncthreads();
  int rounds = 0, cur = 0;
  while (s_left > 0 && rounds < JP_ROUNDS) {
    const int n = s_left;
    const uint2* in = list[cur];
    uint2* out = list[cur ^ 1];
    for (int k = tid; k < n; k += T) {
      const uint32_t ab = in[k].y, a = ab & 0xFFFFu, b = ab >> 16;
      if (a != NO_BODY) s_best[a] = 0u;
      if (b != NO_BODY) s_best[b] = 0u;
    }
    if (tid == 0) {
      s_taken = 0;
      s_fill = 0;
    }
    __syncthreads();
    for (int k = tid; k < n; k += T) {
      const uint2 e = in[k];
      const uint32_t prio = prio32((int)e.x, w);
      const uint32_t a = e.y & 0xFFFFu, b = e.y >> 16;
      if (a != NO_BODY) atomicMax(&s_best[a], prio);
      if (b != NO_BODY) atomicMax(&s_best[b], prio);
    }
    __syncthreads();
    int taken = 0, top = 0;
    for (int k0 = 0; k0 < n; k0 += T) {
      const int k = k0 + tid;
      bool keep = false;
      uint2 e = make_uint2(0u, 0u);
      if (k < n) {
        e = in[k];
        const uint32_t prio = prio32((int)e.x, w);
        const uint32_t a = e.y & 0xFFFFu, b = e.y >> 16;
        const bool sa = a == NO_BODY, sb = b == NO_BODY;
        if ((sa || s_best[a] == prio) && (sb || s_best[b] == prio)) {
          const uint32_t ua = sa ? 0u : used_of(a), ub = sb ? 0u : used_of(b);
          const int l = greedy_colour(ua | ub, rounds);
          colour[e.x] = l;
          if (l < GREEDY_COLOURS) {
            if (!sa) set_used(a, ua | (1u << l));
            if (!sb) set_used(b, ub | (1u << l));
          }
          top = max(top, l + 1);
          ++taken;
        } else {
          keep = true;
        }
      }
      const unsigned long long bal = __ballot(keep);
      int base = 0;
      if ((tid & 63) == 0 && bal) base = atomicAdd(&s_fill, (int)__popcll(bal));
      base = __shfl(base, 0);
      if (keep) out[base + (int)__popcll(bal & below)] = e;
    }
    taken = wave_sum(taken);
    top = wave_max(top);
    if ((tid & 63) == 0) {
      atomicAdd(&s_taken, taken);
      atomicMax(&s_top, top);
    }
    __syncthreads();
    if (tid == 0) s_left = s_fill;
    cur ^= 1;
    ++rounds;
    __syncthreads();
  }
  const bool serial_tail = s_left > 0;
  const int top = s_top;
  for (int k = tid; k < s_left; k += T) colour[list[cur][k].x] = top;
  __syncthreads();
  colour_finish<T>(c, w, p0, np, colour, top, serial_tail, s_off, s_cur);
}

// The LDS accesses of a wavefront before this point are done before any of
// its accesses after it (the wavefront's lanes run together).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Worlds small enough for one wavefront each (world_order[first_world, W),
// at most SWEEP_WAVE_MAX_BODIES bodies, FP32 velocities in the wavefront's
// slice of LDS): the warm pass and the sweeps of sweep_prefetched with 64
// lanes, colour by colour, with the wavefront's own ordering of its LDS
// accesses between colours instead of workgroup barriers; SWEEP_WAVES
// worlds per workgroup.  The same updates in the same colour order as
// sweep_world, so the same results.
__global__ __launch_bounds__(64 * SWEEP_WAVES) void k_sweep_waves(Ctx c, int first_world,
                                                                  int slice_floats) {
  extern __shared__ int lds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int k = first_world + (int)blockIdx.x * SWEEP_WAVES + wave;
  if (k >= c.W) return;  // this wavefront only: no workgroup barrier follows
  const int w = c.world_order[k];
  const int info = c.col_info[w];
  const int ncol = info & 0xFFFF;
  if (ncol == 0) return;
  const bool serial_tail = (info >> 16) & 1;
  const int p0 = c.wpair_off[w];
  const int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  const int* off = c.col_off + (long)w * (COLOR_CAP + 3);
  VelF<true> V;
  V.lds = (float*)lds + wave * slice_floats;
  V.c = &c;
  V.b0 = b0;
  for (int i = lane; i < nb; i += 64) {
    const int g = b0 + i;
    const float v[3] = {(float)c.vx[g], (float)c.vy[g], (float)c.w[g]};
    V.store(g, v);
  }
  wave_sync();
  const PackF P = packf(c);
  sweep_prefetched<true, 64>(c, recw(c, P, p0), ncol, serial_tail, off, lane, V,
                             [] { wave_sync(); });
  for (int i = lane; i < nb; i += 64) {
    float v[3];
    V.load(b0 + i, v);
    c.fv[b0 + i] = make_float4(v[0], v[1], v[2], 0.0f);
  }
}

// Colouring, one workgroup per world, the largest worlds first: in LDS when
// the world's colouring fits (by its pair count, else by its manifold
// count), else with the bodies in LDS and the manifold lists in global
// memory when the bodies fit, else in global memory.
template <int T>
__global__ __launch_bounds__(T) void k_colour(Ctx c, int pool_bytes) {
  extern __shared__ int lds[];
  int* s_off = lds;                      // COLOR_CAP + 3
  int* s_cur = lds + (COLOR_CAP + 3);    // COLOR_CAP + 3
  unsigned char* pool = (unsigned char*)(lds + 2 * (COLOR_CAP + 3) + 2);
  __shared__ int sm[4];
  const int w = c.world_order[blockIdx.x], tid = threadIdx.x;
  // the world's pairs (k_world_ranges computes the same offsets beside
  // this kernel)
  const int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  const int p0 = min(c.off_rank[b0], c.pair_cap), p1 = min(c.off_rank[b0 + nb], c.pair_cap);
  const int np = p1 - p0;
  if (np == 0) {  // nothing to solve
    if (tid == 0) c.col_info[w] = 0;
    return;
  }
  tmark(c, 0);
  // room for every pair as a manifold, or else for the manifolds counted
  int cap = np;
  bool in_lds = colour_fits_lds(nb, np, np, pool_bytes);
  if (!in_lds) {
    cap = count_manifolds<T>(c, p0, np, &sm[0]);
    in_lds = colour_fits_lds(nb, np, cap, pool_bytes);
  }
  if (in_lds)
    colour_world_lds<T>(c, w, p0, np, b0, nb, cap, pool, s_off, s_cur, sm);
  else if (colour_fits_hybrid(nb, np, pool_bytes) && colour_masks_fit(nb, pool_bytes))
    colour_world_hybrid<T, true>(c, w, p0, np, b0, nb, pool, s_off, s_cur, sm);
  else if (colour_fits_hybrid(nb, np, pool_bytes))
    colour_world_hybrid<T, false>(c, w, p0, np, b0, nb, pool, s_off, s_cur, sm);
  else
    colour_world_global<T>(c, w, p0, np, b0, nb, s_off, s_cur, sm);
  tmark(c, 1);
}

// Every coloured pair's <= 2 contacts become its manifold record, over the
// whole batch at once, when the contact kernel ran beside the colouring.
// Every value is loaded before the first store.
__global__ void k_pack_manifolds(Ctx c) {
  const int tp = c.wpair_off[c.W];
  const PackF P = packf(c);
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < tp; p += gridDim.x * blockDim.x) {
    const int M = c.sched_order[p];
    if (M < 0) continue;
    const int k0 = c.coff[p];
    if (k0 >= c.contact_cap) {
      write_null_record(P, M, c.pa[p], c.pb[p]);
      continue;
    }
    const int cap = c.contact_cap - k0, nk = c.ccount[p] < cap ? c.ccount[p] : cap;
    const int ga = c.ga[k0], gb = c.gb[k0];
    const d2 u = row_normal(c, k0);
    const double ima = c.iml[ga], ira = c.imr[ga], imb = c.iml[gb], irb = c.imr[gb];
    double4 x[2];
    double bn[2], mcn[2], mct[2], ln[2], lt[2];
    int warm = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (j >= nk) break;
      const int k = k0 + j;
      x[j] = c.jx[k];
      bn[j] = c.bn[k];
      mcn[j] = c.mcn[k];
      mct[j] = c.mct[k];
      ln[j] = c.ln[k];
      lt[j] = c.lt[k];
      warm |= (c.warm[k] ? 1 : 0) << j;
    }
    write_record_head(c, P, M, ga, gb, u, ima, imb, nk, warm);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (j >= nk) break;
      write_record_rows(P, M, j, ira, irb, x[j].x, x[j].y, x[j].z, x[j].w, bn[j], mcn[j], mct[j]);
      write_record_lambda(P, M, j, ln[j], lt[j]);
    }
  }
}

// Warm-start pass, sweeps and write-back of world w, with FP32 velocities
// in LDS (VLDS) or in global memory.
template <bool VLDS, int T>
__device__ void sweep_world(const Ctx& c, int w, int ncol, bool serial_tail, float* vlds,
                            int* s_off) {
  const int tid = threadIdx.x;
  const int p0 = c.wpair_off[w];
  const int b0 = c.body_off[w], nb = c.body_off[w + 1] - b0;
  const int* off = c.col_off + (long)w * (COLOR_CAP + 3);
  for (int l = tid; l <= ncol; l += T) s_off[l] = off[l];
  VelF<VLDS> V;
  V.lds = vlds;
  V.c = &c;
  V.b0 = b0;
  // FP32 velocities; a lane has the loads of four bodies in flight
  for (int i0 = tid; i0 < nb; i0 += 4 * T) {
    float x[4][3];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int g = b0 + i0 + u * T;
      if (i0 + u * T < nb) {
        x[u][0] = (float)c.vx[g];
        x[u][1] = (float)c.vy[g];
        x[u][2] = (float)c.w[g];
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (i0 + u * T < nb) V.store(b0 + i0 + u * T, x[u]);
  }
  __syncthreads();
  tmark(c, 2);
  tmark(c, 3);
  const PackF P = packf(c);
  sweep_colours<VLDS, T>(c, recw(c, P, p0), ncol, serial_tail, s_off, V);
  tmark(c, 4);
  // (every phase ends with a barrier) the final FP32 velocities to global
  // memory, for k_colour_unpack
  if (VLDS) {
    for (int i = tid; i < nb; i += T) {
      float v[3];
      V.load(b0 + i, v);
      c.fv[b0 + i] = make_float4(v[0], v[1], v[2], 0.0f);
    }
  }
  tmark(c, 5);
}

// One workgroup per world, the largest worlds first; velocities in LDS
// when they fit.
template <int T>
__global__ __launch_bounds__(T) void k_sweep_colours(Ctx c, int pool_bytes) {
  extern __shared__ int lds[];
  __shared__ int s_off[COLOR_CAP + 3];
  const int w = c.world_order[blockIdx.x];
  const int info = c.col_info[w];
  const int ncol = info & 0xFFFF;
  if (ncol == 0) return;
  const bool serial_tail = (info >> 16) & 1;
  const int nb = c.body_off[w + 1] - c.body_off[w];
  if (vel_fits_lds(nb, pool_bytes))
    sweep_world<true, T>(c, w, ncol, serial_tail, (float*)lds, s_off);
  else
    sweep_world<false, T>(c, w, ncol, serial_tail, nullptr, s_off);
}

// After the sweeps, over the whole batch: the lagrangians back to contact
// order, and each swept world's FP32 velocity change into the FP64
// velocities (a body whose world had nothing to solve keeps its own).
__global__ void k_colour_unpack(Ctx c) {
  const long tp = c.wpair_off[c.W];
  const long n = tp > c.NB ? tp : (long)c.NB;
  const PackF P = packf(c);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    if (t < tp) {
      const int M = c.sched_order[t];
      if (M >= 0) {
        const int k0 = c.coff[t];
        const int nk = min(c.ccount[t], c.contact_cap - k0);  // contacts that were stored
        for (int j = 0; j < nk; ++j) {
          const double2 l =
              *(const double2*)(j ? rec2_of(P.rec2, M) + 3 * RB : rec_of(P.rec, M) + 6 * RB);
          c.ln[k0 + j] = l.x;
          c.lt[k0 + j] = l.y;
        }
      }
    }
    if (t < c.NB) {
      const int g = (int)t;
      if (!c.is_static[g] && (c.col_info[c.body_world[g]] & 0xFFFF) != 0) {
        const float4 f = c.fv[g];
        c.vx[g] = c.vx[g] + (double)(f.x - (float)c.vx[g]);
        c.vy[g] = c.vy[g] + (double)(f.y - (float)c.vy[g]);
        c.w[g] = c.w[g] + (double)(f.z - (float)c.w[g]);
      }
    }
  }
}


// LDS of the colouring: colour offsets and cursors, then a pool for the
// per-body priorities and colour masks and the manifold list (pairs per
// world are unknown before a step: budgeted at 8 per body; a world that
// needs more is coloured in global memory); the sweeps use the pool for
// the FP32 velocities.  4 waves per world; a large world gets 8, and its
// colouring 16 (light on registers).
void plan_colored_solver(StepPlan* P, int solver_threads) {
  const long nb = P->max_nb;
  const long np = std::min<long>(32767, 8 * nb + 64);
  const long need = std::max(6 * nb + 2 + 7 * np, 12 * nb);
  const size_t fixed = (2 * (COLOR_CAP + 3) + 2) * 4;
  P->color_lds = std::min<size_t>(fixed + need, LDS_MAX_BYTES);
  P->color_pool = P->color_lds - fixed;
  // 64: worlds of up to SWEEP_WAVE_MAX_BODIES sweep one wavefront each,
  // the others (and the colouring) with 256 lanes
  P->color_threads = solver_threads == 64 ? 256
                     : solver_threads     ? solver_threads
                                          : (nb > 2048 ? 512 : 256);
  // A colouring workgroup of 1024 lanes takes a whole CU; with at least as
  // many worlds as CUs the colouring fills the GPU and gains little beside
  // the contact kernel (1024 piles: 6.60 ms for both plus the packing pass,
  // 6.04 ms in turn with the records from the contact kernel), while a few
  // worlds' colouring costs nothing there (one world: the contact kernel's
  // time) and smaller workgroups share CUs with it.
  P->records_in_contacts = P->color_threads >= 512 && P->W >= RECORDS_IN_CONTACTS_MIN_WORLDS;
}

// k_colour: sm[4]; k_sweep_colours: s_off[COLOR_CAP + 3]
size_t colored_static_lds(bool sweep) { return sweep ? lds16((COLOR_CAP + 3) * 4) : lds16(4 * 4); }

shp_status colored_alloc(shp_batch* b) {
  const StepPlan& P = b->plan;
  const int t = P.color_threads;
  const void* fns[] = {t >= 512 ? (const void*)k_colour<1024> : (const void*)k_colour<256>,
                       t == 1024  ? (const void*)k_sweep_colours<1024>
                       : t == 512 ? (const void*)k_sweep_colours<512>
                                  : (const void*)k_sweep_colours<256>};
  const size_t ldsz[] = {P.color_lds, P.color_pool};
  if (P.g_sweep_waves > 0 &&
      hipFuncSetAttribute((const void*)k_sweep_waves, hipFuncAttributeMaxDynamicSharedMemorySize,
                          (int)std::max<size_t>((size_t)SWEEP_WAVES * P.sweep_slice * 4, 16)) !=
          hipSuccess) {
    shp_set_error("hipFuncSetAttribute (wavefront sweeps LDS) failed");
    return SHP_E_HIP;
  }
  for (int k = 0; k < 2; ++k) {
    if (hipFuncSetAttribute(fns[k], hipFuncAttributeMaxDynamicSharedMemorySize, (int)ldsz[k]) !=
        hipSuccess) {
      shp_set_error("hipFuncSetAttribute (coloured solver LDS) failed");
      return SHP_E_HIP;
    }
    hipFuncAttributes a{};
    if (hipFuncGetAttributes(&a, fns[k]) != hipSuccess ||
        a.sharedSizeBytes > colored_static_lds(k == 1)) {
      shp_set_error("coloured solver: static LDS above the plan's");
      return SHP_E_HIP;
    }
  }
  return SHP_OK;
}

void launch_colored_pack(shp_batch* b) {
  k_pack_manifolds<<<b->plan.g_pack, 256, 0, b->stream>>>(b->ctx);
}

void launch_colouring(shp_batch* b, hipStream_t s) {
  const Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  const int pool = (int)P.color_pool;
  if (P.color_threads >= 512)
    k_colour<1024><<<c.W, 1024, P.color_lds, s>>>(c, pool);
  else
    k_colour<256><<<c.W, 256, P.color_lds, s>>>(c, pool);
}


void launch_colored_sweep(shp_batch* b) {
  const Ctx& c = b->ctx;
  const StepPlan& P = b->plan;
  const int pool = (int)P.color_pool;
  const unsigned big = (unsigned)P.sweep_big;
  if (big > 0) {
    if (P.color_threads == 1024)
      k_sweep_colours<1024><<<big, 1024, P.color_pool, b->stream>>>(c, pool);
    else if (P.color_threads == 512)
      k_sweep_colours<512><<<big, 512, P.color_pool, b->stream>>>(c, pool);
    else
      k_sweep_colours<256><<<big, 256, P.color_pool, b->stream>>>(c, pool);
  }
  if (P.g_sweep_waves > 0)
    k_sweep_waves<<<(unsigned)P.g_sweep_waves, 64 * SWEEP_WAVES,
                    (size_t)SWEEP_WAVES * P.sweep_slice * 4, b->stream>>>(c, (int)big,
                                                                         P.sweep_slice);
}

void launch_colored_unpack(shp_batch* b) {
  k_colour_unpack<<<b->plan.g_unpack, 256, 0, b->stream>>>(b->ctx);
}
