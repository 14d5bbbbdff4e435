// shp_engine.h - device buffer layout of a batch of worlds (internal).
//
// Structure of arrays in HBM.  Bodies of all worlds are concatenated
// (world w owns [body_off[w], body_off[w+1])); hull vertices are CSR
// (body g owns [vert_off[g], vert_off[g+1])).  Pairs and contacts of a
// step are stored world after world, each world's list in the reference's
// descending order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/shapes_gpu.h"
#include "shp_mlp.h"

// LDS of one workgroup on gfx950 (160 KiB per CU); dynamic allocations
// stay 512 B below it for the kernels' own static LDS
#define LDS_BYTES_PER_CU (160 * 1024)
#define LDS_MAX_BYTES (LDS_BYTES_PER_CU - 512)
// coloured solver: colour offsets per world; ordered solver: level bins
#define COLOR_CAP 256
// one wavefront per world in the coloured sweeps (solver_threads 64) for
// worlds of up to this many bodies, SWEEP_WAVES wavefronts per workgroup
#define SWEEP_WAVE_MAX_BODIES 1024
#define SWEEP_WAVES 4
#define LEVEL_CAP 4095
// the coloured solver keeps a world's FP32 velocities (12 B per body) in
// LDS up to this many bodies; beyond, in global memory
#define VEL_LDS_MAX_BODIES ((LDS_MAX_BYTES - (2 * (COLOR_CAP + 3) + 2) * 4) / 12)
// the ordered solver's compact FP64 record (doubles per contact)
#define FLOW_RECORD_DOUBLES 21
// an LDS array's footprint: the compiler places each on a 16-byte boundary
constexpr size_t lds16(size_t bytes) { return (bytes + 15) & ~(size_t)15; }
// coloured solver records: 112 B per manifold and 64 B for its second
// contact, in blocks of 64 manifolds (PC manifolds)
#define COLOR_PACK_FLOATS(PC) (44L * (((PC) + 63) / 64 * 64))

// pair_ab words (coloured solver): world-local body ids a | b << 16, NO_BODY
// for a static body, NO_PAIR for a pair without contacts
#define NO_BODY 0xFFFFu
#define NO_PAIR 0xFFFFFFFFu

// phases of a step on the batch's stream (the colouring, on the second
// stream, is phase SHP_MAIN_PHASES of the timing ABI)
#define SHP_MAIN_PHASES (SHP_NUM_PHASES - 1)

enum { CNT_FLAGS = 0, CNT_WARM = 1, CNT_LEARNED = 2, CNT_LEVELS = 3, CNT_MISS = 4, CNT_MANIFOLDS = 5 };
// grid broadphase: partners per body kept by the count pass for the emit
// pass (bodies with more are visited again)
enum { GRID_SLOTS = 8 };
// worlds from which the coloured solver's records come from the contact
// kernel, after the colouring (with 1024-lane colouring workgroups): the
// CUs of an MI355X
enum { RECORDS_IN_CONTACTS_MIN_WORLDS = 256 };
// solve-quality maxima (doubles stored as their bit patterns): last step,
// then since creation; then the sums of the steps' maxima and the number
// of steps summed
enum {
  Q_DEPTH = 0,
  Q_NRES = 1,
  Q_FRES = 2,
  Q_RUN_DEPTH = 3,
  Q_RUN_NRES = 4,
  Q_RUN_FRES = 5,
  Q_SUM_DEPTH = 6,
  Q_SUM_NRES = 7,
  Q_SUM_FRES = 8,
  Q_STEPS = 9,
  Q_ENTRIES = 10
};
enum {
  SHP_FLAG_PAIR_OVERFLOW = 1,
  SHP_FLAG_CONTACT_OVERFLOW = 2,
  SHP_FLAG_NONFINITE = 4,
  SHP_FLAG_SCHEDULE_OVERFLOW = 8,
  SHP_FLAG_SOLVER_STALL = 16
};

struct Ctx {
  int W, NB, NV;
  int pair_cap, contact_cap;
  double dt, beta, slop, gx, gy;
  int iterations, solver, warm_mode;

  // bodies (NB)
  double *px, *py, *th, *vx, *vy, *w, *iml, *imr, *mu, *e;
  double* rad;  // circle radius, 0 for a hull (circles have no vertices)
  uint8_t* is_static;
  uint8_t* dead;  // 1: an empty slot (World.delete), skipped by every body pass
  int32_t* body_world;
  int32_t* body_off;  // W + 1
  int32_t* world_order;  // W: worlds by descending body count (per-world workgroups)
  // hulls (NV; vert_off NB + 1)
  int32_t* vert_off;
  double *lx, *ly;
  uint16_t* ext;  // per edge: extent vertex indices, min | max << 8
  double4* wv;    // per vertex: world x, y and the unit normal of its edge
  double2* wext;  // per edge: the hull's own extent (min, max) along its normal
  double *bminx, *bmaxx, *bminy, *bmaxy;
  // broadphase
  int32_t *cnt_rank, *off_rank;  // NB + 1, indexed by descending-body rank
  int32_t *pa, *pb;              // pair_cap, global body indices, pa > pb
  // narrowphase slots (2 per pair) and contact counts/offsets (pair_cap+1)
  int32_t *ccount, *coff;
  uint16_t* s_feat;
  uint8_t* s_flip;
  double *s_nx, *s_ny, *s_px, *s_py, *s_depth;
  int32_t *wpair_off, *wcont_off;  // W + 1
  // contacts (contact_cap)
  uint64_t* key;
  int32_t *ga, *gb;
  uint8_t* flip;
  double *cnx, *cny, *cpx, *cpy, *depth;
  double4* jx;      // per contact: the rows' cross terms (jn[2], jn[5], jt[2], jt[5])
  double *bn, *rax, *ray, *rbx, *rby, *mcn, *mct;
  double *ln, *lt;
  uint8_t* warm;
  // warm-start cache (last frame)
  uint64_t* cache_key;
  double *cache_ln, *cache_lt;
  int32_t* cache_woff;  // W + 1
  // coloured schedule: per pair its colour and its manifold slot
  int32_t* sched_level;   // pair_cap
  int32_t* sched_order;   // pair_cap
  // ordered (dataflow) schedule: every (contact, dynamic side) listed in
  // its body's segment, per contact {rank a, degree a, rank b, degree b},
  // per-body offsets/cursors/counters for worlds too large for LDS
  int32_t* flow_list;  // 2 x contact_cap
  int4* flow_rk;       // contact_cap
  int32_t *flow_off, *flow_cur;  // NB
  // level of every contact and the contacts in level order
  // (level-synchronous worlds), contact_cap each
  int32_t *flow_lvl, *flow_order;
  // contact constants packed for the solver (compact FP64 records in
  // contact order for the ordered solver, FP32 records + FP64 lambdas in
  // colour order for the coloured one)
  double* pk;
  int32_t* pk_g;  // bodies a, b: 2 x contact_cap (ordered solver)
  uint8_t* pk_f;  // flags per contact (ordered solver)
  ulonglong2* col_bu;  // per body: round priority, colours taken (large worlds)
  uint2* col_list;     // 2 x pair_cap: uncoloured manifolds, pair index and body word
  uint32_t* pair_ab;             // per pair: world-local bodies for the colouring (k_sat)
  int32_t* col_off;   // per world: COLOR_CAP + 3 colour offsets (manifolds)
  int32_t* col_info;  // per world: colours | serial tail << 16 (0: nothing to solve)
  // solver sub-phase timestamps of world 0 (wall_clock64), timing mode only
  unsigned long long* tstamp;
  unsigned long long* tstamp_store;
  // coloured solver state: FP32 velocities (vx, vy, w, unused) per body,
  // one 16-byte word so that a body's gather reads one cache line
  float4* fv;
  // learned warm start: padded FP32 master / BF16 weights, the step's
  // cache misses (contact index + 14 features) and the training-data log
  const float* mlp_w;
  const uint16_t* mlp_wbf;
  int mlp_ready;
  int collect_misses;
  int features_in_contacts;  // the misses' features from k_contacts (the log needs them)
  int miss_cap;
  int32_t* miss_idx;
  int32_t* miss_rec;  // per miss: its manifold record (slot * 2 + contact), or -1
  float* miss_feat;
  long log_cap;
  float *log_feat, *log_tgt;
  unsigned long long* log_count;
  // uniform-grid broadphase (use_grid): per-world cell size and hash-table
  // offsets, per-body cell, large static bodies listed per world
  int use_grid;
  int records_in_contacts;  // the coloured solver's records from k_contacts
  int grid_buckets;
  int probe;  // 1: contact generation of the current world only (no acceleration)
  double* cell_h;
  uint8_t* is_large;
  int32_t* body_cell;
  int32_t* tab_off;
  int32_t *bucket_cnt, *bucket_start;
  int32_t* body_slot;  // per binned body: its rank within its bucket
  // per binned body in bucket order, 48 B: AABB (min x, max x, min y,
  // max y), cell (x, y), body index, static flag
  double* grid_rec;
  // the count pass's first GRID_SLOTS partners of each body
  int32_t* grid_slots;
  int32_t *large_off, *large_ids;
  int32_t* counters;  // 8 ints: flags, warm hits, learned starts, levels
  unsigned long long* quality;  // Q_ENTRIES entries, see Q_*
};

// ---- launch plan ----------------------------------------------------------
// Every launch of a step takes its grid from the plan, which is computed on
// the host from the worlds and the configuration before anything is
// allocated; shp_batch_plan returns it (and its audit) without a GPU.

// An exclusive sum of n counts: one workgroup (k_scan_small<small_v>) when
// parts == 0, else k_scan_up over `parts` workgroups of SCAN_ITEMS counts,
// k_scan_mid over the parts, k_scan_down.
struct ScanPlan {
  long n = 0;
  int parts = 0;
  int small_v = 0;
};

struct StepPlan {
  long NB = 0, NV = 0, W = 0, max_nb = 0;
  long pair_cap = 0, contact_cap = 0, miss_cap = 0, grid_buckets = 0, n_large = 0;
  int ppb = 8;
  bool use_grid = false, grid_wide = false;  // grid broadphase; 16 lanes per body
  int solver = 0, warm = 0;
  // ordered solver: lanes per world, dynamic LDS, velocities / counters in LDS
  int solve_threads = 256;
  size_t solve_lds = 0;
  bool solve_vlds = false, solve_clds = false;
  // coloured solver: lanes per world, dynamic LDS of the colouring, LDS pool
  int color_threads = 256;
  size_t color_lds = 0, color_pool = 0;
  // the contact kernel writes the manifold records itself, after the
  // colouring, instead of beside it with a packing pass after both
  bool records_in_contacts = false;
  // with solver_threads 64: the worlds of at most SWEEP_WAVE_MAX_BODIES
  // bodies (world_order[sweep_big, W)) sweep one wavefront each, several
  // per workgroup, velocities in the wavefront's LDS slice (floats)
  long sweep_big = 0, g_sweep_waves = 0;
  int sweep_slice = 0;
  ScanPlan scan_buckets, scan_bodies, scan_pairs;
  long scan_parts_max = 0;  // scratch entries of the multi-workgroup scans
  // grids (workgroups) of the step's launches
  long g_reset = 0, g_bodies = 0, g_pairs = 0, g_pairs_count = 0, g_broadphase = 0, g_sat = 0,
       g_ranges = 0;
  long g_contacts = 0, g_infer = 0, g_pack = 0, g_unpack = 0, g_log = 0, g_finish_ni = 0,
       g_finish_nq = 0, g_cache = 0;
  // device memory allocated at creation (bytes)
  long device_bytes = 0;
};

// One kernel launch of a step, for the audit.
struct LaunchRec {
  const char* name;
  long grid;
  int block;
  size_t dyn_lds, static_lds;
};

// The host side of a batch before upload: validated worlds in SoA with -0
// normalised, inverse masses, hull CSR, the grid set-up, and the plan.
struct HostBatch {
  StepPlan P;
  std::vector<double> h[10];  // px py th vx vy w iml imr mu e
  std::vector<double> rad;    // circle radius, 0 for a hull
  std::vector<uint8_t> is_static, is_large, dead;
  std::vector<int32_t> body_world, vert_off, body_off, order, tab_off, large_off, large_ids;
  std::vector<double> lx, ly, cell;
};

shp_status prepare_batch(const shp_world_desc* worlds, int n_worlds, const shp_config& cfg,
                         HostBatch* hb);
shp_status plan_staged(const shp_config& cfg, HostBatch* hb);
void plan_launches(const StepPlan& P, bool infer, bool logging, std::vector<LaunchRec>* out);
size_t mlp_infer_static_lds();
// every launch within the hardware's limits and every 32-bit index in range
shp_status audit_plan(const StepPlan& P, const std::vector<LaunchRec>& L);

struct shp_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  shp_config cfg{};
  Ctx ctx{};
  StepPlan plan{};
  std::vector<int32_t> h_body_off, h_vert_off;
  std::vector<double> h_cell;       // grid cell size per world (grid broadphase)
  std::vector<uint8_t> h_is_large;  // bodies tested explicitly by the grid
  // per world, the slots' free list (Utils/EmptiesVector.hs:17-22): ev_next[s]
  // is -1 for a filled slot, else the empty slot after s (capacity ends the
  // list); ev_head is the first empty slot, ev_filled the filled count
  std::vector<std::vector<int32_t>> ev_next;
  std::vector<int32_t> ev_head, ev_filled;
  int32_t* scan_part = nullptr;  // per-workgroup sums of the multi-workgroup scans
  long device_bytes = 0;         // allocated so far
  size_t table_allocs = 0;       // allocs[0, table_allocs) are the plan's buffers
  long table_bytes = 0;          // ... of this many bytes
  // the coloured solver's colouring runs on side_stream, concurrently with
  // the contact generation on `stream` (fork and join through the events)
  hipStream_t side_stream = nullptr;
  hipEvent_t fork_event = nullptr, join_event = nullptr;
  hipGraphExec_t graph_exec{};
  bool graph_made = false;   // graph_exec holds an instantiated graph
  bool graph_ready = false;  // ... and it matches the current kernel arguments
  int32_t last_flags = 0;
  std::vector<void*> allocs;
  // per-phase event timing (shp_batch_set_timing)
  bool timing = false;
  bool events_made = false;
  hipEvent_t events[SHP_MAIN_PHASES + 1]{};  // main-stream phase boundaries
  hipEvent_t side_events[2]{};                // around the colouring
  MlpParams mlp{};
  bool logging = false;
  double phase_ms[SHP_NUM_PHASES]{};
  double solve_us[8]{};  // accumulated solver sub-phase durations (world 0)
  int32_t timed_steps = 0;
};

void shp_set_error(const std::string& msg);

shp_status batch_alloc(shp_batch* b, const HostBatch& hb);
// device buffers of `bytes` (at least 16) recorded for release at destroy
shp_status dalloc_bytes(shp_batch* b, void** p, size_t bytes);
shp_status solver_alloc(shp_batch* b);
shp_status colored_alloc(shp_batch* b);
// plan values of the solvers (lanes, LDS) from the largest world
void plan_ordered_solver(StepPlan* P, int solver_threads);
void plan_colored_solver(StepPlan* P, int solver_threads);
void launch_colored_pack(shp_batch* b);
// static LDS of the solver kernels at `threads` lanes
size_t ordered_static_lds(int threads);
size_t colored_static_lds(bool sweep);
void launch_ordered_solver(shp_batch* b);
// coloured solver: colouring (needs pairs and contact counts only), then,
// after the contacts, manifold packing and the sweeps
void launch_colouring(shp_batch* b, hipStream_t s);
void launch_colored_sweep(shp_batch* b);
void launch_colored_unpack(shp_batch* b);
hipError_t launch_grid_count(shp_batch* b);
hipError_t scan_grid_buckets(shp_batch* b);
hipError_t launch_grid_emit(shp_batch* b);
