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

// pair_ab words (coloured solver): world-local body ids a | b << 16, NO_BODY
// for a static body, NO_PAIR for a pair without contacts
#define NO_BODY 0xFFFFu
#define NO_PAIR 0xFFFFFFFFu

enum { CNT_FLAGS = 0, CNT_WARM = 1, CNT_LEARNED = 2, CNT_LEVELS = 3, CNT_MISS = 4 };
// grid broadphase: partners per body kept by the count pass for the emit
// pass (bodies with more are visited again)
enum { GRID_SLOTS = 4 };
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
  uint8_t* is_static;
  int32_t* body_world;
  int32_t* body_off;  // W + 1
  int32_t* world_order;  // W: worlds by descending body count (per-world workgroups)
  // hulls (NV; vert_off NB + 1)
  int32_t* vert_off;
  double *lx, *ly;
  uint8_t *ext_min, *ext_max;
  double *wx, *wy, *wnx, *wny;
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
  double *jn, *jt;  // 6 * contact_cap, row-major by component
  double *bn, *rax, *ray, *rbx, *rby, *mcn, *mct;
  double *ln, *lt;
  uint8_t* warm;
  // warm-start cache (last frame)
  uint64_t* cache_key;
  double *cache_ln, *cache_lt;
  int32_t* cache_woff;  // W + 1
  // coloured schedule: per-pair colour, pairs sorted by colour in world
  int32_t* sched_level;   // contact_cap
  int32_t* sched_order;   // contact_cap
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
  int32_t* pk_g;  // 2 x contact_cap: bodies a, b
  uint8_t* pk_f;  // flags: 1 a static, 2 b static, 4 warm, 8 pair continues
  unsigned long long* col_best;  // colouring priorities per body (large worlds)
  uint32_t* col_used;            // colours taken at each body (large worlds)
  uint32_t* pair_ab;             // per pair: world-local bodies for the colouring (k_sat)
  int32_t* col_off;   // per world: COLOR_CAP + 3 colour offsets (manifolds)
  int32_t* col_info;  // per world: colours | serial tail << 16 (0: nothing to solve)
  // solver sub-phase timestamps of world 0 (wall_clock64), timing mode only
  unsigned long long* tstamp;
  unsigned long long* tstamp_store;
  // coloured solver state (FP32 velocities, FP64 accumulators)
  float *fvx, *fvy, *fw;
  // learned warm start: padded FP32 master / BF16 weights, the step's
  // cache misses (contact index + 14 features) and the training-data log
  const float* mlp_w;
  const uint16_t* mlp_wbf;
  int mlp_ready;
  int collect_misses;
  int miss_cap;
  int32_t* miss_idx;
  float* miss_feat;
  long log_cap;
  float *log_feat, *log_tgt;
  unsigned long long* log_count;
  // uniform-grid broadphase (use_grid): per-world cell size and hash-table
  // offsets, per-body cell, large static bodies listed per world
  int use_grid;
  int grid_buckets;
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

struct shp_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  shp_config cfg{};
  Ctx ctx{};
  std::vector<int32_t> h_body_off, h_vert_off;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
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
  hipEvent_t events[SHP_NUM_PHASES + 1]{};
  MlpParams mlp{};
  bool logging = false;
  int ppb = 16;             // pair capacity per body
  bool solve_vlds = false;  // ordered solver keeps velocities in LDS
  bool solve_clds = false;  // ordered solver keeps body counters in LDS
  int solve_threads = 256;  // ordered solver workgroup size
  size_t solve_lds = 0;     // dynamic LDS bytes of the ordered solver
  size_t color_lds = 0;     // dynamic LDS bytes of the coloured solver
  int color_threads = 256;    // coloured solver workgroup size
  double phase_ms[SHP_NUM_PHASES]{};
  double solve_us[8]{};  // accumulated solver sub-phase durations (world 0)
  int32_t timed_steps = 0;
};

void shp_set_error(const std::string& msg);

shp_status batch_alloc(shp_batch* b, const shp_world_desc* worlds, int n_worlds, long NB, long NV,
                       long pair_cap);
shp_status solver_alloc(shp_batch* b);
shp_status colored_alloc(shp_batch* b);
void launch_ordered_solver(shp_batch* b);
// coloured solver: colouring (needs pairs and contact counts only), then,
// after the contacts, manifold packing and the sweeps
void launch_colouring(shp_batch* b, hipStream_t s);
void launch_colored_sweeps(shp_batch* b);
hipError_t launch_grid_count(shp_batch* b);
hipError_t scan_grid_buckets(shp_batch* b);
hipError_t launch_grid_emit(shp_batch* b);
size_t grid_scan_bytes(shp_batch* b);
