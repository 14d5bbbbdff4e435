/* shapes_gpu.h - C ABI of the MI355X (gfx950) contact-physics step.
 *
 * This library replaces the per-step path of the Haskell engine `shapes`
 * that Physics.Engine.Main drives (shapes/src/Physics/Engine/Main.hs:32-98):
 * AABB broadphase, SAT polygon contact generation, and the warm-started
 * sequential-impulse solve of NonPenetration and Friction constraints,
 * followed by integration and shape transforms.  Each entry point below
 * names the reference definition it stands for.
 *
 * Conventions
 *   - Plain C types only; all arrays are caller-owned host memory that is
 *     copied in (create) or filled (read_*).  Nothing is retained.
 *   - Every call returns shp_status; on error shp_last_error() gives a
 *     thread-local message.  No C++ exception crosses this boundary.
 *   - A handle is not thread-safe; distinct handles may be used from
 *     distinct threads.  Each handle owns one HIP stream on the current
 *     device at creation time.
 *   - Body indices are world slot indices, i.e. the reference's fromList
 *     order (shapes/src/Physics/World.hs:118-124).
 */
#ifndef SHAPES_GPU_H
#define SHAPES_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t shp_status;
#define SHP_OK 0
#define SHP_E_INVAL 1       /* bad argument */
#define SHP_E_OOM 2         /* device allocation failed */
#define SHP_E_HIP 3         /* HIP runtime error */
#define SHP_E_UNSUPPORTED 4 /* hulls with > 8 vertices, capacities, ... */
#define SHP_E_OVERFLOW 5    /* pair/contact capacity exceeded */
#define SHP_E_NONFINITE 6   /* a body reached a non-finite state */

/* Solver modes */
#define SHP_SOLVER_ORDERED_F64 0 /* reference order, FP64, bitwise drop-in */
#define SHP_SOLVER_COLORED_F32 1 /* graph-coloured Gauss-Seidel, FP32 math,
                                    FP64 impulse accumulators */

/* Broadphase algorithms (same pair set and order) */
#define SHP_BROADPHASE_AUTO 0 /* grid when a world has more than 2048 bodies */
#define SHP_BROADPHASE_SCAN 1 /* one wavefront per body scans j < i */
#define SHP_BROADPHASE_GRID 2 /* uniform grid + explicit large static bodies;
                                 16 lanes per body up to 131072 bodies, one
                                 lane per body beyond */
#define SHP_BROADPHASE_GRID_1 3  /* the grid, one lane per body */
#define SHP_BROADPHASE_GRID_16 4 /* the grid, 16 lanes per body (one per
                                    candidate cell) */

/* Warm-start modes for contacts absent from last frame's cache */
#define SHP_WARM_ZERO 0    /* the reference: start at 0
                              (shapes/src/Physics/Solvers/Contact.hs:95) */
#define SHP_WARM_LEARNED 1 /* MLP-predicted initial impulses */

/* One world, structure of arrays (n bodies).  Mirrors makePhysicalObj vel
 * rotvel pos rotpos (mass, inertia) (shapes/src/Physics/Engine.hs:27-39),
 * makeWorldObj phys mu bounce shape (World.hs:142-149) and makeHull
 * (Engine.hs:47-48).  mass/inertia 0 = infinite (toInvMass2,
 * Constraint.hs:79-83).  Hull vertices are local-space, counter-clockwise,
 * 3..8 per body, CSR: body b owns hull_x/y[hull_off[b] .. hull_off[b+1]). */
typedef struct shp_world_desc {
  int32_t n_bodies;
  const double* px;
  const double* py;
  const double* theta;
  const double* vx;
  const double* vy;
  const double* omega;
  const double* mass;
  const double* inertia;
  const double* mu;
  const double* bounce;
  const int32_t* hull_off; /* n_bodies + 1 entries */
  const double* hull_x;
  const double* hull_y;
  /* NULL: every body is a hull.  Otherwise radius[i] > 0 makes body i a
   * circle of that radius (makeCircle, Engine.hs:53-54; Circle.hs:19-25)
   * whose hull range must be empty; radius[i] == 0 keeps it a hull. */
  const double* radius;
  /* Slots of the world (World.new capacity, World.hs:61-75): 0 means
   * n_bodies (fromList, :118-124); otherwise >= n_bodies, and slots
   * [n_bodies, capacity) start empty, available to shp_batch_append. */
  int32_t capacity;
} shp_world_desc;

/* One body for shp_batch_append: makePhysicalObj + makeWorldObj as in
 * shp_world_desc, one body; radius > 0 with n_vertices 0 for a circle,
 * else radius 0 and 3..8 local-space counter-clockwise vertices. */
typedef struct shp_body_desc {
  double px, py, theta, vx, vy, omega;
  double mass, inertia, mu, bounce;
  double radius;
  int32_t n_vertices;
  const double* hull_x;
  const double* hull_y;
} shp_body_desc;

/* EngineConfig + ContactBehavior (Engine/Main.hs:34-37, Contact/Types.hs:
 * 20-25), the constant-acceleration External (World/External.hs:23-28)
 * and the sweep count, which the reference fixes at 2 (Main.hs:81-82). */
typedef struct shp_config {
  double dt;
  double baumgarte;
  double slop;
  double gravity_x;
  double gravity_y;
  int32_t iterations;
  int32_t solver;             /* SHP_SOLVER_* */
  int32_t warm_start;         /* SHP_WARM_* */
  int32_t max_pairs_per_body; /* capacity; 0 = default (8).  Pair capacity
                                 = bodies * this + 1024, below 2^30, and
                                 below 2^28 with the coloured solver
                                 (SHP_E_UNSUPPORTED otherwise) */
  int32_t broadphase;         /* SHP_BROADPHASE_* */
  int32_t solver_threads;     /* lanes per world in the solve: 0 = by world
                                 size, or 256 or 512, or for the coloured
                                 solver 1024, or 64: one wavefront per world
                                 (several per workgroup) for worlds of at most
                                 1024 bodies and 256 lanes for larger ones (a
                                 schedule choice; the results do not depend
                                 on it) */
} shp_config;

typedef struct shp_batch shp_batch;

/* Per-step counters of the last step (summed over worlds). */
typedef struct shp_step_stats {
  int64_t pairs;
  int64_t contacts;
  int64_t warm_hits;
  int64_t learned_starts;
  int64_t misses;       /* contacts absent from last frame's cache */
  int32_t max_levels;   /* most colours of a world (coloured solver) */
  int32_t flags;        /* bit 0 pair overflow, 1 contact overflow,
                           2 non-finite, 3 coloured schedule overflow,
                           4 ordered solver stopped waiting (defect) */
  /* Solve quality after the last sweep (FP64 state), maxima over contacts
   * of: penetration depth, |min(lambda_n, Jn.v + b)| and
   * |lambda_t - clamp(lambda_t - Jt.v / mt, +-mu*lambda_n)|; for the last
   * step and since creation. */
  double max_depth;
  double normal_residual;
  double friction_residual;
  double run_max_depth;
  double run_normal_residual;
  double run_friction_residual;
  /* Means over the steps since creation of the three per-step maxima. */
  double mean_max_depth;
  double mean_normal_residual;
  double mean_friction_residual;
  int64_t manifolds;    /* pairs with stored contacts (the coloured solver's
                           manifolds) */
} shp_step_stats;

/* ---- engine lifecycle (initEngine / runEngine, Engine/Main.hs:43-98) ---- */

/* Create n_worlds independent worlds stepped together.  n_worlds = 1 is the
 * reference engine; each world keeps its own warm-start cache
 * (EngineCache, Main.hs:32).  The cache starts empty (initEngine, :43-46). */
shp_status shp_batch_create(const shp_world_desc* worlds, int32_t n_worlds,
                            const shp_config* cfg, shp_batch** out);
shp_status shp_batch_destroy(shp_batch* b);

/* updateWorld repeated n_steps times (stepWorld, Main.hs:71-89). */
shp_status shp_batch_step(shp_batch* b, int32_t n_steps);

/* Enqueue n_steps on the handle's stream without waiting; shp_batch_sync
 * waits and reports deferred errors.  With use_graph != 0 each step is
 * replayed from a captured HIP graph. */
shp_status shp_batch_step_async(shp_batch* b, int32_t n_steps,
                                int32_t use_graph);
shp_status shp_batch_sync(shp_batch* b);

/* ---- readback (readPhysObj, World.hs:93-94; debug as IOWorld.hs:74-81) */

shp_status shp_batch_num_bodies(const shp_batch* b, int32_t world,
                                int32_t* n);
/* px, py, theta, vx, vy, omega of every body of `world` (NULL skips). */
shp_status shp_batch_read_bodies(shp_batch* b, int32_t world, double* px,
                                 double* py, double* theta, double* vx,
                                 double* vy, double* omega);
/* Last step's broadphase pairs (G.culledKeys, Broadphase/Grid.hs:74-78),
 * descending.  Call with ij = NULL to get the count only. */
shp_status shp_batch_read_pairs(shp_batch* b, int32_t world, int32_t* count,
                                int32_t* ij);
/* Last step's contacts in descending ObjectFeatureKey order (prepareFrame,
 * Solvers/Contact.hs:40-52), with the lagrangians accumulated by the last
 * sweep (the next step's cache).  keys: 4 ints (i, j, fi, fj); geom: nx, ny,
 * px, py, depth; lam: normal, friction; warm: 1 if reused from the cache,
 * 2 if started by the learned model.  Any output may be NULL. */
shp_status shp_batch_read_contacts(shp_batch* b, int32_t world,
                                   int32_t* count, int32_t* keys,
                                   int32_t* flip, double* geom, double* lam,
                                   int32_t* warm);
/* World-space vertices and unit edge normals (4 doubles per vertex, CSR
 * order) as left by moveShapes (World.hs:136-140); shp_batch_num_vertices
 * gives the world's vertex count (the hulls' sizes, slot by slot). */
shp_status shp_batch_read_hulls(shp_batch* b, int32_t world, double* out);
shp_status shp_batch_num_vertices(const shp_batch* b, int32_t world,
                                  int32_t* n);
shp_status shp_batch_last_stats(shp_batch* b, shp_step_stats* out);

/* Replace `world`'s warm-start cache (keys descending, 4 ints each;
 * lam: 2 doubles each). */
shp_status shp_batch_set_cache(shp_batch* b, int32_t world, int32_t count,
                               const int32_t* keys, const double* lam);

/* ---- world mutation between steps (World.hs:77-91 over
 * Utils/EmptiesVector.hs:45-85) ------------------------------------------
 * Each world's slots keep the reference's free list: delete pushes the slot
 * onto the list of empty slots, append takes the most recently emptied
 * slot (or the next never-filled one).  Empty slots are skipped by every
 * pass of the step (no acceleration, pairs, advance or placement), keep
 * their last state (read back by shp_batch_read_bodies), and the warm-start
 * cache is kept as it is, so a body appended into a slot emptied since the
 * last step can reuse that slot's cached lagrangians, as in the reference. */

/* World.append (World.hs:77-84, EmptiesVector.append :70-85): the body into
 * the head empty slot, placed by moveShape (World.hs:142-149); *out_slot is
 * its slot.  SHP_E_INVAL when the world has no empty slot (the reference's
 * out-of-bounds read).  A body with as many vertices as the slot's previous
 * occupant (and, with the grid broadphase, no wider than the world's cells)
 * is written in place; otherwise the batch's arrays are laid out again
 * (the cost of a batch creation; state, warm-start caches, run statistics,
 * model and training log carried over; the last step's pairs and contacts
 * read back empty until the next step). */
shp_status shp_batch_append(shp_batch* b, int32_t world,
                            const shp_body_desc* body, int32_t* out_slot);
/* World.delete = EmptiesVector.safeDelete (World.hs:86-87,
 * EmptiesVector.hs:57-68): SHP_E_INVAL when the slot is already empty. */
shp_status shp_batch_delete(shp_batch* b, int32_t world, int32_t slot);
/* ---- debug overlays and dumps (the demo's worldContacts / worldAabbs,
 * shapes-demo/src/Physics/Demo/IOWorld.hs:74-90) -------------------------- */

/* worldContacts (IOWorld.hs:74-81): G.culledKeys and prepareFrame of the
 * worlds as they stand, without stepping (no acceleration, no solve, no
 * advance; the warm-start caches untouched).  shp_batch_read_pairs and
 * shp_batch_read_contacts then return these pairs and contacts (lam: the
 * cached lagrangians joined by applyCachedSlns, 0 for new contacts) until
 * the next step. */
shp_status shp_batch_probe_contacts(shp_batch* b);
/* worldAabbs = B.toAabbs (IOWorld.hs:86-90, Broadphase/Aabb.hs:119-125):
 * the filled slots in slot order and their AABBs (minx, maxx, miny, maxy).
 * Call with slots = boxes = NULL for the count. */
shp_status shp_batch_read_aabbs(shp_batch* b, int32_t world, int32_t* count,
                                int32_t* slots, double* boxes);
/* The EngineCache (Engine/Main.hs:32) the next step joins against: keys
 * descending (4 ints) and the accumulated (normal, friction) lagrangians. */
shp_status shp_batch_read_cache(shp_batch* b, int32_t world, int32_t* count,
                                int32_t* keys, double* lam);

/* One world's dump, a self-describing little-endian byte stream:
 *   header  "SHPD", u16 version (1), u16 sections, i32 world, i32 0,
 *           i64 steps since creation                           (24 bytes)
 *   section char tag[4], u32 record bytes, u64 records, then the records:
 *   "BODY"  filled slots: i32 slot, i32 0, f64 px py theta vx vy omega (56)
 *   "AABB"  as shp_batch_read_aabbs: i32 slot, i32 0, f64 minx maxx miny
 *           maxy                                                      (40)
 *   "PAIR"  as shp_batch_read_pairs: i32 i, i32 j                      (8)
 *   "CONT"  as shp_batch_read_contacts: i32 i, j; u8 fi, fj, flip, warm;
 *           u32 0; f64 nx ny px py depth lambda_n lambda_t             (72)
 *   "CACH"  as shp_batch_read_cache: i32 i, j, fi, fj; f64 lambda_n,
 *           lambda_t                                                   (32)
 * in that order, those selected by `sections` (SHP_DUMP_* bits).  With
 * buf = NULL only *size is set; otherwise capacity must hold *size bytes.
 * Dumps of several worlds or steps concatenate. */
#define SHP_DUMP_VERSION 1
#define SHP_DUMP_BODIES 1u
#define SHP_DUMP_AABBS 2u
#define SHP_DUMP_PAIRS 4u
#define SHP_DUMP_CONTACTS 8u
#define SHP_DUMP_CACHE 16u
#define SHP_DUMP_ALL 31u
shp_status shp_batch_dump(shp_batch* b, int32_t world, uint32_t sections,
                          void* buf, int64_t capacity, int64_t* size);

/* World.filled (World.hs:90-91): the number of filled slots. */
shp_status shp_batch_filled(const shp_batch* b, int32_t world,
                            int32_t* out_filled);
/* EmptiesVector.read (EmptiesVector.hs:87-92) of every slot: 1 filled,
 * 0 empty (shp_batch_num_bodies entries). */
shp_status shp_batch_read_slots(const shp_batch* b, int32_t world,
                                uint8_t* out_filled);

/* ---- learned warm start (no reference counterpart: the reference always
 * zero-starts new contacts, Solvers/Contact.hs:95) ------------------------ */

/* A 14-128-128-128-2 SiLU MLP.  Parameter blob (shp_mlp_param_count() =
 * 35202 FP32 values, weights row-major [in][out]): W1 14x128, b1, W2 128x128,
 * b2, W3 128x128, b3, W4 128x2, b4.  Inputs are 14 features of a contact
 * that missed the cache (normal, depth, relative normal/tangent velocity,
 * inverse masses and inertias, lever arms, mean friction); the two outputs
 * times the inverse effective masses' reciprocals give the initial normal
 * and friction impulses, projected onto lambda_n >= 0, |lambda_t| <=
 * mu*lambda_n.  The device keeps BF16 weights and accumulates in FP32. */
int32_t shp_mlp_param_count(void);
/* Load a model into a batch; used when the batch's warm_start is
 * SHP_WARM_LEARNED (without a model, misses start at zero). */
shp_status shp_batch_set_mlp(shp_batch* b, const float* params);
/* Record (features[14], targets[2]) of every cache miss after each step's
 * solve, up to `capacity` samples (0 disables).  Targets are the final
 * lambdas times the rows' inverse effective masses. */
shp_status shp_batch_set_logging(shp_batch* b, int64_t capacity);
/* Copy up to *count logged samples (feats n x 14, targets n x 2); NULL
 * outputs return the number logged in *count. */
shp_status shp_batch_read_log(shp_batch* b, float* feats, float* targets,
                              int64_t* count);

/* Training and evaluation of the model on the GPU (BF16 matrix cores, FP32
 * accumulation, Adam on FP32 master weights). */
typedef struct shp_mlp shp_mlp;
shp_status shp_mlp_create(const float* params, shp_mlp** out);
shp_status shp_mlp_destroy(shp_mlp* m);
shp_status shp_mlp_get_params(shp_mlp* m, float* params);
/* out[2*i + k] = output k of row i (n x 14 features) */
shp_status shp_mlp_forward(shp_mlp* m, const float* feats, int32_t n,
                           float* out);
/* No counterpart in the reference.  shp_mlp_forward for callers that cannot
 * allocate (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): d_feats (n x 14 FP32) and d_out (n x 2 FP32) are
 * device pointers on the handle's GPU, n <= 65536.  The staging lives in
 * the library's static device scratch (6 MiB per GPU, kept until the process
 * ends) instead of buffers the handle allocates; the call allocates nothing.
 * Enqueued on `stream` (a hipStream_t; NULL: the handle's stream) and
 * returns without waiting.  Every call on a GPU shares the scratch: calls
 * that overlap on one GPU race. */
shp_status shp_mlp_forward_noalloc(shp_mlp* m, const float* d_feats,
                                   int32_t n, float* d_out, void* stream);
/* Gradient of the mean squared error over n rows w.r.t. every parameter
 * (blob layout) and the loss itself (with a Huber threshold set, the
 * gradient is that of twice the mean Huber loss; the loss reported stays
 * the squared error). */
shp_status shp_mlp_gradient(shp_mlp* m, const float* feats,
                            const float* targets, int32_t n, float* grad,
                            float* loss);
/* Training loss: squared error (huber_delta = 0, the default) or Huber with
 * threshold huber_delta (robust to the heavy tail of impact impulses). */
shp_status shp_mlp_set_loss(shp_mlp* m, float huber_delta);
/* Upload a training set of n samples. */
shp_status shp_mlp_set_data(shp_mlp* m, const float* feats,
                            const float* targets, int64_t n);
/* `steps` Adam steps (learning rate lr) on mini-batches of `batch` rows
 * (a multiple of 64) given by row indices (steps x batch) into the training
 * set; loss_out[s] (may be NULL) is step s's batch loss before the update. */
shp_status shp_mlp_train(shp_mlp* m, const int32_t* indices, int32_t steps,
                         int32_t batch, float lr, float* loss_out);

/* ---- per-phase timing (no counterpart in the reference, whose only
 * timing is the criterion benchmark, shapes/bench/Main.hs:19-26) --------- */

/* Phases of one step on the batch's stream, in launch order: 0 reset,
 * 1 broadphase count (with gravity), 2 pair scan, 3 broadphase fill, 4 SAT,
 * 5 contact scan, 6 world ranges, 7 contacts (constraintGen + warm-start
 * join), 8 learned inference, 9 solve set-up (the coloured solver: the wait
 * for the colouring and the manifold packing), 10 sweeps (k_sweep_colours or
 * k_solve_flow), 11 solve write-back (coloured: k_colour_unpack), 12 advance
 * + moveShapes + AABBs + solve quality (k_finish), 13 cache store; and 14
 * the coloured solver's colouring (k_colour), timed on the second stream,
 * where it overlaps phases 5-9. */
#define SHP_NUM_PHASES 15
/* Enable (1) or disable (0) HIP-event timing of each phase on the batch's
 * stream; resets the accumulated times.  Timed steps synchronise once per
 * step and are never replayed from a graph. */
shp_status shp_batch_set_timing(shp_batch* b, int32_t enable);
/* Accumulated milliseconds per phase and the number of timed steps. */
shp_status shp_batch_read_timing(shp_batch* b, double* ms_per_phase,
                                 int32_t* steps);
/* Average microseconds per timed step of the solve's sub-phases in world 0:
 * ordered: 0 per-body ranks, 1 counter clearing and velocity staging,
 * 2 levels + level sort (level-synchronous worlds) + record packing, 3
 * warm-start pass and sweeps, 4 write-back; coloured: 0 colouring, 1 bucketing + packing +
 * velocity staging, 2 warm-start pass, 3 sweeps, 4 write-back. */
shp_status shp_batch_read_solve_profile(shp_batch* b, double* us5);

/* ---- plan and audit of a batch (no counterpart in the reference) -------
 *
 * What shp_batch_create would allocate and launch for these worlds and this
 * configuration, computed on the host only (no GPU needed): capacities,
 * the device footprint, and every kernel launch of one step.  Returns what
 * shp_batch_create would return for the same arguments before touching the
 * device: SHP_OK, or the validation or audit error (a launch beyond the
 * hardware's limits - more than 2^31 - 1 workgroups, 2^32 work-items or
 * 160 KiB of LDS - or an index beyond 32 bits) or a refused configuration.
 * `out` is filled whenever the worlds are valid, also when refused. */
typedef struct shp_plan {
  int64_t bodies;
  int64_t vertices;
  int64_t worlds;
  int64_t max_world_bodies;
  int64_t pair_capacity;
  int64_t contact_capacity;
  int64_t miss_capacity;    /* cache misses the learned warm start can take */
  int64_t grid_buckets;     /* hash buckets of the grid broadphase (0: scan) */
  int64_t device_bytes;     /* device memory allocated at creation */
  int64_t max_workitems;    /* largest grid x workgroup of a step's launches */
  int64_t max_lds_bytes;    /* largest static + dynamic LDS of a launch */
  int32_t launches;         /* kernel launches per step */
  int32_t broadphase;       /* SHP_BROADPHASE_SCAN or SHP_BROADPHASE_GRID */
  int32_t solver_threads;   /* lanes per world of the solve */
  int32_t reserved;
} shp_plan;

typedef struct shp_launch {
  char name[32];
  int64_t workgroups;
  int32_t workgroup_size;
  int32_t lds_bytes; /* static + dynamic */
} shp_launch;

/* launches (may be NULL) receives up to launch_cap launch records. */
shp_status shp_batch_plan(const shp_world_desc* worlds, int32_t n_worlds,
                          const shp_config* cfg, shp_plan* out,
                          shp_launch* launches, int32_t launch_cap);

const char* shp_last_error(void);
const char* shp_version(void);

#ifdef __cplusplus
}
#endif
#endif
