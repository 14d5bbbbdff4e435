/* shapes_oracle.h - CPU FP64 oracle of the shapes contact-physics step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C++ with IEEE
 * binary64 arithmetic and no FMA, the per-step path that
 * Physics.Engine.Main.updateWorld drives in the reference
 * (shapes/src/Physics/Engine/Main.hs:71-86).  It exists to check the HIP
 * implementation and to time a CPU baseline.  Only the test suite, the
 * smoke check and the benchmark's CPU-baseline leg may load it; the product
 * library never links or calls it.
 *
 * All indices are world slot indices (the reference's fromList order,
 * shapes/src/Physics/World.hs:118-124).
 */
#ifndef SHAPES_ORACLE_H
#define SHAPES_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_engine orc_engine;

/* trig modes for the body transform */
#define ORC_TRIG_ENGINE 0 /* correctly rounded shp_sincos (same bits as GPU) */
#define ORC_TRIG_LIBM 1   /* host libm sin/cos, as GHC's sinDouble# calls */

/* solve schedules */
#define ORC_SOLVER_ORDERED 0 /* the reference: sequential, FP64 */
#define ORC_SOLVER_COLORED 1 /* restatement of the GPU's graph-coloured FP32
                                Gauss-Seidel (not a reference function) */

/* broadphase algorithms (the same pair list for finite AABBs) */
#define ORC_BROADPHASE_GRID 0  /* the reference's grid, Grid.hs:67-141 with
                                  the engine's axes (Engine/Main.hs:40-41) */
#define ORC_BROADPHASE_BRUTE 1 /* every pair, Aabb.hs:168-183 */

typedef struct orc_params {
  double dt;
  double baumgarte;   /* ContactBehavior.contactBaumgarte */
  double slop;        /* ContactBehavior.contactPenetrationSlop */
  double gx, gy;      /* constantAccel vector (External.hs:23-28) */
  int32_t iterations; /* improveWorld sweeps; the reference hard-codes 2 */
  int32_t trig;       /* ORC_TRIG_* */
  int32_t solver;     /* ORC_SOLVER_* */
  int32_t world;      /* the world's index in a GPU batch (colouring hash) */
  int32_t broadphase; /* ORC_BROADPHASE_* */
} orc_params;

/* One world.  hull_off has n+1 entries (CSR into hull_x/hull_y); vertices
 * are local-space, counter-clockwise.  mass/inertia 0 means infinite.
 * capacity > n adds empty slots [n, capacity) (World.new capacity). */
orc_engine* orc_engine_create(int32_t n, const double* px, const double* py,
                              const double* theta, const double* vx,
                              const double* vy, const double* omega,
                              const double* mass, const double* inertia,
                              const double* mu, const double* bounce,
                              const int32_t* hull_off, const double* hull_x,
                              const double* hull_y, const double* radius,
                              const orc_params* prm, int32_t capacity);
/* World.append: b10 = px, py, theta, vx, vy, omega, mass, inertia, mu,
 * bounce; a circle when radius > 0 (nv 0).  Returns the slot or -1 (full). */
int32_t orc_engine_append(orc_engine* e, const double* b10, double radius, int32_t nv,
                          const double* hx, const double* hy);
/* World.delete (safeDelete): 0, or -1 when the slot is already empty. */
int32_t orc_engine_delete(orc_engine* e, int32_t slot);
int32_t orc_engine_filled(const orc_engine* e);
/* 1 per filled slot, 0 per empty one (capacity entries) */
void orc_engine_read_slots(const orc_engine* e, uint8_t* out);
void orc_engine_destroy(orc_engine* e);
void orc_engine_step(orc_engine* e, int32_t steps);

/* bodies: out[6*n] = px,py,theta,vx,vy,omega per body (body-major) */
void orc_engine_read_bodies(const orc_engine* e, double* out);
int32_t orc_engine_num_vertices(const orc_engine* e);
/* world-space vertices and unit edge normals, CSR order: 4 doubles each */
void orc_engine_read_hulls(const orc_engine* e, double* out);
/* AABBs: 4 doubles per body (minx, maxx, miny, maxy) */
void orc_engine_read_aabbs(const orc_engine* e, double* out);

/* last step's broadphase pairs (descending), contacts and cache */
int32_t orc_engine_num_pairs(const orc_engine* e);
void orc_engine_read_pairs(const orc_engine* e, int32_t* ij);
int32_t orc_engine_num_contacts(const orc_engine* e);
/* keys[4*c] = (i, j, fi, fj); flip[c]; geom[5*c] = nx, ny, px, py, depth;
 * lam[2*c] = accumulated (normal, friction) after the last sweep; warm[c] =
 * 1 when the contact reused last frame's lagrangians. */
void orc_engine_read_contacts(const orc_engine* e, int32_t* keys,
                              int32_t* flip, double* geom, double* lam,
                              int32_t* warm);

/* worldContacts: pairs and contacts of the world as it stands (no step);
 * read_pairs / read_contacts show them until the next step. */
void orc_engine_probe(orc_engine* e);
/* the warm-start cache (keys descending, 4 ints; lam 2 doubles each) */
int32_t orc_engine_num_cache(const orc_engine* e);
void orc_engine_read_cache(const orc_engine* e, int32_t* keys, double* lam);

/* per-step statistics of the last step */
typedef struct orc_stats {
  int32_t pairs, contacts, warm_hits;
  int32_t dag_levels; /* longest chain of contacts sharing a dynamic body */
  double max_depth;
  double normal_residual;   /* max |min(lambda_n, Jn.v + b)| after sweeps */
  double friction_residual; /* max |lambda_t - clamp(lambda_t - Jt.v/mt)| */
} orc_stats;
void orc_engine_stats(const orc_engine* e, orc_stats* out);

/* Replace the warm-start cache (keys descending) - used to test the join. */
void orc_engine_set_cache(orc_engine* e, int32_t n, const int32_t* keys,
                          const double* lam);

/* ---- unit-level entry points (known-answer tests) ---------------------- */

/* SAT + clipping between two hulls whose vertices are given in world space
 * (listToHull of those vertices, as shapes/bench/Physics/Contact/
 * Benchmark.hs:16-27 builds them).  Returns the number of contacts (0..2);
 * out_keys[2*k] = (fa, fb), out_flip[k], out_geom[5*k] = nx,ny,px,py,depth.
 * out_sep = 1 when a separating axis was found. */
int32_t orc_hull_contacts(int32_t na, const double* ax, const double* ay,
                          int32_t nb, const double* bx, const double* by,
                          int32_t* out_keys, int32_t* out_flip,
                          double* out_geom, int32_t* out_sep);

/* solveConstraint (shapes/src/Physics/Constraint.hs:207-212).
 * j[6], b; obj[6*2] = vx,vy,w per body then ...; im[4] = (ml,mr) a, b.
 * Writes updated velocities into obj. */
/* Contacts between two shapes (Contact.generateContacts, Contact.hs:22-40):
 * a hull of na > 0 world-space vertices, or a circle (na == 0) with
 * ac = {radius, centre x, centre y}; the same for b. */
int32_t orc_shape_contacts(int32_t na, const double* ax, const double* ay, const double* ac,
                           int32_t nb, const double* bx, const double* by, const double* bc,
                           int32_t* out_keys, int32_t* out_flip, double* out_geom);
/* GJK.closestSimplex (GJK.hs:52-146): vertex indices, most recent first. */
int32_t orc_closest_simplex(int32_t n, const double* x, const double* y, double ox, double oy,
                            int32_t* idx);
void orc_solve_constraint(const double* j, double b, double* vel6,
                          const double* im4);

/* unorderedPairs n (shapes/src/Physics/Broadphase/Aabb.hs:155-163):
 * writes n(n-1)/2 pairs into out (2 ints each), returns the count. */
int32_t orc_unordered_pairs(int32_t n, int32_t* out);

void orc_sincos(double x, int32_t trig, double* s, double* c);

/* Batch: run several independent worlds (CPU baseline).  Each world is a
 * separate orc_engine; threads > 1 steps worlds in parallel with OpenMP.
 * Returns wall seconds of the stepping loop only. */
double orc_engines_run(orc_engine** engines, int32_t n_engines,
                       int32_t steps, int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
