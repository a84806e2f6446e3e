/*
 * jdem.h -- C ABI of libjdem_hip.so, the MI355X (gfx950) DEM time step.
 *
 * Plain C: fixed-width integers, plain structs, raw pointers and sizes.  No
 * C++, HIP or framework types cross this boundary; the HIP stream is passed
 * as an opaque `void*` (a hipStream_t, or NULL for the null stream).
 *
 * Ownership: every array pointer in these structs is a DEVICE pointer owned
 * by the caller (torch tensors through ctypes today; XLA buffers under a
 * jax.ffi binding).  The library owns only the scratch memory of a context
 * (`jdem_ctx`), which grows on demand and is reused across calls.
 *
 * Errors: every entry point returns an int status: JDEM_OK (0) or a negative
 * JDEM_E* code; `jdem_last_error(ctx)` gives a message.  Runtime conditions
 * of the reference that are flags and not exceptions (collider overflow,
 * colliders/__init__.py:51-54; _partition.py:96-97, 149) are reported through
 * the `overflow` fields, never as errors.
 *
 * Threading: one context per stream; contexts are independent.  Calls on one
 * context are not thread-safe.
 *
 * Layout: (N, D) arrays are row-major float32 (D = 2 or 3, the spatial
 * dimension); angular arrays are (N, 3) in 3D and (N, 1) in 2D; the
 * quaternion is (N, 4) = [w, x, y, z]; neighbour lists are (N, M) int32
 * padded with -1, as in JaxDEM (cell_list.py:500-595).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * cdelv/JaxDEM).
 */
#ifndef JDEM_H
#define JDEM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JDEM_ABI_VERSION 1

enum {
  JDEM_OK = 0,
  JDEM_EINVAL = -1,       /* bad argument (shape, kind, null pointer)      */
  JDEM_EHIP = -2,         /* a HIP runtime call failed                     */
  JDEM_ENOMEM = -3,       /* scratch allocation failed                     */
  JDEM_EUNSUPPORTED = -4  /* combination not implemented on the GPU path   */
};

/* Domain kinds: domains/free.py, domains/periodic.py, domains/reflect.py. */
enum { JDEM_DOMAIN_FREE = 0, JDEM_DOMAIN_PERIODIC = 1, JDEM_DOMAIN_REFLECT = 2 };

/* Pair laws: forces/spring.py, forces/hertz.py, forces/cundall_strack.py,
 * two history-carrying laws with no reference counterpart, and ROUTED: the
 * composite laws ForceRouter (forces/router.py:26-339) and LawCombiner
 * (forces/law_combiner.py:22-191). */
enum {
  JDEM_LAW_NONE = 0,
  JDEM_LAW_SPRING = 1,
  JDEM_LAW_HERTZ = 2,
  JDEM_LAW_CUNDALL_STRACK = 3,
  JDEM_LAW_SPRING_FRICTION = 4,
  JDEM_LAW_HERTZ_FRICTION = 5,
  JDEM_LAW_ROUTED = 6
};

/* Integrators: integrators/velocity_verlet.py, direct_euler.py,
 * velocity_verlet_spiral.py, spiral.py ("" = no-op). */
enum { JDEM_LIN_NONE = 0, JDEM_LIN_VERLET = 1, JDEM_LIN_EULER = 2 };
enum { JDEM_ROT_NONE = 0, JDEM_ROT_VERLET_SPIRAL = 1, JDEM_ROT_SPIRAL = 2 };

/* Colliders: colliders/naive.py, cell_list.py, neighbor_list.py. */
enum {
  JDEM_COLLIDER_NONE = 0,
  JDEM_COLLIDER_NAIVE = 1,
  JDEM_COLLIDER_CELL_LIST = 2,
  JDEM_COLLIDER_NEIGHBOR_LIST = 3
};

/* Pieces of one step for jdem_step_stage (system.py:60-82). */
enum {
  JDEM_STAGE_DOMAIN_APPLY = 1,
  JDEM_STAGE_LIN_BEFORE = 2,
  JDEM_STAGE_ROT_BEFORE = 3,
  JDEM_STAGE_FORCE_MANAGER = 4,
  JDEM_STAGE_LIN_AFTER = 5,
  JDEM_STAGE_ROT_AFTER = 6
};

/* Particle state: the sphere fields of State (state.py:103-228). */
typedef struct jdem_state {
  int64_t n;          /* number of particles N                         */
  int32_t dim;        /* 2 or 3                                        */
  int32_t bond_width; /* W of bond_id (N, W); 0 if there are no bonds  */
  float* pos;         /* pos_c (N, D), in/out                          */
  float* vel;         /* (N, D), in/out                                */
  float* ang_vel;     /* (N, 3) or (N, 1), in/out                      */
  float* q;           /* (N, 4) [w x y z], in/out                      */
  float* force;       /* (N, D), in/out                                */
  float* torque;      /* (N, 3) or (N, 1), in/out                      */
  const float* rad;     /* (N,)                                        */
  const float* mass;    /* (N,)                                        */
  const float* inertia; /* (N, 3) or (N, 1)                            */
  const int32_t* mat_id;  /* (N,)                                      */
  const int32_t* bond_id; /* (N, W), -1 padded, or NULL                */
  const uint8_t* fixed;   /* (N,) 0/1                                  */
  /* Clumps (rigid bodies of spheres; NULL, NULL for spheres): pos_p (N, D)
   * body-frame offsets from the clump centre pos, clump_id (N,) in [0, N).
   * Every member holds its clump's pos, vel, q, ang_vel, mass (the total)
   * and principal inertia.  Pairs of one clump do not interact; member
   * forces act at pos + R(q) pos_p, and ForceManager sums them over the
   * clump (force_manager.py:359-412).  Colliders NAIVE, CELL_LIST and
   * NEIGHBOR_LIST (on the caller's rows, checked and built on the lab
   * positions); every domain (reflect.py:160-297 for REFLECT). */
  const float* pos_p;
  const int32_t* clump_id;
} jdem_state;

/* Domain parameters (domains/__init__.py:28-120). */
typedef struct jdem_domain {
  int32_t kind;
  int32_t dim;
  float box_size[3];
  float inv_box_size[3];
  float anchor[3];
  float restitution; /* reflect.py:48 */
} jdem_domain;

/* Pair-law parameters.  Material-pair tables are (n_mat, n_mat) float32
 * device arrays, material tables (n_mat,).  Which tables a law reads:
 *   SPRING:          t0 = young_eff                    (spring.py:97-108)
 *   HERTZ:           t0 = (4/3) E*                     (hertz.py:98-117)
 *   CUNDALL_STRACK:  m0 = young, m1 = shear modulus G, m2 = mu, m3 = mu_r,
 *                    t0 = beta(e_eff)          (cundall_strack.py:123-193)
 *   SPRING_FRICTION: t0 = k_n, t1 = k_t, t2 = mu_eff; p0 = 2*damping ratio
 *   HERTZ_FRICTION:  t0 = (4/3) E*, t1 = 8 G*, t2 = mu_eff; p0 = 2*damping
 *   ROUTED:          the pair (i, j) takes the elementary laws
 *                    route_terms[route_off[p] .. route_off[p + 1]) with
 *                    p = species[i] * n_species + species[j] (p = 0 for
 *                    every pair when n_species <= 1), in order: the first
 *                    term's force, torque and energy, plus each later
 *                    term's (a LawCombiner entry starts with a NONE term,
 *                    its zeros, law_combiner.py:72-86; an empty entry is a
 *                    single NONE term).  Terms are kinds NONE .. HERTZ_FRICTION
 *                    with their own tables; an entry holds at most one
 *                    history law, whose tangential spring the pair carries.
 *                    route_off (n_species^2 + 1, or 2) and route_terms
 *                    (n_terms) are host arrays read at each call; species
 *                    is the (N,) device array of State.species_id.
 */
typedef struct jdem_law {
  int32_t kind;
  int32_t n_mat;
  const float* t0;
  const float* t1;
  const float* t2;
  const float* m0;
  const float* m1;
  const float* m2;
  const float* m3;
  float p0;
  float p1;
  /* ROUTED only */
  int32_t n_species;
  int32_t n_terms;
  const int32_t* route_off;
  const struct jdem_law* route_terms;
  const int32_t* species;
} jdem_law;

/* Cell grid of _grid_params (_partition.py:55-99), computed by the host in
 * float32 exactly as the reference does. */
typedef struct jdem_grid {
  int32_t dims[3];
  int32_t strides[3];
  float cell_size[3];   /* per-axis effective cell size                 */
  int32_t search_range; /* stencil half-width (cell_list.py:420-427)    */
  int32_t periodic;
  int64_t n_cells;
} jdem_grid;

/* The resident working set of a NeighborList system between calls (opaque):
 * its work order, build positions, slot-major list, contact-history bits and
 * springs, kept on the device in the step kernel's layouts.  Made by
 * jdem_system_step / jdem_rollout when jdem_nlist.resident_out is set;
 * immutable afterwards, so any number of later calls may start from it. */
typedef struct jdem_resident jdem_resident;

/* NeighborList collider buffers (neighbor_list.py:137-286). */
typedef struct jdem_nlist {
  int32_t max_neighbors; /* M                                            */
  float cutoff;          /* interaction cutoff                           */
  float skin;            /* absolute skin                                */
  int32_t* nl;           /* (N, M) in/out, -1 padded                     */
  float* old_pos;        /* (N, D) in/out: positions at the last build   */
  int32_t n_build_times; /* in/out (host value)                          */
  int32_t overflow;      /* in/out (host value)                          */
  int32_t secondary;     /* JDEM_COLLIDER_CELL_LIST or _NAIVE            */
  jdem_grid grid;        /* secondary cell grid, already inflated so the  *
                          * stencil reaches cutoff + skin                 */
  float* hist_xi;        /* (N, M, D) tangential spring history or NULL  */
  uint32_t* hist_live;   /* (N, ceil(M/32)) slot-carries-history bits     */
  /* in: when not NULL, the list, build positions and history are read from
   * this working set instead of nl / old_pos / hist_* (which may be NULL)  */
  const jdem_resident* resident_in;
  /* out: when not NULL, a step of the resident path stores the working set
   * after the call in a new jdem_resident here and leaves nl / old_pos /
   * hist_* unwritten; jdem_resident_export converts it on demand          */
  jdem_resident** resident_out;
} jdem_nlist;

/* Everything one step reads, i.e. the System pytree (system.py:123-226). */
typedef struct jdem_system {
  jdem_state state;
  jdem_domain domain;
  jdem_law law;
  int32_t lin_kind;
  int32_t rot_kind;
  float dt;
  int32_t interact_same_bond_id;
  /* ForceManager (force_manager.py:341-425), spheres only. */
  float gravity[3];
  const float* ext_force;     /* (N, D) or NULL */
  const float* ext_force_com; /* (N, D) or NULL */
  const float* ext_torque;    /* (N, A) or NULL */
  /* collider */
  int32_t collider_kind;
  jdem_grid grid; /* CellList grid (compute_force)                        */
  jdem_nlist nlist;
} jdem_system;

/* Counters filled by jdem_system_step. */
typedef struct jdem_step_stats {
  int64_t steps;             /* time steps advanced                      */
  int64_t rebuilds;          /* neighbour-list rebuilds performed        */
  int64_t host_syncs;        /* host synchronisations                    */
  int32_t overflow;          /* collider overflow flag after the call    */
  int32_t reserved;
  double free_box[3];        /* FreeDomain box/anchor after the last step */
  double free_anchor[3];
} jdem_step_stats;

/* Arguments of the FIRE optimiser, fire(dt, alpha_init, f_inc, f_dec,
 * f_alpha, N_min, N_bad_max, dt_max_scale, dt_min_scale)
 * (minimizers/optimizers.py:150-160); the reference's defaults are
 * 0.1, 1.1, 0.5, 0.99, 5, 10, 10.0, 1e-3.  Given as Python floats (double):
 * the optimiser scalars are float32, except that dt * dt_{max,min}_scale is
 * a double product rounded once, as in the reference. */
typedef struct jdem_fire {
  double dt;
  double alpha_init;
  double f_inc;
  double f_dec;
  double f_alpha;
  int32_t n_min;
  int32_t n_bad_max;
  double dt_max_scale;
  double dt_min_scale;
} jdem_fire;

/* Result of jdem_minimize. */
typedef struct jdem_minimize_stats {
  int64_t steps;    /* iterations taken (routines.py:150-383 `steps`)      */
  double pe;        /* final potential energy per particle (float64 sum)  */
  double max_grad;  /* final max |force| / |torque| component             */
  double power;     /* FIRE power F . v of the last iteration (float64)   */
  int32_t overflow; /* collider overflow flag after the call               */
  int32_t reserved;
} jdem_minimize_stats;

/* Kernel timing gathered with HIP events on the context's stream while
 * timing is enabled (jdem_ctx_set_timing).  Step kernels that found a
 * neighbour-list rebuild pending and did no work are counted apart. */
typedef struct jdem_timing {
  double step_kernel_ms;   /* fused step kernels that did work          */
  int64_t step_kernels;
  double idle_kernel_ms;   /* step kernels that returned early          */
  int64_t idle_kernels;
  double rebuild_ms;       /* neighbour-list / cell-list rebuilds       */
  int64_t rebuilds;
} jdem_timing;

typedef struct jdem_ctx jdem_ctx;

/* ---- library ---------------------------------------------------------- */

/* ABI version (JDEM_ABI_VERSION). */
int jdem_abi_version(void);

/* gfx target the library was built for, e.g. "gfx950". */
const char* jdem_build_target(void);

/* Create a context bound to `device` and `stream` (hipStream_t or NULL). */
int jdem_ctx_create(int device, void* stream, jdem_ctx** out);
int jdem_ctx_destroy(jdem_ctx* ctx);
int jdem_ctx_set_stream(jdem_ctx* ctx, void* stream);
/* Message for the last failed call on this context. */
const char* jdem_last_error(const jdem_ctx* ctx);
/* Bytes of scratch the context holds now. */
int jdem_ctx_scratch_bytes(const jdem_ctx* ctx, size_t* bytes);
/* Enable (1) or disable (0) HIP-event timing of the kernels that
 * jdem_system_step launches.  Has no counterpart in the reference (JaxDEM
 * times whole jitted calls, benchmarks/run_benchmarks.py:27-51). */
int jdem_ctx_set_timing(jdem_ctx* ctx, int enable);
/* Synchronise the stream, return the accumulated timing and reset it. */
int jdem_ctx_get_timing(jdem_ctx* ctx, jdem_timing* out);

/* ---- (a) cell list ------------------------------------------------------ */

/* Hash + stable sort by cell (a counting sort; an LSD radix sort for grids
 * much larger than N) + cell start/end: the spatial partition of
 * _get_spatial_partition (cell_list.py:37-87).  perm/sorted_hash are (N,)
 * int32; cell_start/cell_end are (grid->n_cells,) int32 (empty cell:
 * start == end). */
int jdem_cell_build(jdem_ctx* ctx, const jdem_state* st, const jdem_domain* dom,
                    const jdem_grid* grid, int32_t* perm, int32_t* sorted_hash,
                    int32_t* cell_start, int32_t* cell_end);

/* jdem_cell_build without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and the same outputs as jdem_cell_build; the
 * sort's scratch lives in one fixed device array of
 * jdem_noalloc_scratch_bytes() bytes (64 MiB) per GPU, held until the
 * process ends.  Sizes whose scratch does not fit (about 1.4 M particles on
 * a counting-sort grid) return JDEM_EINVAL with a message.  All calls on
 * one GPU share that array: calls that overlap on one GPU race. */
int jdem_cell_build_noalloc(jdem_ctx* ctx, const jdem_state* st, const jdem_domain* dom,
                            const jdem_grid* grid, int32_t* perm, int32_t* sorted_hash,
                            int32_t* cell_start, int32_t* cell_end);
/* Bytes of the fixed scratch array of jdem_cell_build_noalloc. */
size_t jdem_noalloc_scratch_bytes(void);

/* Candidate pairs within `cutoff`: DynamicCellList.create_neighbor_list
 * (cell_list.py:500-595) when collider_kind is CELL_LIST (grid already
 * inflated to the cutoff, cell_list.py:537-538), or
 * NaiveSimulator.create_neighbor_list (naive.py:122-185) when NAIVE.
 * nl is (N, M) int32 padded with -1; *overflow (host) receives the flag. */
int jdem_neighbor_list(jdem_ctx* ctx, const jdem_system* sys, int32_t collider_kind,
                       const jdem_grid* grid, float cutoff, int32_t max_neighbors,
                       int32_t* nl, int32_t* overflow);

/* ---- (b) forces ----------------------------------------------------------- */

/* Collider.compute_force (colliders/__init__.py:56-80; naive.py:190-235;
 * cell_list.py:437-464; neighbor_list.py:594-669): writes state.force and
 * state.torque.  For NEIGHBOR_LIST it also performs _check_and_rebuild
 * (neighbor_list.py:58-131) and updates sys->nlist in place. */
int jdem_compute_force(jdem_ctx* ctx, jdem_system* sys);

/* Collider.compute_potential_energy (cell_list.py:469-496; naive.py:76-117):
 * 1/2 sum of pair energies, pair terms in float32, sum in float64.  Result
 * written to *energy (host). */
int jdem_potential_energy(jdem_ctx* ctx, jdem_system* sys, double* energy);

/* Contact virial W = sum_{i<j} r_ij (x) F_ij (contacts.py:89-164) in float64,
 * written to the host array w (dim*dim, row-major). */
int jdem_contact_virial(jdem_ctx* ctx, jdem_system* sys, double* w);

/* Kinetic energy, translational + rotational (utils/thermal.py:83-122), in
 * float64 to *energy (host). */
int jdem_kinetic_energy(jdem_ctx* ctx, const jdem_system* sys, double* energy);

/* The two parts of it, each a float64 sum: translational (thermal.py:83-102)
 * and rotational with the body-frame angular velocity (thermal.py:105-122),
 * to host doubles. */
int jdem_kinetic_energy_parts(jdem_ctx* ctx, const jdem_system* sys, double* translational, double* rotational);

/* ---- (c) integration -------------------------------------------------------- */

/* One stage of _step_once (system.py:60-82) on its own: domain.apply,
 * {linear,rotation}_integrator.step_{before,after}_force, ForceManager.apply.
 * For a FREE domain, DOMAIN_APPLY writes the new box/anchor into
 * sys->domain (host). */
int jdem_step_stage(jdem_ctx* ctx, jdem_system* sys, int32_t stage);

/* System.step(state, system, n) (system.py:701-748): n fused steps.  State
 * arrays are updated in place; sys->nlist is updated in place for the
 * NeighborList collider.  stats may be NULL. */
int jdem_system_step(jdem_ctx* ctx, jdem_system* sys, int64_t n, jdem_step_stats* stats);

/* Frames of a rollout: frame f is the state after steps[0] + ... + steps[f]
 * steps.  Each array holds n_frames consecutive copies of the matching state
 * array (pos/vel/force: n*dim floats a frame; ang_vel/torque: n*(1 or 3);
 * q: n*4), device memory.  steps is a HOST array. */
typedef struct jdem_frames {
  int64_t n_frames;
  const int64_t* steps;
  float* pos;
  float* vel;
  float* ang_vel;
  float* q;
  float* force;
  float* torque;
} jdem_frames;

/* System.trajectory_rollout(state, system, n, stride/strides, save_fn)
 * (system.py:607-700, scan body system.py:101-120) with the state saved at
 * every frame: n_frames segments of System.step, each segment's end state
 * written to its frame without leaving the resident working set in between.
 * NeighborList collider only, not with a FREE domain.  The state and
 * sys->nlist end as after jdem_system_step(sum of steps). */
int jdem_rollout(jdem_ctx* ctx, jdem_system* sys, const jdem_frames* frames, jdem_step_stats* stats);

/* Release a working set made through jdem_nlist.resident_out.  Its device
 * buffers go back to the context's pool for later calls (no free on the
 * hot path).  No counterpart in the reference (JAX frees arrays by
 * reference counting). */
int jdem_resident_release(jdem_ctx* ctx, jdem_resident* res);
/* The NeighborList buffers of a working set in the caller's layouts:
 * nl (N, M) int32 -1 padded, old_pos (N, D), and, when the system carries
 * history, hist_xi (N, M, D) and hist_live (N, ceil(M/32)) -- the arrays
 * jdem_system_step writes when resident_out is NULL
 * (neighbor_list.py:137-286).  hist_xi / hist_live may be NULL. */
int jdem_resident_export(jdem_ctx* ctx, const jdem_resident* res, int32_t* nl, float* old_pos, float* hist_xi,
                         uint32_t* hist_live);
/* Free the device buffers the context keeps in its pool (released working
 * sets, grown scratch).  The next call allocates what it needs again. */
int jdem_ctx_trim(jdem_ctx* ctx);

/* ---- (d) batches of independent systems --------------------------------------- */

/* A segmented batch: n_systems independent systems whose particles are
 * concatenated in sys->state (system k owns particles offsets[k] ..
 * offsets[k+1]-1).  Every array here is a HOST array of n_systems entries
 * (offsets: n_systems + 1).  Per system: its domain (PERIODIC or REFLECT),
 * the grid of its secondary CellList (inflated to cutoff + skin as
 * cell_list.py:537-538), dt, and the NeighborList cutoff and skin. */
typedef struct jdem_batch {
  int32_t n_systems;
  /* how rebuild decisions reach the kernels: 0 = a gated kernel chain after
   * every step, long calls replayed from a captured hipGraph; 1 = the same
   * chain launched step by step; 2 = speculative chunks of step kernels with
   * one host read per chunk, the chain launched only when a system asked for
   * a rebuild (fewer launches when rebuilds are rare); once two rebuilds in
   * a row come spec_min_gap steps or fewer after the previous one, the rest
   * of the call runs as mode 1 */
  int32_t launch_mode;
  int32_t* n_build_times; /* in/out: NeighborList builds of each system       */
  const int64_t* offsets;
  const jdem_domain* domains;
  const jdem_grid* grids;
  const float* dt;
  const float* cutoff;
  const float* skin;
  int32_t tile_width; /* particles per workgroup of the step kernel: 256
                       * (also for 0) or 128 */
  int32_t spec_min_gap; /* launch mode 2: see above; 0 = 2 steps */
} jdem_batch;

/* System.step for a batch of systems of unequal size in one segmented
 * layout -- what vmap over a padded batch of equal-size systems does in the
 * reference (system.py:743-744).  sys carries the concatenated state, the
 * shared law (material tables; mat_id selects a system's material),
 * integrators, gravity and the NeighborList buffers (one max_neighbors for
 * all systems; secondary = CELL_LIST); sys->dt, sys->domain and
 * sys->nlist.{grid,cutoff,skin,n_build_times} are ignored.  Each system
 * rebuilds its list when its own particles moved more than its skin/2
 * (neighbor_list.py:78-85), at the same steps as alone, so every trajectory
 * equals that of the same system stepped alone (jdem_system_step) bit for
 * bit.  sys->nlist.overflow is the OR over the systems. */
int jdem_batch_step(jdem_ctx* ctx, jdem_system* sys, const jdem_batch* batch, int64_t n, jdem_step_stats* stats);

/* ---- (e) energy minimisation -------------------------------------------------- */

/* minimize(state, system, max_steps, pe_tol, pe_diff_tol, force_tol) with the
 * FIRE optimiser (minimizers/routines.py:150-383, optimizers.py:150-339):
 * one force/energy evaluation per iteration; stops when max_steps is
 * reached, |pe|/N <= pe_tol, the relative change of pe drops below
 * pe_diff_tol, or max |grad| <= force_tol.  Positions and orientations are
 * updated in place (fixed particles keep their position), force/torque hold
 * the last evaluation.  The potential energy and the FIRE power are float64
 * sums on the GPU (the reference sums in the state dtype).  Gravity is
 * applied and its energy counted; sys->ext_* must be NULL. */
int jdem_minimize(jdem_ctx* ctx, jdem_system* sys, const jdem_fire* fire, int64_t max_steps, double pe_tol,
                  double pe_diff_tol, double force_tol, jdem_minimize_stats* stats);

/* ---- (f) precision form ---------------------------------------------------------- */

/* Harmonic spheres (the spring law with one spring constant) in a periodic
 * box whose positions are held as an integer cell index plus a float32
 * offset inside the cell: x = cidx * cell + off per axis, 0 <= off < cell.
 * Requirements: dims[a] >= 3 and cell[a] >= the largest diameter.  No
 * counterpart in the reference, whose positions are absolute floats. */
typedef struct jdem_precise {
  int64_t n;
  int32_t dim;
  int32_t dims[3];  /* cells per axis; the box edge is dims * cell     */
  float cell[3];
  float cell_lo[3]; /* low parts: the edge is cell + cell_lo exactly     */
  int32_t* cidx;    /* (N, D) device, in/out                           */
  float* off;       /* (N, D) device, in/out                           */
  const float* rad; /* (N,) device                                     */
  float* force;     /* (N, D) device, out: forces of the last evaluation */
  float k;          /* spring constant                                 */
  int32_t stall_steps; /* stop once neither the largest force component
                          nor the energy (by more than 1e-13 of it) has
                          reached a new low for this many steps (0: never);
                          float32 positions have a residual floor         */
  float skin;            /* Verlet list: pairs within r_i + r_j + skin, rebuilt
                            (with the cell re-sort) once a sphere has moved
                            skin / 2; 0 or max_neighbors 0: the cell stencil
                            every evaluation                               */
  int32_t max_neighbors; /* list row length; a bound on the spheres within
                            2 r_max + skin of a centre, or the call fails  */
} jdem_precise;

/* minimize (minimizers/routines.py:150-383) with FIRE on a jdem_precise
 * system: float32 pair arithmetic, float64 energy, virial, FIRE power and
 * residual.  virial (D*D, may be NULL) receives the float64 contact virial
 * sum over pairs r_ij (x) F_ij of the final configuration. */
int jdem_precise_minimize(jdem_ctx* ctx, jdem_precise* p, const jdem_fire* fire, int64_t max_steps, double pe_tol,
                          double pe_diff_tol, double force_tol, jdem_minimize_stats* stats, double* virial);

/* ---- (f) the plain step in float64 -------------------------------------------
 *
 * The reference computes in float64 when jax_enable_x64 is set (its own test
 * suite does, tests/conftest.py:4).  These structs are jdem_state,
 * jdem_domain, jdem_law, jdem_grid and jdem_system with every floating-point
 * array and scalar in float64, and the entry points run the kernels of the
 * float32 calls above instantiated in double, with the same association of
 * operations.  Colliders: NONE, NAIVE, CELL_LIST and NEIGHBOR_LIST (list,
 * skin and rebuild as jdem_nlist, without the resident working set); laws:
 * SPRING, HERTZ, CUNDALL_STRACK and ROUTED over those (the history laws
 * return JDEM_EUNSUPPORTED); every domain and integrator. */
typedef struct jdem_state_f64 {
  int64_t n;
  int32_t dim;
  int32_t bond_width;
  double* pos;
  double* vel;
  double* ang_vel;
  double* q;
  double* force;
  double* torque;
  const double* rad;
  const double* mass;
  const double* inertia;
  const int32_t* mat_id;
  const int32_t* bond_id;
  const uint8_t* fixed;
  const double* pos_p;
  const int32_t* clump_id;
} jdem_state_f64;

typedef struct jdem_domain_f64 {
  int32_t kind;
  int32_t dim;
  double box_size[3];
  double inv_box_size[3];
  double anchor[3];
  double restitution;
} jdem_domain_f64;

typedef struct jdem_law_f64 {
  int32_t kind;
  int32_t n_mat;
  const double* t0;
  const double* t1;
  const double* t2;
  const double* m0;
  const double* m1;
  const double* m2;
  const double* m3;
  double p0;
  double p1;
  int32_t n_species;
  int32_t n_terms;
  const int32_t* route_off;
  const struct jdem_law_f64* route_terms;
  const int32_t* species;
} jdem_law_f64;

/* _grid_params (_partition.py:55-99) computed by the host in float64. */
typedef struct jdem_grid_f64 {
  int32_t dims[3];
  int32_t strides[3];
  double cell_size[3];
  int32_t search_range;
  int32_t periodic;
  int64_t n_cells;
} jdem_grid_f64;

/* NeighborList buffers (neighbor_list.py:137-286), float64 positions. */
typedef struct jdem_nlist_f64 {
  int32_t max_neighbors;
  double cutoff;
  double skin;
  int32_t* nl;      /* (N, M) in/out, -1 padded                          */
  double* old_pos;  /* (N, D) in/out: positions at the last build        */
  int32_t n_build_times;
  int32_t overflow;
  int32_t secondary; /* JDEM_COLLIDER_CELL_LIST or _NAIVE                 */
  jdem_grid_f64 grid;
} jdem_nlist_f64;

typedef struct jdem_system_f64 {
  jdem_state_f64 state;
  jdem_domain_f64 domain;
  jdem_law_f64 law;
  int32_t lin_kind;
  int32_t rot_kind;
  double dt;
  int32_t interact_same_bond_id;
  double gravity[3];
  const double* ext_force;
  const double* ext_force_com;
  const double* ext_torque;
  int32_t collider_kind;
  jdem_grid_f64 grid;
  jdem_nlist_f64 nlist;
} jdem_system_f64;

/* System.step in float64 (system.py:701-748), as jdem_system_step.  For a
 * FREE domain the box and anchor after the last step are written to
 * sys->domain and stats; for NEIGHBOR_LIST sys->nlist is updated in place
 * (the list is checked after every step). */
int jdem_system_step_f64(jdem_ctx* ctx, jdem_system_f64* sys, int64_t n, jdem_step_stats* stats);
/* Collider.compute_force in float64 (colliders/__init__.py:56-80), as
 * jdem_compute_force (a NeighborList is checked and rebuilt first). */
int jdem_compute_force_f64(jdem_ctx* ctx, jdem_system_f64* sys);
/* Collider.compute_potential_energy in float64 (cell_list.py:469-496;
 * naive.py:76-117): float64 pair terms, float64 sum. */
int jdem_potential_energy_f64(jdem_ctx* ctx, jdem_system_f64* sys, double* energy);
/* Kinetic energy parts in float64 (utils/thermal.py:83-122), as
 * jdem_kinetic_energy_parts. */
int jdem_kinetic_energy_parts_f64(jdem_ctx* ctx, const jdem_system_f64* sys, double* translational,
                                  double* rotational);

#ifdef __cplusplus
}
#endif

#endif /* JDEM_H */
