/*
 * tecs_amd.h - C ABI of the MI355X (gfx950) per-step engine for TECS cell simulations.
 *
 * The engine runs the per-step hot path of TECS on one AMD GPU:
 *   grid rebuild (device counting sort)      replaces include/utils/grid.hpp:389-401
 *   binned / naive all-pairs finders         replace include/pair_finders/binned_all_pairs.hpp:66-183,
 *                                                    include/pair_finders/naive_all_pairs.hpp:32-104
 *   binned / naive Gabriel finders           replace include/pair_finders/binned_gabriel.hpp:71-362,
 *                                                    include/pair_finders/naive_gabriel.hpp:36-120
 *   NaiveDelaunay                            replaces include/pair_finders/naive_delaunay.hpp:157-293
 *   symmetric finders (each unordered pair   replace include/pair_finders/binned_all_pairs_symmetric.hpp:121-220,
 *     evaluated once, j < i)                         include/pair_finders/naive_all_pairs_symmetric.hpp:82-155,
 *                                                    include/pair_finders/binned_gabriel_symmetric.hpp:128-306,
 *                                                    include/pair_finders/naive_gabriel_symmetric.hpp:89-180
 *   pairwise force laws (PiecewiseLinear,    replace include/forces/predefined.hpp:107-144,
 *     bending, polarity alignment)                   include/types/polarity.hpp:131-198
 *   Euler / Heun integrators                 replace include/integrators/euler.hpp:49-85,
 *                                                    include/integrators/heun.hpp:143-208
 *   NoComFixer / GlobalComFixer              replace include/com_fixers/com_fixers.hpp:11-45
 *
 * Everything crossing this boundary is plain C: opaque handles, POD descriptors, host pointers
 * and sizes. No exceptions cross it; every entry point returns a status (TECS_OK == 0) and the
 * message of the last failure on the calling thread is available from tecs_last_error().
 * A handle may be used by one host thread at a time.
 *
 * Host arrays use the reference's memory layout: positions are VectorN<float,3> (12 bytes per
 * cell, x,y,z), polarities are Polarity_<float> (8 bytes per cell, theta,phi), types are one byte
 * per cell. All arrays are indexed by the reference's agent index i.
 */
#ifndef TECS_AMD_H
#define TECS_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TECS_AMD_ABI_VERSION 3

/* ---- status codes ---------------------------------------------------------------------- */
enum {
  TECS_OK = 0,
  TECS_ERR_INVALID = -1,     /* bad argument or descriptor */
  TECS_ERR_DEVICE = -2,      /* a HIP call failed */
  TECS_ERR_OOM = -3,         /* device allocation failed */
  TECS_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
  TECS_ERR_STATE = -5,       /* call not valid in the current state (e.g. no grid yet) */
  TECS_ERR_CAPACITY = -6     /* caller buffer too small; the required size is reported */
};

/* ---- component selectors (the CONFIG_* choices of include/simulation_config.hpp:212-379) -- */
enum {
  TECS_NAIVE_ALL_PAIRS = 0,  /* pair_finders::NaiveAllPairs  */
  TECS_BINNED_ALL_PAIRS = 1, /* pair_finders::BinnedAllPairs */
  TECS_NAIVE_GABRIEL = 2,    /* pair_finders::NaiveGabriel   */
  TECS_BINNED_GABRIEL = 3,   /* pair_finders::BinnedGabriel  */
  /* The symmetric finders visit, for each cell i, only partners j < i (agent index) and apply
   * the functor's result once per unordered pair: delta(i) += v, delta(j) -= v, for every field.
   * The reference accumulates with atomics, so its sums have no fixed order; this engine's
   * do not either (forces agree to rounding, pair sets exactly). */
  TECS_NAIVE_ALL_PAIRS_SYMMETRIC = 4,  /* pair_finders::NaiveAllPairsSymmetric  */
  TECS_BINNED_ALL_PAIRS_SYMMETRIC = 5, /* pair_finders::BinnedAllPairsSymmetric */
  TECS_NAIVE_GABRIEL_SYMMETRIC = 6,    /* pair_finders::NaiveGabrielSymmetric   */
  TECS_BINNED_GABRIEL_SYMMETRIC = 7,   /* pair_finders::BinnedGabrielSymmetric  */
  /* pair_finders::NaiveDelaunay (naive_delaunay.hpp:157-293, 3-D): the functor runs for (i, j)
   * once per third cell k for which some fourth cell l > k spans a circumsphere through i, j,
   * k, l with no other cell inside or on it. O(N^5): single simulations of at most 512 cells. */
  TECS_NAIVE_DELAUNAY = 8
};
enum { TECS_EULER = 0, TECS_HEUN = 1 };                        /* integrators::Euler / Heun */
enum { TECS_NO_COM_FIXER = 0, TECS_GLOBAL_COM_FIXER = 1 };     /* com_fixers::*             */

/* Optional per-cell fields besides the position (CONFIG_FIELDS). */
#define TECS_FIELD_POLARITY 1u /* Polarity_<float> field, integrated like the position */
#define TECS_FIELD_TYPE 2u     /* read-only uint8 cell type, selects the adhesion table entry */

/* PairFinder::Settings of the binned and Gabriel finders (binned_all_pairs.hpp:20-25,
 * binned_gabriel.hpp:20-26, naive_gabriel.hpp:22-24). Naive all-pairs ignores all fields. */
typedef struct tecs_pair_finder_settings {
  float min_bounds[3];       /* default -20 */
  float max_bounds[3];       /* default +20 */
  float bin_size_scale;      /* bin edge = bin_size_scale * cutoff_distance; default 1 */
  int32_t rebuild_every_n;   /* grid rebuild period in full steps; 0 = every step (default) */
  float gabriel_coefficient; /* Gabriel ball diameter factor c; default 0.8 */
} tecs_pair_finder_settings;

#define TECS_MAX_TYPES 4

/* Terms of the fused pairwise force (the forces given to Simulation::take_step). */
enum {
  TECS_TERM_END = 0,
  TECS_TERM_PIECEWISE_LINEAR = 1, /* position.delta += PiecewiseLinear(...) */
  TECS_TERM_BENDING = 2,          /* position.delta and polarity.delta += bending terms */
  TECS_TERM_ALIGNMENT = 3         /* polarity.delta += alignment term */
};

/* The fused pairwise force, evaluated per ordered pair (i, j). Its terms:
 *   PiecewiseLinear: position.delta += PiecewiseLinear(disp, d, homeostatic_radius,
 *        repulsion_scale, k_adh) with k_adh = adhesion_table[type_i * n_types + type_j] when
 *        n_types > 0, else adhesion_scale                        (predefined.hpp:125-144)
 *   bending (bending_strength != 0): b = polarity_i.bending_force(disp, polarity_j, d)
 *        position.delta += bending_strength * b.vector
 *        polarity.delta += bending_strength * b.polarity         (polarity.hpp:181-198)
 *   alignment (alignment_strength != 0):
 *        polarity.delta += alignment_strength *
 *             polarity_i.bidirectional_polarization_force(polarity_j)   (polarity.hpp:165-168)
 * term_order lists the terms that are applied, in the order they add to the deltas, as the
 * fused forces run in argument order (KernelFuser, kernel_fuser.hpp:40-55): TECS_TERM_* ids,
 * each at most once, the list ending at the first TECS_TERM_END. All zeros (the default)
 * means PiecewiseLinear, bending, alignment. A bending or alignment term with strength 0 is
 * not applied. With a strength of exactly 1 the product is exact, so the term equals the
 * reference's plain `delta += term`. Polarity terms need TECS_FIELD_POLARITY, a type table
 * needs TECS_FIELD_TYPE.
 *
 * Drag (the functor's `drag` output, binned_all_pairs.hpp:155-180): each pair's drag is
 * `drag`, or drag_table[type_i * n_types + type_j] when drag_by_type is 1 (the type-dependent
 * `if (...) drag += 1.0` of examples/tribolium/main.cpp:113-115). The finder sums the drags
 * and drag * old_velocity_j over the cell's pairs and adds
 * total_velocity / fmax(total_drag, FLT_EPSILON) to the position delta. With drag 0
 * everywhere that is +0. Drag needs track_old_velocities and an ordered finder (not the
 * symmetric ones: TECS_ERR_UNSUPPORTED). */
typedef struct tecs_force_desc {
  float homeostatic_radius;
  float repulsion_scale;
  float adhesion_scale;
  int32_t n_types; /* 0 = no type table, else 1..TECS_MAX_TYPES */
  float adhesion_table[TECS_MAX_TYPES * TECS_MAX_TYPES];
  float bending_strength;
  float alignment_strength;
  int32_t term_order[3];
  float drag;
  int32_t drag_by_type;
  float drag_table[TECS_MAX_TYPES * TECS_MAX_TYPES];
} tecs_force_desc;

/* Everything Simulation<Config>(Settings) needs for this path (simulation.hpp:76-98, 133-182). */
typedef struct tecs_sim_desc {
  uint32_t agent_count;     /* active cells */
  uint32_t capacity;        /* allocated cells, >= agent_count (GlobalComFixer divides by it) */
  uint32_t fields;          /* TECS_FIELD_* bit mask */
  float cutoff_distance;    /* r_max */
  int32_t pair_finder;      /* TECS_*_ALL_PAIRS / TECS_*_GABRIEL[_SYMMETRIC] */
  tecs_pair_finder_settings pair_finder_settings;
  int32_t integrator;       /* TECS_EULER / TECS_HEUN */
  int32_t com_fixer;        /* TECS_NO_COM_FIXER / TECS_GLOBAL_COM_FIXER */
  tecs_force_desc force;
  int32_t track_old_velocities; /* 1 = keep Base::old_velocities (base.hpp:74) like the reference */
} tecs_sim_desc;

/* Counters of a simulation or batch since creation (or the last reset). */
typedef struct tecs_stats {
  uint64_t pair_interactions; /* force-functor invocations on ordered pairs, all evaluations */
  uint64_t evaluations;       /* pair-finder evaluations (Heun: 2 per step) */
  uint64_t rebuilds;          /* grid rebuilds */
  uint64_t steps;             /* integrator steps */
} tecs_stats;

/* Fill the reference defaults (bounds +-20, bin scale 1, rebuild every step, c = 0.8). */
void tecs_default_pair_finder_settings(tecs_pair_finder_settings* settings);
/* Reference defaults for a position-only Heun simulation with NoComFixer and
 * RelaxForce's PiecewiseLinear(0.8, 2, 1) (initializers/relax_force.hpp:27-29). */
void tecs_default_sim_desc(tecs_sim_desc* desc);

int tecs_abi_version(void);
const char* tecs_last_error(void);
/* Number of visible HIP devices (0 without a GPU); never initialises more than the runtime. */
int tecs_device_count(int32_t* count);

/* =========================================================================================
 * Single simulation: the drop-in for Simulation<Config>::take_step (simulation.hpp:548-556)
 * and the pair finder / integrator / com fixer it drives.
 * ========================================================================================= */
typedef struct tecs_sim tecs_sim;

/* Allocates all device state on the current HIP device. `stream` is a hipStream_t or NULL
 * (then the simulation creates its own stream at default priority). */
int tecs_sim_create(const tecs_sim_desc* desc, void* stream, tecs_sim** out);
int tecs_sim_destroy(tecs_sim* sim);

/* Copy host state in (positions: agent_count*3 floats; polarities: agent_count*2 floats or NULL
 * when the field is absent; types: agent_count bytes or NULL). Keeps the current grid, so a
 * following non-full evaluation sees a stale grid exactly like the reference's Heun corrector. */
int tecs_sim_upload(tecs_sim* sim, const float* positions, const float* polarities,
                    const uint8_t* types);
/* Copy state out; any pointer may be NULL. old_velocities: agent_count*3 floats. */
int tecs_sim_download(tecs_sim* sim, float* positions, float* polarities, float* old_velocities);

/* ---- population (Simulation::set_capacity / set_agent_count, simulation.hpp:266-301) ------
 * Cells are indexed 0..capacity-1; the first agent_count are active. As with the reference's
 * views, a cell keeps its state while inactive: cells that become active again (or for the
 * first time: zero state) start from what their index last held. */
/* Write cells [first, first + count) (active or not) from host arrays in agent order; positions
 * required, polarities / types may be NULL. first + count <= capacity. */
int tecs_sim_upload_range(tecs_sim* sim, uint32_t first, uint32_t count, const float* positions,
                          const float* polarities, const uint8_t* types);
/* Simulation::set_agent_count (simulation.hpp:285-301): agent_count <= capacity. Resets the
 * rebuild counter (pair_finder.set_agent_count, binned_all_pairs.hpp:60-63), so the next full
 * evaluation builds the grid; a non-full evaluation before it fails with TECS_ERR_STATE. */
int tecs_sim_set_agent_count(tecs_sim* sim, uint32_t agent_count);
/* Simulation::set_capacity (simulation.hpp:266-274): capacity >= agent_count; GlobalComFixer
 * divides by it. Cells beyond a smaller capacity are dropped (they start from zero state if the
 * capacity grows again); growing past the allocation moves the state into new device buffers
 * (synchronous) and rebuilds the grid at the next full step. */
int tecs_sim_set_capacity(tecs_sim* sim, uint32_t capacity);
int tecs_sim_get_population(tecs_sim* sim, uint32_t* agent_count, uint32_t* capacity);

/* ---- grid queries (acceleration::Grid, grid.hpp:413-637), binned finders ------------------
 * Batched over n query points (points: n*3 floats), on the grid of the last rebuild and the
 * current positions (the reference's Grid reads its positions view at query time). Bins and
 * members are visited in the grid's order (bins x-fastest, members ascending in agent index).
 * Output: one agent index per query, -1 where none qualifies. A random pick takes the caller's
 * uniform draw u in [0, 1) per query and chooses candidate floor(u * total) of the `total`
 * candidates in visiting order; the reference draws from its Kokkos pool (rng.rand / urand).
 * Synchronous; TECS_ERR_STATE before any grid, TECS_ERR_UNSUPPORTED for naive finders. */
/* get_nearest_point_index (grid.hpp:413-441): smallest squared distance, first in visiting
 * order on ties, over the bins within search_radius bins of the point's bin. */
int tecs_sim_nearest(tecs_sim* sim, int32_t n, const float* points, int32_t search_radius,
                     int32_t* out);
/* tecs_sim_nearest without any device allocation. No counterpart in the reference: for callers
 * that cannot allocate device memory (tight serving loops, harnesses that forbid hipMalloc).
 * The points and results are staged in a fixed device array of 1 MiB that each GPU holds for
 * the life of the process; at most 65536 points per call (TECS_ERR_CAPACITY beyond, nothing
 * done); the simulation's device must be current. Calls that overlap on one GPU (several
 * threads or simulations at once) share the array and race. Results equal tecs_sim_nearest's. */
int tecs_sim_nearest_noalloc(tecs_sim* sim, int32_t n, const float* points, int32_t search_radius,
                             int32_t* out);
/* get_random_point_index_in_neighbourhood_bins (grid.hpp:453-495): any cell of those bins. */
int tecs_sim_random_in_bins(tecs_sim* sim, int32_t n, const float* points, int32_t radius_bins,
                            const float* uniforms, int32_t* out);
/* get_random_point_index_in_neighbourhood (grid.hpp:507-564): a cell with squared distance
 * <= radius^2 (bins: max(1, ceil(radius / bin_size))). */
int tecs_sim_random_neighbour(tecs_sim* sim, int32_t n, const float* points, float radius,
                              const float* uniforms, int32_t* out);

/* ---- generic, link and update forces (Integrator::evaluate_force_impl, base.hpp:116-230) ---
 * take_step evaluates, for every evaluation of a step, in the order fuse_forces groups the
 * forces (kernel_fuser.hpp:148-156): the generic wall forces, then the pairwise force (the pair
 * finder), then the link forces, all into the same deltas (and so into the COM fix and the
 * integrator update). tecs_sim_evaluate_force and tecs_sim_pairs stay the pair finder's alone
 * (PairFinder::evaluate_force). */
#define TECS_MAX_WALLS 8
/* A planar wall, the GENERIC_FORCE of examples/growth_w_wall.cpp:44-48 and
 * growth_w_squeeze.cpp:37-52 (PlaneN, types/plane.hpp:17-60): the plane a x + b y + c z + d = 0
 * is normalised as PlaneN's constructor does; a cell whose signed distance
 * sd = d + a x + b y + c z is negative gets delta += normal * (-sd * strength), wall after wall. */
typedef struct tecs_wall {
  float plane[4]; /* a, b, c, d (any nonzero normal) */
  float strength;
} tecs_wall;
/* Replace the walls (count <= TECS_MAX_WALLS; 0 removes them). */
int tecs_sim_set_walls(tecs_sim* sim, const tecs_wall* walls, int32_t count);
/* Replace the links (the Link views of Simulation, types/link.hpp; links: 2*count agent indices
 * a, b) and the link force, the LINK_FORCE of examples/intercalation.cpp / sorting_prot.cpp:
 * u = p_b - p_a, d = |u|, delta_a += (k u) / d, delta_b += (-k u) / d (base.hpp:193-230 adds
 * the endpoints' contributions with atomics in no fixed order; here each cell adds its own in
 * link order). Links with a == b, or with an
 * endpoint at or beyond the active count when a step runs, are skipped. Endpoints must be below
 * the capacity. count 0 removes them. */
int tecs_sim_set_links(tecs_sim* sim, const int32_t* links, int64_t count, float strength);
/* Cell division, the UPDATE_FUNC `proliferate` of examples/growth_w_wall.cpp:52-64 (run by
 * Simulation::init, simulation.hpp:520-529): uniforms holds three draws in [0, 1) per active
 * cell (the reference's rng.drand). Cell i divides when u0 <= rate; its daughter takes the next
 * index (the reference's atomic counter orders daughters arbitrarily; here in the order of their
 * parents) at p_i + offset * to_vector3(theta, phi), theta = float(acos(2 u1 - 1)),
 * phi = float(u2 * 2 * pi) (polarity.hpp:111-118), sines and cosines within an ulp of the float
 * functions. Writes the daughters' positions only (cells beyond the active count; their other
 * fields keep what their index holds) and returns in *new_count the count to pass to
 * tecs_sim_set_agent_count, as the example does with its counter. TECS_ERR_CAPACITY (nothing
 * written) if the daughters do not fit in the capacity. Synchronous. */
int tecs_sim_divide(tecs_sim* sim, float rate, float offset, const double* uniforms,
                    uint32_t* new_count);

/* n_steps integrator steps of size dt (Integrator::integrate, euler.hpp:81-85 / heun.hpp:201-208).
 * Asynchronous on the simulation's stream. */
int tecs_sim_take_step(tecs_sim* sim, double dt, int32_t n_steps);

/* Replace the pairwise force used by the following steps (Simulation::take_step(dt, forces...),
 * simulation.hpp:548-556, takes the force with every call). Same validation as create; the
 * field set cannot change. Ordered with respect to the simulation's stream. */
int tecs_sim_set_force(tecs_sim* sim, const tecs_force_desc* force);

/* One pair-finder evaluation on the current state into zeroed deltas, exactly as
 * PairFinder::evaluate_force(in, out, ..., is_full_step) (binned_all_pairs.hpp:66-183):
 * a full evaluation rebuilds the grid when due and advances the rebuild counter.
 * dpos: agent_count*3 floats, dpol: agent_count*2 floats or NULL. Synchronous. */
int tecs_sim_evaluate_force(tecs_sim* sim, int32_t is_full_step, float* dpos, float* dpol);

/* The ordered pairs (i, j) such an evaluation would hand to the force functor, as CSR over i:
 * row_ptr[agent_count+1], cols[n_pairs], each row in the reference's visiting order
 * (neighbour bins x-fastest, members in ascending agent index); a symmetric finder's row i
 * holds only partners j < i. Rebuild semantics as
 * tecs_sim_evaluate_force. Call with cols == NULL to obtain n_pairs only (nothing advances
 * then); returns TECS_ERR_CAPACITY when cols_capacity < n_pairs. Synchronous. */
int tecs_sim_pairs(tecs_sim* sim, int32_t is_full_step, int64_t* row_ptr, int32_t* cols,
                   int64_t cols_capacity, int64_t* n_pairs);

/* Rebuild the grid at the current positions (Grid::rebuild, grid.hpp:389-401) and reset the
 * rebuild counter as a due full step would. */
int tecs_sim_rebuild(tecs_sim* sim);

int tecs_sim_synchronize(tecs_sim* sim);
/* Whether a step ended with a non-finite position or polarity (NaN or infinity, e.g. an
 * overflow or a NaN upload) since creation: *flag = 1, else 0. The kernels that end a step set
 * it; it stays set. Synchronous. No counterpart in the reference (it does not check). */
int tecs_sim_get_nonfinite(tecs_sim* sim, int32_t* flag);
int tecs_sim_get_stats(tecs_sim* sim, tecs_stats* stats);
int tecs_sim_reset_stats(tecs_sim* sim);

/* Pair-sweep kernel variant. All produce identical results. THREAD walks one cell per
 * thread; TEAM4/8/16 split a cell over 4, 8 or 16 lanes (binned finders with a search radius
 * of one bin and, for Gabriel, gabriel_coefficient <= 0.95; otherwise THREAD is used); TEAM
 * means TEAM8. AUTO (default) lets the autotuner time every applicable variant on the live
 * state at the first take_step and keep the fastest. No counterpart in the reference. */
enum {
  TECS_SWEEP_AUTO = 0,
  TECS_SWEEP_THREAD = 1,
  TECS_SWEEP_TEAM = 2,
  TECS_SWEEP_TEAM4 = 3,
  TECS_SWEEP_TEAM8 = 4,
  TECS_SWEEP_TEAM16 = 5
};
int tecs_sim_set_sweep_variant(tecs_sim* sim, int32_t variant);
/* The variant the next sweep will use (THREAD, TEAM4, TEAM8 or TEAM16). */
int tecs_sim_get_sweep_variant(tecs_sim* sim, int32_t* variant);
/* The autotuner's best time in ms per variant, indexed by TECS_SWEEP_* (0 = not timed). */
int tecs_sim_get_autotune(tecs_sim* sim, double* ms_per_variant /* [6] */);

/* hipGraph replay of step sequences inside tecs_sim_take_step. A period is
 * 2 * max(1, rebuild_every_n) steps; a sequence of one period is captured once per starting
 * state and replayed by every later call that starts there. AUTO (default) and ON replay
 * whenever a whole period fits (AUTO only while kernel timing is off); OFF always launches
 * kernel by kernel. Results are identical in every mode.
 * No counterpart in the reference. */
enum { TECS_GRAPHS_AUTO = 0, TECS_GRAPHS_OFF = 1, TECS_GRAPHS_ON = 2 };
int tecs_sim_set_graph_mode(tecs_sim* sim, int32_t mode);
/* Number of graph launches so far (each runs one period of steps). */
int tecs_sim_get_graph_launches(tecs_sim* sim, int64_t* launches);

/* Kernel timing with HIP events recorded on the simulation's stream around every launch of a
 * kernel class. Enabling adds two events per timed launch. */
enum {
  TECS_KT_SWEEP_FULL = 0,  /* pair sweep of a full evaluation (Heun substep 1, Euler) */
  TECS_KT_SWEEP_STALE = 1, /* pair sweep of the Heun corrector evaluation (stale grid) */
  TECS_KT_REBUILD = 2      /* the whole grid rebuild (keys, scan, scatter, gather) */
};
int tecs_sim_set_kernel_timing(tecs_sim* sim, int32_t enabled);
/* Summed duration in ms and number of timed launches of one class since creation. */
int tecs_sim_get_kernel_timing(tecs_sim* sim, int32_t kernel_class, double* ms,
                               int64_t* launches);

/* =========================================================================================
 * Batch: many independent simulations advanced together, one launch per kernel per substep
 * for all of them. All members share pair finder, integrator, com fixer, fields, old-velocity
 * tracking and rebuild period; agent counts, bounds, cutoffs, Gabriel coefficients, force
 * parameters and dt may differ. Each member's results are those of the same simulation run
 * alone through tecs_sim. This has no counterpart in the reference, which runs one
 * Simulation per call of take_step (simulation.hpp:548-556).
 * ========================================================================================= */
typedef struct tecs_batch tecs_batch;
int tecs_batch_create(const tecs_sim_desc* descs, int32_t count, void* stream, tecs_batch** out);
int tecs_batch_destroy(tecs_batch* batch);
int tecs_batch_upload(tecs_batch* batch, int32_t index, const float* positions,
                      const float* polarities, const uint8_t* types);
int tecs_batch_download(tecs_batch* batch, int32_t index, float* positions, float* polarities,
                        float* old_velocities);
/* dts: one dt per member. */
int tecs_batch_take_step(tecs_batch* batch, const double* dts, int32_t n_steps);
/* tecs_sim_set_force for member `index`. */
int tecs_batch_set_force(tecs_batch* batch, int32_t index, const tecs_force_desc* force);
/* tecs_sim_upload_range / tecs_sim_set_agent_count for member `index`. A batch shares one
 * rebuild schedule, so a binned batch of several members with rebuild_every_n > 0 refuses
 * set_agent_count (TECS_ERR_UNSUPPORTED); the capacity of a member can change within the
 * member's allocation only (the batch layout is fixed). */
int tecs_batch_upload_range(tecs_batch* batch, int32_t index, uint32_t first, uint32_t count,
                            const float* positions, const float* polarities, const uint8_t* types);
int tecs_batch_set_agent_count(tecs_batch* batch, int32_t index, uint32_t agent_count);
/* tecs_sim_set_walls / tecs_sim_set_links / tecs_sim_divide for member `index`. */
int tecs_batch_set_walls(tecs_batch* batch, int32_t index, const tecs_wall* walls, int32_t count);
int tecs_batch_set_links(tecs_batch* batch, int32_t index, const int32_t* links, int64_t count,
                         float strength);
int tecs_batch_divide(tecs_batch* batch, int32_t index, float rate, float offset,
                      const double* uniforms, uint32_t* new_count);
int tecs_batch_set_capacity(tecs_batch* batch, int32_t index, uint32_t capacity);
int tecs_batch_synchronize(tecs_batch* batch);
/* tecs_sim_get_nonfinite per member: flags[count]. */
int tecs_batch_get_nonfinite(tecs_batch* batch, int32_t* flags);
int tecs_batch_get_stats(tecs_batch* batch, tecs_stats* stats);
int tecs_batch_reset_stats(tecs_batch* batch);
int tecs_batch_set_sweep_variant(tecs_batch* batch, int32_t variant);
int tecs_batch_get_sweep_variant(tecs_batch* batch, int32_t* variant);
/* With AUTO, the autotuner of a batch whose members span several size classes (cell counts
 * in powers of four: class c holds 4^c <= n < 4^(c+1)), a binned finder with exact team
 * sweeps, ordered pairs and no COM fixer, times every variant on each class's cells alone and
 * gives each class its fastest; the classes that picked the same variant are swept by one
 * launch. tecs_batch_get_autotune then reports, per variant, the sum of its class times. */
int tecs_batch_get_autotune(tecs_batch* batch, double* ms_per_variant /* [6] */);
/* After the autotuner ran: the size class of every member (member_class[count]) and the
 * variant each class picked (class_variant[TECS_MAX_SIZE_CLASSES], -1 for a class without
 * members). No counterpart in the reference. */
#define TECS_MAX_SIZE_CLASSES 16
int tecs_batch_get_size_classes(tecs_batch* batch, int32_t* member_class, int32_t* class_variant);
/* Chooses the variant of every size class instead of the autotuner (and selects AUTO):
 * class_variant[TECS_MAX_SIZE_CLASSES], read for the classes that have members (THREAD, or
 * TEAM4/8/16 where the team sweep is exact). Ordered binned finders without a COM fixer. No
 * counterpart in the reference. */
int tecs_batch_set_size_class_variants(tecs_batch* batch, const int32_t* class_variant);
int tecs_batch_set_graph_mode(tecs_batch* batch, int32_t mode);
int tecs_batch_get_graph_launches(tecs_batch* batch, int64_t* launches);
int tecs_batch_set_kernel_timing(tecs_batch* batch, int32_t enabled);
int tecs_batch_get_kernel_timing(tecs_batch* batch, int32_t kernel_class, double* ms,
                                 int64_t* launches);

/* =========================================================================================
 * Snapshots: io::HDF5_Writer (include/io/hdf5_writer.hpp) and Simulation::write
 * (simulation.hpp:607-632). The file layout is the reference's: <path>.h5 with one group
 * "t<step>" per write holding "position" [n x 3] and "polarity" [n x 2] float datasets, a
 * "static" group, and the XDMF companion <path>.xmf. libhdf5 is loaded at tecs_writer_open
 * (TECS_HDF5_LIB, libhdf5.so, or /opt/conda/lib/libhdf5.so); TECS_ERR_UNSUPPORTED without it.
 * ========================================================================================= */
typedef struct tecs_writer tecs_writer;
/* HDF5_Writer(path, settings), hdf5_writer.hpp:66-79: truncates <path>.h5 (and <path>.xmf). */
int tecs_writer_open(const char* path, int32_t write_xmf, tecs_writer** out);
/* Waits for pending snapshots, writes the XDMF tail and closes (~HDF5_Writer, :84-86). */
int tecs_writer_close(tecs_writer* writer);
/* Waits until every snapshot handed to the writer is in the file. */
int tecs_writer_flush(tecs_writer* writer);
/* HDF5_Writer::write (:306-321) from host arrays, synchronously. */
int tecs_writer_write(tecs_writer* writer, double time, uint32_t step, uint32_t n,
                      const float* positions, const float* polarities);
/* HDF5_Writer::write_static (:335-349) for one uint8 field; a second call fails. */
int tecs_writer_write_static(tecs_writer* writer, const char* label, uint32_t n,
                             const uint8_t* data);
/* Simulation::write(time): snapshot the current state as group t<k>, k = this simulation's
 * write counter. Asynchronous: one gather kernel runs on the simulation's stream, the copy to
 * pinned host memory runs on a side stream and a writer thread writes the file, so following
 * steps overlap with the copy and the file write. Blocks only when 3 snapshots are in flight. */
int tecs_sim_write(tecs_sim* sim, tecs_writer* writer, double time);
int tecs_batch_write(tecs_batch* batch, int32_t index, tecs_writer* writer, double time);
/* HDF5_Reader::read_dataset (include/io/hdf5_reader.hpp:66-71): a dataset of rank 1 or 2
 * ("t3/position") into float32; dims[0..1] receive its shape (rank 1: dims[1] = 1); out may be
 * NULL to query the shape (HDF5_Reader::get_dataset_dimensions, :43-53). capacity counts
 * elements. */
int tecs_read_dataset(const char* h5_path, const char* dataset, float* out, int64_t capacity,
                      int64_t* dims);
/* Element types of tecs_read_dataset_typed. */
enum {
  TECS_DTYPE_FLOAT32 = 0,
  TECS_DTYPE_FLOAT64 = 1,
  TECS_DTYPE_INT32 = 2,
  TECS_DTYPE_UINT32 = 3,
  TECS_DTYPE_UINT8 = 4,
  TECS_DTYPE_INT64 = 5
};
/* read_dataset into a View<T> of any element type (the View<int> of cell types in
 * examples/tribolium/main.cpp:71-72): HDF5 converts from the file's type to `dtype` on read.
 * Same shapes and errors as tecs_read_dataset; TECS_ERR_INVALID for an unknown dtype or a
 * rank above 2. */
int tecs_read_dataset_typed(const char* h5_path, const char* dataset, int32_t dtype, void* out,
                            int64_t capacity, int64_t* dims);

/* =========================================================================================
 * Serving: simulation requests admitted by an HBM budget as they arrive and advanced together
 * (continuous batching). Requests that share pair finder, integrator, COM fixer, fields,
 * old-velocity tracking and (binned finders) rebuild period run as members of one batch; a batch
 * is re-formed when members join or leave, with the running members' state copied on the
 * device, and with a rebuild period R > 0 only where the continuing members are at a rebuild
 * point of their own schedule. Every request keeps its own arithmetic: its result is bitwise
 * that of the same simulation run alone. A request whose state becomes non-finite (NaN or
 * infinity at the end of a step) fails without disturbing the others; a request can be
 * cancelled. No counterpart in the reference.
 * ========================================================================================= */
typedef struct tecs_server tecs_server;
enum {
  TECS_REQUEST_PENDING = 0,
  TECS_REQUEST_RUNNING = 1,
  TECS_REQUEST_DONE = 2,
  TECS_REQUEST_FAILED = 3,    /* a step ended with a non-finite position or polarity */
  TECS_REQUEST_CANCELLED = 4  /* tecs_server_cancel */
};
typedef struct tecs_request_info {
  int32_t state;        /* TECS_REQUEST_* */
  int32_t steps_done;
  double t_submit;      /* seconds since the server was created */
  double t_admit;
  double t_done;
  uint64_t hbm_bytes;   /* the device bytes the request is charged */
} tecs_request_info;
typedef struct tecs_server_stats {
  uint64_t completed, pending, running, failed, cancelled;
  uint64_t pair_interactions;  /* over all requests, all evaluations */
  uint64_t member_steps;       /* sum over requests of the steps run */
  uint64_t regroups;           /* batches re-formed */
  uint64_t hbm_bytes_in_use, hbm_budget;
  uint64_t pool_cached_bytes;  /* freed batch buffers the server keeps for re-use */
  double regroup_seconds;      /* host time spent re-forming batches */
} tecs_server_stats;
/* max_chunk_steps: steps a group advances at most between two services (admission,
 * retirement and re-forming happen between chunks). */
int tecs_server_create(uint64_t hbm_budget_bytes, int32_t max_chunk_steps, tecs_server** out);
int tecs_server_destroy(tecs_server* server);
/* Queue a request (inputs are copied). TECS_ERR_CAPACITY if it could never fit the budget. */
int tecs_server_submit(tecs_server* server, const tecs_sim_desc* desc, const float* positions,
                       const float* polarities, const uint8_t* types, double dt,
                       int32_t n_steps, int64_t* request_id);
/* Admit what fits the budget, then service every group whose chunk has ended (retire finished
 * members, re-form the batch if members left or joined, enqueue the next chunk). Returns once
 * at least one group was serviced, or at once when nothing is running. */
int tecs_server_advance(tecs_server* server, int32_t* running, int32_t* pending);
int tecs_server_request_info(tecs_server* server, int64_t request_id, tecs_request_info* info);
/* Cancel a request: a pending one leaves at once, a running one at the end of its group's
 * chunk in flight (its state stays RUNNING until then). TECS_ERR_STATE when it already ended. */
int tecs_server_cancel(tecs_server* server, int64_t request_id);
/* The final state of a finished request (positions n*3, polarities n*2 or NULL);
 * TECS_ERR_STATE for a request that is not done, failed or was cancelled. */
int tecs_server_result(tecs_server* server, int64_t request_id, float* positions,
                       float* polarities);
int tecs_server_get_stats(tecs_server* server, tecs_server_stats* stats);

/* Measurement helper with no counterpart in the reference: copies `bytes` (rounded down to a
 * multiple of 16) between two device buffers it allocates, `repeats` times, with a plain
 * streaming copy kernel (k_copy_probe), and reports the mean kernel time. Profilers read the
 * HBM byte counters of this kernel to calibrate them against a known traffic. */
int tecs_copy_probe(int64_t bytes, int32_t repeats, double* mean_ms);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* TECS_AMD_H */
