/*
 * cosmosgl.h — C ABI of the MI355X force-simulation tick.
 *
 * One context = a set of simulated graphs on one GPU, bound to one HIP
 * stream.  The single-graph entry points (cg_set_positions ... cg_get_positions)
 * act on graph 0 of the set; the cg_graph_* entry points add further graphs,
 * which are packed into the same device arrays so that one launch per force
 * pass advances all of them (batched layouts, continuous batching).  A graph
 * computes the same bits alone and packed with others.
 * A context is not re-entrant; separate contexts may be driven from separate
 * threads.  Inputs are copied on call, outputs are caller-allocated, the
 * context owns all device memory.  No function throws; each returns CG_OK or
 * a negative code, and cg_last_error() holds the message of the last failure.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * cosmos.gl 3.4.1 source tree):
 *   cg_create          Graph constructor + module construction, src/index.ts:422-434
 *   cg_set_positions   Graph.setPointPositions src/index.ts:548 -> Points.updatePositions
 *                      Points/index.ts:390-483, updateExit :1313-1386
 *   cg_set_links       Graph.setLinks src/index.ts:723 + setLinkStrength :836 ->
 *                      GraphData._createAdjacencyLists/_calculateDegrees
 *                      GraphData/index.ts:525-568 -> ForceLink.create x2 ForceLink/index.ts:40-171
 *   cg_set_jitter      ForceManyBody.create random texture, ForceManyBody/index.ts:132-163
 *   cg_set_point_sizes Graph.setPointSizes src/index.ts:627 (collision radii,
 *                      ForceCollision/index.ts:121-148)
 *   cg_set_pinned      Graph.setPinnedPoints src/index.ts:912 -> Points.updatePinnedStatus
 *   cg_set_point_clusters / cg_set_cluster_positions / cg_set_point_cluster_strength
 *                      Graph.setPointClusters / setClusterPositions / setPointClusterStrength
 *                      src/index.ts:854, 871, 888 -> Clusters.create
 *                      Points/index.ts:1264-1304
 *   cg_set_alpha       Graph.start(alpha) / render(alpha) src/index.ts:1457, 936
 *   cg_step            Graph.runSimulationStep src/index.ts:2006-2099 (nticks times)
 *   cg_get_positions   Graph.getPointPositions src/index.ts:1106-1130 (readPixels)
 *   cg_drag_points     Points.drag Points/index.ts:1961-1988 (called from src/index.ts:2303-2313)
 *   cg_get_positions_at  getTrackedPointPositionsArray src/index.ts:1346-1388
 *   cg_get_alpha       Store.alpha, reported through onSimulationTick src/index.ts:2090
 */
#ifndef COSMOSGL_H
#define COSMOSGL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CG_OK 0
#define CG_ERR_INVALID (-1)   /* bad argument */
#define CG_ERR_DEVICE (-2)    /* HIP runtime failure */
#define CG_ERR_BUDGET (-3)    /* would exceed the context's HBM budget */
#define CG_ERR_STATE (-4)     /* call out of order (e.g. step before positions) */

typedef struct cg_ctx cg_ctx;

typedef struct cg_create_params {
    int32_t device;        /* HIP device ordinal */
    uint64_t hbm_budget;   /* max device bytes this context may hold; 0 = no limit */
} cg_create_params;

/* Mirror of the reference's `simulation*` config keys (src/config.ts:368-456,
 * defaults src/variables.ts:68-83) plus the tick gates of runSimulationStep.
 * Parameters are re-read on every cg_step, so a change takes effect on the next
 * tick as in the reference. */
typedef struct cg_sim_params {
    float space_size;                   /* Store.adjustedSpaceSize (Store/index.ts:209-233) */
    float decay;                        /* simulationDecay */
    float gravity;                      /* simulationGravity (0 = pass skipped) */
    float center;                       /* simulationCenter (0 = pass skipped) */
    float repulsion;                    /* simulationRepulsion */
    float link_spring;                  /* simulationLinkSpring */
    float link_distance;                /* simulationLinkDistance */
    float link_dist_random_variation_range[2]; /* simulationLinkDistRandomVariationRange */
    float repulsion_from_mouse;         /* simulationRepulsionFromMouse */
    float friction;                     /* simulationFriction */
    float collision;                    /* simulationCollision (0 = pass skipped) */
    float collision_radius;             /* simulationCollisionRadius; <= 0 = per-point size/2 */
    float collision_padding;            /* simulationCollisionPadding */
    float point_default_size;           /* pointDefaultSize (size of points without one) */
    int32_t enable_simulation;          /* enableSimulation: 0 makes cg_step a no-op */
    int32_t run_forces;                 /* forceExecution || (running && !zoom gate) */
    int32_t mouse_repulsion;            /* right button held && enableRightClickRepulsion */
    float mouse_position[2];            /* Store.mousePosition, space coordinates */
    float cluster;                      /* simulationCluster (cluster pull coefficient) */
} cg_sim_params;

/* Fills the reference defaults (src/variables.ts:10-19, 68-83: spaceSize 4096). */
void cg_default_params(cg_sim_params* p);

int cg_create(cg_ctx** out, const cg_create_params* params);
int cg_destroy(cg_ctx* ctx);

/* xy = [x0, y0, x1, y1, ...], n points.  A point with a NaN coordinate is
 * absent (helper.ts:157-159): it is never binned, never integrated and reads
 * back as NaN.  Resets jitter, sizes and pinned state when n changes, and
 * rebuilds the link CSRs from the last cg_set_links input. */
int cg_set_positions(cg_ctx* ctx, const float* xy, uint32_t n);

/* links = [s0, t0, s1, t1, ...] as the reference's Float32Array, m links.
 * Endpoints that are not integers in [0, n) invalidate the link, which keeps
 * its index (GraphData/index.ts:537-547).  strength: m per-link strengths or
 * NULL (degree-derived default, ForceLink/index.ts:66-73).  rand_in / rand_out:
 * m U(0,1) draws, rand_in[l] for link l's entry in the INCOMING instance and
 * rand_out[l] for its OUTGOING entry (ForceLink/index.ts:76); NULL = drawn
 * from the context's own generator. */
int cg_set_links(cg_ctx* ctx, const float* links, uint32_t m, const float* strength,
                 const float* rand_in, const float* rand_out);

/* Per-point many-body jitter r_i = (U(-1,1)*1e-5, U(-1,1)*1e-5), n pairs. */
int cg_set_jitter(cg_ctx* ctx, const float* xy, uint32_t n);

/* Per-point sizes for collision (NaN = pointDefaultSize); NULL clears. */
int cg_set_point_sizes(cg_ctx* ctx, const float* sizes, uint32_t n);

/* Pinned point indices (indices outside [0, n) are ignored); count 0 clears. */
int cg_set_pinned(cg_ctx* ctx, const uint32_t* indices, uint32_t count);

/* Clusters — Graph.setPointClusters src/index.ts:854, setClusterPositions :871,
 * setPointClusterStrength :888 -> Clusters.create (Clusters/index.ts:54-160).
 * clusters: one index per point, negative = no cluster (the reference's
 * undefined); NULL clears.  positions: [x, y] per cluster, a negative
 * coordinate = pull toward the cluster's centroid; NULL clears.  strength: one
 * coefficient per point (default 1); NULL clears.  The cluster pass runs while
 * clusters or cluster positions are set (src/index.ts:2057-2061). */
int cg_set_point_clusters(cg_ctx* ctx, const int32_t* clusters, uint32_t n);
int cg_set_cluster_positions(cg_ctx* ctx, const float* xy, uint32_t count);
int cg_set_point_cluster_strength(cg_ctx* ctx, const float* strength, uint32_t n);

int cg_set_alpha(cg_ctx* ctx, double alpha);
int cg_get_alpha(cg_ctx* ctx, double* alpha);

/* Runs nticks ticks.  seeds: nticks per-tick near-field seeds in [0,1) (the
 * reference's store.getRandomFloat(0, 1) per tick) or NULL for the context's
 * own generator.  alpha_trace: NULL or nticks doubles receiving alpha after
 * each tick (what onSimulationTick reports).  Asynchronous on the context's
 * stream; cg_get_positions synchronises. */
int cg_step(cg_ctx* ctx, const cg_sim_params* params, uint32_t nticks,
            const float* seeds, double* alpha_trace);

/* Packs graph 0 with these parameters and captures the HIP graphs cg_step
 * will replay for them, without running a tick (no reference counterpart:
 * it moves capture and instantiation out of a caller's first timed steps). */
int cg_prepare(cg_ctx* ctx, const cg_sim_params* params);

/* Copies n positions to xy (2n floats); absent points read back NaN. */
int cg_get_positions(cg_ctx* ctx, float* xy, uint32_t n);

/* Points.drag (Points/index.ts:1961-1988, drag-point.frag:21-33): the listed
 * points take the given positions (the dragged point: the mouse position, not
 * clamped) before the next tick.  Absent points stay absent. */
int cg_drag_points(cg_ctx* ctx, const uint32_t* indices, const float* xy, uint32_t count);

/* getTrackedPointPositionsArray (src/index.ts:1346-1388, Points/index.ts:2267-2337):
 * positions of the listed points in list order; an absent point, or an index
 * at or past the point count, reads NaN. */
int cg_get_positions_at(cg_ctx* ctx, const uint32_t* indices, uint32_t count, float* xy);

const char* cg_last_error(const cg_ctx* ctx);

/* ---------------------------------------------------------------------- */
/* Inspection entry points.  They have no reference counterpart: they expose */
/* one force pass or one intermediate structure so the result can be        */
/* compared with the oracle from an identical input state.                  */
/* ---------------------------------------------------------------------- */

#define CG_FORCE_GRAVITY 0
#define CG_FORCE_CENTER 1
#define CG_FORCE_MOUSE 2
#define CG_FORCE_MANYBODY 3
#define CG_FORCE_LINK_IN 4
#define CG_FORCE_LINK_OUT 5
#define CG_FORCE_COLLISION 6
#define CG_FORCE_CLUSTER 7

/* Velocity that one force pass writes from the current positions (nothing is
 * integrated).  v: 2n floats; absent points get 0. */
int cg_eval_force(cg_ctx* ctx, const cg_sim_params* params, int32_t force, float alpha,
                  float seed, float* v);
/* The same for graph `id` of the set, with its stored parameters, evaluated in
 * place inside the packed set. */
int cg_graph_eval_force(cg_ctx* ctx, uint32_t id, int32_t force, float alpha, float seed, float* v);

/* Allocation-free variant of cg_eval_force(CG_FORCE_MANYBODY).  It has no
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory (code that must not call hipMalloc: capture-adjacent setup, tight
 * serving loops, harnesses that forbid allocation).  The many-body velocity of
 * one graph of n points given here, not through cg_set_positions: xy (2n
 * floats, NaN = absent), jitter (2n floats or NULL = none), result in v (2n
 * floats, 0 for absent points) — the same bits cg_eval_force returns for a
 * context holding these positions and jitter.  Every intermediate lives in one
 * static device array of this library (64 MiB per device, held until the
 * process ends); the call allocates nothing and leaves the context's own
 * memory untouched (ctx gives the device, the stream and cg_last_error).
 * n above CG_NOALLOC_MAX_POINTS is refused with CG_ERR_INVALID.  Calls that
 * overlap on one device (two contexts, two threads) race on the shared array;
 * the call is synchronous. */
#define CG_NOALLOC_MAX_POINTS 262144u
int cg_eval_manybody_noalloc(cg_ctx* ctx, const cg_sim_params* params, float alpha, float seed,
                             const float* xy, const float* jitter, uint32_t n, float* v);

/* Near-field slots of the last many-body evaluation: finest^2 * 8 point
 * indices (row-major cells, ascending ticket, -1 = empty).  finest/levels
 * receive the pyramid geometry. */
int cg_debug_slots(cg_ctx* ctx, int32_t* slots, uint32_t capacity, uint32_t* finest,
                   uint32_t* levels);

/* Level aggregation of the last many-body pass (calculate-level.vert:28-53 per
 * level, the level cells of force-level.frag:66).  level in [0, L): the grid
 * has g = 4 << level cells per side.  counts: g*g present-point counts per cell
 * (row-major cells); centroid: NULL or 2*g*g floats, each non-empty cell's mean
 * position (0, 0 for empty cells).  members: NULL, or at the finest level
 * (level L-1) np graph-local point indices grouped by finest cell in row-major
 * cell order (within a cell the order is unspecified); capacity >= np. */
int cg_debug_cells(cg_ctx* ctx, uint32_t level, uint32_t* counts, float* centroid, uint32_t* members,
                   uint32_t capacity);

/* Spatial hash of the last collision pass (build-grid.vert:46-55).  keys and
 * points: the np (cell, point) pairs of the unshifted grid (row-major cells)
 * in the order the device sorted them, ascending by (cell, point).  cells:
 * NULL or 4*n entries, at [4i + o] the row-major cell of point i in offset
 * grid o (offsets (0,0), (0.5,0), (0,0.5), (0.5,0.5) of a cell) as the cell
 * and force kernels compute it from the current positions, UINT32_MAX for
 * absent points.  counts: NULL or 4*G*G point counts of the four grids' cell
 * aggregates the last pass built (grid o's row-major cells at o*G*G).  grid
 * receives G (cells per side); count receives np. */
int cg_debug_collision_cells(cg_ctx* ctx, uint32_t* keys, uint32_t* points, uint32_t* cells,
                             uint32_t* counts, uint32_t capacity, uint32_t* count, uint32_t* grid);

/* Link CSR of one direction (0 = INCOMING instance, 1 = OUTGOING) in the
 * reference's entry order: offsets[n+1], nbr[E], bias[E], strength[E], and
 * degree[n].  Returns E through *entries. */
int cg_debug_link_csr(cg_ctx* ctx, int32_t dir, uint32_t* offsets, uint32_t* nbr, float* bias,
                      float* strength, uint32_t* degree, uint32_t capacity, uint32_t* entries);

/* Multiple-shooting counters of the chunked link tiers since the last call
 * (then reset), collected while cg_set_profiling has CG_PROFILE_LINK_STATS:
 * out[0..3] for the 64-lane tier and out[4..7] for the 256-lane tier's work
 * units = {sweeps summed over units, units, max sweeps, units stopped at the
 * sweep cap}; out[8..11] for the points split into segments = {rounds summed,
 * points, max rounds, points stopped at the round cap}; out[12..15] zero. */
int cg_debug_link_stats(cg_ctx* ctx, uint32_t* out16);

/* Free implementation parameters of the tick (no reference counterpart):
 * none changes the discretisation.  graph_ticks, the cell tiers and
 * force_tile change no result bit; the link tiers move chains between the
 * serial loop and the multiple-shooting solver (results within the force
 * tolerance).  Chosen per workload by the autotuner (graph_amd/autotune.py);
 * a change takes effect on the next step (the link and cell tiers repack). */
typedef struct cg_tuning {
    uint32_t graph_ticks;      /* ticks per captured HIP graph, 1..256 (default 8) */
    uint32_t cell_items;       /* sorted points per lane of the many-body cell pass: 4 or 8 (8) */
    uint32_t force_tile;       /* level cells staged in LDS by the force kernel: 512 or 1024 (1024) */
    uint32_t link_lane_max;    /* link entries per direction up to which a point runs one lane (64) */
    uint32_t link_group_max;   /* up to which it runs one 64-lane workgroup, <= 256 (256) */
    uint32_t link_hub_entries; /* entries per lane of 256-lane workgroups: 4, 6 or 8 (6) */
} cg_tuning;
void cg_default_tuning(cg_tuning* t);
int cg_set_tuning(cg_ctx* ctx, const cg_tuning* t);
int cg_get_tuning(cg_ctx* ctx, cg_tuning* t);

/* Bytes of device memory the context holds. */
uint64_t cg_device_bytes(const cg_ctx* ctx);

/* Stage timing on the context's stream (no reference counterpart; used by the
 * benchmark).  cg_set_profiling(mask) brackets every launch of the stages in
 * mask (bit s = stage s) with HIP events during later cg_step calls;
 * cg_get_profile synchronises and returns, per stage, the summed
 * milliseconds and the number of bracketed launches since the last call. */
#define CG_STAGE_MOUSE 0
#define CG_STAGE_GRAVITY 1
#define CG_STAGE_CENTER 2
#define CG_STAGE_MB_SORT 3
#define CG_STAGE_MB_CELLS 4
#define CG_STAGE_MB_PYRAMID 5
#define CG_STAGE_MB_FORCE 6
#define CG_STAGE_LINK_IN 7
#define CG_STAGE_LINK_OUT 8
#define CG_STAGE_COLLISION 9
#define CG_STAGE_CLUSTER 10
#define CG_STAGES 11
#define CG_PROFILE_LINK_STATS (1u << 31) /* also count link sweeps (cg_debug_link_stats) */
int cg_set_profiling(cg_ctx* ctx, uint32_t stage_mask);
/* Tick sequences are captured once into HIP graphs and replayed (on by
 * default; off while stage timing is on, and for sets of 8M points or more,
 * whose kernels are long enough to launch directly at the same speed).
 * enable = 0 launches every kernel directly; results are identical either way. */
int cg_set_capture(cg_ctx* ctx, int32_t enable);
int cg_get_profile(cg_ctx* ctx, double* ms, uint32_t* launches, uint32_t nstages);

/* ---------------------------------------------------------------------- */
/* Graph set.  Each call mirrors the single-graph call of the same name on   */
/* graph `id` — in the reference one Graph instance per layout              */
/* (src/index.ts:422-434), each with its own config and Store.alpha; here   */
/* many layouts share the device.  Ids are never reused; id 0 belongs to the */
/* single-graph entry points.  Every graph has its own generator (jitter,    */
/* rest-length draws, near-field seeds), seeded identically unless           */
/* cg_graph_set_seed says otherwise, so a graph gives the same result alone  */
/* and in any mix.  Adding, removing or resizing a graph repacks the set on  */
/* the next step; positions of the other graphs stay on the device.         */
int cg_graph_add(cg_ctx* ctx, uint32_t* id);
int cg_graph_remove(cg_ctx* ctx, uint32_t id);
int cg_graph_set_params(cg_ctx* ctx, uint32_t id, const cg_sim_params* params);
int cg_graph_set_seed(cg_ctx* ctx, uint32_t id, uint64_t seed);
int cg_graph_set_positions(cg_ctx* ctx, uint32_t id, const float* xy, uint32_t n);
int cg_graph_set_links(cg_ctx* ctx, uint32_t id, const float* links, uint32_t m,
                       const float* strength, const float* rand_in, const float* rand_out);
int cg_graph_set_jitter(cg_ctx* ctx, uint32_t id, const float* xy, uint32_t n);
int cg_graph_set_point_sizes(cg_ctx* ctx, uint32_t id, const float* sizes, uint32_t n);
int cg_graph_set_pinned(cg_ctx* ctx, uint32_t id, const uint32_t* indices, uint32_t count);
int cg_graph_set_point_clusters(cg_ctx* ctx, uint32_t id, const int32_t* clusters, uint32_t n);
int cg_graph_set_cluster_positions(cg_ctx* ctx, uint32_t id, const float* xy, uint32_t count);
int cg_graph_set_point_cluster_strength(cg_ctx* ctx, uint32_t id, const float* strength, uint32_t n);
int cg_graph_set_alpha(cg_ctx* ctx, uint32_t id, double alpha);
int cg_graph_get_alpha(cg_ctx* ctx, uint32_t id, double* alpha);
int cg_graph_get_positions(cg_ctx* ctx, uint32_t id, float* xy, uint32_t n);
int cg_graph_drag_points(cg_ctx* ctx, uint32_t id, const uint32_t* indices, const float* xy,
                         uint32_t count);
int cg_graph_get_positions_at(cg_ctx* ctx, uint32_t id, const uint32_t* indices, uint32_t count,
                              float* xy);

/* Advances the listed graphs by nticks ticks each with their stored params
 * (Graph.runSimulationStep per graph); the other graphs of the set keep
 * their state.  Asynchronous like cg_step. */
int cg_step_graphs(cg_ctx* ctx, const uint32_t* ids, uint32_t count, uint32_t nticks);

/* Waits for the context's queued work. */
int cg_sync(cg_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* COSMOSGL_H */
