/*
 * bh.h -- C ABI of the MI355X Barnes-Hut engine (libbh.so).
 *
 * The engine replaces the single-rank `--fc tree` force path of
 * YonatanGM/Barnes-Hut: bounding box -> Morton keys -> key sort -> bottom-up
 * linear octree with monopole centre of mass -> theta-criterion walk, inside
 * the drift-first kick-drift-kick loop (src/main.cpp:191-281).  All entry
 * points return BH_OK (0) or a negative BH_ERR_* code; no exception crosses
 * the ABI and bh_last_error() describes the last failure of a context.
 * Host buffers are caller-owned and sized by the caller; device memory is
 * owned by the context.  One host thread per context; every call is
 * synchronous on return.
 */
#ifndef BH_H
#define BH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
#define BH_OK 0
#define BH_ERR_ARG -1         /* bad argument */
#define BH_ERR_NOMEM -2       /* host or device allocation failed */
#define BH_ERR_HIP -3         /* a HIP runtime call failed */
#define BH_ERR_STATE -4       /* call made in the wrong state (e.g. no bodies) */
#define BH_ERR_IO -5          /* file could not be opened / written */
#define BH_ERR_UNSUPPORTED -6 /* policy/form not provided by this engine */
#define BH_ERR_OVERFLOW -7    /* caller buffer too small */

/* Gravitational constant in AU^3 kg^-1 day^-2 (src/main.cpp:174). */
#define BH_G 1.48812e-34
/* Tree depth of a leaf: 21 bits per axis (src/linear_octree.cpp:10). */
#define BH_MAX_DEPTH 21

/* Body vector, layout-identical to `struct Vec3 {double x,y,z;}`
 * (include/body.h:4-6) so std::vector<Vec3>::data() can be passed as is. */
typedef struct bh_vec3 {
    double x, y, z;
} bh_vec3;

/* Force-exchange policy (include/policy.h:10-15).  One GPU holds every body,
 * so only the full-tree policy exists; LET values are rejected with
 * BH_ERR_UNSUPPORTED. */
typedef enum bh_fc_policy { BH_FC_TREE = 0, BH_FC_LET = 1, BH_FC_LET_DIRECT = 2 } bh_fc_policy;

/* Walk form.
 *  PLAIN   FP64 walk, same decisions and summation order as
 *          src/traversal.cpp:13-103 on the same tree (bit-identical to the
 *          reference built without FMA contraction).
 *  MIXED   FP64 state, drift, box, keys, COM and the central-mass term; the
 *          bulk of the walk in FP32 on leaf-relative offsets with compensated
 *          accumulation.
 *  LEARNED MIXED walk whose leaves are ordered and split into balanced
 *          persistent-kernel chunks by a cost model; forces are bit-identical
 *          to MIXED. */
typedef enum bh_form { BH_FORM_PLAIN = 0, BH_FORM_MIXED = 1, BH_FORM_LEARNED = 2 } bh_form;

typedef struct bh_config {
    int device;       /* HIP device ordinal */
    int fc_policy;    /* bh_fc_policy; BH_FC_TREE */
    int form;         /* bh_form used by bh_step */
    double theta;     /* opening angle (--theta, default 1.05) */
    double G;         /* gravitational constant (BH_G) */
    double soft2;     /* softening eps^2 added to r^2 (0.0 in the reference); the
                         mixed forms' FP32 list terms use at least
                         1e-24 * max(1, (G M)^(2/3)), G M the system's light mass */
} bh_config;

typedef struct bh_ctx bh_ctx;

/* Fill *cfg with the reference defaults: device 0, full tree, PLAIN,
 * theta 1.05 (src/main.cpp:61), G = BH_G, soft2 = 0. */
void bh_config_default(bh_config *cfg);

/* Context lifetime.  bh_create initialises the HIP device; it fails with
 * BH_ERR_HIP when no device is usable and with BH_ERR_UNSUPPORTED for a LET
 * policy. */
int bh_create(bh_ctx **out, const bh_config *cfg);
void bh_destroy(bh_ctx *ctx);
/* Message of the last failing call on ctx ("" if none).  Valid until the
 * next call on ctx.  With ctx == NULL: the last bh_create failure. */
const char *bh_last_error(const bh_ctx *ctx);

/* Upload n bodies (replaces the reference's rank-0 read + scatter,
 * src/main.cpp:117-171).  ids may be NULL (then ids = 0..n-1 in row order,
 * as src/io.cpp:73).  Resets time to 0 and the accelerations to zero
 * (src/main.cpp:175). */
int bh_set_bodies(bh_ctx *ctx, int64_t n, const bh_vec3 *pos, const bh_vec3 *vel,
                  const double *mass, const uint64_t *ids);
/* Number of bodies held by ctx. */
int64_t bh_num_bodies(const bh_ctx *ctx);
/* Download the bodies in their current (last Morton-sorted) order.  Any
 * pointer may be NULL. */
int bh_get_bodies(bh_ctx *ctx, bh_vec3 *pos, bh_vec3 *vel, double *mass, uint64_t *ids, bh_vec3 *acc);

/* The drop-in for one `--fc tree` force evaluation (src/main.cpp:212-263):
 * bounding box, Morton keys, stable key sort that permutes the bodies
 * persistently (include/utility.h:23-52), linear octree with the one-rank
 * exchange round trip (src/linear_octree.cpp:13-107,169-204), and the walk
 * (src/traversal.cpp:13-103).  acc_out (may be NULL) receives the
 * accelerations in the new body order; they are also kept on the device for
 * the next kick. */
int bh_compute_accelerations(bh_ctx *ctx, int form, double theta, double G, double soft2,
                             bh_vec3 *acc_out);

/* nsteps iterations of the drift-first KDK loop (src/main.cpp:191-281):
 * half kick (skipped while t == 0), drift, force evaluation with ctx's form,
 * theta, G and soft2, half kick; t += dt.  Everything stays on the device.
 * Steady steps (t > 0, and for LEARNED once the model is trained) are
 * captured into a HIP graph per parity of the body double buffer and
 * replayed; bh_set_graph(ctx, 0) runs every step eagerly instead.  Graph
 * steps record no phase events (bh_get_step_phase_ms covers eager steps). */
int bh_step(bh_ctx *ctx, int64_t nsteps, double dt);
/* Graph replay of steady steps on (default) or off; *replays (may be NULL)
 * counts the graph steps run so far. */
int bh_set_graph(bh_ctx *ctx, int on);
int bh_get_graph_replays(bh_ctx *ctx, int64_t *replays);
/* Simulation time t accumulated by bh_step. */
double bh_time(const bh_ctx *ctx);
/* Number of iterations `for (t = 0; t < t_end; t += dt)` performs
 * (src/main.cpp:191): e.g. dt = 1h, t_end = 100h gives 101. */
int64_t bh_count_steps(double dt, double t_end);
/* Parse "1h" / "2.5d" / "1y" into days (include/parse_time.h:6-16). */
int bh_parse_time(const char *text, double *days);

/* ---------------------------------------------- exports for tests --- */
/* Sorted Morton keys and box of the last force evaluation. */
int bh_export_keys(bh_ctx *ctx, uint64_t *keys);
int bh_export_bbox(bh_ctx *ctx, double *bb6);
/* Node counts of the last tree: full-chain node count of the reference's
 * map (every (prefix, depth) that holds a body) and compressed node count of
 * the device tree. */
int bh_tree_size(bh_ctx *ctx, int64_t *n_full, int64_t *n_compressed);
/* The reference's node set of the last tree sorted by (depth, prefix) with
 * the (mass, com) the walk reads; arrays sized n_full (com: 3*n_full). */
int bh_export_topology(bh_ctx *ctx, uint64_t *prefix, uint8_t *depth, double *mass, double *com);
/* Per-body counts of the last PLAIN walk: accepted nodes and opened nodes of
 * the full-chain tree (what src/traversal.cpp:63-98 would count), and the
 * number of compressed nodes the body tested (accepted + opened + its own
 * leaf).  Enabled by bh_set_counting(ctx, 1); counting costs walk time.  Any
 * output pointer may be NULL. */
int bh_set_counting(bh_ctx *ctx, int on);
int bh_get_counters(bh_ctx *ctx, int64_t *accepted, int64_t *opened, int64_t *tested);
/* Device-side guards (walk index range, mixed-walk stack capacity) raise
 * bits in a per-context word that no kernel clears: a bit raised in any step
 * of bh_step is reported when the call returns (BH_ERR_STATE), and the check
 * clears it.  This test hook raises `bits` (nonzero) on the context stream as
 * a guard would, so that the reporting path can be tested without corrupt
 * input.  No counterpart in the reference. */
int bh_debug_raise_error(bh_ctx *ctx, uint32_t bits);
/* Device time in ms of the phases of the last force evaluation:
 * [0] kick+drift+box+keys, [1] sort+permute, [2] tree build, [3] walk(+kick),
 * [4] total. */
int bh_get_phase_ms(bh_ctx *ctx, double *ms5);
/* Mean device time per step of the same phases over the steps of the last
 * bh_step call (at most the last 1024 steps); *nsteps gets the count. */
int bh_get_step_phase_ms(bh_ctx *ctx, double *ms5, int64_t *nsteps);

/* Order in which the persistent mixed walk serves its groups (results never
 * depend on it):
 *   KEY_ORDER  dynamic queue in Morton order (default of MIXED)
 *   PREV_COST  heaviest first by the walk work measured in the last step
 *              (list cost + visited nodes / 2 per group, bh_get_group_work)
 *   MODEL      heaviest first by the learned cost model (LEARNED; falls back
 *              to PREV_COST while no model is trained) */
#define BH_SCHED_KEY_ORDER 0
#define BH_SCHED_PREV_COST 1
#define BH_SCHED_MODEL 2
int bh_set_schedule(bh_ctx *ctx, int mode);
/* How PREV_COST and MODEL use the cost estimates: 0 = the dynamic queue
 * serves groups heaviest first (longest processing time first) while there
 * are at most 16 groups per worker wave, and in key order above that (the
 * tail no longer decides the walk, locality does); k > 0 = the
 * groups, in key order, are split into k x (worker waves) runs of equal
 * estimated cost that the dynamic queue serves in key order, which keeps
 * neighbouring groups on the device at the same time.  Default 0. */
int bh_set_schedule_split(bh_ctx *ctx, int chunks_per_worker);

/* Key sort of the step (no counterpart in the reference, whose
 * sortBodiesByMortonKey, include/utility.h:23, is a full stable sort every
 * step).  BH_SORT_NEAR (default): the bodies stay in the previous step's
 * order, so only the keys out of order with their neighbours are sorted (a
 * merge sort) and merged with the rest; when the rest is not in order (or
 * more than half the keys moved) the merge sort takes every key.
 * BH_SORT_FULL: always the full radix sort.  Both give the same, stable
 * order. */
#define BH_SORT_NEAR 0
#define BH_SORT_FULL 1
int bh_set_sort_mode(bh_ctx *ctx, int mode);
/* Path of the last key sort: *near_path = 1 if it took the near-sorted path,
 * *sorted_keys = the keys it radix sorted (n otherwise).  Synchronises. */
int bh_get_sort_stats(bh_ctx *ctx, int *near_path, int64_t *sorted_keys);

/* Tree moments (no counterpart in the reference, whose octree sums them in
 * one recursive pass, src/linear_octree.cpp): one launch per depth, bottom
 * up, except that the branching nodes at depth >= cutoff are summed by a
 * single workgroup, depth by depth, in one launch.  cutoff -1 (default):
 * chosen from the current tree whenever a step graph is captured, one below
 * its deepest branching node (so the empty depths cost one launch, not one
 * each); 0..21 fixes it (21: one launch per depth throughout).  Every cutoff
 * gives the same bits. */
int bh_set_moment_cutoff(bh_ctx *ctx, int cutoff);

/* ------------------------------------------------ learned cost model ---
 * An MLP (8 FP32 group features -> 32 -> 32 -> 1, ReLU, BF16 weights, FP32
 * accumulation) predicts log2(work + 1) / 16 of every walk group (work =
 * list cost + visited nodes / 2, as the walk records it, bh_get_group_work) as
 * feature 5 (the previous step's value) plus its output where that value is
 * ambiguous (feature 5 != feature 6: the group's bodies come from groups of
 * different work) or missing (feature 5 == 0: no previous walk, e.g. after
 * bh_set_bodies), and as feature 5 alone otherwise; the LEARNED form serves
 * groups heaviest-first by its predictions.  It trains on the device from
 * the work the walk records, during the first train_steps LEARNED
 * evaluations (iters_per_step Adam updates of `batch` samples each, the
 * learning rate falling along a cosine to lr / 100, a fifth of the drawn
 * samples stripped of features 5 and 6), and is used from then on.  Features
 * per group:
 *   [0] tree depth / 21         [1] bodies / 64
 *   [2] log2(box extent) / 16   [3] log2(distance of box centre to origin) / 8
 *   [4] (log10(light mass + 1) - 20) / 5
 *   [5] log2(max walk work of its bodies' groups last step + 1) / 16
 *   [6] log2(min walk work of its bodies' groups last step + 1) / 16
 *   [7] log2(z extent / box extent) / 8 */
#define BH_MODEL_FEATURES 8
#define BH_MODEL_PARAMS 1377
/* Defaults: 50 steps, 16 updates per step, batch 2048, learning rate 3e-3. */
int bh_model_configure(bh_ctx *ctx, int train_steps, int iters_per_step, int batch, float lr);
/* Fresh weights from `seed` (He-uniform), empty sample buffer, not ready. */
int bh_model_reset(bh_ctx *ctx, uint64_t seed);
/* FP32 master weights, layout W1[32][8] b1[32] W2[32][32] b2[32] W3[32] b3. */
int bh_model_get_params(bh_ctx *ctx, float *params);
/* Sets the masters (the BF16 copy follows); clears the Adam state. */
int bh_model_set_params(bh_ctx *ctx, const float *params);
/* Model output y for m feature rows (m x 8, row-major). */
int bh_model_forward(bh_ctx *ctx, int64_t m, const float *features, float *out);
/* Gradient of the mean squared error (1/m) sum (y - t)^2 over m rows with
 * respect to every parameter, evaluated with the BF16 weights; loss may be
 * NULL. */
int bh_model_gradients(bh_ctx *ctx, int64_t m, const float *features, const float *targets, float *grad,
                       float *loss);
/* iters Adam updates on mini-batches of `batch` rows drawn from the given
 * rows; marks the model ready.  final_loss may be NULL. */
int bh_model_train(bh_ctx *ctx, int64_t m, const float *features, const float *targets, int iters, int batch,
                   float lr, float *final_loss);
/* ready: model in use; samples: collected; updates: Adam steps so far. */
int bh_model_status(bh_ctx *ctx, int *ready, int64_t *samples, int64_t *updates, float *last_loss);
/* Features of the groups of the last LEARNED evaluation (G x 8).  After the
 * training phase they are only computed while the walk uses the predicted
 * order, i.e. with at most 16 groups per worker wave (see
 * bh_set_schedule_split), and for a group whose features 5 and 6 are equal
 * and nonzero (the network's output is then feature 5) only those two. */
int bh_get_group_features(bh_ctx *ctx, int64_t capacity, float *features);

/* Walk groups of the last MIXED / LEARNED evaluation: up to 64
 * Morton-consecutive bodies walked together.  *n_groups gets the count; when
 * capacity >= count, first_body[g] (count + 1 entries, last = n),
 * interactions[g] (list entries evaluated) and visited[g] (nodes tested)
 * are filled.  Any array may be NULL. */
int bh_get_groups(bh_ctx *ctx, int64_t *n_groups, int64_t capacity, uint32_t *first_body,
                  uint32_t *interactions, uint32_t *visited);

/* Timeline of the last mixed walk: for each group the device wall clock
 * (100 MHz ticks) when a worker wave took it and when it finished, and the
 * worker's index.  Arrays sized by the group count of bh_get_groups. */
int bh_get_group_timeline(bh_ctx *ctx, int64_t capacity, uint64_t *t_start, uint64_t *t_end, uint32_t *worker);

/* Walk work of each group of the last mixed walk: its list cost + visited
 * nodes / 2, the cost unit of the heaviest-first schedules and the learned
 * model's target.  The list cost is its list entries per lane; in a group of
 * 33-63 bodies, whose idle lanes take over stretches of the list, only the
 * entries beyond the body lanes' stretches count (the ordering this gives
 * measured best).  Array sized by the group count of bh_get_groups. */
int bh_get_group_work(bh_ctx *ctx, int64_t capacity, uint32_t *work);

/* Estimated cost of each group of the last mixed walk, as its heaviest-first
 * order used it: the learned model's prediction 2^(16 y) - 1 (LEARNED) or the
 * previous step's work (PREV_COST).  Meaningful only for a walk that served
 * such an order.  Array sized by the group count of bh_get_groups. */
int bh_get_group_costs(bh_ctx *ctx, int64_t capacity, float *cost);

/* FP64 direct sum on the device for m sample bodies (indices into the
 * current body order): a_i = sum_{j != i, r > 0} G m_j d / |d|^3 over the
 * bodies j of i's system, with Neumaier-compensated sums, r^2 = |d|^2 +
 * soft2.  The reference for force errors; acc_out is m x 3. */
int bh_direct_sum(bh_ctx *ctx, int64_t m, const int64_t *idx, double G, double soft2, bh_vec3 *acc_out);

/* ------------------------------------------------ ensembles -----------
 * No counterpart in the reference (one process simulates one system): the
 * batched ensemble form steps many independent systems in one launch
 * sequence.  After bh_set_bodies with the systems' bodies concatenated,
 * bh_set_systems(ctx, S, counts) splits them into S <= BH_MAX_SYSTEMS
 * systems of counts[s] >= 1 bodies (sum = n); system s keeps the index range
 * [off[s], off[s+1]) for good (sorting permutes inside a system only).
 * Every system gets its own box, keys, tree, walk groups and heavy set, so
 * its forces are bit-identical to a run of that system alone.  The MIXED and
 * LEARNED forms run ensembles; PLAIN returns BH_ERR_UNSUPPORTED for S > 1.
 * bh_set_bodies resets to one system.  bh_set_systems must come after
 * bh_set_bodies and before any force evaluation (BH_ERR_STATE otherwise): an
 * evaluation sorts the bodies of a single system over the whole range. */
#define BH_MAX_SYSTEMS 256
int bh_set_systems(bh_ctx *ctx, int64_t n_systems, const int64_t *counts);
/* *n_systems gets S; offsets (S + 1 entries, may be NULL) the body ranges. */
int bh_get_systems(bh_ctx *ctx, int64_t *n_systems, int64_t capacity, int64_t *offsets);

/* Walk-group shape of the mixed forms: at most gmax (1..64) bodies per group
 * and group boxes no wider than efac x the cell side at the first depth
 * whose non-empty cells number at least (system size) / occ.  Defaults 64, 2,
 * 32.  walk_blocks: persistent blocks of the mixed walk (0 = the device's
 * occupancy limit). */
int bh_set_group_params(bh_ctx *ctx, int gmax, double efac, int occ, int walk_blocks);
/* Relative opening test of the mixed walk, on top of the extent test
 * b < theta D: a node with b > 0.8 D is accepted only if also
 * G Q <= alpha |a_h| D^4, Q = sum m |x - com|^2 of its light bodies (the scale
 * of its quadrupole, whose neglect is the monopole's error ~ G Q / D^4), D its
 * distance to the group box, |a_h| the smallest heavy-body acceleration of the
 * group.  Default alpha = 1e-7; 0 turns it off.  It only acts for theta > 0.8
 * and matters for dense, massive clusters (Trojan clouds). */
int bh_set_relative_opening(bh_ctx *ctx, double alpha);
int bh_get_relative_opening(bh_ctx *ctx, double *alpha);
int bh_get_group_params(bh_ctx *ctx, int *gmax, double *efac, int *occ, int *walk_blocks);

/* Autotuner of the mixed walk: times `reps` force evaluations of `form` on
 * the current state (ctx's theta, G, soft2; no kick or drift) for each
 * candidate group shape (gmax 32/48/64 x efac 1/2/4/8) and walk occupancy
 * (full, half), keeps the largest gmax among the candidates within 3% of the
 * fastest and, for it, the fastest within 1% (full grid, then the widest efac,
 * on ties), and reports it.  A LEARNED request before
 * the model is trained is tuned as MIXED with the context's schedule.
 * Call before the first step: the accelerations of the last candidate stay
 * on the device. */
typedef struct bh_tune_result {
    int gmax;
    double efac;
    int occ;
    int walk_blocks;
    double best_ms;     /* median force evaluation of the choice */
    double default_ms;  /* the same with the default shape */
    int n_tried;
} bh_tune_result;
int bh_autotune(bh_ctx *ctx, int form, int reps, bh_tune_result *out);

/* ------------------------------------------- evaluation without allocation
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): one plain FP64
 * force evaluation -- the kernels of bh_compute_accelerations with
 * BH_FORM_PLAIN, the same bits -- on n <= BH_STATIC_MAX_BODIES bodies in
 * the caller's device buffers, with every intermediate buffer in a static
 * device array of the library (about 40 MiB per GPU, held until the process
 * ends).  No context: it runs on the current device, enqueued on `stream` (a
 * hipStream_t; NULL for the null stream), and returns without waiting.
 * pos (n bh_vec3) and mass (n doubles) are device pointers and are not
 * modified; acc_out (n bh_vec3, device) receives the accelerations in the
 * input order (row i = body i; the Morton sort stays inside).  No hipMalloc,
 * hipFree or synchronisation.  Every call on a device shares that device's
 * static array: calls that overlap on one device race.  Argument errors
 * return BH_ERR_ARG (message in bh_static_last_error()); device-side guard
 * bits are not reported. */
#define BH_STATIC_MAX_BODIES 65536
int bh_compute_accelerations_static(int64_t n, const bh_vec3 *pos, const double *mass, double theta, double G,
                                    double soft2, bh_vec3 *acc_out, void *stream);
const char *bh_static_last_error(void);

/* ------------------------------------------------ synthetic inputs --- */
#define BH_GEN_BELT 0   /* central mass + 8 planets + main belt */
#define BH_GEN_KUIPER 1 /* as BELT with 30% of the asteroids in a Kuiper belt */
#define BH_GEN_TROJAN 2 /* 60% belt, 2 x 15% Jupiter Trojans, 10% Hildas */
#define BH_GEN_CENTRAL_MASS 1.989e30

/* n >= 9 bodies: [0] central mass at rest at the origin, [1..8] planets,
 * then asteroids.  Deterministic in (kind, seed, n). */
int bh_generate(int kind, uint64_t seed, int64_t n, bh_vec3 *pos, bh_vec3 *vel, double *mass);
/* count system sizes, log-uniform in [nmin, nmax]. */
int bh_generate_ensemble_sizes(uint64_t seed, int64_t count, int64_t nmin, int64_t nmax, int64_t *sizes);
/* State-vector CSV in the reference format (src/io.cpp:16-81). */
int bh_write_csv(const char *path, int64_t n, const bh_vec3 *pos, const bh_vec3 *vel, const double *mass);
/* Reads at most body_limit bodies (<= 0: all).  With pos/vel/mass NULL only
 * counts (*n_out); otherwise fills up to capacity rows and returns
 * BH_ERR_OVERFLOW when more rows were found.  ids (may be NULL) = row order
 * of the parsed rows; malformed rows are skipped with a warning. */
int bh_read_csv(const char *path, int64_t body_limit, int64_t capacity, int64_t *n_out, bh_vec3 *pos,
                bh_vec3 *vel, double *mass, uint64_t *ids);

#ifdef __cplusplus
}
#endif

#endif /* BH_H */
