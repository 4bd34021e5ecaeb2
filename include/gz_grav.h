/* gz_grav.h -- C ABI of the MI355X tree-gravity library (libgzgrav.so).
 *
 * Drop-in boundary for the shared-memory GIZMO tree (shmem/tree.h): Morton keys + sort, octree
 * build + node moments, the cubic-spline-softened grouped and per-target tree walks, the exact
 * direct sum, and the caller-side target ordering / acceleration write-back / kick
 * (shmem/mfm.cc:679-905, 3699-3813). Plain pointers and sizes only.
 *
 * Conventions
 *   - Every call returns gz_status; GZ_OK is 0. gz_last_error() describes the most recent failure
 *     on the calling thread. Nothing calls exit() or abort() (the reference's CUDA path exits on a
 *     launch error, shmem/grav_cuda.cu:239-243).
 *   - Array arguments are DEVICE pointers (HBM) unless a name says host_. Calls are ordered on the
 *     context's stream; functions that return host-visible results (counts, stats) synchronise it.
 *   - Outputs are caller-allocated. Accelerations INCLUDE the G factor, as shmem::accel and
 *     shmem::accel_grouped return them (tree.cc:580, :880); aold is in no-G units (tree.h:389-391).
 *   - Softening is the cubic-spline kernel EXTENT h per particle (tree.h:52-67).
 *   - Node ids are level-major (breadth-first): the children of a node are contiguous and in octant
 *     order. The reference numbers nodes in build-task order; gz_tree_download also returns the
 *     reference-style threaded next[] so either numbering walks the same tree.
 */
#ifndef GZ_GRAV_H
#define GZ_GRAV_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GZ_OK = 0,
    GZ_ERR_INVALID = 1,     /* bad argument (null pointer, size out of range)             */
    GZ_ERR_HIP = 2,         /* a HIP runtime call failed                                  */
    GZ_ERR_NOMEM = 3,       /* device allocation failed                                   */
    GZ_ERR_UNSUPPORTED = 4, /* a reference option this build does not implement (zeta...) */
    GZ_ERR_ARENA = 5,       /* node arena exhausted and growth was disabled               */
    GZ_ERR_INTERNAL = 6     /* a device-side invariant check failed (reported, not faulted)*/
} gz_status;

typedef struct gz_ctx gz_ctx;   /* one per device + stream */
typedef struct gz_tree gz_tree; /* device-resident tree, owned by the library */
typedef struct gz_model gz_model; /* learned opening-parameter network (device-resident) */
typedef struct gz_walk_plan gz_walk_plan; /* groups of a fixed target list, walked repeatedly */

/* Message for the last failing call on this thread ("" if none). */
const char* gz_last_error(void);
const char* gz_status_string(gz_status s);

/* ---- context ------------------------------------------------------------------------------- */
/* stream: a hipStream_t, or NULL for a stream the context creates and owns. */
gz_status gz_ctx_create(int device, void* stream, gz_ctx** out);
gz_status gz_ctx_destroy(gz_ctx* ctx);
gz_status gz_ctx_synchronize(gz_ctx* ctx);

/* ---- particles (shmem::Particles, tree.h:36-50) --------------------------------------------- */
typedef struct {
    size_t n;
    const double *x, *y, *z; /* positions, FP64                                        */
    const double* m;         /* masses; m <= 0 marks a tombstone left out of the tree   */
    const double* soft;      /* kernel extent per particle; NULL means 0                */
    const double* zeta;      /* adaptive-softening zeta; must be NULL (GZ_ERR_UNSUPPORTED) */
    const uint8_t* type;     /* GIZMO type; NULL = all gas. Only read for sink counts.   */
} gz_particles;

/* ---- build (shmem::build, tree.h:384-385, tree.cc:350-526) ---------------------------------- */
typedef struct {
    long long randomize_seed; /* >= 0: RANDOMIZE_GRAVTREE root shift (tree.cc:381-393); -1 off */
    double shift[3];          /* SHMEM_TREE_SHIFT fractions (tree.cc:400-407); 0 = off         */
    int max_nodes;            /* 0: grow the node arena as needed (the reference's retry,
                                 tree.cc:438-462); > 0: fail with GZ_ERR_ARENA beyond this     */
    /* A forest of independent trees, one per cluster (no counterpart in the reference, which
     * builds one tree): cluster c is particles [cluster_start[c], cluster_start[c+1]) (host
     * array of n_clusters + 1 offsets). Each cluster gets its own root cell (the reference's
     * rule applied to its own particles), nodes 0..n_clusters-1 are the cluster roots, and a
     * walk of a target sees only its own cluster. n_clusters <= 1: one tree. */
    const size_t* cluster_start;
    int n_clusters;
    /* Particle velocities (device, length n), or NULL: the build then also keeps what reusing
     * the tree across steps needs (tree.h:384-385 with want_vcom): per-node max |v| and
     * centre-of-mass velocity, and zeroed incremental moments (rebuild_tree, mfm.cc:1408-1413). */
    const double *vx, *vy, *vz;
} gz_build_opts;

/* Stage timings in ms, the reference's BuildTimes (tree.h:371) measured with HIP events. */
typedef struct {
    double bbox, keys, sort, recurse, links, pack, total;
} gz_build_times;

typedef struct {
    long long n, n_live, nnodes, nleaves, root, depth, arena_grows;
    double root_centre[3], root_side;
    gz_build_times times;
} gz_tree_info;

/* Build a tree over p (device arrays). opts may be NULL (defaults: no randomisation). */
gz_status gz_tree_build(gz_ctx* ctx, const gz_particles* p, const gz_build_opts* opts,
                        gz_tree** out);
gz_status gz_tree_free(gz_tree* tree);
gz_status gz_tree_get_info(const gz_tree* tree, gz_tree_info* info);

/* Copy tree arrays to HOST buffers (any may be NULL). Node arrays have nnodes entries.
 * first = first child or -1; next = the reference's threaded successor (tree.cc:340-348);
 * nchild = number of children; plo/phi = particle range in Morton order; wn_s = size + delta,
 * wn_soft/wn_len = the float fields of the packed walk node (tree.h:195-204). */
gz_status gz_tree_download_nodes(const gz_tree* tree, int* first, int* next, int* nchild,
                                 int* plo, int* phi, int* parent, int* level, double* size,
                                 double* mass, double* cx, double* cy, double* cz, double* soft,
                                 double* delta, double* gcx, double* gcy, double* gcz,
                                 double* wn_s, float* wn_soft, float* wn_len);
/* orderbuf (n_live), rank (n; UINT32_MAX for tombstones), leaf_of (n; -1), skey (n_live). */
gz_status gz_tree_download_particles(const gz_tree* tree, uint32_t* orderbuf, uint32_t* rank,
                                     int* leaf_of, uint64_t* skey);
/* Device pointers owned by the tree (valid until gz_tree_free). */
gz_status gz_tree_device_arrays(const gz_tree* tree, const uint32_t** orderbuf,
                                const uint32_t** rank, const int** leaf_of);

/* Re-read source positions, masses and softenings from p into the tree's Morton-ordered copy,
 * for callers that move particles between a build and a walk. The reference's walk reads source
 * particles from the live arrays (shmem/tree.cc:787-788); this library's walk reads the copy made
 * at build time, for sources and targets alike, so a caller that drifts particles calls this
 * before walking. */
gz_status gz_tree_update_sources(gz_ctx* ctx, gz_tree* tree, const gz_particles* p);

/* ---- reusing a tree across steps (tree.h:250-331) -------------------------------------------
 * Between rebuilds the reference keeps the tree and feeds it each step's velocity changes; before
 * a walk it rewrites the walk nodes from the incremental moments. A caller that drifts particles
 * also calls gz_tree_update_sources before walking (the walk reads the tree's sorted copy). */

/* The kick feed of mfm.cc:3798-3811 for the particles in active (device, n_active): raise every
 * ancestor's speed bound to |v| (Tree::raise_vmax, tree.cc:276-285) and, when v_last is given,
 * add m (v - v_last) and its product with tau to every ancestor's incremental momentum and first
 * moment (Tree::kick_node, tree.cc:211-227), then set v_last = v. The momentum feed needs a tree
 * built with velocities; without them only the speed bound is raised. v_last may be NULL. A tree
 * built with velocities also records v in its sorted velocity copy (for gz_tree_drift); tau < 0
 * takes the tree's age on the device. Under
 * hipGraph capture the call enqueues its kernel only: an index outside the particle set is
 * skipped but not reported. */
gz_status gz_tree_kick_nodes(gz_ctx* ctx, gz_tree* tree, const uint32_t* active, size_t n_active,
                             const double* vx, const double* vy, const double* vz,
                             const double* m, double* vlast_x, double* vlast_y, double* vlast_z,
                             double tau);
/* The drift of the tree's own sorted copy of the sources: every slot's position += v dt, with v the
 * particle's velocity as the build (with velocities) and gz_tree_kick_nodes last saw it. After
 * gz_drift with the same dt and velocities it leaves the sorted copy exactly as
 * gz_tree_update_sources would (the same operations on the same operands), but it streams the
 * copy in slot order instead of gathering every particle through orderbuf. Needs a tree built with
 * velocities, and velocity changes reported through gz_tree_kick_nodes. It also advances the
 * tree's age kept on the device (zero at the build), which gz_tree_refresh_walk_nodes and
 * gz_tree_kick_nodes use when passed tau < 0: a sequence of steps enqueued without the host (a
 * captured graph) then refreshes at the right age on every replay. */
gz_status gz_tree_drift(gz_ctx* ctx, gz_tree* tree, double dt);
/* Tree::refresh_walk_nodes (tree.cc:260-274): every walk node's mass and centre of mass at tau
 * (time since the build) from the incremental moments, and its opening radius and side widened
 * by 2 (vmax tau + xpad). The identity on mass and centre of mass for a tree built without
 * velocities. tau < 0: the tree's age on the device (see gz_tree_drift). */
gz_status gz_tree_refresh_walk_nodes(gz_ctx* ctx, gz_tree* tree, double tau);
/* Tree::raise_soft (tree.cc:311-319) for the particles idx (device, n): raise every ancestor's
 * softening bound to soft[i] (soft: device, indexed by particle). */
gz_status gz_tree_raise_soft(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                             const double* soft);
/* Events between rebuilds that are neither a drift nor a kick, as the reference's serial callers
 * report them (accretion merges and removals, mfm.cc:1676-1687 and :1895-1901; the Hermite
 * corrector, mfm.cc:2350-2356). Each event k names the particle idx[k] whose build-time leaf
 * receives it (particles not in the tree are skipped); the per-event values are device arrays of
 * n, and events are applied in the order of k, so every node's sums round as the reference's
 * serial calls do. An index outside the particle set is an error (unreported under capture).
 *
 * Tree::mass_event (tree.h:321-323, tree.cc:239-251): a particle of mass dm[k] (negative removes
 * one) at (x, y, z)[k] moving at (vx, vy, vz)[k] appears at tau since the build (tau < 0: the
 * tree's age on the device); every ancestor's mass, momentum and first moment take it, with the
 * position unwrapped against that ancestor's current centre of mass in a periodic box of side
 * box (box <= 0: open). dm[k] == 0 is skipped. A no-op on a tree built without velocities, as the
 * reference's is. */
gz_status gz_tree_mass_events(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                              const double* dm, const double* x, const double* y, const double* z,
                              const double* vx, const double* vy, const double* vz, double tau,
                              double box);
/* Tree::shift_node (tree.h:324-325, tree.cc:253-258): a particle of mass m[k] displaced by
 * (dx, dy, dz)[k] without a change of velocity; every ancestor's first moment takes it. A no-op
 * on a tree built without velocities. */
gz_status gz_tree_shift_nodes(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                              const double* m, const double* dx, const double* dy,
                              const double* dz);
/* Tree::raise_pad (tree.h:316-317, tree.cc:323-330): a displacement of d[k] not explained by
 * motion at vmax; the leaf's pad grows by d[k] (rounded up to float) and every ancestor's pad
 * becomes at least the leaf's. d[k] <= 0 (or NaN) is skipped. gz_tree_refresh_walk_nodes widens
 * each node's opening radius and side by 2 xpad, and the neighbour searches widen their node
 * test by it. */
gz_status gz_tree_raise_pad(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                            const double* d);
/* Tree::raise_vmax (tree.h:303-309, tree.cc:276-285) to speed[k] (device floats, >= 0; a negative
 * or NaN speed is an error). A maximum, so the order does not matter. */
gz_status gz_tree_raise_vmax(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                             const float* speed);
/* Per node, to HOST buffers (either may be NULL): the pad xpad (nnodes floats) and the mass change
 * since the build dmass (nnodes doubles, zero for a tree built without velocities). */
gz_status gz_tree_download_pads(const gz_tree* tree, float* xpad, double* dmass);

/* The reuse state to HOST buffers (any may be NULL): vmax (nnodes floats), and vcom, dp, dq
 * (3 x nnodes doubles each, x y z of node i at 3i..3i+2; zero for a tree built without
 * velocities). The refreshed walk-node fields come from gz_tree_download_nodes. */
gz_status gz_tree_download_reuse(const gz_tree* tree, float* vmax, double* vcom, double* dp,
                                 double* dq);

/* ---- walks -------------------------------------------------------------------------------- */
typedef enum {
    GZ_PREC_FP64 = 0,    /* FP64 throughout: the parity walk                                 */
    GZ_PREC_FP32_REL = 1 /* FP32 interactions on group-relative positions; FP64 group centre,
                            node COM and final per-target sum                                  */
} gz_precision;

typedef enum { GZ_LAUNCH_QUEUE = 0, GZ_LAUNCH_PER_GROUP = 1 } gz_launch;

typedef enum {
    GZ_CRIT_REFERENCE = 0, /* shmem open_node (tree.h:109-144)                                 */
    GZ_CRIT_LEARNED = 1    /* per-group theta from the learned model, reference as fallback    */
} gz_criterion;

typedef struct {
    int precision;   /* gz_precision                                                           */
    int criterion;   /* gz_criterion                                                           */
    int group_size;  /* batch <= 0 grouping: 0 = one group per leaf (the reference); G > 0 =
                        targets under the highest node holding <= G particles                  */
    int launch;      /* GZ_LAUNCH_QUEUE (0, default): a persistent grid of wavefronts takes
                        groups from a work queue, which balances dense and sparse regions;
                        GZ_LAUNCH_PER_GROUP (1): one workgroup per group                       */
    int count;       /* 1 = count node visits and interactions into gz_walk_counts (a slower
                        kernel variant); timings are reported whenever counts is non-NULL     */
    int stack_limit; /* 0 = default. > 0 caps the on-chip node stack at this many entries, so
                        groups spill to the global-memory re-walk (for tests of that path)    */
    const void* model; /* gz_model* when criterion == GZ_CRIT_LEARNED                          */
    int waves_per_cu;  /* persistent launches: wavefronts per compute unit, 1..20 (0 = 20, the
                          most the walk's registers and LDS allow)                              */
} gz_walk_opts;

/* SHMEM_COUNT_WALK (tree.cc:671-674, :891-895): nodes tested, multipole interactions, pair
 * interactions, summed over targets; plus launch figures measured on the device. */
typedef struct {
    long long nodes_visited, multipole, pairs, groups, items;
    double walk_ms;    /* walk kernel duration, HIP events on the context stream */
    double prep_ms;    /* grouping + group bounds */
    long long max_stack; /* deepest node stack of any group (counted walks only) */
    long long retries;   /* groups re-walked with a global-memory node stack */
    long long fallback_groups; /* learned criterion: groups that kept the reference tolerance */
} gz_walk_counts;

/* shmem::accel_grouped (tree.h:428-440, tree.cc:626-897) with tidal = lazy = jerk = NULL and
 * sink_direct_radius = 0. targets: n_t particle indices (device); aold: n_t values or NULL;
 * batch <= 0 groups targets by leaf (node-aligned), batch > 0 uses fixed windows.
 * opts may be NULL (FP64, reference criterion, leaf groups). counts may be NULL. */
gz_status gz_tree_accel_grouped(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                const uint32_t* targets, size_t n_t, double theta, double G,
                                int batch, const double* aold, double* ax, double* ay, double* az,
                                const gz_walk_opts* opts, gz_walk_counts* counts);

/* ---- walk plans: batched and graph-captured walks ------------------------------------------
 * No counterpart in the reference, which groups targets inside every accel_grouped call
 * (tree.cc:656-667). A plan forms the groups of a fixed target list once; launching it enqueues
 * only the walk kernels, with no host synchronisation, so a sequence of launches can be captured
 * into a hipGraph and replayed. seg_start (host, n_seg + 1 slot offsets into targets, or NULL for
 * one list) splits the targets into independent lists, e.g. the active sets of several substeps,
 * whose groups never mix: one launch then walks all of them through one work queue. opts as for
 * gz_tree_accel_grouped (counting is not available in plans). The plan keeps pointers to tree,
 * p and targets, which must stay valid while it is used. */
gz_status gz_walk_plan_create(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                              const uint32_t* targets, size_t n_t, const size_t* seg_start,
                              int n_seg, const gz_walk_opts* opts, gz_walk_plan** out);
/* Enqueue the walk on the context's stream: as gz_tree_accel_grouped with batch = 0. */
gz_status gz_walk_plan_launch(gz_ctx* ctx, gz_walk_plan* plan, double theta, double G,
                              const double* aold, double* ax, double* ay, double* az);
/* Synchronise and report the last launch: groups, retries (groups re-walked with a global
 * node stack), fallback_groups. GZ_ERR_INTERNAL if a global node stack overflowed. */
gz_status gz_walk_plan_status(gz_ctx* ctx, gz_walk_plan* plan, gz_walk_counts* counts);
gz_status gz_walk_plan_free(gz_walk_plan* plan);

/* Capture the work enqueued on the context's stream between begin and end into an executable
 * hipGraph (hipStreamBeginCapture / hipGraphInstantiate), and replay it. Only plan launches and
 * the caller-side kernels (order, aold, scatter, kick) may be enqueued while capturing: calls
 * that synchronise fail. */
typedef struct gz_graph gz_graph;
gz_status gz_graph_begin(gz_ctx* ctx);
gz_status gz_graph_end(gz_ctx* ctx, gz_graph** out);
gz_status gz_graph_launch(gz_ctx* ctx, gz_graph* graph);
gz_status gz_graph_free(gz_graph* graph);

/* ---- autotuning the walk ---------------------------------------------------------------------
 * No counterpart in the reference, whose batch size is a measured constant (tree.h:411-413).
 * Times gz_tree_accel_grouped (grouping and walk kernels, HIP events) for candidate launch
 * parameters on the caller's own workload and returns the fastest as *best (base with group_size,
 * stack_limit and waves_per_cu replaced): target-group sizes first (0 and 16/32/64 for FP64,
 * 16/32/64 for FP32), then, at the best of those, the on-chip node-stack depth (stack_limit
 * 384/448/512 entries: a shallower stack spills to global memory sooner) and the wavefronts per
 * compute unit of the persistent launch (10/15/20). Every candidate is walked once untimed and then `reps` times (0 = 2).
 * The choice is cached in the context per workload (precision, criterion and model, theta = 0 or
 * not, and the power-of-two sizes of the target list, of the tree and of its node count); a later
 * call for the same workload returns it without timing unless flags & GZ_TUNE_RETUNE. The walk's
 * outputs go to scratch: the caller's arrays are not written. report may be NULL. */
#define GZ_TUNE_RETUNE 1
#define GZ_TUNE_MAX_POINTS 16
typedef struct {
    int group_size, stack_limit, waves_per_cu;
    double ms; /* mean of the timed walks, grouping included */
} gz_tune_point;
typedef struct {
    int n_points;   /* candidates timed this call (0 when the choice came from the cache) */
    int from_cache; /* 1: the cached choice for this workload was returned                 */
    gz_tune_point points[GZ_TUNE_MAX_POINTS];
    gz_tune_point best;
} gz_tune_report;
gz_status gz_walk_autotune(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                           const uint32_t* targets, size_t n_t, double theta, double G,
                           const double* aold, const gz_walk_opts* base, int reps, int flags,
                           gz_walk_opts* best, gz_tune_report* report);

/* shmem::accel (tree.h:392-394, tree.cc:528-582): per-target walk, a reached leaf is always
 * direct-summed. Targets may come in any order and repeat (a list out of the tree's slot order is
 * walked in slot order); outputs and aold are by list position. Asynchronous on the context's
 * stream once the target list has been checked. */
gz_status gz_tree_accel(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                        const uint32_t* targets, size_t n_t, double theta, double G,
                        const double* aold, double* ax, double* ay, double* az);

/* shmem::accel_grouped with the tidal tensor and the jerk (tree.h:414-440, tree.cc:811-871),
 * FP64 with the reference criterion and grouping: tidal (6 x n_t doubles, xx yy zz xy yz xz,
 * component-major, WITHOUT G as the reference leaves it) and, when jerk and vx/vy/vz (particle
 * velocities, device) are given and the tree was built with velocities, the jerk (3 x n_t, G
 * included); otherwise the jerk is zero. */
gz_status gz_tree_accel_grouped_tidal(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                      const uint32_t* targets, size_t n_t, double theta, double G,
                                      int batch, const double* aold, const double* vx,
                                      const double* vy, const double* vz, double* ax, double* ay,
                                      double* az, double* tidal, double* jerk);

/* shmem::ngb_search (hydro.h:73-75, hydro.cc:37-108) for n_q centres at once (non-periodic): the
 * particles within radius[q] of (cx, cy, cz)[q] (device arrays), read from the tree's sorted copy,
 * with node boxes widened by vmax * tau for a tree aged tau. start (device, n_q + 1) receives the
 * CSR offsets and *total (host) the number found; when found (device) holds at least *total
 * entries it receives, per query, the particle indices in Morton order (the reference's traversal
 * order). Call with found = NULL to size it. Queries in Morton order search fastest. */
gz_status gz_tree_ngb_search(gz_ctx* ctx, const gz_tree* tree, const double* cx,
                             const double* cy, const double* cz, const double* radius,
                             size_t n_q, double tau, size_t* start, uint32_t* found, size_t cap,
                             size_t* total);

/* Tree::init_hmax (tree.h:310-312, tree.cc:332-338): each node's hmax becomes the largest gas
 * kernel radius below it, h rounded up to float (particles 0 .. min(n_h, n_gas) - 1 are gas; h is
 * a device array by particle index; h <= 0 raises nothing). Kept in the tree for
 * gz_tree_ngb_search_reverse; call again when h grows. */
gz_status gz_tree_init_hmax(gz_ctx* ctx, gz_tree* tree, const double* h, size_t n_h,
                            size_t n_gas);

/* Tree::raise_hmax (tree.h:310-311, tree.cc:293-301) for the leaf of each listed particle idx[k]
 * (device arrays), to h[idx[k]] (h by particle; h <= 0 raises nothing): hmax only grows between
 * rebuilds, as the reference's does. Needs gz_tree_init_hmax. */
gz_status gz_tree_raise_hmax(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                             const double* h);

/* shmem::ngb_search_reverse (hydro.h:78-83, hydro.cc:110-148) for n_q centres at once
 * (non-periodic): the gas particles j (j < n_gas) with h_q[q] <= r < h[j], i.e. those whose own
 * kernel reaches centre q although q's does not; needs gz_tree_init_hmax. h (by particle, at
 * least n_gas values), the centres and h_q are device arrays; the output is CSR as
 * gz_tree_ngb_search's, each list in Morton order (the reference's traversal order). */
gz_status gz_tree_ngb_search_reverse(gz_ctx* ctx, const gz_tree* tree, const double* h,
                                     size_t n_gas, const double* cx, const double* cy,
                                     const double* cz, const double* h_q, size_t n_q, double tau,
                                     size_t* start, uint32_t* found, size_t cap, size_t* total);

/* shmem::density (hydro.h:115-119, hydro.cc:150-327), non-periodic, the tree aged tau, without the
 * neighbour cache or lazy drift: for each target particle targets[t] the support radius h[t]
 * solving N_eff(h) = ball_vol(h) sum_j W(r_j, h) = des_ngb (the cubic spline of support h,
 * hydro.h:29-65, in n_dims = 1, 2 or 3 dimensions; the target widened by the face-closure and, with
 * cond_prev (by particle, may be NULL), the condition-number corrections) to within ngb_tol, by
 * bracket-guarded Newton from h_start[t] (NULL: the reference's guess, des_ngb particles at the
 * mean density of the bounding box of all particles), with gas neighbours only (p->type NULL: all
 * gas). Outputs (device, n_t each): h, rho = sum m_j W(r_j, h), the neighbours inside h, and the
 * iterations taken (100: not converged, rho and the count then at the last h). The positions are
 * read from p (the tree's sorted copy must match them: gz_tree_update_sources after a drift). */
gz_status gz_tree_density(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                          const uint32_t* targets, size_t n_t, double des_ngb, double ngb_tol,
                          const double* h_start, int n_dims, const double* cond_prev, double tau,
                          double* h, double* rho, int* nngb, int* iters);

/* shmem::potential (tree.h:446-447, tree.cc:899-944): the spline-softened potential at each
 * target, G included, walked per target like gz_tree_accel (leaves always summed), FP64, with the
 * same target-order rules. */
gz_status gz_tree_potential(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                            const uint32_t* targets, size_t n_t, double theta, double G,
                            const double* aold, double* phi);

/* shmem::accel_brute (tree.h:450-451, tree.cc:946-971): exact O(n_t * N) sum in FP64 with the
 * walk's pair rule, in source index order; includes G. This is the error judge. The reference's
 * device counterpart is shmem_cuda_accel_bruteforce (grav_cuda.cu:186-192), which omits G. */
gz_status gz_direct_sum(gz_ctx* ctx, const gz_particles* p, const uint32_t* targets, size_t n_t,
                        double G, double* ax, double* ay, double* az);
/* The reference's device direct sum, shmem_cuda_accel_bruteforce (grav_cuda.cu:65-247), in both
 * of its arithmetics. fp64 = 0 is its default float-narrowed form (R = float): each separation is
 * differenced in FP64 and narrowed, r, the spline at max(eps_target, eps_source) (1e-30 when both
 * are zero) and the per-target sums are FP32. fp64 != 0 is its SHMEM_CUDA_FP64 form, here the
 * same arithmetic as gz_direct_sum. box > 0 takes the nearest periodic image of each separation
 * in a cube of that side (grav_cuda.cu:104-109); 0 = open boundaries. Unlike the reference, the
 * result includes G, targets and particles are device arrays, and errors are returned. */
gz_status gz_direct_sum_ex(gz_ctx* ctx, const gz_particles* p, const uint32_t* targets,
                           size_t n_t, double G, int fp64, double box, double* ax, double* ay,
                           double* az);

/* ---- learned opening criterion --------------------------------------------------------------
 * No counterpart in the reference, whose opening parameter is one global tolerance
 * (ErrTolForceAcc, shmem/mfm.cc:743-751, used by open_node's relative test, tree.h:134-135).
 * A 16 -> 64 -> 64 -> 1 GELU network (bf16 weight matrices, FP32 biases and accumulation) maps
 * the 16 features of a
 * target group to log10 of that group's tolerance alpha_g; the walk then applies the reference
 * test with aold_min scaled by alpha_g / alpha_ref. A group whose features lie outside the
 * calibrated ranges, or whose prediction lies outside the trained label range, keeps the
 * reference tolerance. Select it with gz_walk_opts.criterion = GZ_CRIT_LEARNED, .model = model.
 * Parameters: W1[64][16], b1[64], W2[64][64], b2[64], w3[64], b3 (GZ_MLP_PARAMS floats). */
#define GZ_MLP_FEATURES 16
#define GZ_MLP_PARAMS 5313
gz_status gz_model_create(gz_ctx* ctx, const float* host_params /* NULL: zeros */,
                          gz_model** out);
gz_status gz_model_free(gz_model* model);
/* Upload FP32 master weights (resets the optimiser state) / read them back. Host arrays. */
gz_status gz_model_set_params(gz_model* model, const float* host_params);
gz_status gz_model_get_params(const gz_model* model, float* host_params);
/* Host arrays of GZ_MLP_FEATURES bounds; labels and margin in log10(alpha). The group tolerance
 * is 10^(prediction - margin); log10_alpha_ref is the tolerance aold was made with. */
gz_status gz_model_set_calibration(gz_model* model, const float* feat_lo, const float* feat_hi,
                                   float label_lo, float label_hi, float margin,
                                   float log10_alpha_ref);
/* Inputs enter the first layer as (x - mean) * scale (host arrays of GZ_MLP_FEATURES, FP32);
 * the default is mean 0, scale 1. */
gz_status gz_model_set_input_map(gz_model* model, const float* mean, const float* scale);
/* Device arrays: X is B x GZ_MLP_FEATURES, out/T are B. */
gz_status gz_model_forward(gz_ctx* ctx, const gz_model* model, const float* X, size_t B,
                           float* out);
/* Training loss: quantile 0 = mean squared error sum_b (y_b - T_b)^2 / B (the default);
 * 0 < quantile < 1 = pinball loss sum_b (T_b - y_b)(quantile - [T_b < y_b]) / B, which fits the
 * conditional quantile of the label. */
gz_status gz_model_set_loss(gz_model* model, float quantile);
/* Gradient of the training loss into grad_out (device, GZ_MLP_PARAMS, may be NULL) and the loss
 * into *loss_out (host, may be NULL). */
gz_status gz_model_grad(gz_ctx* ctx, gz_model* model, const float* X, const float* T, size_t B,
                        float* grad_out, float* loss_out);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the batch's gradient. */
gz_status gz_model_train_step(gz_ctx* ctx, gz_model* model, const float* X, const float* T,
                              size_t B, float lr, float* loss_out);
/* The groups gz_tree_accel_grouped would form for these targets (batch 0, group_size as given)
 * and their features: feats (device, room for n_t x GZ_MLP_FEATURES), group_start (device, room
 * for n_t + 1), *n_groups (host). aold as for the walk (required). */
gz_status gz_tree_group_features(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                 const uint32_t* targets, size_t n_t, int group_size,
                                 const double* aold, float* feats, int* group_start,
                                 size_t* n_groups);

/* ---- caller side: compute_gravity + kick (mfm.cc:679-905, 3699-3813) ---------------------- */
/* Active set in the tree's Morton order (mfm.cc:700-718). active: n_active particle indices in
 * any order (NULL with n_active == N means all active). out_targets: room for n_active. */
gz_status gz_order_targets(gz_ctx* ctx, const gz_tree* tree, const uint32_t* active,
                           size_t n_active, uint32_t* out_targets, size_t* n_out);
/* gz_order_targets with no device allocation. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under hipGraph capture, tight serving loops, harnesses
 * that forbid hipMalloc). Its flags and scan live in a fixed 65 MiB array per GPU, resolved when
 * the first context on that GPU is created and held until the process ends; a tree of more than
 * 8,388,608 (2^23) live particles is refused with GZ_ERR_INVALID. Every call on a GPU shares that
 * array, so calls that overlap on one GPU (several host threads or streams) race: serialise them.
 * Like gz_order_targets it synchronises the context's stream to return *n_out. */
gz_status gz_order_targets_noalloc(gz_ctx* ctx, const gz_tree* tree, const uint32_t* active,
                                   size_t n_active, uint32_t* out_targets, size_t* n_out);
/* aold[k] = err_tol_force_acc * |a_prev[targets[k]]| / G (mfm.cc:743-751); a_prev is N x 3
 * interleaved (Vec3d per particle, the engine's a_grav). */
gz_status gz_aold_from_accel(gz_ctx* ctx, const double* a_prev, const uint32_t* targets,
                             size_t n_t, double err_tol_force_acc, double G, double* aold);
/* a_grav[targets[k]] = (ax[k], ay[k], az[k]) (mfm.cc:828-831); a_grav N x 3 interleaved. */
gz_status gz_scatter_accel(gz_ctx* ctx, const uint32_t* targets, size_t n_t, const double* ax,
                           const double* ay, const double* az, double* a_grav);
/* KDK kick of the active set (mfm.cc:3749-3767, gravity part): for each i = active[k],
 * v[i] += a * (close_dt + open_dt) with a = a_grav[i] (+ a_extra[i] when non-NULL), both N x 3
 * interleaved; close_dt = pending_half_kick[i], open_dt = 0.5 * dt[i], and pending_half_kick[i]
 * becomes open_dt. dt, pending_half_kick and v are indexed by particle, as the reference's dt_of,
 * pending_half_kick and vx/vy/vz are (mfm.cc:3753-3764). */
gz_status gz_kick(gz_ctx* ctx, const uint32_t* active, size_t n_active, const double* a_grav,
                  const double* a_extra, const double* dt, double* pending_half_kick, double* vx,
                  double* vy, double* vz);

/* The drift (drift_particle_to, mfm.cc:1287-1295, gravity part, open boundaries): for every
 * particle with m > 0, x += v * dt; massless tombstones stay where they are. All arrays device, by
 * particle. A tree built before the drift needs gz_tree_update_sources before its next walk. */
gz_status gz_drift(gz_ctx* ctx, size_t n, double dt, const double* vx, const double* vy,
                   const double* vz, const double* m, double* x, double* y, double* z);

#ifdef __cplusplus
}
#endif
#endif /* GZ_GRAV_H */
