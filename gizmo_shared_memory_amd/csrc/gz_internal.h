// Internal declarations shared by the device library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

#include "gz_grav.h"

namespace gz {

constexpr int kLeafMax = 16;  // particles per leaf (shmem/tree.h:357)
constexpr int kMaxLevel = 20; // deepest level; keys carry 20 populated bits per axis (tree.h:358)
constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kMlpIn = 16;    // features per target group
constexpr int kMlpParams = 64 * 16 + 64 + 64 * 64 + 64 + 64 + 1; // 5313

// Packed walk node: everything one opening test and one multipole interaction need, in one
// 64-byte line (the reference's WNode, tree.h:195-204, with the threaded `next` replaced by the
// child count: children are contiguous in this layout, so a walk expands them as a range).
struct alignas(64) WNode {
    double cx, cy, cz; // centre of mass
    double s;          // size + delta: the conservative opening radius
    double mass;
    float soft;        // max softening below the node
    float len;         // side length
    int first;         // first child, or -1 for a leaf
    int nchild;        // number of children (0 for a leaf)
    int plo, phi;      // particle range in Morton order
};
static_assert(sizeof(WNode) == 64, "WNode must be one 64-byte line");

// Opening bounds of one target group (tree.cc:683-702) and its slot range in the target list.
struct GroupInfo {
    double lo[3], hi[3]; // bounding box of the targets
    double emax;         // largest target softening
    double amin;         // smallest aold (0 when there is none)
    double theta_sq;     // geometric opening parameter squared for this group
    int tlo, thi;        // target slots [tlo, thi)
};

// Everything the walk kernel reads. Passed by value.
// Per-node reuse state written by the build (tree.h:250-317); vectors are SoA with stride nn.
struct NodeVel {
    int nn;
    const double *vx, *vy, *vz; // particle velocities, or NULL (vmax = 0, no moments)
    float* vmax;
    float* xpad;
    double *vcom, *m0, *dp, *dq, *dmass;
};

struct WalkArgs {
    const WNode* wn;
    const double4* src;     // (x, y, z, m) per sorted slot
    const double* src_soft; // softening per sorted slot
    const double* node_soft; // FP64 largest softening below each node
    const uint32_t* targets;
    const int* tslot;        // sorted slot of each target (its data in src), -1 if not in the tree
    const double *x, *y, *z, *soft; // target data by particle, for targets outside the tree
    const GroupInfo* groups; // precomputed group bounds, or NULL: bounds from gstart in-kernel
    const int* gstart;       // group g is target slots [gstart[g], gstart[g+1]) (or nt at the end)
    int nt;
    const double* aold;      // per target slot, or NULL
    const float* group_scale; // per group multiplier of aold_min (learned tolerance), or NULL
    double theta_sq;
    int ngroups;
    // the group count on the device, set when the host has only an upper bound (ngroups) for it:
    // the walk then reads the count itself instead of waiting for it to be read back
    const uint32_t* ngroups_dev;
    int root;
    const uint32_t* cluster_start; // forest: particle offsets of the clusters; NULL: one tree
    int nclusters;
    int persistent;
    // tidal tensor (6 x nt, no G) and jerk (3 x nt), FP64 only; NULL when not wanted. The jerk
    // needs particle velocities (pv*, by particle) and node velocities (nvel, 3 x nnodes).
    double* tidal;
    double* jerk;
    const double *pvx, *pvy, *pvz;
    const double* nvel;
    const uint32_t* orderbuf;
    int nnodes;
    int* queue;             // persistent launch: eight work counters, one per XCD (zeroed)
    int* status;            // bit 0: a group's global node stack overflowed too
    long long* counts;      // nodes visited, multipole, pairs, max stack depth (COUNT builds)
    int* n_retry;           // groups whose LDS node stack overflowed ...
    int* retry_list;        // ... and their ids, for the global-stack launch
    int* gstack;            // global node stacks of the retry launch, gstack_cap ints per wave
    int gstack_cap;
    int* spill;             // persistent launches: spill_cap stack entries per wave, or NULL
    int spill_cap;
    int waves_per_cu;       // persistent launches: wavefronts per CU (0 = the most, 20)
    int stack_limit;        // > 0: use at most this many LDS stack entries (exercises the spill
                            // tier, and the retry when below the 512 entries one push can need)
    double G;
    double *ax, *ay, *az;
};

// The search node's half side (tree.cc:496-506) from the node's geometric centre and side: the
// float the reference stores, widened so that it is never smaller than the true one.
__device__ __forceinline__ float snode_half_of(double gx, double gy, double gz, double size,
                                               double snode_eps) {
    const double cmax = fmax(fabs(gx), fmax(fabs(gy), fabs(gz)));
    return (float)(0.5 * size + snode_eps + 4.0 * (double)FLT_EPSILON * cmax);
}

// Everything the per-target walk (shmem::accel, shmem::potential) reads. Passed by value.
// kPtDepth: the walk's stack of sibling ranges, one entry per tree level below the root.
constexpr int kPtDepth = 64;
struct PtArgs {
    const WNode* wn;
    const double4* src;      // (x, y, z, m) per sorted slot
    const double* src_soft;  // softening per sorted slot
    const uint32_t* targets;
    const int* tslot;        // sorted slot of each target, -1 if not in the tree
    const uint32_t* order;   // list positions in walk (slot) order, or NULL: the list's own order
    const double *x, *y, *z, *soft; // target data by particle, for targets outside the tree
    const double* aold;      // per list position, or NULL
    double theta_sq;
    int nt;
    int tpw;                 // targets per wavefront (a power of two, at most 64)
    int root;
    const uint32_t* cluster_start; // forest: particle offsets of the clusters; NULL: one tree
    int nclusters;
    double G;
    double *ax, *ay, *az;    // potential: ax only
};

// ---- errors ---------------------------------------------------------------------------------
void set_error(const char* fmt, ...);
gz_status fail(gz_status s, const char* fmt, ...);
gz_status hip_fail(hipError_t e, const char* what, const char* file, int line);

#define GZ_HIP(call)                                                          \
    do {                                                                      \
        hipError_t gz_e_ = (call);                                            \
        if (gz_e_ != hipSuccess) return ::gz::hip_fail(gz_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define GZ_TRY(expr)                       \
    do {                                   \
        gz_status gz_s_ = (expr);          \
        if (gz_s_ != GZ_OK) return gz_s_;  \
    } while (0)

// Grow-only device allocation.
struct DevMem {
    void* p = nullptr;
    size_t bytes = 0;
    gz_status reserve(size_t want);
    void release();
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// ---- device-wide primitives (gz_prims.hip) --------------------------------------------------
// Exclusive scan of n uint32 values (in may equal out). total_dev, when non-null, receives the
// sum (device). tmp must hold scan_tmp_bytes(n).
size_t scan_tmp_bytes(size_t n);
// starts (optional, for 0/1 flags): starts[out[i]] = i wherever in[i] is set.
gz_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_dev,
                             void* tmp, hipStream_t st, int* starts = nullptr);

// Stable LSD radix sort of (key, value) pairs by the low `bits` bits of the key, 8 bits per pass,
// skipping passes whose digit is the same for every key. The *_alt buffers are scratch of the
// same size; the pointers are swapped so that on return keys/vals hold the sorted pairs.
gz_status radix_sort_pairs(uint64_t*& keys, uint32_t*& vals, uint64_t*& keys_alt,
                           uint32_t*& vals_alt, size_t n, int bits, DevMem& scratch,
                           hipStream_t st);

// Zero n uint32 words with a kernel (no memset node, so the call is safe under stream capture).
gz_status fill_zero_u32(uint32_t* a, size_t n, hipStream_t st);

const char* last_error_cstr();
void clear_error();

// ---- walks (gz_walk.hip) ------------------------------------------------------------------------
gz_status launch_group_prep(const uint32_t* targets, size_t nt, const uint32_t* rank,
                            const uint32_t* orderbuf, size_t n_live,
                            const int* slot_leaf, int* tslot,
                            const int* parent, const WNode* wn, int nnodes, int* leaf_group,
                            int batch, int group_size,
                            size_t n, const uint32_t* seg, int nseg, int* gseg, int* gkey,
                            uint32_t* flag, uint32_t* gid, int* gstart, uint32_t* ngroups_dev,
                            int* status, void* scan_tmp, hipStream_t st);
gz_status launch_group_pack(int ngroups, size_t nt, int group_size, const int* gstart,
                            const int* gkey, const int* parent, const int* gseg, long long* pk,
                            uint32_t* mflag, uint32_t* mid, uint32_t* ngroups2_dev, int* gstart2,
                            void* scan_tmp, hipStream_t st);
gz_status launch_walk(const WalkArgs& a, bool fp64, bool count, hipStream_t st);
gz_status launch_walk_retry(const WalkArgs& a, int n_retry, bool fp64, bool count,
                            hipStream_t st);
// ---- per-target walks (gz_walk_pt.hip) ---------------------------------------------------------
// tslot, and (slot, list position) sort keys of the targets; misc[0] bit 1: a target index outside
// the particle set, misc[1] != 0: the list is not in slot order. misc must start at zero.
gz_status launch_pt_keys(const uint32_t* targets, size_t nt, const gz_tree* t, size_t n,
                         int* tslot, uint64_t* keys, uint32_t* vals, int* misc, hipStream_t st);
gz_status launch_walk_pt(const PtArgs& a, bool potential, hipStream_t st);
gz_status launch_direct(const gz_particles* p, const uint32_t* targets, size_t nt, double G,
                        int fp64, double box, double* ax, double* ay, double* az,
                        hipStream_t st);

// ---- learned opening criterion (gz_learned.hip) ----------------------------------------------------
struct GroupStats;
// The per-group sums behind the learned criterion's features (stats, one per group), and the
// features themselves (feats, ngroups x 16, when non-NULL). With ngroups_dev set, ngroups is an
// upper bound and the kernels read the count from the device (not with feats).
gz_status launch_group_features(int ngroups, int nt, const int* gstart, const uint32_t* targets,
                                const int* tslot, const gz_particles* p, const double* amag,
                                const gz_tree* t, GroupStats* stats, float* feats,
                                hipStream_t st, const uint32_t* ngroups_dev = nullptr);
// scale[g] = the learned tolerance of group g over the reference's (1 on fallback); n_fallback
// counts the groups that fell back; stats is working space for ngroups GroupStats
gz_status launch_group_tolerance(int ngroups, int nt, const int* gstart, const uint32_t* targets,
                                 const int* tslot, const gz_particles* p, const double* aold,
                                 const gz_tree* t, const gz_model* M, GroupStats* stats,
                                 float* scale, int* n_fallback, hipStream_t st,
                                 const uint32_t* ngroups_dev = nullptr);
gz_status launch_mlp_forward(const float* X, int B, const uint16_t* wbf, const float* theta,
                             const float* xnorm, float* out, hipStream_t st);
gz_status launch_mlp_grad(const float* X, const float* T, int B, const uint16_t* wbf,
                          const float* theta, const float* xnorm, float tau, float* grad,
                          float* loss, hipStream_t st);
gz_status launch_adam(float* theta, float* m, float* v, const float* grad, float lr, float b1,
                      float b2, float eps, int step, uint16_t* wbf, hipStream_t st);
gz_status launch_to_bf16(const float* theta, uint16_t* wbf, hipStream_t st);

} // namespace gz

// ---- handle types -----------------------------------------------------------------------------
struct gz_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    gz::DevMem scratch[16];
    // the walk autotuner's choices, one per workload key (gz_walk_autotune)
    std::vector<std::pair<std::vector<long long>, gz_walk_opts>> tuned;
    hipEvent_t ev[16] = {};
    // the walk's node-stack spill slices: allocated once, at a fixed size, on the first walk, so
    // that captured graphs can keep the pointer (walks on one context never overlap)
    int* spill = nullptr;
    int spill_cap = 0;
};

namespace gz {
// The fixed scratch of gz_order_targets_noalloc (gz_caller.hip): resolve its address on `device`,
// which must be current. Called once per device, when a context is created there.
gz_status resolve_noalloc_scratch(int device);

// Tree reuse (gz_reuse.hip). status bit 0: an index outside the particle set.
gz_status launch_kick_nodes(gz_tree* t, const uint32_t* active, size_t na, const double* vx,
                            const double* vy, const double* vz, const double* m, double* vlx,
                            double* vly, double* vlz, double tau, int* status, hipStream_t st);
gz_status launch_refresh(gz_tree* t, double tau, hipStream_t st);
gz_status launch_tree_drift(gz_tree* t, double dt, hipStream_t st);
gz_status launch_node_vel(const gz_tree* t, double* out, hipStream_t st);
gz_status launch_raise_soft(gz_tree* t, const uint32_t* idx, size_t n_idx, const double* soft,
                            int* status, hipStream_t st);
// status bit 1: a negative or NaN speed.
gz_status launch_raise_vmax(gz_tree* t, const uint32_t* idx, size_t n_ev, const float* speed,
                            int* status, hipStream_t st);
// Mass events (v = dm, x/y/z the position, vx/vy/vz the velocity), shifts (v = m, x/y/z the
// displacement) or pads (v = d), applied in the order of the events.
enum { kNodeEventMass = 0, kNodeEventShift = 1, kNodeEventPad = 2 };
gz_status launch_node_events(gz_tree* t, int kind, const uint32_t* idx, size_t n_ev,
                             const double* v, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz, double tau,
                             double box, int* status, hipStream_t st);

// The spill slices of ctx (allocating them on first use) and their size per wave.
gz_status ctx_spill(gz_ctx* ctx, int** spill, int* cap);
} // namespace gz

// A walk whose groups were formed once for a fixed target list; launching it enqueues only
// kernels (no host synchronisation), so launches can be captured into a hipGraph.
struct gz_walk_plan {
    gz_ctx* ctx = nullptr;
    const gz_tree* tree = nullptr;
    gz_particles p = {};
    const uint32_t* targets = nullptr;
    size_t n_t = 0;
    int ngroups = 0;
    int precision = 0;
    int persistent = 1;
    int stack_limit = 0, waves_per_cu = 0;
    gz_model* model = nullptr;
    int gstack_cap = 0, retry_waves = 0;
    void* mem = nullptr; // gstart, counters, retry list, global stacks, learned group scales
    int* gstart = nullptr;
    int* tslot = nullptr; // sorted slot of each target
    float* scale = nullptr;
    gz::GroupStats* stats = nullptr; // learned: the groups' sums (working space of a launch)
    long long* misc = nullptr;
    int* retry_list = nullptr;
    int* gstack = nullptr;
};

// The learned opening-parameter network: FP32 master weights with Adam state, their bf16 copy
// used by every forward pass, and the calibration that decides when a group falls back to the
// reference criterion. One device allocation holds all of it.
struct gz_model {
    gz_ctx* ctx = nullptr;
    void* mem = nullptr;
    float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr;
    float* frange = nullptr; // feature lower bounds [16] then upper bounds [16]
    float* xnorm = nullptr;  // input map x' = (x - mean) * scale: mean [16] then scale [16]
    uint16_t* wbf = nullptr;
    int step = 0;
    float tau = 0.f; // loss: 0 = mean squared error, (0, 1) = pinball loss of that quantile
    float lab_lo = -30.f, lab_hi = 30.f, margin = 0.f, log_alpha_ref = 0.f;
};

struct gz_tree {
    gz_ctx* ctx = nullptr;
    size_t n = 0, n_live = 0;
    int nnodes = 0, nleaves = 0, depth = 0, arena_grows = 0, root = 0;
    int cap = 0;
    double root_box[4] = {0, 0, 0, 0};
    double root_com[3] = {0, 0, 0}, root_mass = 0; // the root node's moments
    std::vector<int> level_begin; // node ids of level L are [level_begin[L], level_begin[L+1])
    int* level_begin_dev = nullptr; // the same table on the device (kMaxLevel + 2 entries)
    int n_clusters = 1;             // independent trees (a forest); their roots are nodes 0..nc-1
    uint32_t* cluster_start_dev = nullptr; // forest: particle offsets of the clusters (nc + 1)
    gz_build_times times = {};

    // particles, Morton order
    uint64_t* skey = nullptr;     // n_live sorted keys
    uint32_t* orderbuf = nullptr; // n_live particle index per sorted slot
    uint32_t* rank = nullptr;     // n: sorted slot per particle, UINT32_MAX if not in the tree
    int* leaf_of = nullptr;       // n: leaf node per particle, -1 if not in the tree
    int* slot_leaf = nullptr;     // n_live: leaf node per sorted slot
    double4* src = nullptr;       // n_live (x, y, z, m) in sorted order
    double* src_soft = nullptr;   // n_live softening in sorted order

    // nodes
    gz::WNode* wn = nullptr;
    int* next = nullptr;   // the reference's threaded successor
    int* parent = nullptr;
    int* nlo = nullptr;    // particle range during the build
    int* nhi = nullptr;
    double* gcx = nullptr; // geometric centre and side
    double* gcy = nullptr;
    double* gcz = nullptr;
    double* size = nullptr;
    double* delta = nullptr;
    double* soft = nullptr; // FP64 max softening

    // reuse between rebuilds (tree.h:250-317), one allocation: vmax and xpad always; with
    // build-time velocities the centre-of-mass velocity, the build moments (mass, COM) that
    // refreshes start from, and the incremental moments. Vectors are SoA with stride nnodes.
    void* reuse_mem = nullptr;
    float* vmax = nullptr;
    float* xpad = nullptr;
    bool has_vel = false;
    double *vcom = nullptr, *m0 = nullptr, *dp = nullptr, *dq = nullptr, *dmass = nullptr;
    // with build-time velocities: each sorted slot's velocity (vx, vy, vz, 0), as the build and
    // the kick feed last saw it, from which gz_tree_drift drifts the sorted copy
    double4* svel = nullptr;
    // the tree's age on the device (t_since_build, tree.h:250), zero at the build and advanced
    // by gz_tree_drift, for refreshes and kick feeds enqueued without the host (tau < 0)
    double* age = nullptr;
    // largest gas kernel radius below each node (tree.h:269-272), set by gz_tree_init_hmax
    float* hmax = nullptr;
};

namespace gz {
gz_status build_tree(gz_ctx* ctx, const gz_particles* p, const gz_build_opts* opts, gz_tree* t);
gz_status regather_sources(gz_ctx* ctx, const gz_particles* p, gz_tree* t);
void tree_release(gz_tree* t);
} // namespace gz
