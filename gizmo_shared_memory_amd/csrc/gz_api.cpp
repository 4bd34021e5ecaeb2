// C ABI of the tree-gravity library (include/gz_grav.h): argument checks, scratch management and
// the order of launches. The device work itself lives in gz_build.hip, gz_walk.hip and
// gz_caller.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "gz_features.h"
#include "gz_internal.h"

using gz::fail;

namespace {

gz_status check_ctx(gz_ctx* ctx) {
    if (!ctx) return fail(GZ_ERR_INVALID, "null context");
    GZ_HIP(hipSetDevice(ctx->device));
    return GZ_OK;
}

gz_status check_particles(const gz_particles* p, const char* who) {
    if (!p) return fail(GZ_ERR_INVALID, "%s: null particles", who);
    if (p->n > 0 && (!p->x || !p->y || !p->z || !p->m))
        return fail(GZ_ERR_INVALID, "%s: null position or mass array", who);
    if (p->zeta) return fail(GZ_ERR_UNSUPPORTED, "%s: zeta corrections are not supported", who);
    if (p->n >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "%s: n=%zu too large", who, p->n);
    return GZ_OK;
}

float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

// Carve consecutive, 256-byte aligned pieces out of one scratch buffer.
struct Carver {
    size_t used = 0;
    template <class T> size_t add(size_t count) {
        const size_t off = used;
        used += (count * sizeof(T) + 255) & ~size_t(255);
        return off;
    }
};

// What a walk call adds beyond the reference's arguments: the learned model, and the
// features-only mode that stops after computing the groups and their features.
struct WalkExtra {
    gz_model* model = nullptr;
    float* feats_out = nullptr; // features-only: n_groups x 16, device
    int* gstart_out = nullptr;  // features-only / groups-only: n_groups + 1 group starts, device
    int* tslot_out = nullptr;   // groups-only: the targets' sorted slots (n_t), device
    size_t* ngroups_out = nullptr;
    bool groups_only = false;   // stop after the grouping (walk plans)
    const uint32_t* seg = nullptr; // device: nseg + 1 slot offsets of independent target lists
    int nseg = 0;
    double* tidal = nullptr;       // tidal tensor out (6 x n_t), FP64 walks
    double* jerk = nullptr;        // jerk out (3 x n_t), with vx/vy/vz
    const double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    int waves_per_cu = 0;          // persistent launch width (0 = the most)
};

// The walk shared by gz_tree_accel_grouped, gz_tree_accel and gz_tree_group_features.
gz_status walk(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p, const uint32_t* targets,
               size_t n_t, double theta, double G, int batch, int group_size, const double* aold,
               int precision, int persistent,
               int count, int stack_limit, int pack, double* ax, double* ay, double* az,
               gz_walk_counts* counts, const char* who, const WalkExtra& ex = WalkExtra()) {
    const bool features_only = ex.feats_out != nullptr;
    const bool groups_only = ex.groups_only;
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "%s: null tree", who);
    GZ_TRY(check_particles(p, who));
    if (p->n != tree->n)
        return fail(GZ_ERR_INVALID, "%s: particle count %zu differs from the tree's %zu", who,
                    p->n, tree->n);
    const bool need_xyz = !features_only && !groups_only;
    if (n_t > 0 && (!targets || (need_xyz && (!ay || !az)) ||
                    (!features_only && !groups_only && !ax)))
        return fail(GZ_ERR_INVALID, "%s: null target or output array", who);
    if ((ex.model || features_only) && !aold)
        return fail(GZ_ERR_INVALID, "%s: the learned criterion needs aold", who);
    if (n_t >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "%s: n_t=%zu too large", who, n_t);
    if (precision != GZ_PREC_FP64 && precision != GZ_PREC_FP32_REL)
        return fail(GZ_ERR_INVALID, "%s: unknown precision %d", who, precision);
    if (group_size < 0 || group_size > 4096)
        return fail(GZ_ERR_INVALID, "%s: group_size %d out of range [0, 4096]", who, group_size);
    if (tree->n_clusters > 1 && batch > 1)
        return fail(GZ_ERR_INVALID, "%s: fixed windows (batch > 1) would mix trees of a forest",
                    who);
    if (counts) std::memset(counts, 0, sizeof *counts);
    if (ex.ngroups_out) *ex.ngroups_out = 0;
    if (n_t == 0) return GZ_OK;
    hipStream_t st = ctx->stream;
    hipEvent_t* ev = ctx->ev;

    // scratch: per-target group keys/flags/ids, group starts, scan temp, counters
    Carver c;
    const size_t o_key = c.add<int>(n_t), o_flag = c.add<uint32_t>(n_t),
                 o_gid = c.add<uint32_t>(n_t), o_gstart = c.add<int>(n_t + 1),
                 o_retry = c.add<int>(n_t), o_pk = c.add<long long>(n_t),
                 o_gseg = c.add<int>(ex.nseg > 1 ? n_t : 1), o_mflag = c.add<uint32_t>(n_t),
                 o_mid = c.add<uint32_t>(n_t), o_gstart2 = c.add<int>(n_t + 1),
                 o_scale = c.add<float>(ex.model ? n_t : 1), o_tslot = c.add<int>(n_t),
                 o_misc = c.add<long long>(16),
                 o_lgroup = c.add<int>(group_size > 0 ? (size_t)std::max(tree->nnodes, 1) : 1),
                 o_tmp = c.add<uint8_t>(gz::scan_tmp_bytes(n_t));
    gz::DevMem& s = ctx->scratch[5];
    GZ_TRY(s.reserve(c.used));
    uint8_t* b = s.as<uint8_t>();
    int* gkey = reinterpret_cast<int*>(b + o_key);
    uint32_t* flag = reinterpret_cast<uint32_t*>(b + o_flag);
    uint32_t* gid = reinterpret_cast<uint32_t*>(b + o_gid);
    int* gstart = reinterpret_cast<int*>(b + o_gstart);
    long long* misc = reinterpret_cast<long long*>(b + o_misc);
    uint32_t* ngroups_dev = reinterpret_cast<uint32_t*>(misc + 4);
    int* queue = reinterpret_cast<int*>(misc + 9); // eight ints: misc[9..12]
    int* status = reinterpret_cast<int*>(misc + 6);
    int* n_retry = reinterpret_cast<int*>(misc + 7);
    int* retry_list = reinterpret_cast<int*>(b + o_retry);

    GZ_HIP(hipEventRecord(ev[8], st));
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(misc), 32, st));
    int* gseg = ex.nseg > 1 ? reinterpret_cast<int*>(b + o_gseg) : nullptr;
    int* tslot = reinterpret_cast<int*>(b + o_tslot);
    GZ_TRY(gz::launch_group_prep(targets, n_t, tree->rank, tree->orderbuf, tree->n_live,
                                 tree->slot_leaf, tslot, tree->parent,
                                 tree->wn, tree->nnodes, reinterpret_cast<int*>(b + o_lgroup),
                                 batch,
                                 group_size, p->n, ex.seg, ex.nseg, gseg, gkey, flag, gid,
                                 gstart, ngroups_dev, status, b + o_tmp, st));
    // the group count (misc[4]) and the status word (misc[6]) in one read-back
    long long head[3] = {0, 0, 0};
    GZ_HIP(hipMemcpyAsync(head, misc + 4, sizeof head, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    uint32_t ngroups0 = 0;
    int prep_status = 0;
    std::memcpy(&ngroups0, &head[0], sizeof ngroups0);
    std::memcpy(&prep_status, &head[2], sizeof prep_status);
    if (prep_status & 2)
        return fail(GZ_ERR_INVALID, "%s: a target index is outside the particle set (n=%zu)", who,
                    p->n);
    uint32_t ngroups = ngroups0;
    // After packing, a persistent walk reads the group count from the device (misc[14]); the host
    // keeps ngroups0 as an upper bound and skips a read-back that would idle the GPU.
    const uint32_t* packed_dev = nullptr;
    if (group_size > 0 && pack && ngroups0 > 1) {
        uint32_t* ng2 = reinterpret_cast<uint32_t*>(misc + 14);
        int* gstart2 = reinterpret_cast<int*>(b + o_gstart2);
        GZ_TRY(gz::launch_group_pack((int)ngroups0, n_t, group_size, gstart, gkey, tree->parent,
                                     gseg, reinterpret_cast<long long*>(b + o_pk),
                                     reinterpret_cast<uint32_t*>(b + o_mflag),
                                     reinterpret_cast<uint32_t*>(b + o_mid), ng2, gstart2,
                                     b + o_tmp, st));
        if (groups_only || features_only || !persistent) {
            GZ_HIP(hipMemcpyAsync(&ngroups, ng2, sizeof ngroups, hipMemcpyDeviceToHost, st));
            GZ_HIP(hipStreamSynchronize(st));
        } else {
            packed_dev = ng2;
        }
        gstart = gstart2;
    }

    // counters, queue and status start at zero (a fill kernel rather than a memset node, so the
    // sequence stays valid under stream capture); misc[14], the packed group count, is kept
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(misc), 28, st));
    int* nfb = nullptr;
    if (groups_only) {
        GZ_HIP(hipMemcpyAsync(ex.gstart_out, gstart, ngroups * sizeof(int),
                              hipMemcpyDeviceToDevice, st));
        GZ_HIP(hipMemcpyAsync(ex.tslot_out, tslot, n_t * sizeof(int), hipMemcpyDeviceToDevice,
                              st));
        const int end = (int)n_t;
        GZ_HIP(hipMemcpyAsync(ex.gstart_out + ngroups, &end, sizeof(int), hipMemcpyHostToDevice,
                              st));
        GZ_HIP(hipStreamSynchronize(st));
        *ex.ngroups_out = ngroups;
        return GZ_OK;
    }
    if (features_only) {
        // the groups' features for training (the walk computes the same ones for itself)
        gz::DevMem& fm = ctx->scratch[6];
        GZ_TRY(fm.reserve((size_t)ngroups * sizeof(gz::GroupStats)));
        GZ_TRY(gz::launch_group_features((int)ngroups, (int)n_t, gstart, targets, tslot, p, aold,
                                         tree, fm.as<gz::GroupStats>(), ex.feats_out, st));
        GZ_HIP(hipMemcpyAsync(ex.gstart_out, gstart, ngroups * sizeof(int),
                              hipMemcpyDeviceToDevice, st));
        const int end = (int)n_t;
        GZ_HIP(hipMemcpyAsync(ex.gstart_out + ngroups, &end, sizeof(int), hipMemcpyHostToDevice,
                              st));
        GZ_HIP(hipStreamSynchronize(st));
        *ex.ngroups_out = ngroups;
        return GZ_OK;
    }
    float* scale = nullptr;
    if (ex.model) {
        // the learned tolerance of every group (features + network), before the walk
        nfb = reinterpret_cast<int*>(misc + 8);
        scale = reinterpret_cast<float*>(b + o_scale);
        gz::DevMem& fm = ctx->scratch[6];
        GZ_TRY(fm.reserve((size_t)ngroups * sizeof(gz::GroupStats)));
        GZ_TRY(gz::launch_group_tolerance((int)ngroups, (int)n_t, gstart, targets, tslot, p, aold,
                                          tree, ex.model, fm.as<gz::GroupStats>(), scale, nfb,
                                          st, packed_dev));
    }
    GZ_HIP(hipEventRecord(ev[9], st));

    gz::WalkArgs a{};
    a.wn = tree->wn;
    a.src = tree->src;
    a.src_soft = tree->src_soft;
    a.node_soft = tree->soft;
    a.targets = targets;
    a.tslot = tslot;
    a.x = p->x; a.y = p->y; a.z = p->z; a.soft = p->soft;
    a.groups = nullptr; // bounds are computed by each group's wavefront
    a.gstart = gstart;
    a.nt = (int)n_t;
    a.aold = aold;
    a.group_scale = scale;
    a.theta_sq = theta * theta;
    a.ngroups = (int)ngroups;
    a.ngroups_dev = packed_dev;
    a.root = tree->root;
    a.cluster_start = tree->cluster_start_dev;
    a.nclusters = tree->n_clusters;
    a.persistent = persistent;
    if (ex.tidal) {
        if (precision != GZ_PREC_FP64)
            return fail(GZ_ERR_UNSUPPORTED, "%s: the tidal tensor is an FP64 walk", who);
        a.tidal = ex.tidal;
        a.orderbuf = tree->orderbuf;
        a.nnodes = tree->nnodes;
        const bool jerk = ex.jerk && ex.vx && ex.vy && ex.vz && tree->has_vel;
        if (ex.jerk && !jerk) // the reference leaves the jerk zero without velocities
            GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(ex.jerk), 6 * n_t, st));
        if (jerk) {
            gz::DevMem& nv = ctx->scratch[9];
            GZ_TRY(nv.reserve(3 * (size_t)std::max(tree->nnodes, 1) * sizeof(double)));
            GZ_TRY(gz::launch_node_vel(tree, nv.as<double>(), st));
            a.jerk = ex.jerk;
            a.nvel = nv.as<double>();
            a.pvx = ex.vx;
            a.pvy = ex.vy;
            a.pvz = ex.vz;
        }
    }
    a.queue = queue;
    a.status = status;
    a.counts = misc;
    a.n_retry = n_retry;
    a.retry_list = retry_list;
    a.gstack = nullptr;
    a.gstack_cap = 0;
    a.stack_limit = stack_limit;
    a.waves_per_cu = ex.waves_per_cu;
    if (persistent) GZ_TRY(gz::ctx_spill(ctx, &a.spill, &a.spill_cap));
    a.G = G;
    a.ax = ax; a.ay = ay; a.az = az;
    const bool fp64 = precision == GZ_PREC_FP64, cnt = count && counts != nullptr && !ex.tidal;
    GZ_TRY(gz::launch_walk(a, fp64, cnt, st));
    GZ_HIP(hipEventRecord(ev[10], st));

    long long host_misc[16] = {0};
    GZ_HIP(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    const int n_fallback = (int)(host_misc[8] & 0xFFFFFFFF);
    const int retries = (int)(host_misc[7] & 0xFFFFFFFF);
    if (retries > 0) {
        // groups whose node stack outgrew LDS: walk them again with global-memory stacks. A
        // group's stack never holds more entries than the tree has nodes.
        const int cap = std::max(1024, std::min(tree->nnodes, 1 << 22));
        const int per_launch = std::max(1, (1 << 28) / cap); // at most 1 GiB of stacks
        gz::DevMem& gs = ctx->scratch[4];
        GZ_TRY(gs.reserve((size_t)std::min(per_launch, retries) * cap * sizeof(int)));
        gz::WalkArgs r = a;
        r.gstack = gs.as<int>();
        r.gstack_cap = cap;
        GZ_TRY(gz::launch_walk_retry(r, std::min(per_launch, retries), fp64, cnt, st));
        GZ_HIP(hipEventRecord(ev[10], st)); // the walk's time includes the re-walk
        GZ_HIP(hipMemcpyAsync(host_misc, misc, sizeof host_misc, hipMemcpyDeviceToHost, st));
        GZ_HIP(hipStreamSynchronize(st));
        if (host_misc[6] & 1)
            return fail(GZ_ERR_INTERNAL, "%s: a group's node stack outgrew %d entries", who, cap);
    }
    if (counts) {
        counts->fallback_groups = n_fallback;
        counts->retries = retries;
        counts->max_stack = host_misc[3];
        counts->nodes_visited = host_misc[0];
        counts->multipole = host_misc[1];
        counts->pairs = host_misc[2];
        const uint32_t ng = packed_dev ? (uint32_t)(host_misc[14] & 0xFFFFFFFF) : ngroups;
        counts->groups = ng;
        counts->items = ng;
        counts->prep_ms = ev_ms(ev[8], ev[9]);
        counts->walk_ms = ev_ms(ev[9], ev[10]);
    }
    return GZ_OK;
}

// shmem::accel and shmem::potential (tree.cc:528-582, :899-944): every target walked on its own,
// 64 Morton-neighbouring targets per wavefront (gz_walk_pt.hip). A target list that is not in
// the tree's slot order is walked in that order (sorted on the device); outputs stay by list
// position.
gz_status walk_per_target(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                          const uint32_t* targets, size_t n_t, double theta, double G,
                          const double* aold, bool potential, double* ax, double* ay, double* az,
                          const char* who) {
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "%s: null tree", who);
    GZ_TRY(check_particles(p, who));
    if (p->n != tree->n)
        return fail(GZ_ERR_INVALID, "%s: particle count %zu differs from the tree's %zu", who,
                    p->n, tree->n);
    if (n_t > 0 && (!targets || !ax || (!potential && (!ay || !az))))
        return fail(GZ_ERR_INVALID, "%s: null target or output array", who);
    if (n_t >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "%s: n_t=%zu too large", who, n_t);
    if (n_t == 0) return GZ_OK;
    if (tree->depth + 2 > gz::kPtDepth)
        return fail(GZ_ERR_INTERNAL, "%s: tree depth %d exceeds the walk's stack", who,
                    tree->depth);
    hipStream_t st = ctx->stream;
    Carver c;
    const size_t o_slot = c.add<int>(n_t), o_k0 = c.add<uint64_t>(n_t),
                 o_k1 = c.add<uint64_t>(n_t), o_v0 = c.add<uint32_t>(n_t),
                 o_v1 = c.add<uint32_t>(n_t), o_misc = c.add<int>(64);
    gz::DevMem& s = ctx->scratch[5];
    GZ_TRY(s.reserve(c.used));
    uint8_t* b = s.as<uint8_t>();
    int* tslot = reinterpret_cast<int*>(b + o_slot);
    uint64_t* k0 = reinterpret_cast<uint64_t*>(b + o_k0);
    uint64_t* k1 = reinterpret_cast<uint64_t*>(b + o_k1);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(b + o_v0);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(b + o_v1);
    int* misc = reinterpret_cast<int*>(b + o_misc);
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(misc), 4, st));
    GZ_TRY(gz::launch_pt_keys(targets, n_t, tree, p->n, tslot, k0, v0, misc, st));
    int head[2] = {0, 0};
    GZ_HIP(hipMemcpyAsync(head, misc, sizeof head, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (head[0] & 2)
        return fail(GZ_ERR_INVALID, "%s: a target index is outside the particle set (n=%zu)", who,
                    p->n);
    const uint32_t* order = nullptr;
    if (head[1]) {
        int bits = 1;
        while (bits < 40 && (1ull << bits) <= (unsigned long long)tree->n_live) ++bits;
        GZ_TRY(gz::radix_sort_pairs(k0, v0, k1, v1, n_t, bits, ctx->scratch[3], st));
        order = v0;
    }
    gz::PtArgs a{};
    a.wn = tree->wn;
    a.src = tree->src;
    a.src_soft = tree->src_soft;
    a.targets = targets;
    a.tslot = tslot;
    a.order = order;
    a.x = p->x; a.y = p->y; a.z = p->z; a.soft = p->soft;
    a.aold = aold;
    a.theta_sq = theta * theta;
    a.nt = (int)n_t;
    // 64 targets per wave while that gives 2,048 waves or more; fewer (down to one) for small
    // target sets, whose targets are often far apart
    a.tpw = 64;
    while (a.tpw > 1 && (n_t + a.tpw - 1) / a.tpw < 2048) a.tpw >>= 1;
    a.root = tree->root;
    a.cluster_start = tree->cluster_start_dev;
    a.nclusters = tree->n_clusters;
    a.G = G;
    a.ax = ax; a.ay = ay; a.az = az;
    return gz::launch_walk_pt(a, potential, st);
}

} // namespace

extern "C" {

const char* gz_last_error(void) { return gz::last_error_cstr(); }

const char* gz_status_string(gz_status s) {
    switch (s) {
    case GZ_OK: return "GZ_OK";
    case GZ_ERR_INVALID: return "GZ_ERR_INVALID";
    case GZ_ERR_HIP: return "GZ_ERR_HIP";
    case GZ_ERR_NOMEM: return "GZ_ERR_NOMEM";
    case GZ_ERR_UNSUPPORTED: return "GZ_ERR_UNSUPPORTED";
    case GZ_ERR_ARENA: return "GZ_ERR_ARENA";
    case GZ_ERR_INTERNAL: return "GZ_ERR_INTERNAL";
    }
    return "unknown gz_status";
}

gz_status gz_ctx_create(int device, void* stream, gz_ctx** out) {
    gz::clear_error();
    if (!out) return fail(GZ_ERR_INVALID, "gz_ctx_create: null out");
    *out = nullptr;
    int ndev = 0;
    GZ_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return fail(GZ_ERR_INVALID, "gz_ctx_create: device %d of %d", device, ndev);
    GZ_HIP(hipSetDevice(device));
    GZ_TRY(gz::resolve_noalloc_scratch(device));
    gz_ctx* c = new (std::nothrow) gz_ctx();
    if (!c) return fail(GZ_ERR_NOMEM, "gz_ctx_create: out of host memory");
    c->device = device;
    if (stream) {
        c->stream = static_cast<hipStream_t>(stream);
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return gz::hip_fail(e, "hipStreamCreateWithFlags", __FILE__, __LINE__);
        }
        c->own_stream = true;
    }
    for (auto& e : c->ev) {
        hipError_t r = hipEventCreate(&e);
        if (r != hipSuccess) {
            gz_ctx_destroy(c);
            return gz::hip_fail(r, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    *out = c;
    return GZ_OK;
}

gz_status gz_ctx_destroy(gz_ctx* ctx) {
    if (!ctx) return GZ_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& m : ctx->scratch) m.release();
    if (ctx->spill) (void)hipFree(ctx->spill);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GZ_OK;
}

gz_status gz_ctx_synchronize(gz_ctx* ctx) {
    GZ_TRY(check_ctx(ctx));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    return GZ_OK;
}

gz_status gz_tree_build(gz_ctx* ctx, const gz_particles* p, const gz_build_opts* opts,
                        gz_tree** out) {
    gz::clear_error();
    if (!out) return fail(GZ_ERR_INVALID, "gz_tree_build: null out");
    *out = nullptr;
    GZ_TRY(check_ctx(ctx));
    GZ_TRY(check_particles(p, "gz_tree_build"));
    gz_tree* t = new (std::nothrow) gz_tree();
    if (!t) return fail(GZ_ERR_NOMEM, "gz_tree_build: out of host memory");
    t->ctx = ctx;
    gz_status s = gz::build_tree(ctx, p, opts, t);
    if (s != GZ_OK) {
        (void)hipStreamSynchronize(ctx->stream);
        gz::tree_release(t);
        delete t;
        return s;
    }
    *out = t;
    return GZ_OK;
}

gz_status gz_tree_free(gz_tree* tree) {
    if (!tree) return GZ_OK;
    if (tree->ctx) {
        (void)hipSetDevice(tree->ctx->device);
        (void)hipStreamSynchronize(tree->ctx->stream);
    }
    gz::tree_release(tree);
    delete tree;
    return GZ_OK;
}

gz_status gz_tree_get_info(const gz_tree* tree, gz_tree_info* info) {
    if (!tree || !info) return fail(GZ_ERR_INVALID, "gz_tree_get_info: null argument");
    info->n = (long long)tree->n;
    info->n_live = (long long)tree->n_live;
    info->nnodes = tree->nnodes;
    info->nleaves = tree->nleaves;
    info->root = tree->root;
    info->depth = tree->depth;
    info->arena_grows = tree->arena_grows;
    for (int k = 0; k < 3; ++k) info->root_centre[k] = tree->root_box[k];
    info->root_side = tree->root_box[3];
    info->times = tree->times;
    return GZ_OK;
}

gz_status gz_tree_download_nodes(const gz_tree* tree, int* first, int* next, int* nchild,
                                 int* plo, int* phi, int* parent, int* level, double* size,
                                 double* mass, double* cx, double* cy, double* cz, double* soft,
                                 double* delta, double* gcx, double* gcy, double* gcz,
                                 double* wn_s, float* wn_soft, float* wn_len) {
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_download_nodes: null tree");
    GZ_TRY(check_ctx(tree->ctx));
    hipStream_t st = tree->ctx->stream;
    const size_t nn = (size_t)tree->nnodes;
    if (nn == 0) return GZ_OK;
    std::vector<gz::WNode> w(nn);
    GZ_HIP(hipMemcpyAsync(w.data(), tree->wn, nn * sizeof(gz::WNode), hipMemcpyDeviceToHost, st));
    auto dl = [&](void* dst, const void* src, size_t bytes) -> gz_status {
        if (dst) GZ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        return GZ_OK;
    };
    GZ_TRY(dl(next, tree->next, nn * sizeof(int)));
    GZ_TRY(dl(parent, tree->parent, nn * sizeof(int)));
    GZ_TRY(dl(size, tree->size, nn * sizeof(double)));
    GZ_TRY(dl(soft, tree->soft, nn * sizeof(double)));
    GZ_TRY(dl(delta, tree->delta, nn * sizeof(double)));
    GZ_TRY(dl(gcx, tree->gcx, nn * sizeof(double)));
    GZ_TRY(dl(gcy, tree->gcy, nn * sizeof(double)));
    GZ_TRY(dl(gcz, tree->gcz, nn * sizeof(double)));
    GZ_HIP(hipStreamSynchronize(st));
    for (size_t i = 0; i < nn; ++i) {
        const gz::WNode& q = w[i];
        if (first) first[i] = q.first;
        if (nchild) nchild[i] = q.nchild;
        if (plo) plo[i] = q.plo;
        if (phi) phi[i] = q.phi;
        if (mass) mass[i] = q.mass;
        if (cx) cx[i] = q.cx;
        if (cy) cy[i] = q.cy;
        if (cz) cz[i] = q.cz;
        if (wn_s) wn_s[i] = q.s;
        if (wn_soft) wn_soft[i] = q.soft;
        if (wn_len) wn_len[i] = q.len;
    }
    if (level) {
        for (size_t L = 0; L + 1 < tree->level_begin.size(); ++L)
            for (int i = tree->level_begin[L]; i < tree->level_begin[L + 1]; ++i) level[i] = (int)L;
    }
    return GZ_OK;
}

gz_status gz_tree_kick_nodes(gz_ctx* ctx, gz_tree* tree, const uint32_t* active, size_t n_active,
                             const double* vx, const double* vy, const double* vz,
                             const double* m, double* vlast_x, double* vlast_y, double* vlast_z,
                             double tau) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_kick_nodes: null tree");
    if (n_active == 0) return GZ_OK;
    if (!active || !vx || !vy || !vz)
        return fail(GZ_ERR_INVALID, "gz_tree_kick_nodes: null index or velocity array");
    const bool vl = vlast_x || vlast_y || vlast_z;
    if (vl && !(vlast_x && vlast_y && vlast_z))
        return fail(GZ_ERR_INVALID, "gz_tree_kick_nodes: v_last needs all three components");
    if (vl && !m) return fail(GZ_ERR_INVALID, "gz_tree_kick_nodes: the momentum feed needs m");
    hipStream_t st = ctx->stream;
    gz::DevMem& sm = ctx->scratch[8];
    GZ_TRY(sm.reserve(256));
    int* status = sm.as<int>();
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(status), 1, st));
    GZ_TRY(gz::launch_kick_nodes(tree, active, n_active, vx, vy, vz, m, vlast_x, vlast_y, vlast_z,
                                 tau, status, st));
    // Under stream capture the index check cannot be read back: the kernel still skips an index
    // outside the particle set, but the call cannot report it.
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    GZ_HIP(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone) return GZ_OK;
    int h = 0;
    GZ_HIP(hipMemcpyAsync(&h, status, sizeof h, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (h & 1)
        return fail(GZ_ERR_INVALID, "gz_tree_kick_nodes: an index is outside the particle set");
    return GZ_OK;
}

gz_status gz_tree_refresh_walk_nodes(gz_ctx* ctx, gz_tree* tree, double tau) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_refresh_walk_nodes: null tree");
    return gz::launch_refresh(tree, tau, ctx->stream);
}

gz_status gz_tree_drift(gz_ctx* ctx, gz_tree* tree, double dt) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_drift: null tree");
    if (!tree->has_vel || !tree->svel)
        return fail(GZ_ERR_INVALID, "gz_tree_drift: the tree was built without velocities");
    if (!std::isfinite(dt)) return fail(GZ_ERR_INVALID, "gz_tree_drift: dt is not finite");
    return gz::launch_tree_drift(tree, dt, ctx->stream);
}

gz_status gz_tree_raise_soft(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                             const double* soft) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_raise_soft: null tree");
    if (n == 0) return GZ_OK;
    if (!idx || !soft) return fail(GZ_ERR_INVALID, "gz_tree_raise_soft: null array");
    hipStream_t st = ctx->stream;
    gz::DevMem& sm = ctx->scratch[8];
    GZ_TRY(sm.reserve(256));
    int* status = sm.as<int>();
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(status), 1, st));
    GZ_TRY(gz::launch_raise_soft(tree, idx, n, soft, status, st));
    int h = 0;
    GZ_HIP(hipMemcpyAsync(&h, status, sizeof h, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (h & 1) return fail(GZ_ERR_INVALID, "gz_tree_raise_soft: an index is outside the particles");
    return GZ_OK;
}

namespace {
// The per-call status word of the reuse kernels: cleared before the launch, read back after it
// unless the stream is being captured (then the kernels still skip bad entries, unreported).
struct ReuseStatus {
    int* dev = nullptr;
    gz_status begin(gz_ctx* ctx) {
        gz::DevMem& sm = ctx->scratch[8];
        GZ_TRY(sm.reserve(256));
        dev = sm.as<int>();
        return gz::fill_zero_u32(reinterpret_cast<uint32_t*>(dev), 1, ctx->stream);
    }
    gz_status end(gz_ctx* ctx, int* bits) {
        *bits = 0;
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        GZ_HIP(hipStreamIsCapturing(ctx->stream, &cap));
        if (cap != hipStreamCaptureStatusNone) return GZ_OK;
        GZ_HIP(hipMemcpyAsync(bits, dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        GZ_HIP(hipStreamSynchronize(ctx->stream));
        return GZ_OK;
    }
};

gz_status node_events(const char* name, gz_ctx* ctx, gz_tree* tree, int kind,
                      const uint32_t* idx, size_t n, const double* v, const double* x,
                      const double* y, const double* z, const double* vx, const double* vy,
                      const double* vz, double tau, double box) {
    ReuseStatus rs;
    GZ_TRY(rs.begin(ctx));
    GZ_TRY(gz::launch_node_events(tree, kind, idx, n, v, x, y, z, vx, vy, vz, tau, box, rs.dev,
                                  ctx->stream));
    int bits = 0;
    GZ_TRY(rs.end(ctx, &bits));
    if (bits & 1)
        return fail(GZ_ERR_INVALID, "%s: an index is outside the particle set", name);
    return GZ_OK;
}
} // namespace

gz_status gz_tree_mass_events(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                              const double* dm, const double* x, const double* y, const double* z,
                              const double* vx, const double* vy, const double* vz, double tau,
                              double box) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_mass_events: null tree");
    if (n == 0 || !tree->has_vel) return GZ_OK; // the reference's no-op without vcom
    if (!idx || !dm || !x || !y || !z || !vx || !vy || !vz)
        return fail(GZ_ERR_INVALID, "gz_tree_mass_events: null array");
    if (!std::isfinite(box)) return fail(GZ_ERR_INVALID, "gz_tree_mass_events: box is not finite");
    return node_events("gz_tree_mass_events", ctx, tree, gz::kNodeEventMass, idx, n, dm, x, y, z,
                       vx, vy, vz, tau, box);
}

gz_status gz_tree_shift_nodes(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                              const double* m, const double* dx, const double* dy,
                              const double* dz) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_shift_nodes: null tree");
    if (n == 0 || !tree->has_vel) return GZ_OK; // the reference's no-op without vcom
    if (!idx || !m || !dx || !dy || !dz)
        return fail(GZ_ERR_INVALID, "gz_tree_shift_nodes: null array");
    return node_events("gz_tree_shift_nodes", ctx, tree, gz::kNodeEventShift, idx, n, m, dx, dy,
                       dz, nullptr, nullptr, nullptr, 0.0, 0.0);
}

gz_status gz_tree_raise_pad(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                            const double* d) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_raise_pad: null tree");
    if (n == 0) return GZ_OK;
    if (!idx || !d) return fail(GZ_ERR_INVALID, "gz_tree_raise_pad: null array");
    return node_events("gz_tree_raise_pad", ctx, tree, gz::kNodeEventPad, idx, n, d, nullptr,
                       nullptr, nullptr, nullptr, nullptr, nullptr, 0.0, 0.0);
}

gz_status gz_tree_raise_vmax(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                             const float* speed) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_raise_vmax: null tree");
    if (n == 0) return GZ_OK;
    if (!idx || !speed) return fail(GZ_ERR_INVALID, "gz_tree_raise_vmax: null array");
    ReuseStatus rs;
    GZ_TRY(rs.begin(ctx));
    GZ_TRY(gz::launch_raise_vmax(tree, idx, n, speed, rs.dev, ctx->stream));
    int bits = 0;
    GZ_TRY(rs.end(ctx, &bits));
    if (bits & 1)
        return fail(GZ_ERR_INVALID, "gz_tree_raise_vmax: an index is outside the particle set");
    if (bits & 2) return fail(GZ_ERR_INVALID, "gz_tree_raise_vmax: a speed is negative or NaN");
    return GZ_OK;
}

gz_status gz_tree_download_pads(const gz_tree* tree, float* xpad, double* dmass) {
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_download_pads: null tree");
    GZ_TRY(check_ctx(tree->ctx));
    hipStream_t st = tree->ctx->stream;
    const size_t nn = (size_t)tree->nnodes;
    if (nn == 0) return GZ_OK;
    if (xpad)
        GZ_HIP(hipMemcpyAsync(xpad, tree->xpad, nn * sizeof(float), hipMemcpyDeviceToHost, st));
    if (dmass) {
        if (tree->dmass)
            GZ_HIP(hipMemcpyAsync(dmass, tree->dmass, nn * sizeof(double), hipMemcpyDeviceToHost,
                                  st));
        else
            std::memset(dmass, 0, nn * sizeof(double));
    }
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

gz_status gz_tree_download_reuse(const gz_tree* tree, float* vmax, double* vcom, double* dp,
                                 double* dq) {
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_download_reuse: null tree");
    GZ_TRY(check_ctx(tree->ctx));
    hipStream_t st = tree->ctx->stream;
    const size_t nn = (size_t)tree->nnodes;
    if (nn == 0) return GZ_OK;
    if (vmax) GZ_HIP(hipMemcpyAsync(vmax, tree->vmax, nn * sizeof(float), hipMemcpyDeviceToHost, st));
    std::vector<double> h(3 * nn);
    const double* srcs[3] = {tree->vcom, tree->dp, tree->dq};
    double* dsts[3] = {vcom, dp, dq};
    for (int k = 0; k < 3; ++k) {
        if (!dsts[k]) continue;
        if (!srcs[k]) {
            std::memset(dsts[k], 0, 3 * nn * sizeof(double));
            continue;
        }
        GZ_HIP(hipMemcpyAsync(h.data(), srcs[k], 3 * nn * sizeof(double), hipMemcpyDeviceToHost,
                              st));
        GZ_HIP(hipStreamSynchronize(st));
        for (size_t i = 0; i < nn; ++i)
            for (int d = 0; d < 3; ++d) dsts[k][3 * i + d] = h[d * nn + i];
    }
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

gz_status gz_tree_download_particles(const gz_tree* tree, uint32_t* orderbuf, uint32_t* rank,
                                     int* leaf_of, uint64_t* skey) {
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_download_particles: null tree");
    GZ_TRY(check_ctx(tree->ctx));
    hipStream_t st = tree->ctx->stream;
    if (orderbuf && tree->n_live)
        GZ_HIP(hipMemcpyAsync(orderbuf, tree->orderbuf, tree->n_live * sizeof(uint32_t),
                              hipMemcpyDeviceToHost, st));
    if (skey && tree->n_live)
        GZ_HIP(hipMemcpyAsync(skey, tree->skey, tree->n_live * sizeof(uint64_t),
                              hipMemcpyDeviceToHost, st));
    if (rank && tree->n)
        GZ_HIP(hipMemcpyAsync(rank, tree->rank, tree->n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              st));
    if (leaf_of && tree->n)
        GZ_HIP(hipMemcpyAsync(leaf_of, tree->leaf_of, tree->n * sizeof(int),
                              hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

gz_status gz_tree_device_arrays(const gz_tree* tree, const uint32_t** orderbuf,
                                const uint32_t** rank, const int** leaf_of) {
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_device_arrays: null tree");
    if (orderbuf) *orderbuf = tree->orderbuf;
    if (rank) *rank = tree->rank;
    if (leaf_of) *leaf_of = tree->leaf_of;
    return GZ_OK;
}

gz_status gz_tree_update_sources(gz_ctx* ctx, gz_tree* tree, const gz_particles* p) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree) return fail(GZ_ERR_INVALID, "gz_tree_update_sources: null tree");
    GZ_TRY(check_particles(p, "gz_tree_update_sources"));
    return gz::regather_sources(ctx, p, tree);
}

gz_status gz_tree_accel_grouped(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                const uint32_t* targets, size_t n_t, double theta, double G,
                                int batch, const double* aold, double* ax, double* ay, double* az,
                                const gz_walk_opts* opts, gz_walk_counts* counts) {
    gz::clear_error();
    int precision = GZ_PREC_FP64, group_size = 0, persistent = 1, count = 0, stack_limit = 0;
    WalkExtra ex;
    if (opts) {
        precision = opts->precision;
        group_size = opts->group_size;
        if (opts->launch != GZ_LAUNCH_QUEUE && opts->launch != GZ_LAUNCH_PER_GROUP)
            return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped: unknown launch %d", opts->launch);
        persistent = opts->launch == GZ_LAUNCH_QUEUE;
        count = opts->count;
        stack_limit = opts->stack_limit;
        if (opts->waves_per_cu < 0 || opts->waves_per_cu > 20)
            return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped: waves_per_cu %d out of [0, 20]",
                        opts->waves_per_cu);
        ex.waves_per_cu = opts->waves_per_cu;
        if (opts->criterion == GZ_CRIT_LEARNED) {
            if (!opts->model)
                return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped: learned criterion without a model");
            ex.model = static_cast<gz_model*>(const_cast<void*>(opts->model));
        } else if (opts->criterion != GZ_CRIT_REFERENCE)
            return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped: unknown criterion %d",
                        opts->criterion);
    }
    if (batch > 0 && group_size > 0)
        return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped: batch > 0 and group_size > 0");
    return walk(ctx, tree, p, targets, n_t, theta, G, batch, group_size, aold, precision,
                persistent, count, stack_limit, 1, ax, ay, az, counts,
                "gz_tree_accel_grouped", ex);
}

// ---- walk autotuner ----------------------------------------------------------------------------
extern "C++" { // helpers with C++ types, inside the C-linkage block
namespace {

int ilog2_bucket(long long v) {
    int b = 0;
    while (v > 1) {
        v >>= 1;
        ++b;
    }
    return b;
}

// The workload a tuned choice is kept for: what decides which launch parameters win.
std::vector<long long> tune_key(const gz_tree* tree, size_t n_t, double theta,
                                const gz_walk_opts& o) {
    return {o.precision, o.criterion, (long long)reinterpret_cast<uintptr_t>(o.model),
            o.launch, theta > 0.0 ? 1 : 0, ilog2_bucket((long long)n_t),
            ilog2_bucket((long long)tree->n_live), ilog2_bucket((long long)tree->nnodes),
            tree->n_clusters > 1 ? 1 : 0};
}

} // namespace
} // extern "C++"

gz_status gz_walk_autotune(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                           const uint32_t* targets, size_t n_t, double theta, double G,
                           const double* aold, const gz_walk_opts* base, int reps, int flags,
                           gz_walk_opts* best, gz_tune_report* report) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!tree || !p || !best) return fail(GZ_ERR_INVALID, "gz_walk_autotune: null argument");
    if (n_t > 0 && !targets) return fail(GZ_ERR_INVALID, "gz_walk_autotune: null targets");
    gz_walk_opts b0{};
    if (base) b0 = *base;
    b0.count = 0;
    if (b0.criterion == GZ_CRIT_LEARNED && !aold)
        return fail(GZ_ERR_INVALID, "gz_walk_autotune: the learned criterion needs aold");
    if (reps <= 0) reps = 2;
    if (report) std::memset(report, 0, sizeof *report);
    const std::vector<long long> key = tune_key(tree, n_t, theta, b0);
    if (!(flags & GZ_TUNE_RETUNE)) {
        for (const auto& e : ctx->tuned)
            if (e.first == key) {
                *best = e.second;
                best->model = b0.model;
                if (report) {
                    report->from_cache = 1;
                    report->best = {best->group_size, best->stack_limit, best->waves_per_cu, 0.0};
                }
                return GZ_OK;
            }
    }
    *best = b0;
    if (n_t == 0) return GZ_OK;
    // scratch outputs: the caller's arrays are left alone
    gz::DevMem& out = ctx->scratch[12];
    GZ_TRY(out.reserve(3 * n_t * sizeof(double)));
    double* ax = out.as<double>();
    int np = 0;
    gz_tune_point pts[GZ_TUNE_MAX_POINTS];
    auto time_one = [&](const gz_walk_opts& o, double* ms) -> gz_status {
        gz_walk_counts c;
        double sum = 0.0;
        for (int r = 0; r <= reps; ++r) {
            GZ_TRY(gz_tree_accel_grouped(ctx, tree, p, targets, n_t, theta, G, 0, aold, ax,
                                         ax + n_t, ax + 2 * n_t, &o, &c));
            if (r > 0) sum += c.walk_ms + c.prep_ms;
        }
        *ms = sum / reps;
        if (np < GZ_TUNE_MAX_POINTS)
            pts[np++] = {o.group_size, o.stack_limit, o.waves_per_cu, *ms};
        return GZ_OK;
    };
    gz_walk_opts cur = b0;
    cur.stack_limit = 0;
    cur.waves_per_cu = 0;
    double best_ms = 1e300;
    // 1. target-group size at the default stack and width
    std::vector<int> sizes = {16, 32, 64};
    if (b0.precision == GZ_PREC_FP64) sizes.insert(sizes.begin(), 0);
    int best_gs = sizes[0];
    for (int gs : sizes) {
        gz_walk_opts o = cur;
        o.group_size = gs;
        double ms = 0;
        GZ_TRY(time_one(o, &ms));
        if (ms < best_ms) {
            best_ms = ms;
            best_gs = gs;
        }
    }
    cur.group_size = best_gs;
    // 2. the on-chip node-stack depth, then 3. the persistent launch width (queue launches)
    if (cur.launch == GZ_LAUNCH_QUEUE) {
        int best_sl = 0;
        for (int sl : {384, 448}) {
            gz_walk_opts o = cur;
            o.stack_limit = sl;
            double ms = 0;
            GZ_TRY(time_one(o, &ms));
            if (ms < best_ms) {
                best_ms = ms;
                best_sl = sl;
            }
        }
        cur.stack_limit = best_sl;
        int best_w = 0;
        for (int w : {10, 15}) {
            gz_walk_opts o = cur;
            o.waves_per_cu = w;
            double ms = 0;
            GZ_TRY(time_one(o, &ms));
            if (ms < best_ms) {
                best_ms = ms;
                best_w = w;
            }
        }
        cur.waves_per_cu = best_w;
    }
    *best = cur;
    bool stored = false;
    for (auto& e : ctx->tuned)
        if (e.first == key) {
            e.second = cur;
            stored = true;
        }
    if (!stored) ctx->tuned.emplace_back(key, cur);
    if (report) {
        report->n_points = np;
        std::memcpy(report->points, pts, (size_t)np * sizeof(gz_tune_point));
        report->best = {cur.group_size, cur.stack_limit, cur.waves_per_cu, best_ms};
    }
    return GZ_OK;
}

gz_status gz_tree_group_features(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                 const uint32_t* targets, size_t n_t, int group_size,
                                 const double* aold, float* feats, int* group_start,
                                 size_t* n_groups) {
    gz::clear_error();
    if (!feats || !group_start || !n_groups)
        return fail(GZ_ERR_INVALID, "gz_tree_group_features: null output");
    WalkExtra ex;
    ex.feats_out = feats;
    ex.gstart_out = group_start;
    ex.ngroups_out = n_groups;
    return walk(ctx, tree, p, targets, n_t, 0.0, 1.0, 0, group_size, aold, GZ_PREC_FP32_REL,
                0, 0, 0, 1, nullptr, nullptr, nullptr, nullptr, "gz_tree_group_features", ex);
}

// ---- walk plans ---------------------------------------------------------------------------------
gz_status gz_walk_plan_create(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                              const uint32_t* targets, size_t n_t, const size_t* seg_start,
                              int n_seg, const gz_walk_opts* opts, gz_walk_plan** out) {
    gz::clear_error();
    if (!out) return fail(GZ_ERR_INVALID, "gz_walk_plan_create: null out");
    *out = nullptr;
    GZ_TRY(check_ctx(ctx));
    if (!tree || !p) return fail(GZ_ERR_INVALID, "gz_walk_plan_create: null tree or particles");
    if (n_seg < 0 || (n_seg > 1 && !seg_start))
        return fail(GZ_ERR_INVALID, "gz_walk_plan_create: bad segments");
    int precision = GZ_PREC_FP64, group_size = 0, persistent = 1;
    gz_model* model = nullptr;
    if (opts) {
        precision = opts->precision;
        group_size = opts->group_size;
        persistent = opts->launch == GZ_LAUNCH_QUEUE;
        if (opts->count) return fail(GZ_ERR_UNSUPPORTED, "gz_walk_plan_create: counting");
        if (opts->waves_per_cu < 0 || opts->waves_per_cu > 20 || opts->stack_limit < 0)
            return fail(GZ_ERR_INVALID, "gz_walk_plan_create: waves_per_cu or stack_limit out of "
                                        "range");
        if (opts->criterion == GZ_CRIT_LEARNED) {
            if (!opts->model) return fail(GZ_ERR_INVALID, "gz_walk_plan_create: no model");
            model = static_cast<gz_model*>(const_cast<void*>(opts->model));
        }
    }
    if (n_seg > 1) {
        if (seg_start[0] != 0 || seg_start[n_seg] != n_t)
            return fail(GZ_ERR_INVALID, "gz_walk_plan_create: segments must cover [0, n_t)");
        for (int k = 0; k < n_seg; ++k)
            if (seg_start[k + 1] < seg_start[k])
                return fail(GZ_ERR_INVALID, "gz_walk_plan_create: segments must be ordered");
    }
    gz_walk_plan* P = new (std::nothrow) gz_walk_plan();
    if (!P) return fail(GZ_ERR_NOMEM, "gz_walk_plan_create: out of host memory");
    P->ctx = ctx;
    P->tree = tree;
    P->p = *p;
    P->targets = targets;
    P->n_t = n_t;
    P->precision = precision;
    P->persistent = persistent;
    P->model = model;
    P->stack_limit = opts ? opts->stack_limit : 0;
    P->waves_per_cu = opts ? opts->waves_per_cu : 0;
    // device memory: group starts (n_t + 1), counters (16 x i64), retry list (n_t), segment
    // offsets, and the global stacks of the re-walk
    P->gstack_cap = std::max(1024, std::min(tree->nnodes, 1 << 18));
    P->retry_waves = 64;
    Carver c;
    const size_t o_gs = c.add<int>(n_t + 1), o_misc = c.add<long long>(16),
                 o_rl = c.add<int>(n_t + 1), o_seg = c.add<uint32_t>(n_seg + 2),
                 o_sc = c.add<float>(model ? n_t : 1), o_ts = c.add<int>(n_t),
                 o_stk = c.add<int>((size_t)P->retry_waves * P->gstack_cap);
    hipError_t e = hipMalloc(&P->mem, c.used);
    if (e != hipSuccess) {
        delete P;
        return gz::hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    uint8_t* b = static_cast<uint8_t*>(P->mem);
    P->gstart = reinterpret_cast<int*>(b + o_gs);
    P->misc = reinterpret_cast<long long*>(b + o_misc);
    P->retry_list = reinterpret_cast<int*>(b + o_rl);
    P->gstack = reinterpret_cast<int*>(b + o_stk);
    P->scale = reinterpret_cast<float*>(b + o_sc);
    P->tslot = reinterpret_cast<int*>(b + o_ts);
    uint32_t* seg = reinterpret_cast<uint32_t*>(b + o_seg);
    gz_status s = GZ_OK;
    if (n_seg > 1) {
        std::vector<uint32_t> h(n_seg + 1);
        for (int k = 0; k <= n_seg; ++k) h[k] = (uint32_t)seg_start[k];
        if (hipMemcpy(seg, h.data(), h.size() * sizeof(uint32_t), hipMemcpyHostToDevice) !=
            hipSuccess)
            s = fail(GZ_ERR_HIP, "gz_walk_plan_create: upload failed");
    }
    // the spill slices are allocated here, never during a launch (which may be captured)
    if (s == GZ_OK && persistent) {
        int* sp = nullptr;
        int cap = 0;
        s = gz::ctx_spill(ctx, &sp, &cap);
    }
    if (s == GZ_OK && n_t > 0) {
        WalkExtra ex;
        ex.groups_only = true;
        ex.gstart_out = P->gstart;
        ex.tslot_out = P->tslot;
        size_t ng = 0;
        ex.ngroups_out = &ng;
        ex.seg = seg;
        ex.nseg = n_seg;
        s = walk(ctx, tree, p, targets, n_t, 0.0, 1.0, 0, group_size, nullptr, precision,
                 persistent, 0, 0, 1, nullptr, nullptr, nullptr, nullptr,
                 "gz_walk_plan_create", ex);
        P->ngroups = (int)ng;
    }
    if (s == GZ_OK && model && P->ngroups > 0) {
        void* f = nullptr;
        if (hipMalloc(&f, (size_t)P->ngroups * sizeof(gz::GroupStats)) != hipSuccess)
            s = fail(GZ_ERR_NOMEM, "gz_walk_plan_create: no device memory for group features");
        P->stats = static_cast<gz::GroupStats*>(f);
    }
    if (s != GZ_OK) {
        gz_walk_plan_free(P);
        return s;
    }
    *out = P;
    return GZ_OK;
}

gz_status gz_walk_plan_launch(gz_ctx* ctx, gz_walk_plan* P, double theta, double G,
                              const double* aold, double* ax, double* ay, double* az) {
    gz::clear_error();
    if (!ctx || !P) return fail(GZ_ERR_INVALID, "gz_walk_plan_launch: null argument");
    if (P->n_t > 0 && (!ax || !ay || !az))
        return fail(GZ_ERR_INVALID, "gz_walk_plan_launch: null output");
    if (P->model && !aold)
        return fail(GZ_ERR_INVALID, "gz_walk_plan_launch: the learned criterion needs aold");
    if (P->n_t == 0) return GZ_OK;
    hipStream_t st = ctx->stream;
    const gz_tree* tree = P->tree;
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(P->misc), 32, st));
    gz::WalkArgs a{};
    a.wn = tree->wn;
    a.src = tree->src;
    a.src_soft = tree->src_soft;
    a.node_soft = tree->soft;
    a.targets = P->targets;
    a.tslot = P->tslot;
    a.x = P->p.x; a.y = P->p.y; a.z = P->p.z; a.soft = P->p.soft;
    a.groups = nullptr;
    a.gstart = P->gstart;
    a.nt = (int)P->n_t;
    a.aold = aold;
    a.group_scale = nullptr;
    a.theta_sq = theta * theta;
    a.ngroups = P->ngroups;
    a.root = tree->root;
    a.cluster_start = tree->cluster_start_dev;
    a.nclusters = tree->n_clusters;
    a.persistent = P->persistent;
    a.stack_limit = P->stack_limit;
    a.waves_per_cu = P->waves_per_cu;
    a.queue = reinterpret_cast<int*>(P->misc + 9); // eight ints: misc[9..12]
    a.status = reinterpret_cast<int*>(P->misc + 6);
    a.counts = P->misc;
    a.n_retry = reinterpret_cast<int*>(P->misc + 7);
    a.retry_list = P->retry_list;
    if (P->persistent) GZ_TRY(gz::ctx_spill(ctx, &a.spill, &a.spill_cap));
    a.G = G;
    a.ax = ax; a.ay = ay; a.az = az;
    if (P->model) {
        a.group_scale = P->scale;
        GZ_TRY(gz::launch_group_tolerance(P->ngroups, (int)P->n_t, P->gstart, P->targets, P->tslot,
                                          &P->p, aold, tree, P->model, P->stats, P->scale,
                                          reinterpret_cast<int*>(P->misc + 8), st));
    }
    const bool fp64 = P->precision == GZ_PREC_FP64;
    GZ_TRY(gz::launch_walk(a, fp64, false, st));
    // the re-walk of groups whose node stack outgrew LDS; it exits at once when there are none
    gz::WalkArgs r = a;
    r.gstack = P->gstack;
    r.gstack_cap = P->gstack_cap;
    GZ_TRY(gz::launch_walk_retry(r, P->retry_waves, fp64, false, st));
    return GZ_OK;
}

gz_status gz_walk_plan_status(gz_ctx* ctx, gz_walk_plan* P, gz_walk_counts* counts) {
    if (!ctx || !P) return fail(GZ_ERR_INVALID, "gz_walk_plan_status: null argument");
    long long h[16] = {0};
    GZ_HIP(hipMemcpyAsync(h, P->misc, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    if (counts) {
        std::memset(counts, 0, sizeof *counts);
        counts->groups = P->ngroups;
        counts->items = P->ngroups;
        counts->retries = h[7] & 0xFFFFFFFF;
        counts->fallback_groups = h[8] & 0xFFFFFFFF;
    }
    if (h[6] & 1)
        return fail(GZ_ERR_INTERNAL, "gz_walk_plan: a group's node stack outgrew %d entries",
                    P->gstack_cap);
    return GZ_OK;
}

gz_status gz_walk_plan_free(gz_walk_plan* P) {
    if (!P) return GZ_OK;
    if (P->mem || P->stats) (void)hipStreamSynchronize(P->ctx->stream);
    if (P->mem) (void)hipFree(P->mem);
    if (P->stats) (void)hipFree(P->stats);
    delete P;
    return GZ_OK;
}

// ---- graphs ---------------------------------------------------------------------------------------
} // extern "C"
struct gz_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};
extern "C" {

gz_status gz_graph_begin(gz_ctx* ctx) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    GZ_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    return GZ_OK;
}

gz_status gz_graph_end(gz_ctx* ctx, gz_graph** out) {
    gz::clear_error();
    if (!out) return fail(GZ_ERR_INVALID, "gz_graph_end: null out");
    *out = nullptr;
    GZ_TRY(check_ctx(ctx));
    gz_graph* g = new (std::nothrow) gz_graph();
    if (!g) return fail(GZ_ERR_NOMEM, "gz_graph_end: out of host memory");
    hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        gz_graph_free(g);
        return gz::hip_fail(e, "graph capture", __FILE__, __LINE__);
    }
    *out = g;
    return GZ_OK;
}

gz_status gz_graph_launch(gz_ctx* ctx, gz_graph* g) {
    gz::clear_error();
    if (!ctx || !g) return fail(GZ_ERR_INVALID, "gz_graph_launch: null argument");
    GZ_HIP(hipGraphLaunch(g->exec, ctx->stream));
    return GZ_OK;
}

gz_status gz_graph_free(gz_graph* g) {
    if (!g) return GZ_OK;
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    delete g;
    return GZ_OK;
}

// ---- the learned model ------------------------------------------------------------------------------
gz_status gz_model_create(gz_ctx* ctx, const float* host_params, gz_model** out) {
    gz::clear_error();
    if (!out) return fail(GZ_ERR_INVALID, "gz_model_create: null out");
    *out = nullptr;
    GZ_TRY(check_ctx(ctx));
    gz_model* M = new (std::nothrow) gz_model();
    if (!M) return fail(GZ_ERR_NOMEM, "gz_model_create: out of host memory");
    M->ctx = ctx;
    const size_t nf = (size_t)gz::kMlpParams;
    const size_t bytes = 4 * nf * sizeof(float) + 256 + 4 * gz::kMlpIn * sizeof(float) +
                         nf * sizeof(uint16_t) + 1024;
    hipError_t e = hipMalloc(&M->mem, bytes);
    if (e != hipSuccess) {
        delete M;
        return gz::hip_fail(e, "hipMalloc", __FILE__, __LINE__);
    }
    float* f = static_cast<float*>(M->mem);
    M->theta = f;
    M->m = f + nf;
    M->v = f + 2 * nf;
    M->grad = f + 3 * nf;
    M->loss = f + 4 * nf;
    M->frange = f + 4 * nf + 64;
    M->xnorm = M->frange + 2 * gz::kMlpIn;
    M->wbf = reinterpret_cast<uint16_t*>(M->xnorm + 2 * gz::kMlpIn);
    hipStream_t st = ctx->stream;
    gz_status s = gz::fill_zero_u32(reinterpret_cast<uint32_t*>(f), 4 * nf + 64, st);
    if (s == GZ_OK && host_params) s = gz_model_set_params(M, host_params);
    if (s == GZ_OK) {
        float fr[4 * gz::kMlpIn]; // open feature ranges, identity input map
        for (int k = 0; k < gz::kMlpIn; ++k) {
            fr[k] = -1e30f;
            fr[gz::kMlpIn + k] = 1e30f;
            fr[2 * gz::kMlpIn + k] = 0.f;
            fr[3 * gz::kMlpIn + k] = 1.f;
        }
        if (hipMemcpyAsync(M->frange, fr, sizeof fr, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            s = fail(GZ_ERR_HIP, "gz_model_create: upload failed");
    }
    if (s != GZ_OK) {
        gz_model_free(M);
        return s;
    }
    *out = M;
    return GZ_OK;
}

gz_status gz_model_free(gz_model* M) {
    if (!M) return GZ_OK;
    if (M->mem) {
        (void)hipStreamSynchronize(M->ctx->stream);
        (void)hipFree(M->mem);
    }
    delete M;
    return GZ_OK;
}

gz_status gz_model_set_params(gz_model* M, const float* host_params) {
    if (!M || !host_params) return fail(GZ_ERR_INVALID, "gz_model_set_params: null argument");
    hipStream_t st = M->ctx->stream;
    const size_t nf = (size_t)gz::kMlpParams;
    GZ_HIP(hipMemcpyAsync(M->theta, host_params, nf * sizeof(float), hipMemcpyHostToDevice, st));
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(M->m), 2 * nf, st)); // Adam state
    GZ_TRY(gz::launch_to_bf16(M->theta, M->wbf, st));
    GZ_HIP(hipStreamSynchronize(st));
    M->step = 0;
    return GZ_OK;
}

gz_status gz_model_get_params(const gz_model* M, float* host_params) {
    if (!M || !host_params) return fail(GZ_ERR_INVALID, "gz_model_get_params: null argument");
    hipStream_t st = M->ctx->stream;
    GZ_HIP(hipMemcpyAsync(host_params, M->theta, gz::kMlpParams * sizeof(float),
                          hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

gz_status gz_model_set_calibration(gz_model* M, const float* feat_lo, const float* feat_hi,
                                   float label_lo, float label_hi, float margin,
                                   float log10_alpha_ref) {
    if (!M || !feat_lo || !feat_hi)
        return fail(GZ_ERR_INVALID, "gz_model_set_calibration: null argument");
    float fr[2 * gz::kMlpIn];
    for (int k = 0; k < gz::kMlpIn; ++k) {
        fr[k] = feat_lo[k];
        fr[gz::kMlpIn + k] = feat_hi[k];
    }
    hipStream_t st = M->ctx->stream;
    GZ_HIP(hipMemcpyAsync(M->frange, fr, sizeof fr, hipMemcpyHostToDevice, st));
    GZ_HIP(hipStreamSynchronize(st));
    M->lab_lo = label_lo;
    M->lab_hi = label_hi;
    M->margin = margin;
    M->log_alpha_ref = log10_alpha_ref;
    return GZ_OK;
}

gz_status gz_model_set_input_map(gz_model* M, const float* mean, const float* scale) {
    if (!M || !mean || !scale) return fail(GZ_ERR_INVALID, "gz_model_set_input_map: null argument");
    float nm[2 * gz::kMlpIn];
    for (int k = 0; k < gz::kMlpIn; ++k) {
        nm[k] = mean[k];
        nm[gz::kMlpIn + k] = scale[k];
    }
    hipStream_t st = M->ctx->stream;
    GZ_HIP(hipMemcpyAsync(M->xnorm, nm, sizeof nm, hipMemcpyHostToDevice, st));
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

gz_status gz_model_set_loss(gz_model* M, float quantile) {
    if (!M) return fail(GZ_ERR_INVALID, "gz_model_set_loss: null model");
    if (!(quantile >= 0.f && quantile < 1.f))
        return fail(GZ_ERR_INVALID, "gz_model_set_loss: quantile %g outside [0, 1)", quantile);
    M->tau = quantile;
    return GZ_OK;
}

gz_status gz_model_forward(gz_ctx* ctx, const gz_model* M, const float* X, size_t B, float* out) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!M || (B && (!X || !out))) return fail(GZ_ERR_INVALID, "gz_model_forward: null argument");
    if (B >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "gz_model_forward: B too large");
    return gz::launch_mlp_forward(X, (int)B, M->wbf, M->theta, M->xnorm, out, ctx->stream);
}

gz_status gz_model_grad(gz_ctx* ctx, gz_model* M, const float* X, const float* T, size_t B,
                        float* grad_out, float* loss_out) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    if (!M || (B && (!X || !T))) return fail(GZ_ERR_INVALID, "gz_model_grad: null argument");
    if (B == 0 || B >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "gz_model_grad: bad B");
    hipStream_t st = ctx->stream;
    GZ_TRY(gz::fill_zero_u32(reinterpret_cast<uint32_t*>(M->grad), gz::kMlpParams + 64, st));
    GZ_TRY(gz::launch_mlp_grad(X, T, (int)B, M->wbf, M->theta, M->xnorm, M->tau, M->grad,
                               M->loss, st));
    if (grad_out)
        GZ_HIP(hipMemcpyAsync(grad_out, M->grad, gz::kMlpParams * sizeof(float),
                              hipMemcpyDeviceToDevice, st));
    if (loss_out) {
        GZ_HIP(hipMemcpyAsync(loss_out, M->loss, sizeof(float), hipMemcpyDeviceToHost, st));
        GZ_HIP(hipStreamSynchronize(st));
    }
    return GZ_OK;
}

gz_status gz_model_train_step(gz_ctx* ctx, gz_model* M, const float* X, const float* T, size_t B,
                              float lr, float* loss_out) {
    GZ_TRY(gz_model_grad(ctx, M, X, T, B, nullptr, nullptr));
    M->step += 1;
    GZ_TRY(gz::launch_adam(M->theta, M->m, M->v, M->grad, lr, 0.9f, 0.999f, 1e-8f, M->step,
                           M->wbf, ctx->stream));
    if (loss_out) {
        GZ_HIP(hipMemcpyAsync(loss_out, M->loss, sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
        GZ_HIP(hipStreamSynchronize(ctx->stream));
    }
    return GZ_OK;
}

gz_status gz_tree_accel(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                        const uint32_t* targets, size_t n_t, double theta, double G,
                        const double* aold, double* ax, double* ay, double* az) {
    gz::clear_error();
    return walk_per_target(ctx, tree, p, targets, n_t, theta, G, aold, false, ax, ay, az,
                           "gz_tree_accel");
}

gz_status gz_tree_accel_grouped_tidal(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                      const uint32_t* targets, size_t n_t, double theta, double G,
                                      int batch, const double* aold, const double* vx,
                                      const double* vy, const double* vz, double* ax, double* ay,
                                      double* az, double* tidal, double* jerk) {
    gz::clear_error();
    if (n_t > 0 && !tidal) return fail(GZ_ERR_INVALID, "gz_tree_accel_grouped_tidal: null tidal");
    WalkExtra ex;
    ex.tidal = tidal;
    ex.jerk = jerk;
    ex.vx = vx;
    ex.vy = vy;
    ex.vz = vz;
    return walk(ctx, tree, p, targets, n_t, theta, G, batch, 0, aold, GZ_PREC_FP64, 1, 0, 0, 0,
                ax, ay, az, nullptr, "gz_tree_accel_grouped_tidal", ex);
}

gz_status gz_tree_potential(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                            const uint32_t* targets, size_t n_t, double theta, double G,
                            const double* aold, double* phi) {
    gz::clear_error();
    if (n_t > 0 && !phi) return fail(GZ_ERR_INVALID, "gz_tree_potential: null output");
    return walk_per_target(ctx, tree, p, targets, n_t, theta, G, aold, true, phi, nullptr,
                           nullptr, "gz_tree_potential");
}

gz_status gz_direct_sum(gz_ctx* ctx, const gz_particles* p, const uint32_t* targets, size_t n_t,
                        double G, double* ax, double* ay, double* az) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    GZ_TRY(check_particles(p, "gz_direct_sum"));
    if (n_t > 0 && (!targets || !ax || !ay || !az))
        return fail(GZ_ERR_INVALID, "gz_direct_sum: null target or output array");
    return gz::launch_direct(p, targets, n_t, G, 1, 0.0, ax, ay, az, ctx->stream);
}

gz_status gz_direct_sum_ex(gz_ctx* ctx, const gz_particles* p, const uint32_t* targets,
                           size_t n_t, double G, int fp64, double box, double* ax, double* ay,
                           double* az) {
    gz::clear_error();
    GZ_TRY(check_ctx(ctx));
    GZ_TRY(check_particles(p, "gz_direct_sum_ex"));
    if (n_t > 0 && (!targets || !ax || !ay || !az))
        return fail(GZ_ERR_INVALID, "gz_direct_sum_ex: null target or output array");
    if (!(box >= 0.0) || std::isinf(box))
        return fail(GZ_ERR_INVALID, "gz_direct_sum_ex: box must be finite and >= 0");
    return gz::launch_direct(p, targets, n_t, G, fp64 ? 1 : 0, box, ax, ay, az, ctx->stream);
}

} // extern "C"
