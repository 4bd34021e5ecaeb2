// Reusing a tree across steps instead of rebuilding it (shmem/tree.cc:211-319, tree.h:250-331):
// the kick feed that raises each node's speed bound and accumulates its momentum change
// (mfm.cc:3798-3811: Tree::raise_vmax and Tree::kick_node), the refresh of the walk nodes from the
// incremental moments (Tree::refresh_walk_nodes), raising a node's softening and speed bounds
// (Tree::raise_soft, Tree::raise_vmax), and the events that are neither a drift nor a kick
// (Tree::mass_event, Tree::shift_node, Tree::raise_pad).
//
// The reference climbs the parent chain of each active particle's leaf with atomics on the CPU;
// here one thread per particle does the same climb with device atomics. Maxima (vmax, softening)
// come out the same in any order; the momentum sums are FP64 atomic additions, whose rounding
// depends on the order, as the reference's OpenMP atomics do. The reference applies the other
// events one at a time in serial loops; here they are applied in the caller's order too, so their
// sums round as the reference's do.
#include <algorithm>
#include <cmath>

#include "gz_internal.h"

namespace gz {

// Sum of v over the lanes of this lane's run (runs: maximal stretches of consecutive lanes with
// the same run id), left in the run's first lane; exact to the order of the doubling steps.
__device__ __forceinline__ double run_sum(double v, int run, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double o = __shfl_down(v, d);
        const int r = __shfl_down(run, d);
        if (lane + d < 64 && r == run) v += o;
    }
    return v;
}

// One thread per active particle. The speed bound climbs each particle's own parent chain with
// atomicMax, stopping at the first ancestor that already covers it (tree.cc:276-285). The momentum
// feed (tree.cc:211-227) climbs level by level, all lanes of the wave together: at level L the
// lanes whose chain has reached level L hold the node there; lanes holding the same node in
// consecutive lanes (in a Morton-ordered active list, all lanes under one node are consecutive)
// add their contributions in registers first, and one lane per run adds the run's total to the
// node atomically. Any order of `active` is correct; Morton order makes the runs long.
__global__ __launch_bounds__(256) void k_kick_nodes(
    const uint32_t* __restrict__ active, size_t na, size_t n, const int* __restrict__ leaf_of,
    const int* __restrict__ parent, const int* __restrict__ level_begin,
    const double* __restrict__ vx, const double* __restrict__ vy, const double* __restrict__ vz,
    const double* __restrict__ m, double* __restrict__ vlx, double* __restrict__ vly,
    double* __restrict__ vlz, double tau, uint32_t* __restrict__ vmax, double* __restrict__ dp,
    double* __restrict__ dq, size_t nn, const uint32_t* __restrict__ rank,
    double4* __restrict__ svel, const double* __restrict__ age, int* __restrict__ status) {
    if (age) tau = *age; // the tree's own age, kept on the device
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    int leaf = -1;
    uint32_t i = 0;
    if (k < na) {
        i = active[k];
        if (i >= n) atomicOr(status, 1);
        else leaf = leaf_of[i];
    }
    double ux = 0, uy = 0, uz = 0;
    if (leaf >= 0) {
        ux = vx[i]; uy = vy[i]; uz = vz[i];
        if (svel) svel[rank[i]] = make_double4(ux, uy, uz, 0.0); // the sorted velocity copy
        const float speed = (float)sqrt(ux * ux + uy * uy + uz * uz);
        const uint32_t want = __float_as_uint(speed);
        for (int no = leaf; no >= 0; no = parent[no])
            if (atomicMax(&vmax[no], want) >= want) break;
    }
    if (!vlx) return;
    if (dp) {
        double c[6] = {0, 0, 0, 0, 0, 0};
        int lvl = -1, node = leaf;
        if (leaf >= 0) {
            const double mi = m[i];
            c[0] = (ux - vlx[i]) * mi;
            c[1] = (uy - vly[i]) * mi;
            c[2] = (uz - vlz[i]) * mi;
            c[3] = c[0] * tau;
            c[4] = c[1] * tau;
            c[5] = c[2] * tau;
            lvl = 0; // node ids are level-major: the leaf's level is where its id falls
            while (lvl < kMaxLevel && level_begin[lvl + 1] <= leaf) ++lvl;
        }
        for (int L = kMaxLevel; L >= 0; --L) {
            const bool here = (lvl == L);
            const uint64_t on = __ballot(here);
            if (!on) continue;
            // runs: consecutive participating lanes at the same node
            const int prev = __shfl_up(here ? node : -1, 1);
            const bool head = here && (lane == 0 || prev != node);
            const uint64_t heads = __ballot(head);
            const uint64_t below = (lane == 63) ? heads : (heads & ((2ull << lane) - 1));
            const int run = here ? (int)__popcll(below) : -1 - lane;
            double t[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) t[q] = run_sum(here ? c[q] : 0.0, run, lane);
            if (head) {
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    atomicAdd(&dp[q * nn + node], t[q]);
                    atomicAdd(&dq[q * nn + node], t[3 + q]);
                }
            }
            if (here) {
                node = parent[node];
                lvl = (node >= 0) ? L - 1 : -1;
            }
        }
    }
    if (leaf >= 0) {
        vlx[i] = ux;
        vly[i] = uy;
        vlz[i] = uz;
    }
}

// refresh_walk_nodes (tree.cc:260-274) with node_mass and node_com (tree.h:318-320,
// tree.cc:230-237), expression for expression.
__global__ __launch_bounds__(256) void k_refresh(int nn, WNode* __restrict__ wn,
                                                 const double* __restrict__ size,
                                                 const double* __restrict__ delta,
                                                 const float* __restrict__ vmax,
                                                 const float* __restrict__ xpad,
                                                 const double* __restrict__ vcom,
                                                 const double* __restrict__ m0,
                                                 const double* __restrict__ dp,
                                                 const double* __restrict__ dq,
                                                 const double* __restrict__ dmass, double tau,
                                                 const double* __restrict__ age) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    if (age) tau = *age; // the tree's own age, kept on the device
    const size_t n = (size_t)nn;
    WNode w = wn[i];
    if (m0) {
        const double mass = m0[i];
        const double M = mass + dmass[i];
        const double cx = m0[n + i], cy = m0[2 * n + i], cz = m0[3 * n + i];
        if (M > 0) {
            w.cx = (mass * cx + (mass * vcom[i] + dp[i]) * tau - dq[i]) / M;
            w.cy = (mass * cy + (mass * vcom[n + i] + dp[n + i]) * tau - dq[n + i]) / M;
            w.cz = (mass * cz + (mass * vcom[2 * n + i] + dp[2 * n + i]) * tau - dq[2 * n + i]) / M;
        } else {
            w.cx = cx;
            w.cy = cy;
            w.cz = cz;
        }
        w.mass = M;
    }
    // every particle may have moved vmax*tau, and the COM with them
    const double grow = 2.0 * ((double)vmax[i] * tau + (double)xpad[i]);
    w.s = size[i] + delta[i] + grow;
    w.len = (float)(size[i] + grow);
    wn[i] = w;
}

// The softening bound rounded up to float, so the node's bound stays one-sided (tree.cc:289-292).
__device__ __forceinline__ float soft_as_bound(double h) {
    const float f = (float)h;
    return ((double)f < h) ? nextafterf(f, 1e30f) : f;
}

// raise_soft (tree.cc:311-319): the float bound of the walk node, and the FP64 node softening.
__global__ __launch_bounds__(256) void k_raise_soft(const uint32_t* __restrict__ idx, size_t n_idx,
                                                    size_t n, const int* __restrict__ leaf_of,
                                                    const int* __restrict__ parent,
                                                    const double* __restrict__ soft,
                                                    WNode* __restrict__ wn,
                                                    double* __restrict__ soft64,
                                                    int* __restrict__ status) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_idx) return;
    const uint32_t i = idx[k];
    if (i >= n) {
        atomicOr(status, 1);
        return;
    }
    const int leaf = leaf_of[i];
    if (leaf < 0) return;
    const double h = soft[i];
    const uint32_t want = __float_as_uint(soft_as_bound(h));
    const unsigned long long want64 = (unsigned long long)__double_as_longlong(h);
    for (int no = leaf; no >= 0; no = parent[no]) {
        if (atomicMax(reinterpret_cast<uint32_t*>(&wn[no].soft), want) >= want) break;
        atomicMax(reinterpret_cast<unsigned long long*>(&soft64[no]), want64);
    }
}

// Tree::node_vel (tree.h:323-331): each node's centre-of-mass velocity corrected for the kicks
// since the build, (m0 vcom + dp) / M, for the jerk's multipole terms. Out: 3 x nn SoA.
__global__ __launch_bounds__(256) void k_node_vel(int nn, const double* __restrict__ vcom,
                                                  const double* __restrict__ m0,
                                                  const double* __restrict__ dp,
                                                  const double* __restrict__ dmass,
                                                  double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const size_t n = (size_t)nn;
    const double mass = m0[i];
    const double M = mass + dmass[i];
    for (int d = 0; d < 3; ++d)
        out[d * n + i] = (M > 0) ? (mass * vcom[d * n + i] + dp[d * n + i]) / M : vcom[d * n + i];
}

gz_status launch_node_vel(const gz_tree* t, double* out, hipStream_t st) {
    if (t->nnodes == 0 || !t->has_vel) return GZ_OK;
    k_node_vel<<<(unsigned)((t->nnodes + 255) / 256), 256, 0, st>>>(t->nnodes, t->vcom, t->m0,
                                                                    t->dp, t->dmass, out);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_kick_nodes(gz_tree* t, const uint32_t* active, size_t na, const double* vx,
                            const double* vy, const double* vz, const double* m, double* vlx,
                            double* vly, double* vlz, double tau, int* status, hipStream_t st) {
    if (na == 0 || t->nnodes == 0) return GZ_OK;
    k_kick_nodes<<<(unsigned)((na + 255) / 256), 256, 0, st>>>(
        active, na, t->n, t->leaf_of, t->parent, t->level_begin_dev, vx, vy, vz, m, vlx, vly,
        vlz, tau,
        reinterpret_cast<uint32_t*>(t->vmax), t->dp, t->dq, (size_t)t->nnodes, t->rank, t->svel,
        tau < 0.0 ? t->age : nullptr, status);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

// The drift applied to the tree's sorted copy of the sources, from the sorted velocity copy: for
// every slot, x += v dt with the same operands (and so the same rounding) as gz_drift followed by
// a regather, but streamed in slot order instead of gathered through orderbuf.
__global__ __launch_bounds__(256) void k_tree_drift(size_t nl, double dt,
                                                    const double4* __restrict__ svel,
                                                    double4* __restrict__ src,
                                                    double* __restrict__ age) {
    const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s == 0) *age += dt; // read by later kernels on the stream, never by this one
    if (s >= nl) return;
    const double4 v = svel[s];
    double4 q = src[s];
    q.x = q.x + v.x * dt;
    q.y = q.y + v.y * dt;
    q.z = q.z + v.z * dt;
    src[s] = q;
}

gz_status launch_tree_drift(gz_tree* t, double dt, hipStream_t st) {
    k_tree_drift<<<(unsigned)((std::max<size_t>(t->n_live, 1) + 255) / 256), 256, 0, st>>>(
        t->n_live, dt, t->svel, t->src, t->age);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_refresh(gz_tree* t, double tau, hipStream_t st) {
    if (t->nnodes == 0) return GZ_OK;
    k_refresh<<<(unsigned)((t->nnodes + 255) / 256), 256, 0, st>>>(
        t->nnodes, t->wn, t->size, t->delta, t->vmax, t->xpad, t->vcom, t->m0, t->dp, t->dq,
        t->dmass, tau, tau < 0.0 ? t->age : nullptr);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_raise_soft(gz_tree* t, const uint32_t* idx, size_t n_idx, const double* soft,
                            int* status, hipStream_t st) {
    if (n_idx == 0 || t->nnodes == 0) return GZ_OK;
    k_raise_soft<<<(unsigned)((n_idx + 255) / 256), 256, 0, st>>>(
        idx, n_idx, t->n, t->leaf_of, t->parent, soft, t->wn, t->soft, status);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

// raise_vmax (tree.cc:276-285) with the caller's speeds, one thread per event: the same climb as
// the kick feed's, stopping at the first ancestor that already covers the speed. A maximum, so
// the order of the events does not matter. Negative or NaN speeds are skipped and flagged.
__global__ __launch_bounds__(256) void k_raise_vmax(const uint32_t* __restrict__ idx, size_t n_ev,
                                                    size_t n, const int* __restrict__ leaf_of,
                                                    const int* __restrict__ parent,
                                                    const float* __restrict__ speed,
                                                    uint32_t* __restrict__ vmax,
                                                    int* __restrict__ status) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_ev) return;
    const uint32_t i = idx[k];
    if (i >= n) {
        atomicOr(status, 1);
        return;
    }
    const float s = speed[k];
    if (!(s >= 0.0f)) {
        atomicOr(status, 2);
        return;
    }
    const int leaf = leaf_of[i];
    const uint32_t want = __float_as_uint(s);
    for (int no = leaf; no >= 0; no = parent[no])
        if (atomicMax(&vmax[no], want) >= want) break;
}

gz_status launch_raise_vmax(gz_tree* t, const uint32_t* idx, size_t n_ev, const float* speed,
                            int* status, hipStream_t st) {
    if (n_ev == 0 || t->nnodes == 0) return GZ_OK;
    k_raise_vmax<<<(unsigned)((n_ev + 255) / 256), 256, 0, st>>>(
        idx, n_ev, t->n, t->leaf_of, t->parent, speed, reinterpret_cast<uint32_t*>(t->vmax),
        status);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

// The minimum-image separation (shmem/vec.h:25-30): folded into [-box/2, box/2] when box > 0.
__device__ __forceinline__ double min_image_sep(double d, double box) {
    if (box > 0) {
        if (d > 0.5 * box) d -= box;
        else if (d < -0.5 * box) d += box;
    }
    return d;
}

// Events between rebuilds that are not a drift or a kick, in the caller's order (the reference
// calls them one at a time from serial loops): Tree::mass_event (tree.cc:239-251),
// Tree::shift_node (tree.cc:253-258) and Tree::raise_pad (tree.cc:323-330). Every event climbs
// from its leaf to the root, and a mass event reads each ancestor's current centre of mass (to
// unwrap its position), so the events at one node must be applied in order. One wave does it:
// lane L owns the ancestor at level L (node ids are level-major, so a node's level is its depth),
// which makes every node's updates those of the serial loop, in the same order and with the same
// operations, while the ~20 levels of one event proceed together. Per 64 events, each lane first
// climbs one event's parent chain into LDS.
struct NodeEventArgs {
    const uint32_t* idx;
    size_t n_ev, n, nn;
    const int* leaf_of;
    const int* parent;
    const int* level_begin;
    const double* a;                      // dm (mass event), m (shift), d (pad)
    const double *x, *y, *z;              // position (mass event) or displacement (shift)
    const double *vx, *vy, *vz;           // velocity (mass event)
    double tau, box;
    const double* age;                    // the tree's age on the device, used when non-null
    const double* m0;                     // build moments: mass, then COM x, y, z (stride nn)
    const double* vcom;                   // build COM velocity (stride nn)
    double *dp, *dq, *dmass;
    float* xpad;
    int* status;
};

template <int KIND>
__global__ __launch_bounds__(64) void k_node_events(NodeEventArgs a) {
    __shared__ int chain[64][kMaxLevel + 1];
    __shared__ int depth[64];
    const int lane = threadIdx.x;
    const double tau = a.age ? *a.age : a.tau;
    const size_t nn = a.nn;
    for (size_t base = 0; base < a.n_ev; base += 64) {
        const size_t k0 = base + lane;
        int lvl = -1;
        if (k0 < a.n_ev) {
            const uint32_t i = a.idx[k0];
            int leaf = -1;
            if (i >= a.n) atomicOr(a.status, 1);
            else leaf = a.leaf_of[i];
            bool skip = leaf < 0;
            if (KIND == kNodeEventMass) skip = skip || a.a[k0] == 0.0;
            if (KIND == kNodeEventPad) skip = skip || !(a.a[k0] > 0);
            if (!skip) {
                lvl = 0;
                while (lvl < kMaxLevel && a.level_begin[lvl + 1] <= leaf) ++lvl;
                int no = leaf;
                for (int L = lvl; L >= 0; --L) {
                    chain[lane][L] = no;
                    no = a.parent[no];
                }
            }
        }
        depth[lane] = lvl;
        __syncthreads();
        const int cnt = (int)min((size_t)64, a.n_ev - base);
        for (int e = 0; e < cnt; ++e) {
            const int le = depth[e]; // the same for every lane
            if (le < 0) continue;
            const size_t k = base + e;
            const int node = (lane <= le) ? chain[e][lane] : -1;
            if (KIND == kNodeEventPad) {
                // the leaf's pad grows by d (rounded up to float); the ancestors take it as a
                // maximum: the reference's climb stops at the first that already covers it, and
                // every node above that one covers it too
                float pad = 0.0f;
                if (lane == le) pad = soft_as_bound((double)a.xpad[node] + a.a[k]);
                pad = __shfl(pad, le);
                if (node >= 0 && a.xpad[node] < pad) a.xpad[node] = pad;
            } else if (KIND == kNodeEventShift) {
                if (node >= 0) {
                    const double m = a.a[k];
                    a.dq[node] -= m * a.x[k];
                    a.dq[nn + node] -= m * a.y[k];
                    a.dq[2 * nn + node] -= m * a.z[k];
                }
            } else if (node >= 0) {
                // node_com (tree.cc:229-237) at tau, then the moments of a particle of mass dm at
                // x (unwrapped against that centre) moving at v
                const double mass = a.m0[node];
                const double M = mass + a.dmass[node];
                double c[3];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    c[d] = (M > 0) ? (mass * a.m0[(d + 1) * nn + node] +
                                      (mass * a.vcom[d * nn + node] + a.dp[d * nn + node]) * tau -
                                      a.dq[d * nn + node]) / M
                                   : a.m0[(d + 1) * nn + node];
                }
                const double dm = a.a[k];
                const double xs[3] = {a.x[k], a.y[k], a.z[k]};
                const double vs[3] = {a.vx[k], a.vy[k], a.vz[k]};
                a.dmass[node] += dm;
#pragma unroll
                for (int d = 0; d < 3; ++d) a.dp[d * nn + node] += dm * vs[d];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    const double xr = c[d] + min_image_sep(xs[d] - c[d], a.box);
                    a.dq[d * nn + node] -= dm * (xr - vs[d] * tau);
                }
            }
        }
        __syncthreads();
    }
}

gz_status launch_node_events(gz_tree* t, int kind, const uint32_t* idx, size_t n_ev,
                             const double* v, const double* x, const double* y, const double* z,
                             const double* vx, const double* vy, const double* vz, double tau,
                             double box, int* status, hipStream_t st) {
    if (n_ev == 0 || t->nnodes == 0) return GZ_OK;
    NodeEventArgs a{};
    a.idx = idx;
    a.n_ev = n_ev;
    a.n = t->n;
    a.nn = (size_t)t->nnodes;
    a.leaf_of = t->leaf_of;
    a.parent = t->parent;
    a.level_begin = t->level_begin_dev;
    a.a = v;
    a.x = x;
    a.y = y;
    a.z = z;
    a.vx = vx;
    a.vy = vy;
    a.vz = vz;
    a.tau = tau;
    a.box = box;
    a.age = tau < 0.0 ? t->age : nullptr;
    a.m0 = t->m0;
    a.vcom = t->vcom;
    a.dp = t->dp;
    a.dq = t->dq;
    a.dmass = t->dmass;
    a.xpad = t->xpad;
    a.status = status;
    if (kind == kNodeEventMass) k_node_events<kNodeEventMass><<<1, 64, 0, st>>>(a);
    else if (kind == kNodeEventShift) k_node_events<kNodeEventShift><<<1, 64, 0, st>>>(a);
    else k_node_events<kNodeEventPad><<<1, 64, 0, st>>>(a);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // namespace gz
