// Reusing a tree across steps instead of rebuilding it (shmem/tree.cc:211-319, tree.h:250-331):
// the kick feed that raises each node's speed bound and accumulates its momentum change
// (mfm.cc:3798-3811: Tree::raise_vmax and Tree::kick_node), the refresh of the walk nodes from the
// incremental moments (Tree::refresh_walk_nodes), and raising a node's softening bound
// (Tree::raise_soft).
//
// The reference climbs the parent chain of each active particle's leaf with atomics on the CPU;
// here one thread per particle does the same climb with device atomics. Maxima (vmax, softening)
// come out the same in any order; the momentum sums are FP64 atomic additions, whose rounding
// depends on the order, as the reference's OpenMP atomics do.
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

} // namespace gz
