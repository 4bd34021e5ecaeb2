// Tree build on the device: bounding box and root cell, Morton keys, radix sort, level-synchronous
// octree topology, bottom-up node moments and the packed walk nodes.
//
// Follows shmem::build (shmem/tree.cc:350-526) and shmem::build_node (tree.cc:70-197). The
// reference builds recursively with OpenMP tasks and an atomic node counter; here every level of
// the tree is one pass over all nodes of that level, children are allocated by a prefix sum, and
// moments are combined bottom-up one level per launch. The resulting tree is the reference's tree
// node for node (same ranges, same geometric cells, same threaded next[]), numbered level-major.
//
// This file is compiled with -ffp-contract=off: key quantisation and the moment sums must round
// exactly as the reference's do (no fused multiply-add), so keys, topology and every node moment
// are bit-identical to the reference built the same way.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "gz_internal.h"

namespace gz {

constexpr int kThreads = 256;

static inline unsigned grid_for(size_t n, int threads = kThreads) {
    return (unsigned)((n + threads - 1) / threads);
}

__device__ __forceinline__ double dmin_(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax_(double a, double b) { return (a < b) ? b : a; }

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = dmin_(v, __shfl_xor(v, d));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = dmax_(v, __shfl_xor(v, d));
    return v;
}

// ---- bounding box + root cell (tree.cc:361-407) ------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_bbox_partial(const double* __restrict__ x,
                                                           const double* __restrict__ y,
                                                           const double* __restrict__ z,
                                                           const double* __restrict__ m, size_t n,
                                                           double* __restrict__ partial) {
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kThreads) {
        if (!(m[i] > 0)) continue;
        const double a = x[i], b = y[i], c = z[i];
        lo0 = dmin_(lo0, a); hi0 = dmax_(hi0, a);
        lo1 = dmin_(lo1, b); hi1 = dmax_(hi1, b);
        lo2 = dmin_(lo2, c); hi2 = dmax_(hi2, c);
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    __shared__ double s[4][6];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s[w][0] = lo0; s[w][1] = lo1; s[w][2] = lo2; s[w][3] = hi0; s[w][4] = hi1; s[w][5] = hi2;
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int k = threadIdx.x;
        double v = s[0][k];
        for (int j = 1; j < 4; ++j) v = (k < 3) ? dmin_(v, s[j][k]) : dmax_(v, s[j][k]);
        partial[(size_t)blockIdx.x * 6 + k] = v;
    }
}

// splitmix64 draw in [0,1) for RANDOMIZE_GRAVTREE (tree.cc:381-393)
__device__ double splitmix_unit(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}

// root[0..3] = centre x, y, z and side; root[4..9] = live bounding box lo, hi.
__global__ void k_root_cell(const double* __restrict__ partial, int nb, long long seed, double fx,
                            double fy, double fz, double* __restrict__ root) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) {
            lo[k] = dmin_(lo[k], partial[(size_t)b * 6 + k]);
            hi[k] = dmax_(hi[k], partial[(size_t)b * 6 + 3 + k]);
        }
    double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    double side = dmax_(hi[0] - lo[0], dmax_(hi[1] - lo[1], hi[2] - lo[2])) * 1.0000001;
    if (side <= 0) side = 1.0;
    if (seed >= 0) {
        uint64_t s = (uint64_t)seed * 0x9E3779B97F4A7C15ull;
        cx += side * (splitmix_unit(s) - 0.5);
        cy += side * (splitmix_unit(s) - 0.5);
        cz += side * (splitmix_unit(s) - 0.5);
        side *= 2.0;
    }
    if (fx != 0.0 || fy != 0.0 || fz != 0.0) { // the reference's SHMEM_TREE_SHIFT (tree.cc:400-407)
        cx += fx * side; cy += fy * side; cz += fz * side;
        side *= 1.0 + 2.0 * (fabs(fx) + fabs(fy) + fabs(fz));
    }
    root[0] = cx; root[1] = cy; root[2] = cz; root[3] = side;
    for (int k = 0; k < 3; ++k) { root[4 + k] = lo[k]; root[7 + k] = hi[k]; }
}

// ---- independent clusters (a forest) -------------------------------------------------------------
// Cluster c is particles [cstart[c], cstart[c+1]); each gets its own root cell from its own live
// bounding box, exactly as the single-tree root (tree.cc:361-372), and its own tree.
__device__ __forceinline__ int cluster_of(uint32_t i, const uint32_t* __restrict__ cstart, int nc) {
    int lo = 0, hi = nc; // last c with cstart[c] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cstart[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kThreads) void k_bbox_clusters(const double* __restrict__ x,
                                                            const double* __restrict__ y,
                                                            const double* __restrict__ z,
                                                            const double* __restrict__ m,
                                                            const uint32_t* __restrict__ cstart,
                                                            double* __restrict__ roots) {
    const int c = blockIdx.x;
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (uint32_t i = cstart[c] + threadIdx.x; i < cstart[c + 1]; i += kThreads) {
        if (!(m[i] > 0)) continue;
        lo0 = dmin_(lo0, x[i]); hi0 = dmax_(hi0, x[i]);
        lo1 = dmin_(lo1, y[i]); hi1 = dmax_(hi1, y[i]);
        lo2 = dmin_(lo2, z[i]); hi2 = dmax_(hi2, z[i]);
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    __shared__ double s[4][6];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        s[w][0] = lo0; s[w][1] = lo1; s[w][2] = lo2; s[w][3] = hi0; s[w][4] = hi1; s[w][5] = hi2;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double lo[3], hi[3];
    for (int k = 0; k < 3; ++k) {
        lo[k] = s[0][k];
        hi[k] = s[0][3 + k];
        for (int j = 1; j < 4; ++j) {
            lo[k] = dmin_(lo[k], s[j][k]);
            hi[k] = dmax_(hi[k], s[j][3 + k]);
        }
    }
    double side = dmax_(hi[0] - lo[0], dmax_(hi[1] - lo[1], hi[2] - lo[2])) * 1.0000001;
    if (side <= 0) side = 1.0;
    double* r = roots + (size_t)c * 4;
    r[0] = 0.5 * (lo[0] + hi[0]);
    r[1] = 0.5 * (lo[1] + hi[1]);
    r[2] = 0.5 * (lo[2] + hi[2]);
    r[3] = side;
}

__global__ __launch_bounds__(kThreads) void k_cluster_keys(const uint32_t* __restrict__ v,
                                                           size_t nl,
                                                           const uint32_t* __restrict__ cstart,
                                                           int nc, uint64_t* __restrict__ ck) {
    const size_t r = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (r < nl) ck[r] = (uint64_t)cluster_of(v[r], cstart, nc);
}

__global__ __launch_bounds__(kThreads) void k_cluster_counts(const uint32_t* __restrict__ v,
                                                             size_t nl,
                                                             const uint32_t* __restrict__ cstart,
                                                             int nc, int* __restrict__ cnt) {
    const size_t r = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (r < nl) atomicAdd(&cnt[cluster_of(v[r], cstart, nc)], 1);
}

__global__ __launch_bounds__(kThreads) void k_init_roots(int nc, const double* __restrict__ roots,
                                                         const int* __restrict__ rstart, int* nlo,
                                                         int* nhi, double* gcx, double* gcy,
                                                         double* gcz, double* size, int* parent,
                                                         int* next) {
    const int c = blockIdx.x * kThreads + threadIdx.x;
    if (c >= nc) return;
    nlo[c] = rstart[c];
    nhi[c] = rstart[c + 1];
    gcx[c] = roots[4 * c];
    gcy[c] = roots[4 * c + 1];
    gcz[c] = roots[4 * c + 2];
    size[c] = roots[4 * c + 3];
    parent[c] = -1;
    next[c] = -1;
}

// ---- live filter + keys (tree.cc:355-359, 409-418) ----------------------------------------------
__global__ __launch_bounds__(kThreads) void k_live_flags(const double* __restrict__ m, size_t n,
                                                         uint32_t* __restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) flag[i] = (m[i] > 0) ? 1u : 0u;
}

__global__ __launch_bounds__(kThreads) void k_live_compact(const double* __restrict__ m, size_t n,
                                                           const uint32_t* __restrict__ pos,
                                                           uint32_t* __restrict__ order) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n && m[i] > 0) order[pos[i]] = (uint32_t)i;
}

__device__ __forceinline__ uint64_t spread3(uint32_t v) {
    uint64_t x = v & 0x1FFFFFu;
    x = (x | (x << 32)) & 0x1F00000000FFFFull;
    x = (x | (x << 16)) & 0x1F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kThreads) void k_keys(const uint32_t* __restrict__ order, size_t nl,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const double* __restrict__ roots,
                                                   const uint32_t* __restrict__ cstart, int nc,
                                                   uint64_t* __restrict__ keys) {
    const size_t r = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= nl) return;
    const double* root = roots + (nc > 1 ? (size_t)cluster_of(order[r], cstart, nc) * 4 : 0);
    const double side = root[3];
    const double scale = (double)((1u << kMaxLevel) - 1) / side;
    const double bx = root[0] - 0.5 * side, by = root[1] - 0.5 * side, bz = root[2] - 0.5 * side;
    const uint32_t i = order[r];
    const uint32_t a = (uint32_t)((x[i] - bx) * scale);
    const uint32_t b = (uint32_t)((y[i] - by) * scale);
    const uint32_t c = (uint32_t)((z[i] - bz) * scale);
    keys[r] = spread3(a) | (spread3(b) << 1) | (spread3(c) << 2);
}

// Sorted copy of the source data, and the inverse permutation (tree.cc:483-485).
__global__ __launch_bounds__(kThreads) void k_gather_sorted(
    const uint32_t* __restrict__ orderbuf, size_t nl, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    const double* __restrict__ soft, double4* __restrict__ src, double* __restrict__ src_soft,
    uint32_t* __restrict__ rank) {
    const size_t r = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= nl) return;
    const uint32_t p = orderbuf[r];
    src[r] = make_double4(x[p], y[p], z[p], m[p]);
    src_soft[r] = soft ? soft[p] : 0.0;
    rank[p] = (uint32_t)r;
}

// Sorted copy of the build-time velocities.
__global__ __launch_bounds__(kThreads) void k_gather_vel(const uint32_t* __restrict__ orderbuf,
                                                         size_t nl, const double* __restrict__ vx,
                                                         const double* __restrict__ vy,
                                                         const double* __restrict__ vz,
                                                         double4* __restrict__ svel) {
    const size_t r = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (r >= nl) return;
    const uint32_t p = orderbuf[r];
    svel[r] = make_double4(vx[p], vy[p], vz[p], 0.0);
}

__global__ __launch_bounds__(kThreads) void k_fill_u32(uint32_t* __restrict__ a, size_t n,
                                                       uint32_t v) {
    const size_t i = (size_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void k_init_root(const double* __restrict__ root, int nl, int* nlo, int* nhi,
                            double* gcx, double* gcy, double* gcz, double* size, int* parent,
                            int* next) {
    nlo[0] = 0;
    nhi[0] = nl;
    gcx[0] = root[0];
    gcy[0] = root[1];
    gcz[0] = root[2];
    size[0] = root[3];
    parent[0] = -1;
    next[0] = -1;
}

// ---- topology, one level per launch (tree.cc:86-143) -------------------------------------------
// For each node of the level: leaf test, then the start of each octant's run of keys by binary
// search on the 3-bit digit (keys in a node share every higher digit, so the digit is sorted).
__global__ __launch_bounds__(kThreads) void k_split(int lb, int count, int level,
                                                    const int* __restrict__ nlo,
                                                    const int* __restrict__ nhi,
                                                    const uint64_t* __restrict__ skey,
                                                    int* __restrict__ oct_start,
                                                    uint32_t* __restrict__ nkids) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const int node = lb + i;
    const int lo = nlo[node], hi = nhi[node];
    if ((hi - lo <= kLeafMax) || (level >= kMaxLevel)) {
        nkids[i] = 0;
        return;
    }
    const int shift = 3 * (kMaxLevel - level - 1);
    int* os = oct_start + (size_t)i * 8;
    os[0] = lo;
    int a = lo;
    for (int o = 1; o < 8; ++o) {
        int b = hi;
        while (a < b) { // first index in [a, hi) with digit >= o
            const int mid = a + ((b - a) >> 1);
            if ((int)((skey[mid] >> shift) & 7ull) < o) a = mid + 1;
            else b = mid;
        }
        os[o] = a;
    }
    uint32_t nk = 0;
    for (int o = 0; o < 8; ++o) {
        const int e = (o < 7) ? os[o + 1] : hi;
        nk += (e > os[o]) ? 1u : 0u;
    }
    nkids[i] = nk;
}

// Allocate each node's children at next_lb + off[i] in octant order, with the child cell from
// halving the parent's (tree.cc:101, 116-118) and the threaded successor (tree.cc:340-348): the
// next sibling, or the parent's successor for the last child.
__global__ __launch_bounds__(kThreads) void k_emit(
    int lb, int count, int next_lb, const uint32_t* __restrict__ nkids,
    const uint32_t* __restrict__ off, const int* __restrict__ oct_start, int* __restrict__ nlo,
    int* __restrict__ nhi, double* __restrict__ gcx, double* __restrict__ gcy,
    double* __restrict__ gcz, double* __restrict__ size, int* __restrict__ parent,
    int* __restrict__ next, WNode* __restrict__ wn) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= count) return;
    const int node = lb + i;
    const int nk = (int)nkids[i];
    if (nk == 0) {
        wn[node].first = -1;
        wn[node].nchild = 0;
        return;
    }
    const int base = next_lb + (int)off[i];
    wn[node].first = base;
    wn[node].nchild = nk;
    const int hi = nhi[node];
    const double sz = size[node];
    const double h = sz * 0.5, q = sz * 0.25;
    const double cx = gcx[node], cy = gcy[node], cz = gcz[node];
    const int succ = next[node];
    const int* os = oct_start + (size_t)i * 8;
    int k = 0;
    for (int o = 0; o < 8; ++o) {
        const int s = os[o];
        const int e = (o < 7) ? os[o + 1] : hi;
        if (e <= s) continue;
        const int c = base + k;
        nlo[c] = s;
        nhi[c] = e;
        gcx[c] = cx + ((o & 1) ? q : -q);
        gcy[c] = cy + ((o & 2) ? q : -q);
        gcz[c] = cz + ((o & 4) ? q : -q);
        size[c] = h;
        parent[c] = node;
        next[c] = (k < nk - 1) ? c + 1 : succ;
        ++k;
    }
}

// ---- moments, deepest level first (tree.cc:153-196) + walk-node pack (tree.cc:507-523) ---------
__global__ __launch_bounds__(kThreads) void k_moments(
    int lb, int count, WNode* __restrict__ wn, const int* __restrict__ nlo,
    const int* __restrict__ nhi, const double* __restrict__ gcx, const double* __restrict__ gcy,
    const double* __restrict__ gcz, const double* __restrict__ size,
    const double4* __restrict__ src, const double* __restrict__ src_soft,
    const uint32_t* __restrict__ orderbuf, int* __restrict__ leaf_of,
    int* __restrict__ slot_leaf, NodeVel nv, double* __restrict__ delta_out,
    double* __restrict__ soft_out, int* __restrict__ nleaf) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    const bool live = i < count;
    const int node = lb + (live ? i : 0);
    const int nk = live ? wn[node].nchild : 1;
    const uint64_t leaves = __ballot(live && nk == 0);
    if (__lane_id() == 0 && leaves) atomicAdd(nleaf, (int)__popcll(leaves));
    if (!live) return;
    const int first = wn[node].first;
    const int lo = nlo[node], hi = nhi[node];
    double M = 0, sx = 0, sy = 0, sz = 0, smax = 0;
    double vmax = 0, pvx = 0, pvy = 0, pvz = 0; // with velocities: max |v| and momentum
    if (nk == 0) {
        for (int k = lo; k < hi; ++k) {
            const double4 p = src[k];
            const double mm = p.w;
            M += mm;
            sx += mm * p.x;
            sy += mm * p.y;
            sz += mm * p.z;
            const double s = src_soft[k];
            if (s > smax) smax = s;
            if (nv.vx) { // tree.cc:161-166
                const uint32_t q = orderbuf[k];
                const double ux = nv.vx[q], uy = nv.vy[q], uz = nv.vz[q];
                const double v = sqrt(ux * ux + uy * uy + uz * uz);
                if (v > vmax) vmax = v;
                pvx += mm * ux;
                pvy += mm * uy;
                pvz += mm * uz;
            }
            leaf_of[orderbuf[k]] = node;
            slot_leaf[k] = node;
        }
    } else {
        for (int c = first; c < first + nk; ++c) {
            const double mm = wn[c].mass;
            M += mm;
            sx += mm * wn[c].cx;
            sy += mm * wn[c].cy;
            sz += mm * wn[c].cz;
            const double s = soft_out[c];
            if (s > smax) smax = s;
            if (nv.vx) { // tree.cc:177-181
                if (nv.vmax[c] > vmax) vmax = nv.vmax[c];
                pvx += mm * nv.vcom[c];
                pvy += mm * nv.vcom[nv.nn + c];
                pvz += mm * nv.vcom[2 * (size_t)nv.nn + c];
            }
        }
    }
    if (M > 0) {
        sx /= M;
        sy /= M;
        sz /= M;
    }
    const double dx = sx - gcx[node], dy = sy - gcy[node], dz = sz - gcz[node];
    const double delta = sqrt(dx * dx + dy * dy + dz * dz);
    const double side = size[node];
    WNode w;
    w.cx = sx;
    w.cy = sy;
    w.cz = sz;
    w.s = side + delta;
    w.mass = M;
    w.soft = (float)smax;
    w.len = (float)side;
    w.first = first;
    w.nchild = nk;
    w.plo = lo;
    w.phi = hi;
    wn[node] = w;
    delta_out[node] = delta;
    soft_out[node] = smax;
    nv.vmax[node] = (float)vmax; // 0 without velocities
    if (nv.vx) {
        // centre-of-mass velocity (tree.cc:188-190), the build moments that refreshes start
        // from, and zeroed incremental moments
        const size_t nn = (size_t)nv.nn;
        nv.vcom[node] = (M > 0) ? pvx / M : 0.0;
        nv.vcom[nn + node] = (M > 0) ? pvy / M : 0.0;
        nv.vcom[2 * nn + node] = (M > 0) ? pvz / M : 0.0;
        nv.m0[node] = M;
        nv.m0[nn + node] = sx;
        nv.m0[2 * nn + node] = sy;
        nv.m0[3 * nn + node] = sz;
        for (int d = 0; d < 3; ++d) {
            nv.dp[d * nn + node] = 0.0;
            nv.dq[d * nn + node] = 0.0;
        }
        nv.dmass[node] = 0.0;
    }
    nv.xpad[node] = 0.0f;
}

// ---- host driver ---------------------------------------------------------------------------------
static gz_status dalloc(void** p, size_t bytes) {
    GZ_HIP(hipMalloc(p, std::max<size_t>(bytes, 64)));
    return GZ_OK;
}

template <class T>
static gz_status grow_array(T*& a, size_t used, size_t cap, hipStream_t st) {
    T* b = nullptr;
    GZ_TRY(dalloc(reinterpret_cast<void**>(&b), cap * sizeof(T)));
    if (a && used) GZ_HIP(hipMemcpyAsync(b, a, used * sizeof(T), hipMemcpyDeviceToDevice, st));
    if (a) {
        GZ_HIP(hipStreamSynchronize(st));
        GZ_HIP(hipFree(a));
    }
    a = b;
    return GZ_OK;
}

static gz_status grow_nodes(gz_tree* t, int used, int cap, hipStream_t st) {
    GZ_TRY(grow_array(t->wn, used, cap, st));
    GZ_TRY(grow_array(t->next, used, cap, st));
    GZ_TRY(grow_array(t->parent, used, cap, st));
    GZ_TRY(grow_array(t->nlo, used, cap, st));
    GZ_TRY(grow_array(t->nhi, used, cap, st));
    GZ_TRY(grow_array(t->gcx, used, cap, st));
    GZ_TRY(grow_array(t->gcy, used, cap, st));
    GZ_TRY(grow_array(t->gcz, used, cap, st));
    GZ_TRY(grow_array(t->size, used, cap, st));
    GZ_TRY(grow_array(t->delta, used, cap, st));
    GZ_TRY(grow_array(t->soft, used, cap, st));
    t->cap = cap;
    return GZ_OK;
}

// The reuse arrays of a tree with nn nodes, in one allocation: vmax and xpad (floats), and with
// velocities vcom (3 nn), the build moments (4 nn: mass, then COM), dp and dq (3 nn each) and
// dmass (nn).
static gz_status alloc_reuse(gz_tree* t, int nn, bool vel) {
    if (t->reuse_mem) {
        (void)hipFree(t->reuse_mem);
        t->reuse_mem = nullptr;
    }
    const size_t n = (size_t)std::max(nn, 1);
    const size_t fl = (2 * n * sizeof(float) + 255) & ~size_t(255);
    const size_t db = vel ? 14 * n * sizeof(double) : 0;
    GZ_TRY(dalloc(&t->reuse_mem, fl + db));
    uint8_t* b = static_cast<uint8_t*>(t->reuse_mem);
    t->vmax = reinterpret_cast<float*>(b);
    t->xpad = t->vmax + n;
    t->has_vel = vel;
    t->vcom = t->m0 = t->dp = t->dq = t->dmass = nullptr;
    if (vel) {
        double* d = reinterpret_cast<double*>(b + fl);
        t->vcom = d;
        t->m0 = d + 3 * n;
        t->dp = d + 7 * n;
        t->dq = d + 10 * n;
        t->dmass = d + 13 * n;
    }
    return GZ_OK;
}

void tree_release(gz_tree* t) {
    void* ptrs[] = {t->skey, t->orderbuf, t->rank, t->leaf_of, t->slot_leaf, t->src, t->src_soft,
                    t->wn, t->next, t->parent, t->nlo, t->nhi, t->gcx, t->gcy, t->gcz, t->size,
                    t->delta, t->soft, t->level_begin_dev, t->cluster_start_dev, t->reuse_mem,
                    t->hmax, t->svel, t->age};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}

gz_status build_tree(gz_ctx* ctx, const gz_particles* p, const gz_build_opts* opts,
                     gz_tree* t) {
    hipStream_t st = ctx->stream;
    const size_t n = p->n;
    if (n >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "gz_tree_build: n=%zu too large", n);
    if (n > 0 && (!p->x || !p->y || !p->z || !p->m))
        return fail(GZ_ERR_INVALID, "gz_tree_build: null position or mass array");
    if (p->zeta) return fail(GZ_ERR_UNSUPPORTED, "gz_tree_build: zeta corrections not supported");
    const long long seed = opts ? opts->randomize_seed : -1;
    const double fx = opts ? opts->shift[0] : 0.0, fy = opts ? opts->shift[1] : 0.0,
                 fz = opts ? opts->shift[2] : 0.0;
    const int max_nodes = opts ? opts->max_nodes : 0;
    const int nc = (opts && opts->n_clusters > 1) ? opts->n_clusters : 1;
    if (nc > 1) {
        if (!opts->cluster_start)
            return fail(GZ_ERR_INVALID, "gz_tree_build: n_clusters > 1 without cluster_start");
        if (seed >= 0 || fx != 0.0 || fy != 0.0 || fz != 0.0)
            return fail(GZ_ERR_UNSUPPORTED, "gz_tree_build: root randomisation with clusters");
        if (opts->cluster_start[0] != 0 || opts->cluster_start[nc] != n)
            return fail(GZ_ERR_INVALID, "gz_tree_build: clusters must cover [0, n)");
        for (int c = 0; c < nc; ++c)
            if (opts->cluster_start[c + 1] < opts->cluster_start[c])
                return fail(GZ_ERR_INVALID, "gz_tree_build: cluster offsets must not decrease");
    }
    t->n = n;
    t->n_clusters = nc;
    hipEvent_t* ev = ctx->ev;
    GZ_HIP(hipEventRecord(ev[0], st));

    // --- bbox + root cell (one per cluster for a forest)
    DevMem& s0 = ctx->scratch[0];
    const unsigned nbb = (unsigned)std::max<size_t>(1, std::min<size_t>(grid_for(n), 1024));
    GZ_TRY(s0.reserve(nbb * 6 * sizeof(double) + 16 * sizeof(double) +
                      (nc > 1 ? (size_t)nc * (4 * sizeof(double) + 2 * sizeof(int)) + 512 : 0)));
    double* partial = s0.as<double>();
    double* droot = partial + (size_t)nbb * 6;
    double* croots = droot + 16;                                     // nc x 4 (forest)
    int* ccount = reinterpret_cast<int*>(croots + (size_t)4 * nc);   // nc + 1 (forest)
    uint32_t* cstart_dev = nullptr;
    if (nc > 1) {
        GZ_TRY(dalloc(reinterpret_cast<void**>(&t->cluster_start_dev), (nc + 1) * sizeof(uint32_t)));
        std::vector<uint32_t> hc(nc + 1);
        for (int c = 0; c <= nc; ++c) hc[c] = (uint32_t)opts->cluster_start[c];
        GZ_HIP(hipMemcpy(t->cluster_start_dev, hc.data(), hc.size() * sizeof(uint32_t),
                         hipMemcpyHostToDevice));
        cstart_dev = t->cluster_start_dev;
        k_bbox_clusters<<<nc, kThreads, 0, st>>>(p->x, p->y, p->z, p->m, cstart_dev, croots);
        // droot keeps the first cluster's cell for the tree's info
        GZ_HIP(hipMemcpyAsync(droot, croots, 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
        k_bbox_partial<<<nbb, kThreads, 0, st>>>(p->x, p->y, p->z, p->m, n, partial);
        k_root_cell<<<1, 64, 0, st>>>(partial, (int)nbb, seed, fx, fy, fz, droot);
    }
    GZ_HIP(hipGetLastError());
    double hroot[10];
    GZ_HIP(hipMemcpyAsync(hroot, droot, sizeof hroot, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipEventRecord(ev[1], st));

    // --- live filter + keys
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->rank), n * sizeof(uint32_t)));
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->leaf_of), n * sizeof(int)));
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->slot_leaf), n * sizeof(int)));
    DevMem& s1 = ctx->scratch[1];
    const size_t nalloc = std::max<size_t>(n, 1);
    const size_t flag_bytes = (nalloc + 64) * sizeof(uint32_t);
    GZ_TRY(s1.reserve(flag_bytes + scan_tmp_bytes(nalloc) + 64));
    uint32_t* flags = s1.as<uint32_t>();
    uint32_t* dcount = flags + nalloc;
    void* stmp = reinterpret_cast<uint8_t*>(s1.p) + flag_bytes;
    // order / keys ping-pong buffers
    DevMem& s2 = ctx->scratch[2];
    GZ_TRY(s2.reserve(nalloc * (2 * sizeof(uint64_t) + 2 * sizeof(uint32_t)) + 256));
    uint64_t* k0 = s2.as<uint64_t>();
    uint64_t* k1 = k0 + nalloc;
    uint32_t* v0 = reinterpret_cast<uint32_t*>(k1 + nalloc);
    uint32_t* v1 = v0 + nalloc;
    uint32_t n_live32 = 0;
    if (n > 0) {
        k_live_flags<<<grid_for(n), kThreads, 0, st>>>(p->m, n, flags);
        GZ_TRY(exclusive_scan_u32(flags, flags, n, dcount, stmp, st));
        k_live_compact<<<grid_for(n), kThreads, 0, st>>>(p->m, n, flags, v0);
        GZ_HIP(hipGetLastError());
        GZ_HIP(hipMemcpyAsync(&n_live32, dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    }
    GZ_HIP(hipStreamSynchronize(st));
    for (int k = 0; k < 4; ++k) t->root_box[k] = hroot[k];
    const size_t nl = n_live32;
    t->n_live = nl;
    const double* key_roots = nc > 1 ? croots : droot;
    if (nl > 0) {
        k_keys<<<grid_for(nl), kThreads, 0, st>>>(v0, nl, p->x, p->y, p->z, key_roots, cstart_dev,
                                                  nc, k0);
        GZ_HIP(hipGetLastError());
    }
    GZ_HIP(hipEventRecord(ev[2], st));

    // --- sort (a forest: by key, then stably by cluster, so each cluster is one sorted block)
    GZ_TRY(radix_sort_pairs(k0, v0, k1, v1, nl, 3 * kMaxLevel, ctx->scratch[3], st));
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->skey), nl * sizeof(uint64_t)));
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->orderbuf), nl * sizeof(uint32_t)));
    std::vector<int> rstart_h(nc + 1, 0);
    if (nc > 1 && nl > 0) {
        k_cluster_keys<<<grid_for(nl), kThreads, 0, st>>>(v0, nl, cstart_dev, nc, k1);
        GZ_HIP(hipGetLastError());
        int bits = 1;
        while ((1 << bits) < nc) ++bits;
        GZ_TRY(radix_sort_pairs(k1, v0, k0, v1, nl, bits, ctx->scratch[3], st));
        k_keys<<<grid_for(nl), kThreads, 0, st>>>(v0, nl, p->x, p->y, p->z, key_roots,
                                                  cstart_dev, nc, t->skey);
        GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(ccount), nc, st));
        k_cluster_counts<<<grid_for(nl), kThreads, 0, st>>>(v0, nl, cstart_dev, nc, ccount);
        GZ_HIP(hipGetLastError());
        std::vector<int> cnt(nc);
        GZ_HIP(hipMemcpyAsync(cnt.data(), ccount, nc * sizeof(int), hipMemcpyDeviceToHost, st));
        GZ_HIP(hipStreamSynchronize(st));
        for (int c = 0; c < nc; ++c) rstart_h[c + 1] = rstart_h[c] + cnt[c];
    } else if (nl) {
        GZ_HIP(hipMemcpyAsync(t->skey, k0, nl * sizeof(uint64_t), hipMemcpyDeviceToDevice, st));
    }
    if (nl)
        GZ_HIP(hipMemcpyAsync(t->orderbuf, v0, nl * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    GZ_HIP(hipEventRecord(ev[3], st));

    // --- sorted source copy, rank, leaf_of reset
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->src), nl * sizeof(double4)));
    GZ_TRY(dalloc(reinterpret_cast<void**>(&t->src_soft), nl * sizeof(double)));
    if (n) {
        k_fill_u32<<<grid_for(n), kThreads, 0, st>>>(t->rank, n, 0xFFFFFFFFu);
        k_fill_u32<<<grid_for(n), kThreads, 0, st>>>(reinterpret_cast<uint32_t*>(t->leaf_of), n,
                                                     0xFFFFFFFFu);
    }
    if (nl)
        k_gather_sorted<<<grid_for(nl), kThreads, 0, st>>>(t->orderbuf, nl, p->x, p->y, p->z,
                                                           p->m, p->soft, t->src, t->src_soft,
                                                           t->rank);
    GZ_HIP(hipGetLastError());
    GZ_HIP(hipEventRecord(ev[4], st));

    // --- topology, level by level
    int cap = (int)std::min<size_t>((2 * nl) / kLeafMax * 3 + 1024, 0x7FFFFFF0ull);
    if (max_nodes > 0) cap = std::min(cap, max_nodes);
    GZ_TRY(grow_nodes(t, 0, cap, st));
    t->level_begin.clear();
    t->level_begin.push_back(0);
    t->arena_grows = 0;
    int nnodes = 0;
    {
        // root: range [0, nl), the root cell, no parent, no successor. An empty particle set
        // still gets a root (a leaf of mass 0), as the reference's build_node does for [0, 0).
        if (nc > 1) {
            // one root per cluster at level 0, each over its block of the sorted order
            int* rstart_d = ccount; // reuse: nc + 1 sorted offsets
            GZ_HIP(hipMemcpyAsync(rstart_d, rstart_h.data(), (nc + 1) * sizeof(int),
                                  hipMemcpyHostToDevice, st));
            if (nc > cap) {
                GZ_TRY(grow_nodes(t, 0, nc + 1024, st));
                cap = nc + 1024;
            }
            k_init_roots<<<grid_for(nc), kThreads, 0, st>>>(nc, croots, rstart_d, t->nlo, t->nhi,
                                                            t->gcx, t->gcy, t->gcz, t->size,
                                                            t->parent, t->next);
            GZ_HIP(hipStreamSynchronize(st)); // rstart_h is a host vector read by the copy
        } else {
            k_init_root<<<1, 1, 0, st>>>(droot, (int)nl, t->nlo, t->nhi, t->gcx, t->gcy, t->gcz,
                                         t->size, t->parent, t->next);
        }
        GZ_HIP(hipGetLastError());
        nnodes = nc;
        t->level_begin.push_back(nc);
        DevMem& s4 = ctx->scratch[4];
        for (int level = 0;; ++level) {
            const int lb = t->level_begin[level], le = t->level_begin[level + 1];
            const int count = le - lb;
            if (count == 0) break;
            const size_t per = (size_t)count;
            const size_t ob = per * 8 * sizeof(int), cb = (per + 64) * sizeof(uint32_t);
            GZ_TRY(s4.reserve(ob + 2 * cb + scan_tmp_bytes(per) + 256));
            int* oct = s4.as<int>();
            uint32_t* nk = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(s4.p) + ob);
            uint32_t* off = reinterpret_cast<uint32_t*>(reinterpret_cast<uint8_t*>(nk) + cb);
            uint32_t* tot = off + per;
            void* tmp = reinterpret_cast<uint8_t*>(off) + cb;
            k_split<<<grid_for(per), kThreads, 0, st>>>(lb, count, level, t->nlo, t->nhi, t->skey,
                                                        oct, nk);
            GZ_HIP(hipGetLastError());
            GZ_TRY(exclusive_scan_u32(nk, off, per, tot, tmp, st));
            uint32_t total = 0;
            GZ_HIP(hipMemcpyAsync(&total, tot, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            GZ_HIP(hipStreamSynchronize(st));
            const long long need = (long long)le + (long long)total;
            if (need > 0x7FFFFFF0ll) return fail(GZ_ERR_NOMEM, "gz_tree_build: node count overflow");
            if (need > cap) {
                if (max_nodes > 0)
                    return fail(GZ_ERR_ARENA,
                                "gz_tree_build: tree needs more than max_nodes=%d nodes", max_nodes);
                const long long want = std::min<long long>(2 * need + 1024, 0x7FFFFFF0ll);
                GZ_TRY(grow_nodes(t, le, (int)want, st));
                cap = (int)want;
                ++t->arena_grows;
            }
            k_emit<<<grid_for(per), kThreads, 0, st>>>(lb, count, le, nk, off, oct, t->nlo, t->nhi,
                                                       t->gcx, t->gcy, t->gcz, t->size, t->parent,
                                                       t->next, t->wn);
            GZ_HIP(hipGetLastError());
            t->level_begin.push_back(le + (int)total);
            nnodes = le + (int)total;
            if (total == 0) break;
        }
    }
    t->nnodes = nnodes;
    t->depth = (int)t->level_begin.size() - 2;
    {
        // the level table on the device, padded with nnodes so any level index is valid
        int lb[kMaxLevel + 2];
        for (int L = 0; L < kMaxLevel + 2; ++L)
            lb[L] = (L < (int)t->level_begin.size()) ? t->level_begin[L] : nnodes;
        GZ_TRY(dalloc(reinterpret_cast<void**>(&t->level_begin_dev), sizeof lb));
        GZ_HIP(hipMemcpy(t->level_begin_dev, lb, sizeof lb, hipMemcpyHostToDevice));
    }
    GZ_HIP(hipEventRecord(ev[5], st));

    // --- reuse state (tree.h:250-317): vmax and xpad always, the velocity moments with velocities
    NodeVel nv{};
    GZ_TRY(alloc_reuse(t, nnodes, opts && opts->vx && opts->vy && opts->vz));
    if (!t->age) GZ_TRY(dalloc(reinterpret_cast<void**>(&t->age), sizeof(double)));
    k_fill_u32<<<1, kThreads, 0, st>>>(reinterpret_cast<uint32_t*>(t->age), 2, 0u);
    nv.nn = nnodes;
    nv.vmax = t->vmax;
    nv.xpad = t->xpad;
    if (t->has_vel) {
        if (t->svel) (void)hipFree(t->svel);
        t->svel = nullptr;
        GZ_TRY(dalloc(reinterpret_cast<void**>(&t->svel),
                      std::max<size_t>(t->n_live, 1) * sizeof(double4)));
        if (t->n_live)
            k_gather_vel<<<grid_for(t->n_live), kThreads, 0, st>>>(t->orderbuf, t->n_live,
                                                                   opts->vx, opts->vy, opts->vz,
                                                                   t->svel);
        GZ_HIP(hipGetLastError());
        nv.vx = opts->vx;
        nv.vy = opts->vy;
        nv.vz = opts->vz;
        nv.vcom = t->vcom;
        nv.m0 = t->m0;
        nv.dp = t->dp;
        nv.dq = t->dq;
        nv.dmass = t->dmass;
    }

    // --- moments + pack, deepest level first
    int* dleaf = reinterpret_cast<int*>(droot + 12);
    k_fill_u32<<<1, kThreads, 0, st>>>(reinterpret_cast<uint32_t*>(dleaf), 1, 0u);
    for (int level = (int)t->level_begin.size() - 2; level >= 0; --level) {
        const int lb = t->level_begin[level], le = t->level_begin[level + 1];
        if (le <= lb) continue;
        k_moments<<<grid_for(le - lb), kThreads, 0, st>>>(lb, le - lb, t->wn, t->nlo, t->nhi,
                                                          t->gcx, t->gcy, t->gcz, t->size, t->src,
                                                          t->src_soft, t->orderbuf, t->leaf_of,
                                                          t->slot_leaf, nv, t->delta, t->soft,
                                                          dleaf);
        GZ_HIP(hipGetLastError());
    }
    GZ_HIP(hipEventRecord(ev[6], st));
    int nleaf = 0;
    GZ_HIP(hipMemcpyAsync(&nleaf, dleaf, sizeof(int), hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    t->nleaves = nleaf;
    t->root = 0;
    {
        WNode r;
        GZ_HIP(hipMemcpy(&r, t->wn, sizeof r, hipMemcpyDeviceToHost));
        t->root_com[0] = r.cx;
        t->root_com[1] = r.cy;
        t->root_com[2] = r.cz;
        t->root_mass = r.mass;
    }
    t->times.bbox = ev_ms(ev[0], ev[1]);
    t->times.keys = ev_ms(ev[1], ev[2]);
    t->times.sort = ev_ms(ev[2], ev[3]);
    t->times.links = ev_ms(ev[3], ev[4]);
    t->times.recurse = ev_ms(ev[4], ev[5]);
    t->times.pack = ev_ms(ev[5], ev[6]);
    t->times.total = ev_ms(ev[0], ev[6]);
    return GZ_OK;
}

// Refresh the tree's sorted copy of the source particles (positions, masses, softenings) from p,
// for callers that move particles between a build and a walk. The reference walk reads source
// particles from the live arrays (tree.cc:787-788); this library's walk reads the sorted copy.
gz_status regather_sources(gz_ctx* ctx, const gz_particles* p, gz_tree* t) {
    if (p->n != t->n) return fail(GZ_ERR_INVALID, "particle count %zu differs from the tree's %zu",
                                  p->n, t->n);
    if (t->n_live)
        k_gather_sorted<<<grid_for(t->n_live), kThreads, 0, ctx->stream>>>(
            t->orderbuf, t->n_live, p->x, p->y, p->z, p->m, p->soft, t->src, t->src_soft,
            t->rank);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // namespace gz
