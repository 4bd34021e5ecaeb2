// Tree walks on the device: target grouping, per-group opening bounds, the group-coherent walk
// with an LDS interaction list (FP64 and FP32 group-relative), and the exact direct sum.
//
// Semantics follow shmem::accel_grouped (shmem/tree.cc:626-897) without lazy drift or a
// sink-direct radius; the tidal tensor and jerk (tree.cc:811-871, the TJ instantiation) are the
// same walk with other sums (the per-target walks, shmem::accel and shmem::potential, are in
// gz_walk_pt.hip):
//   * groups are runs of Morton-ordered targets that share a leaf (tree.cc:656-667), or fixed
//     windows of `batch` targets, or (group_size > 0) runs under the highest node that holds at most
//     group_size particles;
//   * each group is tested against every node with the bounding box of its targets, their largest
//     softening and their smallest aold (tree.cc:683-702), using the reference's open_node
//     (tree.h:109-144) evaluated in FP64 exactly as written there;
//   * an accepted node (leaves included) is a monopole, an opened leaf is summed pair by pair
//     (tree.cc:779-874), both with the cubic-spline kernel of the larger softening (tree.h:59-67,
//     tree.cc:39-53, 58-68).
//
// How it runs on CDNA: one wavefront (64 lanes) owns one group. The traversal is wave-coherent:
// up to 64 candidate nodes are popped from an LDS stack at once, one per lane, their opening tests
// run in parallel (the decision only depends on the group, so it is the same for every target),
// accepted nodes and the particles of opened leaves are appended to an LDS interaction list, and
// the children of opened interior nodes (contiguous in this tree's level-major numbering) are
// pushed back. When the list fills, the whole wave evaluates it: lanes are (target, sub-slice)
// pairs, each reading list entries as LDS broadcasts. The set of interactions each target receives
// is exactly the reference's; only the order of the summation differs.
#include <algorithm>
#include <cmath>
#include <cstdio>

#include "gz_features.h"
#include "gz_grav_math.h"
#include "gz_wave.h"

namespace gz {

constexpr int kWalkStack = 512;   // node ids per wave in LDS
constexpr int kSpillKeep = 256;   // LDS entries kept when the bottom of the stack is spilled
constexpr int kXcdChunk = 64;     // groups per chunk of the persistent launch's XCD interleave
constexpr int kMonoF = 120;       // monopole entries per wave, FP32 (16 B each)
constexpr int kPairF = 88;        // particle entries per wave, FP32 (20 B each)
constexpr int kMonoD = 64;        // monopole entries per wave, FP64 (32 B each)
constexpr int kPairD = 64;        // particle entries per wave, FP64 (40 B each)
constexpr int kMaxChunk = 64;     // targets evaluated together (one wave)
constexpr int kPadF = 16;         // FP32 lists: zero-mass entries after the last (two slice steps)
// A batch appends up to 64 entries to a list right after flushing it (64 accepted nodes, or 64
// leaf particles), so every list holds at least 64; the stack takes one batch's 512 children.
static_assert(kMonoF >= 64 && kPairF >= 64 && kMonoD >= 64 && kPairD >= 64,
              "a list must take one batch of 64 entries after a flush");
static_assert(kWalkStack >= 512, "the LDS stack must take the 512 children of one batch");

// ---- grouping ------------------------------------------------------------------------------------
// Group key of each target slot: its leaf (batch <= 0, group_size == 0), the highest ancestor of the
// leaf holding at most group_size particles (group_size > 0), or the window index (batch > 0).
// A target outside the tree (leaf_of < 0) always starts a group of its own (tree.cc:663-664).
// The group node of every leaf for group_size > 0: its highest ancestor holding at most
// group_size particles. One thread per node (interior nodes are skipped), so each leaf climbs
// once however many targets it holds.
__global__ __launch_bounds__(256) void k_leaf_group(int nnodes, const WNode* __restrict__ wn,
                                                    const int* __restrict__ parent,
                                                    int group_size, int* __restrict__ leaf_group) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nnodes || wn[i].first >= 0) return;
    int key = i, p = parent[i];
    while (p >= 0 && wn[p].phi - wn[p].plo <= group_size) {
        key = p;
        p = parent[p];
    }
    leaf_group[i] = key;
}

__global__ __launch_bounds__(256) void k_group_flags(const uint32_t* __restrict__ targets,
                                                     size_t nt, const uint32_t* __restrict__ rank,
                                                     const uint32_t* __restrict__ orderbuf,
                                                     size_t n_live,
                                                     const int* __restrict__ slot_leaf,
                                                     int* __restrict__ tslot,
                                                     const int* __restrict__ parent,
                                                     const WNode* __restrict__ wn,
                                                     const int* __restrict__ leaf_group,
                                                     int batch, int group_size, size_t n,
                                                     const uint32_t* __restrict__ seg, int nseg,
                                                     int* __restrict__ gseg,
                                                     int* __restrict__ gkey,
                                                     int* __restrict__ status) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nt) return;
    if (nseg > 1) { // the segment (independent target list) this slot belongs to
        int lo = 0, hi = nseg; // last s with seg[s] <= i
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (seg[mid] <= i) lo = mid;
            else hi = mid;
        }
        gseg[i] = lo;
    }
    if (targets[i] >= n) { // reported to the caller before anything reads through it
        atomicOr(status, 2);
        gkey[i] = -1;
        tslot[i] = -1;
        return;
    }
    // the target's sorted slot: its data is read from the tree's sorted copy, which for a
    // Morton-ordered target list is a near-sequential read (the particle arrays are gathers).
    // A target that is orderbuf[i] sits in slot i (an all-active list is orderbuf itself), which
    // spares the random read of rank.
    const uint32_t r = (i < n_live && orderbuf[i] == targets[i]) ? (uint32_t)i : rank[targets[i]];
    const int slot = (r == 0xFFFFFFFFu) ? -1 : (int)r;
    tslot[i] = slot;
    int key;
    if (batch > 0) {
        key = (int)(i / (size_t)batch);
    } else {
        key = (slot >= 0) ? slot_leaf[slot] : -1;
        if (key >= 0 && group_size > 0) {
            if (leaf_group) {
                key = leaf_group[key];
            } else {
                int p = parent[key];
                while (p >= 0 && wn[p].phi - wn[p].plo <= group_size) {
                    key = p;
                    p = parent[p];
                }
            }
        }
    }
    gkey[i] = key;
}

__global__ __launch_bounds__(256) void k_group_flags2(size_t nt, int batch,
                                                      const int* __restrict__ gkey,
                                                      const int* __restrict__ gseg,
                                                      uint32_t* __restrict__ flag) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nt) return;
    if (i == 0) {
        flag[i] = 1u;
        return;
    }
    const int k = gkey[i], kp = gkey[i - 1];
    bool f = (batch > 0) ? (k != kp) : (k != kp || k < 0);
    if (gseg && gseg[i] != gseg[i - 1]) f = true; // groups never span two segments
    flag[i] = f ? 1u : 0u;
}

// Sibling packing (group_size > 0 only): groups whose key nodes are children of the same parent
// are merged greedily, in octant order, while the merged group holds at most group_size targets.
// A merged group's box lies inside the parent's cell, so it stays compact, and the wavefront that
// walks it has more of its 64 lanes busy in the evaluation.
__global__ __launch_bounds__(256) void k_pack_parent(int ngroups, size_t nt,
                                                     const int* __restrict__ gstart,
                                                     const int* __restrict__ gkey,
                                                     const int* __restrict__ parent,
                                                     const int* __restrict__ gseg,
                                                     long long* __restrict__ pk) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ngroups) return;
    const int key = gkey[gstart[i]];
    // merge partner: same parent node, same segment; a target outside the tree never merges
    const long long sg = gseg ? (long long)gseg[gstart[i]] : 0;
    pk[i] = (key >= 0 && parent[key] >= 0) ? ((sg << 32) | (long long)parent[key]) : -2 - i;
}

__global__ __launch_bounds__(256) void k_pack_greedy(int ngroups, size_t nt, int group_size,
                                                     const int* __restrict__ gstart,
                                                     const long long* __restrict__ pk,
                                                     uint32_t* __restrict__ mflag) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ngroups) return;
    const long long p = pk[i];
    const bool seg_start = (i == 0) || (p < 0) || (pk[i - 1] != p);
    if (!seg_start) return;
    int acc = 0;
    for (int j = i; j < ngroups && (j == i || (p >= 0 && pk[j] == p)); ++j) {
        const int hi = (j + 1 < ngroups) ? gstart[j + 1] : (int)nt;
        const int cnt = hi - gstart[j];
        if (j == i || acc + cnt > group_size) {
            mflag[j] = 1u;
            acc = cnt;
        } else {
            mflag[j] = 0u;
            acc += cnt;
        }
    }
}

__global__ __launch_bounds__(256) void k_pack_starts(int ngroups, const int* __restrict__ gstart,
                                                     const uint32_t* __restrict__ mflag,
                                                     const uint32_t* __restrict__ mid,
                                                     int* __restrict__ gstart2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= ngroups) return;
    if (mflag[i]) gstart2[mid[i]] = gstart[i];
}


// A value known to be the same in every lane, moved to scalar registers.
__device__ __forceinline__ double uniform_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xFFFFFFFFll));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Position and softening of target slot i: from the tree's sorted copy when the target is in the
// tree, else from the particle arrays.
__device__ __forceinline__ void target_data(const WalkArgs& a, int i, double& px, double& py,
                                            double& pz, double& pe) {
    const int s = a.tslot[i];
    if (s >= 0) {
        const double4 q = a.src[s];
        px = q.x; py = q.y; pz = q.z;
        pe = a.src_soft[s];
    } else {
        const uint32_t p = a.targets[i];
        px = a.x[p]; py = a.y[p]; pz = a.z[p];
        pe = a.soft ? a.soft[p] : 0.0;
    }
}

// The opening bounds of group g (tree.cc:683-702: bounding box, largest softening, smallest aold),
// computed by the group's own wavefront. Min and max are exact, so every lane ends with the same
// values whatever the reduction order. A learned tolerance enters as a factor on aold_min.
__device__ __forceinline__ GroupInfo group_bounds(const WalkArgs& a, int g, int ngroups,
                                                  int lane) {
    const int lo = a.gstart[g];
    const int hi = (g + 1 < ngroups) ? a.gstart[g + 1] : a.nt;
    double bx0 = 1e300, by0 = 1e300, bz0 = 1e300, bx1 = -1e300, by1 = -1e300, bz1 = -1e300;
    double emax = 0, amin = a.aold ? 1e300 : 0.0;
    for (int i = lo + lane; i < hi; i += 64) {
        double px, py, pz, pe;
        target_data(a, i, px, py, pz, pe);
        bx0 = dmin_w(bx0, px); bx1 = dmax_w(bx1, px);
        by0 = dmin_w(by0, py); by1 = dmax_w(by1, py);
        bz0 = dmin_w(bz0, pz); bz1 = dmax_w(bz1, pz);
        emax = dmax_w(emax, pe);
        if (a.aold) amin = dmin_w(amin, a.aold[i]);
    }
    GroupInfo o;
    o.lo[0] = wave_min_d(bx0); o.lo[1] = wave_min_d(by0); o.lo[2] = wave_min_d(bz0);
    o.hi[0] = wave_max_d(bx1); o.hi[1] = wave_max_d(by1); o.hi[2] = wave_max_d(bz1);
    o.emax = wave_max_d(emax);
    o.amin = wave_min_d(amin);
    o.theta_sq = a.theta_sq;
    o.tlo = lo;
    o.thi = hi;
    if (a.group_scale) o.amin *= (double)a.group_scale[g];
    return o;
}

// ---- the walk --------------------------------------------------------------------------------------
template <bool FP64> struct WalkTypes;
template <> struct WalkTypes<true> {
    using real = double;
    using real4 = double4;
};
template <> struct WalkTypes<false> {
    using real = float;
    using real4 = float4;
};

// Set bits of m below this lane (mbcnt: two VALU instructions, no cross-lane traffic).
__device__ __forceinline__ int lanes_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// FP32 cubic-spline force/r inside the kernel, 0 < r <= h (tree.h:59-67 in single precision),
// given r^2, inv = 1/r and g3 = 1/r^3 from the Newtonian path, and h^2. Both branches are
// evaluated and one selected, with the hardware 1/sqrt(h^2) in place of the division; the outer
// branch's -1/(15 u^3) term times h^-3 is -1/(15 r^3) = -g3/15, so it needs no second reciprocal.
__device__ __forceinline__ float spline_inside_h(float r2, float inv, float g3, float hinv,
                                                 float hinv3) {
    const float u = (r2 * inv) * hinv;
    const float u2 = u * u;
    // 32/3 + u^2 (32 u - 38.4) and 64/3 - 48 u + 38.4 u^2 - 32/3 u^3, in Horner form
    const float lo = __builtin_fmaf(u2, __builtin_fmaf(32.0f, u, -38.4f), 10.666666666666667f);
    const float hi = __builtin_fmaf(
        u, __builtin_fmaf(u, __builtin_fmaf(u, -10.666666666666667f, 38.4f), -48.0f),
        21.333333333333332f);
    const bool inner = u < 0.5f;
    return __builtin_fmaf(-0.06666666666666667f, inner ? 0.0f : g3, hinv3 * (inner ? lo : hi));
}

// The same for a kernel extent h given as h^2 (hinv = 1/h from the hardware reciprocal square
// root, hinv3 = hinv^3).
__device__ __forceinline__ float spline_inside_f(float r2, float inv, float g3, float h2) {
    const float hinv = __builtin_amdgcn_rsqf(h2);
    return spline_inside_h(r2, inv, g3, hinv, hinv * hinv * hinv);
}

// The walks run five waves per SIMD: at most 96 VGPRs (the wave-uniform group data lives in
// scalar registers) and under 8 KB of LDS per wave (the node stack, both lists; FP64 7.2 KB with
// 64-entry monopole lists, which measured faster at five waves than 96 entries at four); the
// counting FP64 walk is given the registers of four, and the tidal walk's extra accumulators
// leave room for three.
// TJ (FP64 only): the tidal tensor and, when velocities are given, the jerk, accumulated in the
// same walk (tree.cc:811-871).
template <bool FP64, bool COUNT, bool GSTACK, bool TJ = false>
__global__ __launch_bounds__(64, TJ ? 3 : ((FP64 && COUNT) ? 4 : 5)) void k_walk(WalkArgs a) {
    static_assert(!TJ || FP64, "the tidal tensor and jerk are FP64 walks");
    using real = typename WalkTypes<FP64>::real;
    using real4 = typename WalkTypes<FP64>::real4;
    constexpr int kMono = FP64 ? kMonoD : kMonoF;
    constexpr int kPair = FP64 ? kPairD : kPairF;
    // The node stack lives in LDS; a group whose stack would outgrow it is handed to a second
    // launch of this kernel (GSTACK) that keeps the stack in a per-wave slice of global memory.
    __shared__ int s_stack[GSTACK ? 1 : kWalkStack];
    // Interaction lists. An accepted node always lies outside every target's softening (the
    // first arm of open_node), so monopoles are Newtonian and need no softening (the FP32 walk
    // also lists there the particles of opened leaves that lie outside every target's kernel);
    // particle entries carry theirs (squared in the FP32 walk).
    // The FP32 lists end in up to kPadF zero-mass entries far away, so that every lane steps
    // through whole slice steps with no per-lane bounds test (the loop stays wave-uniform).
    constexpr int kPad = FP64 ? 0 : kPadF;
    __shared__ real4 s_mono[kMono + kPad];
    __shared__ real4 s_pair[kPair + kPad];
    __shared__ real s_psoft[kPair + kPad];
    // TJ: node velocity and softening per monopole, source velocity per particle entry
    __shared__ double4 s_mext[TJ ? kMono : 1];
    __shared__ double4 s_pext[TJ ? kPair : 1];
    __shared__ int s_lpre[65];
    __shared__ int s_lplo[64];
    int* const stk = GSTACK ? a.gstack + (size_t)blockIdx.x * (size_t)a.gstack_cap : s_stack;
    const int scap = GSTACK ? a.gstack_cap
                            : (a.stack_limit > 0 ? min(a.stack_limit, kWalkStack) : kWalkStack);
    // Persistent waves extend the LDS stack downwards into a slice of global memory: when a push
    // would overflow LDS, the bottom of the LDS stack moves to the slice, and it comes back when
    // LDS runs empty. The logical stack, and so the traversal order, is that of an unbounded stack.
    int* const spill = (!GSTACK && a.persistent && a.spill)
                           ? a.spill + (size_t)blockIdx.x * (size_t)a.spill_cap
                           : nullptr;

    // the group count: read from the device when the host holds only an upper bound
    const int ngroups = a.ngroups_dev ? (int)*a.ngroups_dev : a.ngroups;
    const int lane = threadIdx.x;
    long long n_visit = 0, n_multi = 0, n_pair = 0;
    int max_sp = 0;

    for (int item = 0;; ++item) {
        long long g_visit = 0, g_multi = 0, g_pair = 0; // this group's share of the counts
        int g;
        if (GSTACK) {
            const int k = blockIdx.x + item * gridDim.x;
            if (k >= *a.n_retry) break;
            g = a.retry_list[k];
        } else if (a.persistent) {
            // Eight queues, one per XCD (block b runs on XCD b % 8). The Morton-ordered groups
            // are cut into chunks of kXcdChunk, dealt to the XCDs in turn, so that the waves of
            // one XCD walk compact runs of neighbouring groups (their L2 then holds the nodes
            // those groups share) while all XCDs sweep the Morton order at the same pace. A
            // wave whose queue is empty helps the others. Every group is handed out exactly
            // once; a wave stops when all eight queues are empty.
            int v = -1;
            if (lane == 0) {
                const int x0 = blockIdx.x & 7;
                constexpr int C = kXcdChunk;
                const int nch = (ngroups + C - 1) / C; // chunks; chunk k belongs to XCD k % 8
                for (int k = 0; k < 8 && v < 0; ++k) {
                    const int x = (x0 + k) & 7;
                    const int myc = (nch > x) ? (nch - x + 7) / 8 : 0; // chunks of XCD x
                    const int n = myc * C;
                    const int c = atomicAdd(&a.queue[x], 1);
                    if (c < n) {
                        const int gg = ((c / C) * 8 + x) * C + (c % C);
                        if (gg < ngroups) v = gg;
                        else k = -1; // the tail of the last chunk: claim again (the loop ends)
                    }
                }
            }
            g = __shfl(v, 0);
            if (g < 0) break;
        } else {
            // Block b runs on XCD b % 8; give each XCD one contiguous share of the Morton-ordered
            // groups so that its L2 serves one compact region of the tree.
            if (item > 0) break;
            const int b = blockIdx.x;
            const int nb = gridDim.x;
            const int xcd = b & 7, slot = b >> 3;
            const int per = nb >> 3, rem = nb & 7;
            g = xcd * per + min(xcd, rem) + slot; // a bijection of [0, nb)
            if (g >= ngroups) break;
        }

        GroupInfo gi = a.groups ? a.groups[g] : group_bounds(a, g, ngroups, lane);
        // every field is the same in all lanes: keep them in scalar registers
        for (int k = 0; k < 3; ++k) {
            gi.lo[k] = uniform_d(gi.lo[k]);
            gi.hi[k] = uniform_d(gi.hi[k]);
        }
        gi.emax = uniform_d(gi.emax);
        gi.amin = uniform_d(gi.amin);
        gi.theta_sq = uniform_d(gi.theta_sq);
        gi.tlo = __builtin_amdgcn_readfirstlane(gi.tlo);
        gi.thi = __builtin_amdgcn_readfirstlane(gi.thi);
        const double gcx = 0.5 * (gi.lo[0] + gi.hi[0]);
        const double gcy = 0.5 * (gi.lo[1] + gi.hi[1]);
        const double gcz = 0.5 * (gi.lo[2] + gi.hi[2]);
        const int nt_all = gi.thi - gi.tlo;
        // FP32: the group box's half-sides about the centre, widened past the rounding of the
        // group-relative coordinates, and the largest target softening. A leaf particle farther
        // from the box than max(emax, its own softening) lies outside the kernel of every target
        // (u > 1), where the pair force is exactly the monopole form, so it joins the monopole
        // list instead of the particle list.
        const double gslack = 1e-12 * (fabs(gcx) + fabs(gcy) + fabs(gcz));
        const float bhx = (float)(0.5 * (gi.hi[0] - gi.lo[0]) * (1.0 + 1e-5) + gslack);
        const float bhy = (float)(0.5 * (gi.hi[1] - gi.lo[1]) * (1.0 + 1e-5) + gslack);
        const float bhz = (float)(0.5 * (gi.hi[2] - gi.lo[2]) * (1.0 + 1e-5) + gslack);
        const float emaxf = (float)gi.emax;

        for (int c0 = 0; c0 < nt_all; c0 += kMaxChunk) {
            const int tlo = gi.tlo + c0;
            const int nt = min(kMaxChunk, nt_all - c0);
            // Lanes are (target t, slice sub) pairs: lane = sub * nt + t. Each of the L slices
            // takes every L-th list entry; lanes beyond L * nt sit out the evaluation.
            // (the FP32 lists' padding covers two steps of up to 8 slices; FP64 loops test bounds)
            const int L = min(64 / nt, FP64 ? 64 : 8);
            const int t = lane % nt;
            const int sub = lane / nt;
            const bool tvalid = sub < L;
            const int jstart = tvalid ? sub : (1 << 30);
            real tx = 0, ty = 0, tz = 0, te = 0;
            double pe_t = 1e300; // the target's softening (FP64), for the uniform-kernel test
            if (tvalid) {
                double px, py, pz, pe;
                target_data(a, tlo + t, px, py, pz, pe);
                pe_t = pe;
                if (FP64) {
                    tx = (real)px; ty = (real)py; tz = (real)pz;
                } else {
                    tx = (real)(px - gcx); ty = (real)(py - gcy); tz = (real)(pz - gcz);
                }
                te = (real)pe;
            }
            double accx = 0, accy = 0, accz = 0;
            double tt[6] = {0}, jj[3] = {0}; // tidal (xx yy zz xy yz xz), jerk (TJ only)
            double tvx = 0, tvy = 0, tvz = 0;
            if (TJ && tvalid && a.jerk) {
                const uint32_t p = a.targets[tlo + t];
                tvx = a.pvx[p]; tvy = a.pvy[p]; tvz = a.pvz[p];
            }
            int mcount = 0, pcount = 0;
            bool overflow = false;
            // FP32: when no particle below the group's root is softer than the chunk's least
            // softened target, every pair's kernel extent max(h_t, h_s) is the target's own, so
            // the particle loop takes h = h_t and its reciprocal powers once per lane
            const int groot = root_of(a.targets[gi.tlo], a.cluster_start, a.nclusters, a.root);
            const bool uni = !FP64 && a.node_soft[groot] <= wave_min_d(pe_t);

            // Evaluate both lists against every target of the chunk.
            auto flush = [&]() {
                if (!FP64 && lane < 2 * L) {
                    const real4 far = {(real)1e18, (real)1e18, (real)1e18, (real)0};
                    s_mono[mcount + lane] = far;
                    s_pair[pcount + lane] = far;
                    s_psoft[pcount + lane] = (real)0;
                }
                __syncthreads();
                if (TJ) {
                    // tree.cc:811-871: g1 and g2 are the first and second kernel-derivative
                    // factors; the tidal tensor is -g1 delta + g2 d d (no G), the jerk
                    // g1 dv - (dv . d) g2 d
                    auto tj = [&](double dx, double dy, double dz, double g1, double g2,
                                  const double4& v) {
                        accx += dx * g1;
                        accy += dy * g1;
                        accz += dz * g1;
                        tt[0] += -g1 + dx * dx * g2;
                        tt[1] += -g1 + dy * dy * g2;
                        tt[2] += -g1 + dz * dz * g2;
                        tt[3] += dx * dy * g2;
                        tt[4] += dy * dz * g2;
                        tt[5] += dx * dz * g2;
                        if (a.jerk) {
                            const double dvx = v.x - tvx, dvy = v.y - tvy, dvz = v.z - tvz;
                            const double vdotr = dvx * dx + dvy * dy + dvz * dz;
                            jj[0] += g1 * dvx - vdotr * g2 * dx;
                            jj[1] += g1 * dvy - vdotr * g2 * dy;
                            jj[2] += g1 * dvz - vdotr * g2 * dz;
                        }
                    };
                    // an accepted node lies outside both softenings (u >= 1), where the two
                    // factors are 1/r^3 and 3/r^5 (tree.h:63, :90): both from rsqrt_d, a few ulp
                    // from the reference's square root and divisions
                    for (int j = jstart; j < mcount; j += L) {
                        const real4 s = s_mono[j];
                        const double4 v = s_mext[j];
                        const double dx = (double)s.x - (double)tx, dy = (double)s.y - (double)ty,
                                     dz = (double)s.z - (double)tz;
                        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                        if (r2 <= 0) continue;
                        const double inv = rsqrt_d(r2), inv2 = inv * inv;
                        const double g1 = (double)s.w * (inv2 * inv);
                        tj(dx, dy, dz, g1, 3.0 * g1 * inv2, v);
                    }
                    // particles: the same beyond the kernel (r^2 >= h^2), the reference's
                    // spline factors inside it (r from r^2 and 1/r)
                    for (int j = jstart; j < pcount; j += L) {
                        const real4 s = s_pair[j];
                        const double es = (double)s_psoft[j];
                        const double dx = (double)s.x - (double)tx, dy = (double)s.y - (double)ty,
                                     dz = (double)s.z - (double)tz;
                        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                        if (r2 <= 0) continue;
                        const double h = dmax_w(dmax_w((double)te, es), 1e-300);
                        const double inv = rsqrt_d(r2), inv2 = inv * inv;
                        double f1 = inv2 * inv, f2 = 3.0 * f1 * inv2;
                        if (r2 < h * h) {
                            const double r = r2 * inv;
                            f1 = spline_force_over_r_d(r, h);
                            f2 = tidal_factor_d(r, h);
                        }
                        tj(dx, dy, dz, (double)s.w * f1, (double)s.w * f2, s_pext[j]);
                    }
                } else if (FP64) {
                    for (int j = jstart; j < mcount; j += L) {
                        // u >= 1 for every accepted node, so the spline is 1/r^3 (tree.h:63),
                        // here from 1/sqrt(r^2) to a few ulp instead of a square root and a
                        // division (r^2 > 0: an accepted node lies outside every target's box)
                        const real4 s = s_mono[j];
                        const double dx = (double)s.x - (double)tx, dy = (double)s.y - (double)ty,
                                     dz = (double)s.z - (double)tz;
                        // (the library is built without contraction; the walk's sums need
                        // not round as the reference's do, so multiply-adds are fused here)
                        const double inv =
                            rsqrt_d(__builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz)));
                        const double f = (double)s.w * (inv * inv * inv);
                        accx = __builtin_fma(dx, f, accx);
                        accy = __builtin_fma(dy, f, accy);
                        accz = __builtin_fma(dz, f, accz);
                    }
                    for (int j = jstart; j < pcount; j += L) {
                        const real4 s = s_pair[j];
                        const double es = (double)s_psoft[j];
                        const double dx = (double)s.x - (double)tx, dy = (double)s.y - (double)ty,
                                     dz = (double)s.z - (double)tz;
                        const double r2 = __builtin_fma(dx, dx, __builtin_fma(dy, dy, dz * dz));
                        if (r2 <= 0) continue;
                        if (COUNT && tvalid) ++g_pair;
                        const double h = dmax_w(dmax_w((double)te, es), 1e-300);
                        const double f = (double)s.w * spline_force_over_r_r2(r2, h);
                        accx = __builtin_fma(dx, f, accx);
                        accy = __builtin_fma(dy, f, accy);
                        accz = __builtin_fma(dz, f, accz);
                    }
                } else {
                    float fx = 0.f, fy = 0.f, fz = 0.f;
                    const float tef = (float)te, te2 = tef * tef;
                    const float txf = (float)tx, tyf = (float)ty, tzf = (float)tz;
                    // lanes outside the slices repeat slice 0 (their sums are never used)
                    const int js = tvalid ? sub : 0;
                    auto mono = [&](const real4 s) {
                        const float dx = (float)s.x - txf, dy = (float)s.y - tyf,
                                    dz = (float)s.z - tzf;
                        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        const float f = ((float)s.w * inv) * (inv * inv);
                        fx = __builtin_fmaf(dx, f, fx);
                        fy = __builtin_fmaf(dy, f, fy);
                        fz = __builtin_fmaf(dz, f, fz);
                    };
                    // es2: the entry's squared softening (FP32 lists hold squares)
                    auto pair = [&](const real4 s, const float es2, int j) {
                        const float dx = (float)s.x - txf, dy = (float)s.y - tyf,
                                    dz = (float)s.z - tzf;
                        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        const float h2 = __builtin_fmaxf(te2, es2);
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        float g3 = inv * inv * inv;
                        // inside the kernel (u <= 1, where the spline meets 1/r^3), or the target
                        // itself (r2 = 0, skipped as tree.cc:799-802 does)
                        if (r2 <= h2) {
                            const float sp = spline_inside_f(r2, inv, g3, h2);
                            g3 = (r2 > 0.f) ? sp : 0.f; // a select: sp is NaN at r2 = 0
                        }
                        if (COUNT && tvalid && j < pcount && r2 > 0.f) ++g_pair;
                        const float f = (float)s.w * g3;
                        fx = __builtin_fmaf(dx, f, fx);
                        fy = __builtin_fmaf(dy, f, fy);
                        fz = __builtin_fmaf(dz, f, fz);
                    };
                    // uniform kernel extent: h = the target's, 1/h and 1/h^3 once per lane
                    const float hinvt = __builtin_amdgcn_rsqf(te2);
                    const float hinv3t = hinvt * hinvt * hinvt;
                    auto pair_u = [&](const real4 s, int j) {
                        const float dx = (float)s.x - txf, dy = (float)s.y - tyf,
                                    dz = (float)s.z - tzf;
                        const float r2 = __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
                        const float inv = __builtin_amdgcn_rsqf(r2);
                        float g3 = inv * inv * inv;
                        if (r2 <= te2) {
                            const float sp = spline_inside_h(r2, inv, g3, hinvt, hinv3t);
                            g3 = (r2 > 0.f) ? sp : 0.f;
                        }
                        if (COUNT && tvalid && j < pcount && r2 > 0.f) ++g_pair;
                        const float f = (float)s.w * g3;
                        fx = __builtin_fmaf(dx, f, fx);
                        fy = __builtin_fmaf(dy, f, fy);
                        fz = __builtin_fmaf(dz, f, fz);
                    };
                    // two slice steps per iteration, so the LDS reads of two entries overlap;
                    // the padding covers the second step past the end
                    if (uni && L == 1) {
                        for (int j0 = 0; j0 < mcount; j0 += 2) {
                            const real4 s0 = s_mono[j0], s1 = s_mono[j0 + 1];
                            mono(s0);
                            mono(s1);
                        }
                        for (int j0 = 0; j0 < pcount; j0 += 2) {
                            const real4 s0 = s_pair[j0], s1 = s_pair[j0 + 1];
                            pair_u(s0, j0);
                            pair_u(s1, j0 + 1);
                        }
                    } else if (uni) {
                        for (int j0 = 0; j0 < mcount; j0 += 2 * L) {
                            const real4 s0 = s_mono[j0 + js], s1 = s_mono[j0 + L + js];
                            mono(s0);
                            mono(s1);
                        }
                        for (int j0 = 0; j0 < pcount; j0 += 2 * L) {
                            const int j1 = j0 + L;
                            pair_u(s_pair[j0 + js], j0 + js);
                            pair_u(s_pair[j1 + js], j1 + js);
                        }
                    } else if (L == 1) {
                        // one slice (more than 32 targets): every lane reads the same entries,
                        // the second at a fixed offset from the first
                        for (int j0 = 0; j0 < mcount; j0 += 2) {
                            const real4 s0 = s_mono[j0], s1 = s_mono[j0 + 1];
                            mono(s0);
                            mono(s1);
                        }
                        for (int j0 = 0; j0 < pcount; j0 += 2) {
                            const real4 s0 = s_pair[j0], s1 = s_pair[j0 + 1];
                            const float e0 = s_psoft[j0], e1 = s_psoft[j0 + 1];
                            pair(s0, e0, j0);
                            pair(s1, e1, j0 + 1);
                        }
                    } else {
                        for (int j0 = 0; j0 < mcount; j0 += 2 * L) {
                            const real4 s0 = s_mono[j0 + js], s1 = s_mono[j0 + L + js];
                            mono(s0);
                            mono(s1);
                        }
                        for (int j0 = 0; j0 < pcount; j0 += 2 * L) {
                            const int j1 = j0 + L;
                            pair(s_pair[j0 + js], s_psoft[j0 + js], j0 + js);
                            pair(s_pair[j1 + js], s_psoft[j1 + js], j1 + js);
                        }
                    }
                    accx += (double)fx;
                    accy += (double)fy;
                    accz += (double)fz;
                }
                __syncthreads();
                mcount = 0;
                pcount = 0;
            };

            // Node batches: the next batch is popped from the stack while this batch's lists
            // are built, and its nodes are loaded once they are.
            int sp = 0;
            int gsp = 0; // entries below the LDS stack, in the spill slice
            int k = 1;
            int node = (lane == 0) ? groot : -1;
            WNode w;
            if (node >= 0) w = a.wn[node];
            while (k > 0 && !overflow) {
                bool accept = false, open_leaf = false;
                int nch = 0, first = -1, plo = 0, phi = 0;
                if (node >= 0) {
                    const double dx = dmax_w(0.0, dmax_w(gi.lo[0] - w.cx, w.cx - gi.hi[0]));
                    const double dy = dmax_w(0.0, dmax_w(gi.lo[1] - w.cy, w.cy - gi.hi[1]));
                    const double dz = dmax_w(0.0, dmax_w(gi.lo[2] - w.cz, w.cz - gi.hi[2]));
                    const double rmin2 = dx * dx + dy * dy + dz * dz;
                    accept = !open_node_d(rmin2, (double)w.len, w.s, w.mass, (double)w.soft,
                                          gi.emax, gi.amin, gi.theta_sq, dx, dy, dz);
                    first = w.first;
                    if (!accept) {
                        if (first < 0) {
                            open_leaf = true;
                            plo = w.plo;
                            phi = w.phi;
                        } else {
                            nch = w.nchild;
                        }
                    }
                }
                if (COUNT && c0 == 0) g_visit += (lane == 0) ? k : 0;

                // monopole entries (kept in registers until the list has room)
                real4 me;
                if (accept) {
                    if (FP64) {
                        me.x = (real)w.cx; me.y = (real)w.cy; me.z = (real)w.cz;
                    } else {
                        me.x = (real)(w.cx - gcx); me.y = (real)(w.cy - gcy);
                        me.z = (real)(w.cz - gcz);
                    }
                    me.w = (real)w.mass;
                }

                // children of opened interior nodes go on the stack, then the next batch is
                // popped and its loads are issued
                int ctot;
                const int coff = wave_excl_scan(nch, ctot);
                if (sp + ctot > scap) {
                    // keep the top `keep` entries in LDS, move the rest to the spill slice
                    const int keep = min(min(sp, kSpillKeep), scap - ctot);
                    if (!spill || keep < 0 || gsp + (sp - keep) > a.spill_cap) {
                        overflow = true;
                        break;
                    }
                    const int ms = sp - keep;
                    for (int i = lane; i < ms; i += 64) spill[gsp + i] = stk[i];
                    int hold[kSpillKeep / 64];
#pragma unroll
                    for (int r = 0; r < kSpillKeep / 64; ++r) {
                        const int i = lane + 64 * r;
                        hold[r] = (i < keep) ? stk[ms + i] : 0;
                    }
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < kSpillKeep / 64; ++r) {
                        const int i = lane + 64 * r;
                        if (i < keep) stk[i] = hold[r];
                    }
                    gsp += ms;
                    sp = keep;
                }
                for (int c = 0; c < nch; ++c) stk[sp + coff + c] = first + c;
                sp += ctot;
                if (COUNT) max_sp = max(max_sp, gsp + sp);
                if (GSTACK) __threadfence_block();
                __syncthreads();
                if (sp < 64 && gsp > 0) {
                    // fewer than a batch left in LDS: slide them up and bring back the top of
                    // the spilled part beneath them, so the batch popped next is the same top
                    // 64 entries an unbounded stack would give
                    const int r = min(gsp, min(scap / 2, scap - sp));
                    const int held = (lane < sp) ? stk[lane] : 0;
                    __threadfence_block();
                    __syncthreads();
                    if (lane < sp) stk[r + lane] = held;
                    for (int i = lane; i < r; i += 64) stk[i] = spill[gsp - r + i];
                    gsp -= r;
                    sp += r;
                    __syncthreads();
                }
                k = min(sp, 64);
                const int base = sp - k;
                const int next_node = (lane < k) ? stk[base + lane] : -1;
                sp = base;

                const uint64_t mb = __ballot(accept);
                const int nm = __popcll(mb);
                if (nm > 0) {
                    if (mcount + nm > kMono) flush();
                    if (accept) {
                        const int slot_m = mcount + lanes_below(mb);
                        s_mono[slot_m] = me;
                        if (TJ) {
                            double4 ext = make_double4(0.0, 0.0, 0.0, (double)w.soft);
                            if (a.jerk) {
                                const size_t nn = (size_t)a.nnodes;
                                ext.x = a.nvel[node];
                                ext.y = a.nvel[nn + node];
                                ext.z = a.nvel[2 * nn + node];
                            }
                            s_mext[slot_m] = ext;
                        }
                    }
                    if (COUNT && tvalid && sub == 0) g_multi += nm;
                    mcount += nm;
                }

                // particles of opened leaves become pair entries
                const uint64_t lb = __ballot(open_leaf);
                if (lb) {
                    const int nleaf = __popcll(lb);
                    int ptot;
                    const int sz = open_leaf ? (phi - plo) : 0;
                    const int poff = wave_excl_scan(sz, ptot);
                    if (open_leaf) {
                        const int li = lanes_below(lb);
                        s_lpre[li] = poff;
                        s_lplo[li] = plo;
                    }
                    if (lane == 0) s_lpre[nleaf] = ptot;
                    __syncthreads();
                    if (!FP64) {
                        // one entry per lane: near particles to the particle list, far ones to
                        // the monopole list (both counted as pair interactions)
                        for (int done = 0; done < ptot; done += 64) {
                            const int idx = done + lane;
                            const bool have = idx < ptot;
                            real4 e = {(real)0, (real)0, (real)0, (real)0};
                            float es = 0.f;
                            bool far = false;
                            if (have) {
                                int lo_ = 0, hi_ = nleaf; // last li with s_lpre[li] <= idx
                                while (hi_ - lo_ > 1) {
                                    const int mid = (lo_ + hi_) >> 1;
                                    if (s_lpre[mid] <= idx) lo_ = mid;
                                    else hi_ = mid;
                                }
                                const int q = s_lplo[lo_] + (idx - s_lpre[lo_]);
                                const double4 sp4 = a.src[q];
                                e.x = (real)(sp4.x - gcx); e.y = (real)(sp4.y - gcy);
                                e.z = (real)(sp4.z - gcz); e.w = (real)sp4.w;
                                es = (float)a.src_soft[q];
                                const float ox = fabsf((float)e.x) - bhx;
                                const float oy = fabsf((float)e.y) - bhy;
                                const float oz = fabsf((float)e.z) - bhz;
                                const float qx = (ox > 0.f) ? ox : 0.f, qy = (oy > 0.f) ? oy : 0.f,
                                            qz = (oz > 0.f) ? oz : 0.f;
                                const float H = (emaxf < es) ? es : emaxf;
                                far = qx * qx + qy * qy + qz * qz > H * H;
                            }
                            const uint64_t fb = __ballot(have && far);
                            const uint64_t nb = __ballot(have && !far);
                            const int nf = __popcll(fb), nn = __popcll(nb);
                            if (mcount + nf > kMono || pcount + nn > kPair) flush();
                            if (have) {
                                if (far) {
                                    s_mono[mcount + lanes_below(fb)] = e;
                                } else {
                                    const int k = pcount + lanes_below(nb);
                                    s_pair[k] = e;
                                    s_psoft[k] = (real)(es * es);
                                }
                            }
                            if (COUNT && tvalid && sub == 0) g_pair += nf;
                            mcount += nf;
                            pcount += nn;
                        }
                    } else {
                        for (int done = 0; done < ptot;) {
                            if (pcount == kPair) flush();
                            const int ncopy = min(kPair - pcount, ptot - done);
                            for (int j = lane; j < ncopy; j += 64) {
                                const int idx = done + j;
                                int lo_ = 0, hi_ = nleaf; // last li with s_lpre[li] <= idx
                                while (hi_ - lo_ > 1) {
                                    const int mid = (lo_ + hi_) >> 1;
                                    if (s_lpre[mid] <= idx) lo_ = mid;
                                    else hi_ = mid;
                                }
                                const int q = s_lplo[lo_] + (idx - s_lpre[lo_]);
                                const double4 sp4 = a.src[q];
                                real4 e; // FP64 entries keep global coordinates
                                e.x = (real)sp4.x; e.y = (real)sp4.y; e.z = (real)sp4.z;
                                e.w = (real)sp4.w;
                                s_pair[pcount + j] = e;
                                s_psoft[pcount + j] = (real)a.src_soft[q];
                                if (TJ) {
                                    double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
                                    if (a.jerk) {
                                        const uint32_t pq = a.orderbuf[q];
                                        v.x = a.pvx[pq]; v.y = a.pvy[pq]; v.z = a.pvz[pq];
                                    }
                                    s_pext[pcount + j] = v;
                                }
                            }
                            pcount += ncopy;
                            done += ncopy;
                        }
                    }
                }
                __syncthreads();
                // the next batch's nodes are loaded here, after this batch's lists are built:
                // loading them earlier keeps 16 more registers live through the list building
                // and the flushes (measured 2% slower), and other waves hide the latency
                node = next_node;
                if (node >= 0) w = a.wn[node];
            }
            if (overflow) {
                if (!GSTACK && lane == 0) {
                    const int slot = atomicAdd(a.n_retry, 1);
                    a.retry_list[slot] = g;
                }
                if (GSTACK && lane == 0) atomicOr(a.status, 1);
                g_visit = g_multi = g_pair = 0; // the retry counts this group
                __syncthreads();
                break; // the whole group is redone by the global-stack launch
            }
            if (mcount > 0 || pcount > 0) flush();

            // combine the slices of each target pairwise (log2 L steps), then write (tree.cc:880)
            for (int step = 1; step < L; step <<= 1) {
                const int src = min(lane + step * nt, 63);
                const bool take = (sub % (2 * step) == 0) && (sub + step < L);
                const double ux = __shfl(accx, src), uy = __shfl(accy, src),
                             uz = __shfl(accz, src);
                if (take) {
                    accx += ux;
                    accy += uy;
                    accz += uz;
                }
                if (TJ) {
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        const double u = __shfl(tt[c], src);
                        if (take) tt[c] += u;
                    }
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const double u = __shfl(jj[c], src);
                        if (take) jj[c] += u;
                    }
                }
            }
            if (TJ && tvalid && sub == 0) {
                const size_t ntt = (size_t)a.nt, slot = (size_t)(tlo + t);
#pragma unroll
                for (int c = 0; c < 6; ++c) a.tidal[c * ntt + slot] = tt[c];
                if (a.jerk) {
#pragma unroll
                    for (int c = 0; c < 3; ++c) a.jerk[c * ntt + slot] = a.G * jj[c];
                }
            }
            if (tvalid && sub == 0) {
                a.ax[tlo + t] = a.G * accx;
                a.ay[tlo + t] = a.G * accy;
                a.az[tlo + t] = a.G * accz;
            }
        }
        if (COUNT) {
            n_visit += g_visit;
            n_multi += g_multi;
            n_pair += g_pair;
        }
    }
    if (COUNT) {
        // wave totals, one atomic per wave
        for (int d = 32; d >= 1; d >>= 1) {
            n_visit += __shfl_xor(n_visit, d);
            n_multi += __shfl_xor(n_multi, d);
            n_pair += __shfl_xor(n_pair, d);
        }
        if (lane == 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[0]),
                      (unsigned long long)n_visit);
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[1]),
                      (unsigned long long)n_multi);
            atomicAdd(reinterpret_cast<unsigned long long*>(&a.counts[2]),
                      (unsigned long long)n_pair);
            atomicMax(reinterpret_cast<unsigned long long*>(&a.counts[3]),
                      (unsigned long long)max_sp);
        }
    }
}

// Nearest periodic image of a separation in a cubic box of side L (grav_cuda.cu:104-109): one
// wrap either way, for separations within (-1.5 L, 1.5 L).
__device__ __forceinline__ double min_image(double d, double L) {
    const double hb = 0.5 * L;
    if (d > hb) return d - L;
    if (d < -hb) return d + L;
    return d;
}

// ---- direct sum (tree.cc:946-971; the reference's device form is grav_cuda.cu:65-141) -----------
// One target per thread, sources staged through LDS 256 at a time in index order, so each target
// sums its sources in exactly the reference's order with the reference's FP64 expressions.
constexpr int kDsThreads = 256;

__global__ __launch_bounds__(kDsThreads) void k_direct_fp64(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, const double* __restrict__ soft, size_t n,
    const uint32_t* __restrict__ targets, size_t nt, double G, double box,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double sx[kDsThreads], sy[kDsThreads], sz[kDsThreads], sm[kDsThreads],
        ss[kDsThreads];
    const size_t t = (size_t)blockIdx.x * kDsThreads + threadIdx.x;
    const bool live = t < nt;
    uint32_t p = 0;
    double px = 0, py = 0, pz = 0, ep = 0;
    bool bad = false;
    if (live) {
        p = targets[t];
        if (p >= n) { // an index outside the particle set gets NaN, never a stray read
            bad = true;
            p = 0;
        }
        px = x[p]; py = y[p]; pz = z[p];
        ep = soft ? soft[p] : 0.0;
    }
    double ax_ = 0, ay_ = 0, az_ = 0;
    for (size_t base = 0; base < n; base += kDsThreads) {
        const size_t k = base + threadIdx.x;
        if (k < n) {
            sx[threadIdx.x] = x[k]; sy[threadIdx.x] = y[k]; sz[threadIdx.x] = z[k];
            sm[threadIdx.x] = m[k]; ss[threadIdx.x] = soft ? soft[k] : 0.0;
        }
        __syncthreads();
        const int ntile = (int)min((size_t)kDsThreads, n - base);
        if (live) {
            for (int j = 0; j < ntile; ++j) {
                if (base + (size_t)j == (size_t)p) continue;
                double dx = sx[j] - px, dy = sy[j] - py, dz = sz[j] - pz;
                if (box > 0) {
                    dx = min_image(dx, box);
                    dy = min_image(dy, box);
                    dz = min_image(dz, box);
                }
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r <= 0) continue;
                const double fac =
                    sm[j] * spline_force_over_r_d(r, dmax_w(dmax_w(ep, ss[j]), 1e-300));
                ax_ += dx * fac;
                ay_ += dy * fac;
                az_ += dz * fac;
            }
        }
        __syncthreads();
    }
    if (live) {
        const double nan = __builtin_nan("");
        ax[t] = bad ? nan : G * ax_;
        ay[t] = bad ? nan : G * ay_;
        az[t] = bad ? nan : G * az_;
    }
}

// The float-narrowed direct sum, the reference device form's default arithmetic
// (grav_cuda.cu:65-141 with R = float): each separation is differenced in FP64 and then narrowed,
// and r, the spline (at the larger of the two softenings, 1e-30 when both are zero), the mass
// factor and the per-target sums are FP32. Sources are summed in index order, self and r^2 <= 0
// skipped, the result widened and multiplied by G (the reference leaves G to its caller,
// grav_cuda.cu:186-192).
__device__ __forceinline__ float spline_force_over_r_f(float r, float h) {
    const float h_inv = 1.0f / h;
    const float h_inv3 = h_inv * h_inv * h_inv;
    const float u = r * h_inv;
    if (u >= 1.0f) return 1.0f / (r * r * r);
    if (u < 0.5f) return h_inv3 * (10.666666666666667f + u * u * (32.0f * u - 38.4f));
    return h_inv3 * (21.333333333333332f - 48.0f * u + 38.4f * u * u -
                     10.666666666666667f * u * u * u - 0.06666666666666667f / (u * u * u));
}

__global__ __launch_bounds__(kDsThreads) void k_direct_f32(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const double* __restrict__ m, const double* __restrict__ soft, size_t n,
    const uint32_t* __restrict__ targets, size_t nt, double G, double box,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az) {
    __shared__ double sx[kDsThreads], sy[kDsThreads], sz[kDsThreads];
    __shared__ float sm[kDsThreads], ss[kDsThreads];
    const size_t t = (size_t)blockIdx.x * kDsThreads + threadIdx.x;
    const bool live = t < nt;
    uint32_t p = 0;
    double px = 0, py = 0, pz = 0;
    float ep = 1.0f;
    bool bad = false;
    if (live) {
        p = targets[t];
        if (p >= n) {
            bad = true;
            p = 0;
        }
        px = x[p]; py = y[p]; pz = z[p];
        ep = soft ? (float)soft[p] : 0.0f;
    }
    float ax_ = 0.0f, ay_ = 0.0f, az_ = 0.0f;
    for (size_t base = 0; base < n; base += kDsThreads) {
        const size_t k = base + threadIdx.x;
        if (k < n) {
            sx[threadIdx.x] = x[k]; sy[threadIdx.x] = y[k]; sz[threadIdx.x] = z[k];
            sm[threadIdx.x] = (float)m[k];
            ss[threadIdx.x] = soft ? (float)soft[k] : 0.0f;
        }
        __syncthreads();
        const int ntile = (int)min((size_t)kDsThreads, n - base);
        if (live) {
            for (int j = 0; j < ntile; ++j) {
                if (base + (size_t)j == (size_t)p) continue;
                double dxd = sx[j] - px, dyd = sy[j] - py, dzd = sz[j] - pz;
                if (box > 0) {
                    dxd = min_image(dxd, box);
                    dyd = min_image(dyd, box);
                    dzd = min_image(dzd, box);
                }
                const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
                const float r2 = dx * dx + dy * dy + dz * dz;
                if (r2 <= 0.0f) continue;
                const float r = __builtin_sqrtf(r2);
                const float es = ss[j];
                float e = ep > es ? ep : es;
                if (!(e > 0.0f)) e = 1e-30f;
                const float fac = sm[j] * spline_force_over_r_f(r, e);
                ax_ += fac * dx;
                ay_ += fac * dy;
                az_ += fac * dz;
            }
        }
        __syncthreads();
    }
    if (live) {
        const double nan = __builtin_nan("");
        ax[t] = bad ? nan : G * (double)ax_;
        ay[t] = bad ? nan : G * (double)ay_;
        az[t] = bad ? nan : G * (double)az_;
    }
}

// ---- host launchers ----------------------------------------------------------------------------
// Wavefronts of a persistent walk launch: 20 per CU (five per SIMD, the occupancy the walk's
// registers and LDS allow).
static gz_status persistent_waves(int* waves) {
    int dev = 0, ncu = 0;
    GZ_HIP(hipGetDevice(&dev));
    GZ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    *waves = std::max(1, ncu) * 20;
    return GZ_OK;
}

// A group's logical stack never exceeds 448 entries per tree level plus one batch of children
// (each batch pops 64 nodes and pushes at most 512), under 10k for kMaxLevel + 1 levels.
constexpr int kSpillCap = 16384;

gz_status ctx_spill(gz_ctx* ctx, int** spill, int* cap) {
    if (!ctx->spill) {
        int waves = 0;
        GZ_TRY(persistent_waves(&waves));
        void* p = nullptr;
        GZ_HIP(hipMalloc(&p, (size_t)waves * kSpillCap * sizeof(int)));
        ctx->spill = static_cast<int*>(p);
        ctx->spill_cap = kSpillCap;
    }
    *spill = ctx->spill;
    *cap = ctx->spill_cap;
    return GZ_OK;
}

gz_status launch_walk(const WalkArgs& a, bool fp64, bool count, hipStream_t st) {
    if (a.ngroups <= 0) return GZ_OK;
    const bool tj = a.tidal != nullptr;
    unsigned grid;
    if (a.persistent) {
        int waves = 0;
        GZ_TRY(persistent_waves(&waves));
        if (a.waves_per_cu > 0 && a.waves_per_cu < 20) waves = waves / 20 * a.waves_per_cu;
        grid = (unsigned)std::min<long long>(std::max(waves, 1), a.ngroups);
    } else {
        grid = (unsigned)a.ngroups;
    }
    if (tj) {
        k_walk<true, false, false, true><<<grid, 64, 0, st>>>(a);
    } else if (fp64) {
        if (count) k_walk<true, true, false><<<grid, 64, 0, st>>>(a);
        else k_walk<true, false, false><<<grid, 64, 0, st>>>(a);
    } else {
        if (count) k_walk<false, true, false><<<grid, 64, 0, st>>>(a);
        else k_walk<false, false, false><<<grid, 64, 0, st>>>(a);
    }
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

// Groups whose node stack outgrew LDS, again, with the stack in global memory: `waves`
// wavefronts (each with gstack_cap entries of a.gstack) take the groups listed in a.retry_list
// until *a.n_retry is exhausted, so the launch needs no host-side count.
gz_status launch_walk_retry(const WalkArgs& a, int waves, bool fp64, bool count,
                            hipStream_t st) {
    if (waves <= 0) return GZ_OK;
    const int n_retry = waves;
    if (a.tidal) {
        k_walk<true, false, true, true><<<n_retry, 64, 0, st>>>(a);
    } else if (fp64) {
        if (count) k_walk<true, true, true><<<n_retry, 64, 0, st>>>(a);
        else k_walk<true, false, true><<<n_retry, 64, 0, st>>>(a);
    } else {
        if (count) k_walk<false, true, true><<<n_retry, 64, 0, st>>>(a);
        else k_walk<false, false, true><<<n_retry, 64, 0, st>>>(a);
    }
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_group_prep(const uint32_t* targets, size_t nt, const uint32_t* rank,
                            const uint32_t* orderbuf, size_t n_live,
                            const int* slot_leaf, int* tslot, const int* parent, const WNode* wn,
                            int nnodes, int* leaf_group, int batch, int group_size,
                            size_t n, const uint32_t* seg, int nseg, int* gseg, int* gkey,
                            uint32_t* flag, uint32_t* gid, int* gstart, uint32_t* ngroups_dev,
                            int* status, void* scan_tmp, hipStream_t st) {
    const unsigned grid = (unsigned)((nt + 255) / 256);
    // every leaf's group node once, when targets outnumber nodes (else each target climbs)
    const bool per_leaf = batch <= 0 && group_size > 0 && leaf_group && nnodes > 0 &&
                          (size_t)nnodes < nt;
    if (per_leaf)
        k_leaf_group<<<(unsigned)((nnodes + 255) / 256), 256, 0, st>>>(nnodes, wn, parent,
                                                                       group_size, leaf_group);
    k_group_flags<<<grid, 256, 0, st>>>(targets, nt, rank, orderbuf, n_live, slot_leaf, tslot,
                                        parent, wn, per_leaf ? leaf_group : nullptr, batch,
                                        group_size, n,
                                        seg, nseg, gseg, gkey, status);
    k_group_flags2<<<grid, 256, 0, st>>>(nt, batch, gkey, nseg > 1 ? gseg : nullptr, flag);
    GZ_HIP(hipGetLastError());
    // group ids, and each group's first target slot (gstart[gid[i]] = i where a group starts)
    GZ_TRY(exclusive_scan_u32(flag, gid, nt, ngroups_dev, scan_tmp, st, gstart));
    return GZ_OK;
}

gz_status launch_group_pack(int ngroups, size_t nt, int group_size, const int* gstart,
                            const int* gkey, const int* parent, const int* gseg, long long* pk,
                            uint32_t* mflag, uint32_t* mid, uint32_t* ngroups2_dev, int* gstart2,
                            void* scan_tmp, hipStream_t st) {
    if (ngroups <= 0) return GZ_OK;
    const unsigned grid = (unsigned)((ngroups + 255) / 256);
    k_pack_parent<<<grid, 256, 0, st>>>(ngroups, nt, gstart, gkey, parent, gseg, pk);
    k_pack_greedy<<<grid, 256, 0, st>>>(ngroups, nt, group_size, gstart, pk, mflag);
    GZ_HIP(hipGetLastError());
    GZ_TRY(exclusive_scan_u32(mflag, mid, (size_t)ngroups, ngroups2_dev, scan_tmp, st));
    k_pack_starts<<<grid, 256, 0, st>>>(ngroups, gstart, mflag, mid, gstart2);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_direct(const gz_particles* p, const uint32_t* targets, size_t nt, double G,
                        int fp64, double box, double* ax, double* ay, double* az,
                        hipStream_t st) {
    if (nt == 0) return GZ_OK;
    const unsigned grid = (unsigned)((nt + kDsThreads - 1) / kDsThreads);
    if (fp64)
        k_direct_fp64<<<grid, kDsThreads, 0, st>>>(p->x, p->y, p->z, p->m, p->soft, p->n,
                                                   targets, nt, G, box, ax, ay, az);
    else
        k_direct_f32<<<grid, kDsThreads, 0, st>>>(p->x, p->y, p->z, p->m, p->soft, p->n,
                                                  targets, nt, G, box, ax, ay, az);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // namespace gz
