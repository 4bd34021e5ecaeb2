// Per-target tree walks on the device: shmem::accel (shmem/tree.cc:528-582) and shmem::potential
// (tree.cc:899-944). Every target meets the tree on its own: a node is accepted or opened by the
// reference's open_node (tree.h:109-144) with the target's own position, softening and aold, an
// accepted node adds its monopole with the spline of the larger softening, and every reached leaf
// (leaves are always opened in these two walks) is summed particle by particle, the target itself
// skipped.
//
// How it runs on CDNA: one wavefront takes 64 targets that are neighbours in Morton order (fewer,
// down to one, when there are too few targets to occupy the GPU that way), one per lane, and walks the union of their 64 traversals depth first. The wave visits one node at a
// time; the node's index is wave-uniform, so its 64-byte record comes through the scalar cache
// and every lane still waiting on it tests it against its own target. Lanes that accept it add
// its monopole; the lanes that open it are the mask under which its children are visited. Each
// lane therefore meets exactly the nodes of its own reference traversal and adds the same terms
// (in the same depth-first order); lanes not waiting on a node sit it out. The stack holds
// sibling ranges (next child, children left, lane mask), one per tree level, in lane-indexed
// registers (entry e in lane e, read back with v_readlane): no LDS, and no barrier anywhere. A
// leaf's particles are read once per wave, again through the scalar cache.
//
// One lane per target replaces a wave per target: a target's walk is a chain of dependent node
// visits, and a wave that walks a single target has too little parallel work per visit to hide
// that chain (most of its 64 lanes idle while the first levels are opened).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gz_features.h"
#include "gz_grav_math.h"
#include "gz_internal.h"

namespace gz {

static_assert(kPtDepth >= kMaxLevel + 2, "the range stack must hold one entry per tree level");

// Sort keys of the targets: their sorted slot (n_live for a target outside the tree, so those go
// last), with the list position as the value; tslot as the grouped walk computes it. status bit
// 1: a target index outside the particle set; misc[1] != 0: the list is not in slot order.
__global__ __launch_bounds__(256) void k_pt_keys(const uint32_t* __restrict__ targets, size_t nt,
                                                 const uint32_t* __restrict__ rank,
                                                 const uint32_t* __restrict__ orderbuf,
                                                 size_t n_live, size_t n, int* __restrict__ tslot,
                                                 uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals,
                                                 int* __restrict__ misc) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nt) return;
    auto slot_of = [&](size_t k) -> uint64_t {
        const uint32_t p = targets[k];
        if (p >= n) return (uint64_t)n_live;
        const uint32_t r = (k < n_live && orderbuf[k] == p) ? (uint32_t)k : rank[p];
        return (r == 0xFFFFFFFFu) ? (uint64_t)n_live : (uint64_t)r;
    };
    if (targets[i] >= n) atomicOr(misc, 2);
    const uint64_t key = slot_of(i);
    tslot[i] = (key == (uint64_t)n_live) ? -1 : (int)key;
    keys[i] = key;
    vals[i] = (uint32_t)i;
    if (i > 0 && slot_of(i - 1) > key) atomicOr(misc + 1, 1);
}

// POT: the potential (else the acceleration); REL: per-target aold are given (else the relative
// arm reduces to theta_sq <= 0 for every target, as it does for aold = 0).
template <bool POT, bool REL>
__global__ __launch_bounds__(64) void k_walk_pt(PtArgs a) {
    const int lane = threadIdx.x;
    // Block b runs on XCD b % 8: each XCD gets one contiguous share of the Morton-ordered waves,
    // so its L2 serves one compact region of the tree.
    const int b = blockIdx.x, nb = gridDim.x;
    const int xcd = b & 7, per = nb >> 3, rem = nb & 7;
    const int wv = xcd * per + min(xcd, rem) + (b >> 3);
    const int k = wv * a.tpw + lane;
    const bool valid = lane < a.tpw && k < a.nt;
    int i = 0, my = -1, croot = 0;
    double px = 0, py = 0, pz = 0, pe = 0, ao = 0;
    if (valid) {
        i = a.order ? (int)a.order[k] : k;
        const uint32_t part = a.targets[i];
        my = a.tslot[i];
        if (my >= 0) {
            const double4 q = a.src[my];
            px = q.x; py = q.y; pz = q.z;
            pe = a.src_soft[my];
        } else {
            px = a.x[part]; py = a.y[part]; pz = a.z[part];
            pe = a.soft ? a.soft[part] : 0.0;
        }
        if (REL) ao = a.aold[i];
        croot = root_of(part, a.cluster_start, a.nclusters, a.root);
    }
    const double theta_sq = a.theta_sq;
    const double pe1 = dmax_w(pe, 1e-300); // a pair's kernel extent is max(this, the source's)
    double sx = 0, sy = 0, sz = 0;

    // One pair term for this lane's target, added when `use` (else zero times a finite value is
    // added, so every lane runs the same instructions; only pairs inside the kernel branch off).
    auto pair = [&](const double4 q, double es, bool use) {
        const double ex = q.x - px, ey = q.y - py, ez = q.z - pz;
        const double r2 = __builtin_fma(ex, ex, __builtin_fma(ey, ey, ez * ez));
        const bool ok = use & (r2 > 0);
        const double r2s = ok ? r2 : 1.0;
        const double h = dmax_w(pe1, es);
        const double inv = rsqrt_d(r2s);
        const double m = ok ? q.w : 0.0;
        if (POT) {
            double v = -inv;
            if (ok & (r2s < h * h)) v = pair_potential_inside(r2s, inv, h);
            sx = __builtin_fma(m, v, sx);
        } else {
            double f = inv * inv * inv;
            if (ok & (r2s < h * h)) f = pair_force_over_r_inside(r2s, inv, h);
            f *= m;
            sx = __builtin_fma(ex, f, sx);
            sy = __builtin_fma(ey, f, sy);
            sz = __builtin_fma(ez, f, sz);
        }
    };

    uint64_t todo = __ballot(valid);
    while (todo != 0) {
        // the lanes whose targets lie under one root (a forest has a root per cluster)
        const int r = __builtin_amdgcn_readlane(croot, (int)__builtin_ctzll(todo));
        uint64_t cm = __ballot(valid && croot == r);
        todo &= ~cm;
        // the current sibling range (next node, nodes left, lanes waiting on it) and the stack of
        // the ranges above it: entry e in lane e of stf/stc/stl/sth. A range is pushed only when
        // a node is opened, so the stack holds at most one range per tree level (the host checks
        // the tree's depth against kPtDepth).
        int cf = r, cc = 1, sp = 0;
        int stf = 0, stc = 0, stl = 0, sth = 0;
        for (;;) {
            if (cc == 0) {
                if (sp == 0) break;
                --sp;
                cf = __builtin_amdgcn_readlane(stf, sp);
                cc = __builtin_amdgcn_readlane(stc, sp);
                cm = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(sth, sp) << 32) |
                     (uint64_t)(uint32_t)__builtin_amdgcn_readlane(stl, sp);
            }
            const int node = cf;
            ++cf;
            --cc;
            const WNode w = a.wn[node];
            const bool act = (cm >> lane) & 1ull;
            uint64_t om = 0; // the lanes that open an interior node
            if (w.first < 0) {
                // tree.cc:558-570 / :921-931: a reached leaf is summed particle by particle by
                // every waiting lane, the target itself skipped; two particles per step, their
                // records loaded together
                const int phi = w.phi;
                const double4* qp = a.src + w.plo;
                const double* ep = a.src_soft + w.plo;
                int s = w.plo;
                for (; s + 1 < phi; s += 2, qp += 2, ep += 2) {
                    const double4 q0 = qp[0], q1 = qp[1];
                    const double e0 = ep[0], e1 = ep[1];
                    pair(q0, e0, act & (s != my));
                    pair(q1, e1, act & (s + 1 != my));
                }
                if (s < phi) pair(qp[0], ep[0], act & (s != my));
            } else {
                // tree.cc:552-557 (accel) and :915-920 (potential): the offset to the node's
                // centre of mass, its squared norm, and open_node with the target's own bounds
                const double dx = w.cx - px, dy = w.cy - py, dz = w.cz - pz;
                const double r_sq = dx * dx + dy * dy + dz * dz;
                const bool open = act & open_node_d(r_sq, (double)w.len, w.s, w.mass,
                                                    (double)w.soft, pe, REL ? ao : 0.0,
                                                    theta_sq, fabs(dx), fabs(dy), fabs(dz));
                // an accepted node lies outside both softenings (open_node's first arm), where
                // the spline is exactly Newtonian: -m/r, or m/r^3 times the offset
                const bool acc = act & !open & (r_sq > 0);
                const double inv = rsqrt_d(r_sq);
                if (POT) {
                    sx = __builtin_fma(-w.mass, acc ? inv : 0.0, sx);
                } else {
                    const double f = acc ? w.mass * (inv * inv * inv) : 0.0;
                    sx = __builtin_fma(dx, f, sx);
                    sy = __builtin_fma(dy, f, sy);
                    sz = __builtin_fma(dz, f, sz);
                }
                om = __ballot(open);
            }
            // descend into an opened node: its remaining siblings (if any) wait on the stack
            // below its children. The stack registers change in this one place of the loop.
            const bool down = om != 0;
            const bool mine = down & (lane == sp);
            stf = mine ? cf : stf;
            stc = mine ? cc : stc;
            stl = mine ? (int)(uint32_t)cm : stl;
            sth = mine ? (int)(uint32_t)(cm >> 32) : sth;
            sp += (down && cc > 0) ? 1 : 0;
            if (down) {
                cf = w.first;
                cc = w.nchild;
                cm = om;
            }
        }
    }
    if (valid) {
        a.ax[i] = a.G * sx;
        if (!POT) {
            a.ay[i] = a.G * sy;
            a.az[i] = a.G * sz;
        }
    }
}

gz_status launch_pt_keys(const uint32_t* targets, size_t nt, const gz_tree* t, size_t n,
                         int* tslot, uint64_t* keys, uint32_t* vals, int* misc, hipStream_t st) {
    if (nt == 0) return GZ_OK;
    k_pt_keys<<<(unsigned)((nt + 255) / 256), 256, 0, st>>>(targets, nt, t->rank, t->orderbuf,
                                                            t->n_live, n, tslot, keys, vals, misc);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_walk_pt(const PtArgs& a, bool potential, hipStream_t st) {
    if (a.nt <= 0) return GZ_OK;
    const unsigned grid = (unsigned)((a.nt + a.tpw - 1) / a.tpw);
    const bool rel = a.aold != nullptr;
    if (potential && rel) k_walk_pt<true, true><<<grid, 64, 0, st>>>(a);
    else if (potential) k_walk_pt<true, false><<<grid, 64, 0, st>>>(a);
    else if (rel) k_walk_pt<false, true><<<grid, 64, 0, st>>>(a);
    else k_walk_pt<false, false><<<grid, 64, 0, st>>>(a);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // namespace gz
