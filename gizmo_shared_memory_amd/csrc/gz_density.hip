// The density / smoothing-length solve of the shared-memory hydro (shmem::density, hydro.h:115-119,
// hydro.cc:150-327), non-periodic and without the neighbour cache or lazy drift.
//
// The reference iterates each target privately: search the tree within h, sum the kernel over the
// gas neighbours, and take a bracket-guarded Newton step on N_eff(h) = ball_vol(h) sum_j W until
// N_eff is within the (widening) tolerance of the neighbour target. Here every still-unconverged
// target advances one step per round, and a round is one kernel (k_density_round): each target
// has a lane, a wavefront holds up to 64 targets that are neighbours in Morton order, and the wave
// walks the union of their own searches depth first, one node at a time (the nodes' search boxes
// packed into 32-byte records per call). A lane waits on a node only while the reference's node
// test (hydro.cc:55-72, on the float search nodes) has admitted its sphere at every ancestor; at
// a reached leaf, each waiting lane tests every particle against its sphere (hydro.cc:80-87) and
// notes a neighbour in its column of an LDS buffer; a lane adds its noted neighbours' kernel
// terms whenever some column fills, and at the end. Depth first with children in octant order is
// the tree's slot order, which is the order of the reference's neighbour list, so each lane adds
// the reference's terms in the reference's order without a stored list. The lane then applies
// the reference's update, and the targets that are done leave the active list (a stable
// compaction, so the waves keep their Morton locality).
//
// Two cases leave this path for k_density_far, where a wavefront takes a single target, tests a
// node's children together and a leaf's particles together, and adds the terms in the same order.
// A target whose radius exceeds kFarLeaf times its leaf's side would make its wave of 64 walk a
// large part of the tree one node and one particle at a time: it takes a wave of its own after the
// round. And once kLoopTargets or fewer targets remain, one launch gives each a wave that iterates
// it to the end, instead of a round (and a host round trip) per step.
//
// The one difference from the reference's arithmetic is pow(dx, 2) in the 3D face-closure test,
// taken here as dx * dx (correctly rounded; glibc's pow is within half an ulp of it, not always
// equal).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "gz_internal.h"

namespace gz {
namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kMaxIter = 100; // hydro.cc:201
constexpr int kWavesWanted = 2048; // a round's targets are spread thinner below this many waves
constexpr int kNbBuf = 32;          // neighbours each lane holds before their terms are added
constexpr double kFarLeaf = 4.0;    // radius over leaf side beyond which a target takes the lists
constexpr uint32_t kLoopTargets = 131072; // active targets few enough to iterate in one launch

// kernel_norm, ball_vol, kernel_w, kernel_dwdh (hydro.h:29-65): cubic spline of support h in dim
// dimensions, operand for operand.
__device__ __forceinline__ double knorm(double h, int dim) {
    if (dim == 1) return (4.0 / 3.0) / h;
    if (dim == 2) return 40.0 / (7.0 * kPi * h * h);
    return 8.0 / (kPi * h * h * h);
}
__device__ __forceinline__ double ball_vol(double h, int dim) {
    if (dim == 1) return 2.0 * h;
    if (dim == 2) return kPi * h * h;
    return 4.0 * kPi / 3.0 * h * h * h;
}
__device__ __forceinline__ double kernel_w(double r, double h, int dim) {
    const double q = r / h;
    if (q >= 1.0) return 0.0;
    const double norm = knorm(h, dim);
    if (q < 0.5) return norm * (1.0 - 6.0 * q * q + 6.0 * q * q * q);
    const double u = 1.0 - q;
    return norm * 2.0 * u * u * u;
}
__device__ __forceinline__ double kernel_dwdh(double r, double h, int dim) {
    const double q = r / h;
    if (q >= 1.0) return 0.0;
    const double norm = knorm(h, dim);
    double f, fp;
    if (q < 0.5) {
        f = 1.0 - 6.0 * q * q + 6.0 * q * q * q;
        fp = -12.0 * q + 18.0 * q * q;
    } else {
        const double u = 1.0 - q;
        f = 2.0 * u * u * u;
        fp = -6.0 * u * u;
    }
    return -((double)dim / h) * norm * f - (q / h) * norm * fp;
}

// invert_moments (shmem/vec.h:43-63) of the symmetric moments e (xx yy zz xy yz xz) into b
// (row-major 3x3, zero outside the live block); false when degenerate. 3D is Mat3::invert
// (math_types/mat3.h:41-67): cofactors times 1 / det, singular when det is 0 or NaN.
__device__ __forceinline__ bool invert_moments(const double* e, double* b, int dim) {
#pragma unroll
    for (int k = 0; k < 9; ++k) b[k] = 0.0;
    if (dim == 1) {
        if (fabs(e[0]) < 1e-300) return false;
        b[0] = 1.0 / e[0];
        return true;
    }
    if (dim == 2) {
        const double det = e[0] * e[1] - e[3] * e[3];
        if (fabs(det) < 1e-300) return false;
        b[0] = e[1] / det;
        b[1] = -e[3] / det;
        b[3] = -e[3] / det;
        b[4] = e[0] / det;
        return true;
    }
    const double m00 = e[0], m01 = e[3], m02 = e[5], m10 = e[3], m11 = e[1], m12 = e[4],
                 m20 = e[5], m21 = e[4], m22 = e[2];
    const double d = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) +
                     m02 * (m10 * m21 - m11 * m20);
    if (d == 0.0 || d != d) return false;
    const double id = 1.0 / d;
    b[0] = (m11 * m22 - m12 * m21) * id;
    b[1] = (m02 * m21 - m01 * m22) * id;
    b[2] = (m01 * m12 - m02 * m11) * id;
    b[3] = (m12 * m20 - m10 * m22) * id;
    b[4] = (m00 * m22 - m02 * m20) * id;
    b[5] = (m02 * m10 - m00 * m12) * id;
    b[6] = (m10 * m21 - m11 * m20) * id;
    b[7] = (m01 * m20 - m00 * m21) * id;
    b[8] = (m00 * m11 - m01 * m10) * id;
    return true;
}

// One search node as the fused rounds read it: one 32-byte scalar load per visit. The centre and
// half side are the float search node's (tree.cc:496-506); xpad and vmax as the tree holds them;
// link is the first child, or ~plo for a leaf, and span the child count, or phi for a leaf.
struct alignas(32) SRec {
    float cx, cy, cz, half;
    float xpad, vmax;
    int link, span;
};

// A read of memory no kernel of the solve writes (the tree's records), through the constant
// address space: at a wave-uniform address it is a scalar load, served by the scalar cache.
template <class T>
__device__ __forceinline__ T ld_const(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "read as whole dwords");
    const __attribute__((address_space(4))) uint32_t* q =
        (const __attribute__((address_space(4))) uint32_t*)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (unsigned i = 0; i < sizeof(T) / 4; ++i) w[i] = q[i];
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

struct DensArgs {
    const uint32_t* act; // active targets (positions in the target list), in slot order
    int n_act;
    int tpw;             // targets per wavefront (a power of two, at most 64)
    const uint32_t* targets;
    const double *x, *y, *z, *m;
    const uint8_t* type; // null: all gas
    double des_ngb, ngb_tol;
    int n_dims;
    const double* cond_prev; // by particle, or null
    double *h, *h_lo, *h_hi, *rho;
    int *iter, *nngb, *iters;
    uint32_t* keep; // 1: the target stays active
    // the rounds: a target whose radius exceeds kFarLeaf times its leaf's side is left to a wave
    // of its own (appended to far_list)
    const int* tslot;     // sorted slot per target (list position), -1 outside the tree
    const int* slot_leaf;
    uint32_t* far_list;
    uint32_t* n_far;
    // the tree: its search nodes (built per call), the side of each node (the far test), the
    // sorted copy of the particles and their indices, the roots, and the time since the build
    // (the vmax pad)
    const SRec* srec;
    const WNode* wn;
    const double* size;
    const double4* src;
    const uint32_t* orderbuf;
    int root, nclusters;
    double tau;
};

// The terms one gas neighbour j adds to a target at (p0, p1, p2) with support h (hydro.cc:206-217):
// v = {W, dW/dh, m W, dx dx W, dy dy W, dz dz W, dx dy W, dy dz W, dx dz W, dx W, dy W, dz W}, and
// whether r < h. Non-gas neighbours add nothing (returns false).
constexpr int kTerms = 12;
__device__ __forceinline__ bool density_terms(const DensArgs& a, uint32_t j, double p0, double p1,
                                              double p2, double h, int dim, double* v,
                                              bool* inside) {
    if (a.type && a.type[j] != 0) return false; // gas neighbours only
    const double d0 = a.x[j] - p0, d1 = a.y[j] - p1, d2 = a.z[j] - p2;
    const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double w = kernel_w(r, h, dim);
    v[0] = w;
    v[1] = kernel_dwdh(r, h, dim);
    v[2] = a.m[j] * w;
    v[3] = d0 * d0 * w;
    v[4] = d1 * d1 * w;
    v[5] = d2 * d2 * w;
    v[6] = d0 * d1 * w;
    v[7] = d1 * d2 * w;
    v[8] = d0 * d2 * w;
    v[9] = d0 * w;
    v[10] = d1 * w;
    v[11] = d2 * w;
    *inside = r < h;
    return true;
}

// A target's sums over its neighbours, added in list (traversal) order.
struct DensSums {
    double s[kTerms];
    int n_inside;
};

// The rest of one step (hydro.cc:218-296) from the target's sums: the final pass's outputs when
// out of iterations, else the corrections, the convergence test and the Newton / bisection step.
__device__ void density_update(const DensArgs& a, int k, uint32_t t, uint32_t p, double h, int it,
                               const DensSums& S) {
    const int dim = a.n_dims;
    if (it >= kMaxIter) {
        // out of iterations: rho and the count at the last h (hydro.cc:285-295)
        a.rho[t] = S.s[2];
        a.nngb[t] = S.n_inside;
        a.iters[t] = kMaxIter;
        a.keep[k] = 0;
        return;
    }
    const double weight_sum = S.s[0], dweight_dh = S.s[1];
    const double* mom = S.s + 3;
    const double* face = S.s + 9;
    const double n_eff = ball_vol(h, dim) * weight_sum;
    // the condition-number and face-closure corrections widen the neighbour target
    // (hydro.cc:220-262)
    double ncorr = 1.0;
    if (a.cond_prev) {
        const double cn = a.cond_prev[p];
        if (cn > 100.0) ncorr = fmin(sqrt(1.0 + (cn - 100.0) / 1000.0), 2.0);
    }
    double des_eff = a.des_ngb * ncorr, tol_eff = a.ngb_tol * ncorr;
    if (weight_sum > 0) {
        double bm[9];
        if (invert_moments(mom, bm, dim)) {
            const double vol = 1.0 / weight_sum;
            const double dx_i = sqrt(vol * (mom[0] + mom[1] + mom[2]));
            double sum_abs = 0.0;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double lk =
                    (bm[3 * c] * face[0] + bm[3 * c + 1] * face[1] + bm[3 * c + 2] * face[2]) *
                    (2.0 * vol);
                sum_abs += fabs(lk) / dim;
            }
            const double pw = (dim == 3) ? dx_i * dx_i : (dim == 2 ? dx_i : 1.0);
            const double denom = 2.0 * dim * pw;
            if (denom > 0) {
                const double fce = sum_abs / denom;
                if (fce > 0.35) ncorr = fmax(ncorr, fmin(fce / 0.35, 2.0));
                des_eff = a.des_ngb * ncorr;
                tol_eff = a.ngb_tol * ncorr;
            }
        }
    }
    // the tolerance widens as the solve iterates (hydro.cc:263-267)
    if (it > 1)
        tol_eff = fmin(0.25 * des_eff, tol_eff * exp(0.1 * log(des_eff / (16.0 * tol_eff)) * it));
    const double residual = n_eff - des_eff;
    if (fabs(residual) < tol_eff) {
        a.rho[t] = S.s[2];
        a.nngb[t] = S.n_inside;
        a.iters[t] = it;
        a.keep[k] = 0;
        return;
    }
    double h_lo = a.h_lo[t], h_hi = a.h_hi[t];
    if (residual > 0) h_hi = h;
    else h_lo = h;
    // Newton on N_eff(h), guarded by the bracket (hydro.cc:271-281)
    const double dneff_dh = ball_vol(h, dim) * ((double)dim / h * weight_sum + dweight_dh);
    double h_next = (dneff_dh > 0) ? h - residual / dneff_dh : 0.0;
    if (h_next <= h_lo || (h_hi > 0 && h_next >= h_hi) || h_next <= 0) {
        if (h_hi > 0) h_next = (h_lo > 0) ? sqrt(h_lo * h_hi) : 0.5 * h_hi;
        else h_next = h * ((residual > 0) ? 0.7 : 1.6);
    }
    a.h[t] = h_next;
    a.h_lo[t] = h_lo;
    a.h_hi[t] = h_hi;
    a.iter[t] = it + 1;
    a.keep[k] = 1;
}

// One step for each of its targets (a.act[e] for block e), one wavefront per target walking that
// target's own search depth first. The lanes test a node's children together (a forest's roots up
// to 64 at a time), so the wave takes one step per admitted interior node; the admitted leaves
// among a node's children up to its next admitted interior child are taken together, their
// particles (contiguous slot ranges) spread over the lanes in slot order, tested against the
// sphere, and the neighbours' terms computed into LDS in that order. Lane c (c < 12) then adds
// term c of each, in order, and lane 12 counts those inside h, so every sum is the round's, term
// by term. For the targets whose spheres reach far beyond their leaf, and for the last few
// targets: with LOOP, the wave repeats its target's steps until the target is done.
template <bool LOOP>
__global__ __launch_bounds__(64) void k_density_far(DensArgs a) {
    __shared__ double s_v[kTerms][65]; // rows padded so the 12 adding lanes hit distinct banks
    __shared__ int s_in[64];
    const int lane = threadIdx.x;
    const int e = blockIdx.x;
    const int dim = a.n_dims;
    const uint32_t t = a.act[e];
    const uint32_t p = a.targets[t];
    const double qx = a.x[p], qy = a.y[p], qz = a.z[p];
    const int r_first = (a.nclusters > 1) ? 0 : a.root;
    const int r_count = (a.nclusters > 1) ? a.nclusters : 1;
    for (;;) {
        const double h = a.h[t];
        const int it = a.iter[t];
        const double r2 = h * h; // the search's radius_sq (hydro.cc:50)
        // the node test of the forward search (hydro.cc:55-72), as the rounds evaluate it
        auto admits = [&](const SRec& nr) -> bool {
            double dist = h + (double)nr.half + (double)nr.xpad;
            if (a.tau > 0) dist += (double)nr.vmax * a.tau;
            const double dx = (double)nr.cx - qx, dy = (double)nr.cy - qy,
                         dz = (double)nr.cz - qz;
            if ((dx > dist || -dx > dist) || (dy > dist || -dy > dist) ||
                (dz > dist || -dz > dist))
                return false;
            const double ds = dist + 0.7320508075688772 * (double)nr.half;
            return !(dx * dx + dy * dy + dz * dz > ds * ds);
        };
        double acc = 0.0; // lane c < 12: term c's sum
        int n_in = 0;     // lane 12: the count inside h
        int nbuf = 0;     // neighbours waiting in s_v (the same in every lane)
        auto flush = [&]() {
            __syncthreads();
            if (lane < kTerms) {
                for (int i = 0; i < nbuf; ++i) acc += s_v[lane][i];
            } else if (lane == kTerms) {
                for (int i = 0; i < nbuf; ++i) n_in += s_in[i];
            }
            __syncthreads();
            nbuf = 0;
        };
        // one particle per lane (slot s, or none for s < 0): a neighbour when within the radius
        // (the tree's copy of the position, as the search tests it), its terms packed into s_v
        // in lane order
        auto take = [&](int s) {
            double v[kTerms];
            bool inside = false, ok = false;
            if (s >= 0) {
                const double4 q = a.src[s];
                const double ex = q.x - qx, ey = q.y - qy, ez = q.z - qz;
                if (ex * ex + ey * ey + ez * ez < r2)
                    ok = density_terms(a, a.orderbuf[s], qx, qy, qz, h, dim, v, &inside);
            }
            const uint64_t om = __ballot(ok);
            if (om == 0) return;
            const int n_ok = __builtin_popcountll(om);
            if (nbuf + n_ok > 64) flush();
            if (ok) {
                const int pos = nbuf + __builtin_popcountll(om & ((1ull << lane) - 1));
#pragma unroll
                for (int c = 0; c < kTerms; ++c) s_v[c][pos] = v[c];
                s_in[pos] = inside ? 1 : 0;
            }
            nbuf += n_ok;
        };
        // the admitted leaves run (bit j: child cf + j), their particles in slot order
        auto leaves = [&](int cf, uint64_t run) {
            int plo = 0, cnt = 0, off = 0, total = 0;
            if ((run >> lane) & 1ull) {
                const SRec c = a.srec[cf + lane];
                plo = ~c.link;
                cnt = c.span - plo;
            }
            for (uint64_t r = run; r != 0; r &= r - 1) {
                const int j = (int)__builtin_ctzll(r);
                if (lane == j) off = total;
                total += __builtin_amdgcn_readlane(cnt, j);
            }
            for (int g0 = 0; g0 < total; g0 += 64) {
                const int g = g0 + lane;
                int s = -1;
                for (uint64_t r = run; r != 0; r &= r - 1) {
                    const int j = (int)__builtin_ctzll(r);
                    const int o = __builtin_amdgcn_readlane(off, j);
                    if (g >= o && g < o + __builtin_amdgcn_readlane(cnt, j))
                        s = __builtin_amdgcn_readlane(plo, j) + (g - o);
                }
                take(s);
            }
        };
        // A node whose box (its centre and half side as the node test reads them, so the cells
        // below it too) lies inside the sphere, by a margin above the float rounding of node
        // centres: every node below it is admitted, so the walk would test each of its particles
        // in slot order. They are a contiguous slot range, taken 64 at a time instead.
        auto inside_sphere = [&](const SRec& nr) -> bool {
            const double hf = (double)nr.half;
            const double ax = fabs((double)nr.cx - qx) + hf, ay = fabs((double)nr.cy - qy) + hf,
                         az = fabs((double)nr.cz - qz) + hf;
            const double m = 1e-6 * (fabs((double)nr.cx) + fabs((double)nr.cy) +
                                     fabs((double)nr.cz));
            const double hm = h * (1.0 - 1e-12) - m;
            return hm > 0 && ax * ax + ay * ay + az * az < hm * hm;
        };
        for (int r0 = 0; r0 < r_count; r0 += 64) {
            // the children of the node being walked (first index; lane masks of those admitted
            // and not yet visited, and of those that are leaves) and the stack of the same above
            // it, entry d in lane d
            int cf = r_first + r0;
            uint64_t cm, lm;
            {
                bool hit = false, leaf = false;
                if (lane < min(64, r_count - r0)) {
                    const SRec c = a.srec[cf + lane];
                    hit = admits(c);
                    leaf = c.link < 0;
                }
                cm = __ballot(hit);
                lm = __ballot(hit && leaf);
            }
            int sp = 0, stf = 0, sml = 0, smh = 0, sll = 0, slh = 0;
            for (;;) {
                if (cm == 0) {
                    if (sp == 0) break;
                    --sp;
                    cf = __builtin_amdgcn_readlane(stf, sp);
                    cm = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(smh, sp) << 32) |
                         (uint64_t)(uint32_t)__builtin_amdgcn_readlane(sml, sp);
                    lm = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(slh, sp) << 32) |
                         (uint64_t)(uint32_t)__builtin_amdgcn_readlane(sll, sp);
                    continue;
                }
                const uint64_t inner = cm & ~lm;
                if ((cm & (0 - cm) & lm) != 0) {
                    // the next child is a leaf: it and the leaves after it up to the next
                    // interior child
                    const uint64_t run = inner ? (cm & ((inner & (0 - inner)) - 1)) : cm;
                    cm &= ~run;
                    leaves(cf, run);
                    continue;
                }
                // an interior child: its admitted children are walked before its remaining
                // siblings, which wait on the stack
                const int node = cf + (int)__builtin_ctzll(cm);
                cm &= cm - 1;
                const SRec nr = ld_const(a.srec + node);
                if (inside_sphere(nr)) {
                    const WNode w = ld_const(a.wn + node);
                    for (int s0 = w.plo; s0 < w.phi; s0 += 64)
                        take(s0 + lane < w.phi ? s0 + lane : -1);
                    continue;
                }
                bool hit = false, leaf = false;
                if (lane < nr.span) {
                    const SRec c = a.srec[nr.link + lane];
                    hit = admits(c);
                    leaf = c.link < 0;
                }
                const uint64_t om = __ballot(hit);
                if (om == 0) continue;
                if (cm != 0) {
                    const bool top = lane == sp;
                    stf = top ? cf : stf;
                    sml = top ? (int)(uint32_t)cm : sml;
                    smh = top ? (int)(uint32_t)(cm >> 32) : smh;
                    sll = top ? (int)(uint32_t)lm : sll;
                    slh = top ? (int)(uint32_t)(lm >> 32) : slh;
                    ++sp;
                }
                cf = nr.link;
                cm = om;
                lm = __ballot(hit && leaf);
            }
        }
        flush();
        DensSums S;
#pragma unroll
        for (int c = 0; c < kTerms; ++c) S.s[c] = __shfl(acc, c);
        S.n_inside = __shfl(n_in, kTerms);
        if (lane == 0) density_update(a, e, t, p, h, it, S);
        if (!LOOP) break;
        __syncthreads(); // the update's state, written by lane 0, for every lane
        if (a.keep[e] == 0) break;
    }
}

// One step for every active target (see the top of the file): tpw targets per wavefront, lane l
// holding active entry wave * tpw + l. A target whose radius reaches far beyond its own leaf would
// make its wave walk a large part of the tree one node at a time; it is queued for a wave of its
// own (k_density_far) and takes no part here.
__global__ __launch_bounds__(64) void k_density_round(DensArgs a) {
    const int lane = threadIdx.x;
    // Block b runs on XCD b % 8: each XCD gets one contiguous share of the Morton-ordered waves.
    const int b = blockIdx.x, nb = gridDim.x;
    const int xcd = b & 7, per = nb >> 3, rem = nb & 7;
    const int wv = xcd * per + min(xcd, rem) + (b >> 3);
    const int k = wv * a.tpw + lane;
    const bool live = lane < a.tpw && k < a.n_act;
    const int dim = a.n_dims;
    uint32_t t = 0, p = 0;
    double qx = 0, qy = 0, qz = 0, h = 0;
    int it = 0;
    bool mine = false; // this lane steps its target here
    bool far = false;  // this lane's target is left to a wave of its own
    if (live) {
        t = a.act[k];
        p = a.targets[t];
        qx = a.x[p];
        qy = a.y[p];
        qz = a.z[p];
        h = a.h[t];
        it = a.iter[t];
        mine = true;
        const int ts = a.tslot[t];
        if (ts >= 0 && h > kFarLeaf * a.size[a.slot_leaf[ts]]) {
            far = true;
            mine = false;
        }
    }
    // Each lane's neighbours (particle indices, in traversal order) wait in its column of s_nb
    // until some column is full or the walk ends; then every lane adds its waiting neighbours'
    // terms, in order. A lane reads back only what it wrote itself, so no barrier is needed.
    __shared__ uint32_t s_nb[kNbBuf][64];
    DensSums S;
    int nb_n = 0;
    auto flush = [&]() {
        for (int e = 0; __ballot(e < nb_n) != 0; ++e) {
            if (e < nb_n) {
                double v[kTerms];
                bool inside = false;
                if (density_terms(a, s_nb[e][lane], qx, qy, qz, h, dim, v, &inside)) {
#pragma unroll
                    for (int c = 0; c < kTerms; ++c) S.s[c] += v[c];
                    if (inside) ++S.n_inside;
                }
            }
        }
        nb_n = 0;
    };

    // (the far targets are listed at the end: no store precedes the walk's reads)
    const uint64_t lanes = __ballot(mine);
    if (lanes == 0) {
        if (far) a.far_list[atomicAdd(a.n_far, 1u)] = (uint32_t)k;
        return;
    }
    const double r2 = h * h; // the search's radius_sq (hydro.cc:50)
#pragma unroll
    for (int c = 0; c < kTerms; ++c) S.s[c] = 0.0;
    S.n_inside = 0;
    // The current sibling range (next node, nodes left, lanes waiting on it) and the
    // ranges above it, entry e in lane e of stf/stc/stl/sth: one per tree level, below the
    // roots' range (a forest's roots are nodes 0..nc-1, searched in that order).
    int cf = (a.nclusters > 1) ? 0 : a.root;
    int cc = (a.nclusters > 1) ? a.nclusters : 1;
    uint64_t cm = lanes;
    int sp = 0;
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
        const bool wait = (cm >> lane) & 1ull;
        // the node test of the forward search, per lane (hydro.cc:55-72), exactly as the
        // device search evaluates it
        const SRec nr = ld_const(a.srec + node);
        double dist = h + (double)nr.half + (double)nr.xpad;
        if (a.tau > 0) dist += (double)nr.vmax * a.tau;
        const double dx = (double)nr.cx - qx, dy = (double)nr.cy - qy, dz = (double)nr.cz - qz;
        bool hit = false;
        if (wait && !(dx > dist || -dx > dist) && !(dy > dist || -dy > dist) &&
            !(dz > dist || -dz > dist)) {
            const double ds = dist + 0.7320508075688772 * (double)nr.half;
            hit = !(dx * dx + dy * dy + dz * dz > ds * ds);
        }
        const uint64_t om = __ballot(hit);
        if (om == 0) continue;
        if (nr.link < 0) {
            // the leaf's particles in slot order: a neighbour when within the radius (the
            // tree's copy of the position, as the search tests it)
            // (four records loaded together, then noted one by one in slot order)
            for (int s0 = ~nr.link; s0 < nr.span; s0 += 4) {
                double4 q[4];
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (s0 + j < nr.span) q[j] = ld_const(a.src + s0 + j);
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (s0 + j >= nr.span) break;
                    const double ex = q[j].x - qx, ey = q[j].y - qy, ez = q[j].z - qz;
                    const bool in = hit && ex * ex + ey * ey + ez * ez < r2;
                    if (__ballot(in) == 0) continue;
                    if (__ballot(nb_n == kNbBuf) != 0) flush();
                    if (in) s_nb[nb_n++][lane] = ld_const(a.orderbuf + s0 + j);
                }
            }
            continue;
        }
        // descend: the node's remaining siblings (if any) wait on the stack below its
        // children
        const bool top = lane == sp;
        stf = top ? cf : stf;
        stc = top ? cc : stc;
        stl = top ? (int)(uint32_t)cm : stl;
        sth = top ? (int)(uint32_t)(cm >> 32) : sth;
        sp += (cc > 0) ? 1 : 0;
        cf = nr.link;
        cc = nr.span;
        cm = om;
    }
    flush();
    if (mine) density_update(a, k, t, p, h, it, S);
    if (far) a.far_list[atomicAdd(a.n_far, 1u)] = (uint32_t)k;
}

// The search nodes of the fused rounds, one per tree node.
__global__ __launch_bounds__(256) void k_density_srec(int nn, const WNode* __restrict__ wn,
                                                      const double* __restrict__ gcx,
                                                      const double* __restrict__ gcy,
                                                      const double* __restrict__ gcz,
                                                      const double* __restrict__ size,
                                                      const float* __restrict__ xpad,
                                                      const float* __restrict__ vmax,
                                                      double snode_eps, SRec* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nn) return;
    const WNode w = wn[i];
    SRec r;
    r.cx = (float)gcx[i];
    r.cy = (float)gcy[i];
    r.cz = (float)gcz[i];
    r.half = snode_half_of(gcx[i], gcy[i], gcz[i], size[i], snode_eps);
    r.xpad = xpad[i];
    r.vmax = vmax[i];
    r.link = (w.first >= 0) ? w.first : ~w.plo;
    r.span = (w.first >= 0) ? w.nchild : w.phi;
    out[i] = r;
}

// Start state: every target active at its first guess; a target outside the particle set flags
// status bit 0 (the call then fails before any search).
__global__ __launch_bounds__(256) void k_density_init(int nt, const uint32_t* __restrict__ targets,
                                                      size_t n, const double* __restrict__ h_start,
                                                      double h_default, double* __restrict__ h,
                                                      double* __restrict__ h_lo,
                                                      double* __restrict__ h_hi,
                                                      int* __restrict__ iter,
                                                      uint32_t* __restrict__ act,
                                                      int* __restrict__ status) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= nt) return;
    if (targets[t] >= n) atomicOr(status, 1);
    h[t] = h_start ? h_start[t] : h_default;
    h_lo[t] = 0.0;
    h_hi[t] = 0.0;
    iter[t] = 0;
    act[t] = (uint32_t)t;
}

// The far targets' entries in the active list: their target positions, for the list path.
__global__ __launch_bounds__(256) void k_density_far_act(int nf, const uint32_t* __restrict__ far,
                                                         const uint32_t* __restrict__ act,
                                                         uint32_t* __restrict__ far_act) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < nf) far_act[j] = act[far[j]];
}

// The list path's keep flags back into the active list's.
__global__ __launch_bounds__(256) void k_density_far_keep(int nf, const uint32_t* __restrict__ far,
                                                          const uint32_t* __restrict__ far_keep,
                                                          uint32_t* __restrict__ keep) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < nf) keep[far[j]] = far_keep[j];
}

// next[j] = cur[pos[j]] for the kept entries (pos from the flags' scan).
__global__ __launch_bounds__(256) void k_density_compact(int nk, const int* __restrict__ pos,
                                                         const uint32_t* __restrict__ cur,
                                                         uint32_t* __restrict__ next) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < nk) next[j] = cur[pos[j]];
}

// The bounding box of all particles (hydro.cc:172-178), as order-preserving 64-bit keys of the
// doubles: min x y z in box[0..2], max in box[3..5] (the max stored complemented so that both
// are atomicMin).
__device__ __forceinline__ unsigned long long dkey(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__global__ __launch_bounds__(256) void k_density_bbox(size_t n, const double* __restrict__ x,
                                                      const double* __restrict__ y,
                                                      const double* __restrict__ z,
                                                      unsigned long long* __restrict__ box) {
    unsigned long long lo[3] = {~0ull, ~0ull, ~0ull}, hi[3] = {~0ull, ~0ull, ~0ull};
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const double v[3] = {x[i], y[i], z[i]};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const unsigned long long kk = dkey(v[d]);
            lo[d] = min(lo[d], kk);
            hi[d] = min(hi[d], ~kk);
        }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        for (int s = 32; s >= 1; s >>= 1) {
            lo[d] = min(lo[d], (unsigned long long)__shfl_xor((long long)lo[d], s));
            hi[d] = min(hi[d], (unsigned long long)__shfl_xor((long long)hi[d], s));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            atomicMin(&box[d], lo[d]);
            atomicMin(&box[3 + d], hi[d]);
        }
    }
}

double key_to_double(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double v;
    std::memcpy(&v, &b, sizeof v);
    return v;
}

} // namespace
} // namespace gz

using namespace gz;

extern "C" gz_status gz_tree_density(gz_ctx* ctx, const gz_tree* tree, const gz_particles* p,
                                     const uint32_t* targets, size_t n_t, double des_ngb,
                                     double ngb_tol, const double* h_start, int n_dims,
                                     const double* cond_prev, double tau, double* h, double* rho,
                                     int* nngb, int* iters) {
    clear_error();
    if (!ctx || !tree || !p) return fail(GZ_ERR_INVALID, "gz_tree_density: null argument");
    GZ_HIP(hipSetDevice(ctx->device));
    if (n_t == 0) return GZ_OK;
    if (!targets || !h || !rho || !nngb || !iters || !p->x || !p->y || !p->z || !p->m)
        return fail(GZ_ERR_INVALID, "gz_tree_density: null array");
    if (n_dims < 1 || n_dims > 3)
        return fail(GZ_ERR_INVALID, "gz_tree_density: n_dims=%d (1, 2 or 3)", n_dims);
    if (!(des_ngb > 0) || !(ngb_tol >= 0))
        return fail(GZ_ERR_INVALID, "gz_tree_density: des_ngb must be > 0 and ngb_tol >= 0");
    if (n_t >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "gz_tree_density: n_t too large");
    if (p->n != tree->n)
        return fail(GZ_ERR_INVALID, "gz_tree_density: %zu particles, the tree holds %zu", p->n,
                    (size_t)tree->n);
    hipStream_t st = ctx->stream;
    const int nt = (int)n_t;

    // state: h_lo, h_hi (doubles); iter, the active list, the compaction's output, the keep
    // flags, their scan and the kept positions; the targets' slots and the sort's keys and
    // values (twice); the box keys, the scan total and the status words; the scan's temporary
    // space
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += (bytes + 255) & ~size_t(255);
        return o;
    };
    const size_t o_lo = take(n_t * 8), o_hi = take(n_t * 8), o_iter = take(n_t * 4),
                 o_act = take(n_t * 4), o_nxt = take(n_t * 4), o_keep = take((n_t + 1) * 4),
                 o_scan = take((n_t + 1) * 4), o_kept = take((n_t + 1) * 4),
                 o_slot = take(n_t * 4), o_k0 = take(n_t * 8), o_k1 = take(n_t * 8),
                 o_v0 = take(n_t * 4), o_v1 = take(n_t * 4), o_far = take(n_t * 4),
                 o_fact = take(n_t * 4), o_fkeep = take(n_t * 4),
                 o_srec = take((size_t)std::max(tree->nnodes, 1) * sizeof(SRec)),
                 o_misc = take(256), o_tmp = take(scan_tmp_bytes(n_t + 1));
    DevMem& mem = ctx->scratch[13];
    GZ_TRY(mem.reserve(off));
    uint8_t* base = mem.as<uint8_t>();
    double* d_lo = reinterpret_cast<double*>(base + o_lo);
    double* d_hi = reinterpret_cast<double*>(base + o_hi);
    int* d_iter = reinterpret_cast<int*>(base + o_iter);
    uint32_t* d_act = reinterpret_cast<uint32_t*>(base + o_act);
    uint32_t* d_nxt = reinterpret_cast<uint32_t*>(base + o_nxt);
    uint32_t* d_keep = reinterpret_cast<uint32_t*>(base + o_keep);
    uint32_t* d_scan = reinterpret_cast<uint32_t*>(base + o_scan);
    int* d_kept = reinterpret_cast<int*>(base + o_kept);
    unsigned long long* d_box = reinterpret_cast<unsigned long long*>(base + o_misc);
    uint32_t* d_total = reinterpret_cast<uint32_t*>(base + o_misc + 64);
    int* d_status = reinterpret_cast<int*>(base + o_misc + 128);
    int* d_sort = reinterpret_cast<int*>(base + o_misc + 192); // bit 1: bad index; [1]: unsorted
    uint32_t* d_nfar = reinterpret_cast<uint32_t*>(base + o_misc + 224);
    uint32_t* d_far = reinterpret_cast<uint32_t*>(base + o_far);
    uint32_t* d_fact = reinterpret_cast<uint32_t*>(base + o_fact);
    uint32_t* d_fkeep = reinterpret_cast<uint32_t*>(base + o_fkeep);
    GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(base + o_misc), 64, st));

    // the default first guess: des_ngb particles at the mean density of the bounding box of all
    // particles (hydro.cc:168-186), the box reduced on the device and the guess formed here
    double h_default = 0.0;
    if (!h_start) {
        if (p->n == 0) return fail(GZ_ERR_INVALID, "gz_tree_density: no particles");
        const unsigned long long init[6] = {~0ull, ~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
        GZ_HIP(hipMemcpyAsync(d_box, init, sizeof init, hipMemcpyHostToDevice, st));
        const unsigned gb = (unsigned)std::min<size_t>(1024, (p->n + 255) / 256);
        k_density_bbox<<<gb, 256, 0, st>>>(p->n, p->x, p->y, p->z, d_box);
        GZ_HIP(hipGetLastError());
        unsigned long long kb[6];
        GZ_HIP(hipMemcpyAsync(kb, d_box, sizeof kb, hipMemcpyDeviceToHost, st));
        GZ_HIP(hipStreamSynchronize(st));
        double lo[3], hi[3];
        for (int d = 0; d < 3; ++d) {
            lo[d] = key_to_double(kb[d]);
            hi[d] = key_to_double(~kb[3 + d]);
        }
        const double bbox_volume = (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        const double np = (double)p->n;
        h_default = std::cbrt(3.0 * des_ngb * bbox_volume / (4.0 * kPi * np));
        if (n_dims == 1) h_default = 0.5 * des_ngb * (hi[0] - lo[0]) / np;
        if (n_dims == 2)
            h_default = std::sqrt(des_ngb * (hi[0] - lo[0]) * (hi[1] - lo[1]) / (kPi * np));
    }
    k_density_init<<<(unsigned)((n_t + 255) / 256), 256, 0, st>>>(
        nt, targets, p->n, h_start, h_default, h, d_lo, d_hi, d_iter, d_act, d_status);
    GZ_HIP(hipGetLastError());
    // the targets' sorted slots: a list out of slot order is walked in slot order (the waves'
    // targets are then neighbours), the results staying by list position
    uint64_t* k0 = reinterpret_cast<uint64_t*>(base + o_k0);
    uint64_t* k1 = reinterpret_cast<uint64_t*>(base + o_k1);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(base + o_v0);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(base + o_v1);
    GZ_TRY(launch_pt_keys(targets, n_t, tree, p->n, reinterpret_cast<int*>(base + o_slot), k0, v0,
                          d_sort, st));
    int head[3] = {0, 0, 0};
    GZ_HIP(hipMemcpyAsync(head, d_status, sizeof(int), hipMemcpyDeviceToHost, st));
    GZ_HIP(hipMemcpyAsync(head + 1, d_sort, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if ((head[0] & 1) || (head[1] & 2))
        return fail(GZ_ERR_INVALID, "gz_tree_density: a target is outside the particle set");
    if (head[2]) {
        int bits = 1;
        while (bits < 40 && (1ull << bits) <= (unsigned long long)tree->n_live) ++bits;
        GZ_TRY(radix_sort_pairs(k0, v0, k1, v1, n_t, bits, ctx->scratch[3], st));
        GZ_HIP(hipMemcpyAsync(d_act, v0, n_t * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
    }
    if (tree->depth + 2 > kPtDepth)
        return fail(GZ_ERR_INTERNAL, "gz_tree_density: tree depth %d exceeds the walk's stack",
                    tree->depth);

    DensArgs a{};
    a.targets = targets;
    a.x = p->x;
    a.y = p->y;
    a.z = p->z;
    a.m = p->m;
    a.type = p->type;
    a.des_ngb = des_ngb;
    a.ngb_tol = ngb_tol;
    a.n_dims = n_dims;
    a.cond_prev = cond_prev;
    a.h = h;
    a.h_lo = d_lo;
    a.h_hi = d_hi;
    a.rho = rho;
    a.iter = d_iter;
    a.nngb = nngb;
    a.iters = iters;
    a.keep = d_keep;
    a.tslot = reinterpret_cast<const int*>(base + o_slot);
    a.slot_leaf = tree->slot_leaf;
    a.far_list = d_far;
    a.n_far = d_nfar;
    a.srec = reinterpret_cast<const SRec*>(base + o_srec);
    if (tree->nnodes > 0) {
        k_density_srec<<<(unsigned)((tree->nnodes + 255) / 256), 256, 0, st>>>(
            tree->nnodes, tree->wn, tree->gcx, tree->gcy, tree->gcz, tree->size, tree->xpad,
            tree->vmax, 1e-6 * tree->root_box[3], reinterpret_cast<SRec*>(base + o_srec));
        GZ_HIP(hipGetLastError());
    }
    a.wn = tree->wn;
    a.size = tree->size;
    a.src = tree->src;
    a.orderbuf = tree->orderbuf;
    a.root = tree->root;
    a.nclusters = tree->n_clusters;
    a.tau = tau;

    // rounds: a target takes at most 100 steps and one last search
    uint32_t n_act = (uint32_t)n_t;
    for (int round = 0; n_act > 0; ++round) {
        if (round > kMaxIter + 1)
            return fail(GZ_ERR_INTERNAL, "gz_tree_density: the solve did not end");
        // 64 targets per wave while that gives enough waves; fewer (down to one) when the active
        // targets are too few to occupy the GPU that way (and often far apart, as the last
        // unconverged targets are). The last few iterate to the end in one launch.
        int tpw = 64;
        while (tpw > 1 && (n_act + tpw - 1) / tpw < (uint32_t)kWavesWanted) tpw >>= 1;
        a.act = d_act;
        a.n_act = (int)n_act;
        a.tpw = tpw;
        const bool loop = n_act <= kLoopTargets;
        GZ_TRY(fill_zero_u32(d_nfar, 1, st));
        if (loop) k_density_far<true><<<n_act, 64, 0, st>>>(a);
        else k_density_round<<<(n_act + tpw - 1) / tpw, 64, 0, st>>>(a);
        GZ_HIP(hipGetLastError());
        uint32_t n_far = 0;
        if (!loop) {
            GZ_HIP(hipMemcpyAsync(&n_far, d_nfar, sizeof n_far, hipMemcpyDeviceToHost, st));
            GZ_HIP(hipStreamSynchronize(st));
        }
        if (n_far > 0) {
            // the far targets: a wave each, in the round's order (their keep flags are then
            // moved to the round's positions)
            const unsigned gf = (n_far + 255) / 256;
            k_density_far_act<<<gf, 256, 0, st>>>((int)n_far, d_far, d_act, d_fact);
            GZ_HIP(hipGetLastError());
            DensArgs f = a;
            f.act = d_fact;
            f.n_act = (int)n_far;
            f.keep = d_fkeep;
            k_density_far<false><<<n_far, 64, 0, st>>>(f);
            GZ_HIP(hipGetLastError());
            k_density_far_keep<<<gf, 256, 0, st>>>((int)n_far, d_far, d_fkeep, d_keep);
            GZ_HIP(hipGetLastError());
        }
        // the targets still active, in their order: kept positions, then their target indices
        GZ_TRY(exclusive_scan_u32(d_keep, d_scan, n_act, d_total, base + o_tmp, st, d_kept));
        uint32_t n_next = 0;
        GZ_HIP(hipMemcpyAsync(&n_next, d_total, sizeof n_next, hipMemcpyDeviceToHost, st));
        GZ_HIP(hipStreamSynchronize(st));
        if (n_next > 0) {
            k_density_compact<<<(n_next + 255) / 256, 256, 0, st>>>((int)n_next, d_kept, d_act,
                                                                     d_nxt);
            GZ_HIP(hipGetLastError());
            std::swap(d_act, d_nxt);
        }
        n_act = n_next;
    }
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}
