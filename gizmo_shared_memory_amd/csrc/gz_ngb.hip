// Neighbour searches over the tree (shmem/hydro.cc:37-148, the SNode of tree.h:206-226), for many
// centres at once: every particle within radius_q of centre q (ngb_search), and every gas particle
// whose own kernel reaches centre q although q's does not (ngb_search_reverse, with the node
// hmax of Tree::init_hmax / raise_hmax, tree.cc:286-338).
//
// One wavefront per group of 64 consecutive queries. Up to 64 stacked nodes are tested at once,
// one per lane: first against the group's bounding box with the group's largest radius (for the
// reverse search: the node's hmax, and only if it exceeds the group's smallest kernel radius),
// then with the reference's own node test for each query of the group in turn, a node staying
// only if some query admits it — so the wave walks the union of its queries' traversals even when
// the group straddles a jump of the Morton order. Node boxes are the reference's SNode boxes
// (half side widened for float centres) grown by vmax * tau + xpad. The particles of kept leaves
// are staged through LDS 64 at a time and each lane tests its own query against every staged
// particle with the reference's FP64 expression, so the sets are exact. Two passes: counts, then
// (with the offsets) the sorted slots of each query's neighbours, which map to particle indices
// in Morton order, the reference's traversal order.
#include <cfloat>

#include "gz_features.h"
#include "gz_internal.h"
#include "gz_wave.h"

namespace gz {

struct NgbArgs {
    const WNode* wn;
    const double* gcx;
    const double* gcy;
    const double* gcz;
    const double* size;
    const float* vmax;
    const float* xpad;
    const double4* src;
    const uint32_t* orderbuf;
    const double *cx, *cy, *cz, *radius;
    int nq, ngroups, root, nclusters;
    int qpw; // queries per wave (a power of two, at most 64)
    double snode_eps, tau;
    int* queue;
    int* stack; // stack_cap ints per wave
    int stack_cap;
    int* status; // bit 0: a stack outgrew stack_cap
    uint32_t* count;     // pass 1: neighbours per query
    const size_t* start; // pass 2: CSR offsets
    uint32_t* slots;     // pass 2: the neighbours' sorted slots
    // reverse search: node hmax, the gas kernel radii by particle, the gas count
    const float* hmax;
    const double* h;
    size_t n_gas;
};

// The search node's half side (tree.cc:496-506): float-safe, never smaller than the true one.
__device__ __forceinline__ float snode_half(const NgbArgs& a, int i) {
    return snode_half_of(a.gcx[i], a.gcy[i], a.gcz[i], a.size[i], a.snode_eps);
}

// REV: the reverse search (hydro.cc:110-148). radius[q] is then the query's own kernel radius h_q,
// a node is a candidate only when its hmax exceeds the smallest h_q of the group, its reach is
// hmax instead of the query radius, and a leaf particle j counts when it is gas and
// h_q <= r < h[j].
template <bool FILL, bool REV>
__global__ __launch_bounds__(64) void k_ngb(NgbArgs a) {
    __shared__ int s_lpre[64], s_lplo[64]; // kept leaves: first staged index, first slot
    __shared__ double s_px[64], s_py[64], s_pz[64], s_h2[64];
    __shared__ int s_slot[64];
    const int lane = threadIdx.x;
    int* const stk = a.stack + (size_t)blockIdx.x * (size_t)a.stack_cap;
    for (;;) {
        int v = 0;
        if (lane == 0) v = atomicAdd(a.queue, 1);
        const int g = __shfl(v, 0);
        if (g >= a.ngroups) break;
        const int q = g * a.qpw + lane;
        const bool live = lane < a.qpw && q < a.nq;
        double qx = 0, qy = 0, qz = 0, r = 0;
        if (live) {
            qx = a.cx[q]; qy = a.cy[q]; qz = a.cz[q];
            r = a.radius[q];
        }
        const double lo0 = wave_min_d(live ? qx : 1e300), hi0 = wave_max_d(live ? qx : -1e300);
        const double lo1 = wave_min_d(live ? qy : 1e300), hi1 = wave_max_d(live ? qy : -1e300);
        const double lo2 = wave_min_d(live ? qz : 1e300), hi2 = wave_max_d(live ? qz : -1e300);
        const double rmax = wave_max_d(live ? r : 0.0);
        const double rmin = REV ? wave_min_d(live ? r : 1e300) : 0.0;
        const double r2 = r * r;
        uint32_t found = 0;
        size_t out = live && FILL ? a.start[q] : 0;

        // The roots, one per tree of a forest, form a virtual range that is fed onto the stack
        // whenever it holds less than a batch: a forest of any size then needs no more stack
        // than one tree plus a batch (the lists are sorted afterwards, so the order in which
        // the trees are searched does not matter).
        const int nroots = (a.nclusters > 1) ? a.nclusters : 1;
        int rnext = 0;
        int sp = 0;
        bool overflow = false;
        while (sp > 0 || rnext < nroots) {
            if (sp < 64 && rnext < nroots) {
                const int add = min(64 - sp, nroots - rnext);
                if (lane < add) stk[sp + lane] = (a.nclusters > 1) ? rnext + lane : a.root;
                sp += add;
                rnext += add;
                __threadfence_block();
                __syncthreads();
            }
            const int k = min(sp, 64);
            const int base = sp - k;
            const int node = (lane < k) ? stk[base + lane] : -1;
            sp = base;
            bool keep = false, leaf = false;
            // the node as the per-query test below needs it (kept nodes only)
            double nc0 = 0, nc1 = 0, nc2 = 0, nhalf = 0, nxpad = 0, nvmax = 0, nhm = 0;
            int first = -1, nch = 0, plo = 0, phi = 0;
            if (node >= 0) {
                const WNode w = a.wn[node];
                const float half = snode_half(a, node);
                const double hm = REV ? (double)a.hmax[node] : 0.0;
                double dist = (REV ? hm : rmax) + (double)half + (double)a.xpad[node];
                if (a.tau > 0) dist += (double)a.vmax[node] * a.tau;
                const double c0 = (double)(float)a.gcx[node], c1 = (double)(float)a.gcy[node],
                             c2 = (double)(float)a.gcz[node];
                const double d0 = fmax(0.0, fmax(lo0 - c0, c0 - hi0));
                const double d1 = fmax(0.0, fmax(lo1 - c1, c1 - hi1));
                const double d2 = fmax(0.0, fmax(lo2 - c2, c2 - hi2));
                if ((!REV || hm > rmin) && d0 <= dist && d1 <= dist && d2 <= dist) {
                    const double ds = dist + 0.7320508075688772 * (double)half;
                    keep = d0 * d0 + d1 * d1 + d2 * d2 <= ds * ds;
                }
                if (keep) {
                    nc0 = c0; nc1 = c1; nc2 = c2;
                    nhalf = (double)half;
                    nxpad = (double)a.xpad[node];
                    nvmax = (double)a.vmax[node];
                    nhm = hm;
                    first = w.first;
                    if (first < 0) {
                        leaf = true;
                        plo = w.plo;
                        phi = w.phi;
                    } else {
                        nch = w.nchild;
                    }
                }
            }
            // A kept node is dropped when the reference's node test (hydro.cc:55-72, reverse
            // :120-133) rejects it for every query of the group: the group's box and largest
            // radius admit it, but no single query needs it (a group that straddles a jump of
            // the Morton order has a wide box). The wave then walks the union of its queries'
            // own traversals.
            {
                uint64_t rem = __ballot(keep), need = 0;
                while (rem) {
                    const int sl = __ffsll((unsigned long long)rem) - 1;
                    rem &= rem - 1;
                    const double b0 = __shfl(nc0, sl), b1 = __shfl(nc1, sl), b2 = __shfl(nc2, sl);
                    const double bh = __shfl(nhalf, sl), bx = __shfl(nxpad, sl);
                    const double bv = __shfl(nvmax, sl), bm = REV ? __shfl(nhm, sl) : 0.0;
                    bool hit = false;
                    if (live && (!REV || bm > r)) {
                        double dist = (REV ? bm : r) + bh + bx;
                        if (a.tau > 0) dist += bv * a.tau;
                        const double dx = b0 - qx, dy = b1 - qy, dz = b2 - qz;
                        if (!(dx > dist || -dx > dist) && !(dy > dist || -dy > dist) &&
                            !(dz > dist || -dz > dist)) {
                            dist += 0.7320508075688772 * bh;
                            hit = !(dx * dx + dy * dy + dz * dz > dist * dist);
                        }
                    }
                    if (__ballot(hit)) need |= 1ull << sl;
                }
                if (!((need >> lane) & 1ull)) {
                    keep = false;
                    leaf = false;
                    nch = 0;
                }
            }
            // children of kept interior nodes go on the stack
            int ctot = 0;
            const int coff = wave_excl_scan(nch, ctot);
            if (sp + ctot > a.stack_cap) {
                overflow = true;
                break;
            }
            for (int c = 0; c < nch; ++c) stk[sp + coff + c] = first + c;
            sp += ctot;
            // Particles of the kept leaves: the wave loads up to 64 at a time, one per lane, into
            // LDS, and every lane then tests its own query against each staged particle.
            const uint64_t lb = __ballot(leaf);
            if (lb) {
                const int nleaf = __popcll(lb);
                int ptot = 0;
                const int poff = wave_excl_scan(leaf ? phi - plo : 0, ptot);
                if (leaf) {
                    const int li = (int)__builtin_amdgcn_mbcnt_hi(
                        (uint32_t)(lb >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)lb, 0u));
                    s_lpre[li] = poff;
                    s_lplo[li] = plo;
                }
                __syncthreads();
                for (int base = 0; base < ptot; base += 64) {
                    const int idx = base + lane;
                    if (idx < ptot) {
                        int lo_ = 0, hi_ = nleaf; // the last leaf li with s_lpre[li] <= idx
                        while (hi_ - lo_ > 1) {
                            const int mid = (lo_ + hi_) >> 1;
                            if (s_lpre[mid] <= idx) lo_ = mid;
                            else hi_ = mid;
                        }
                        const int s = s_lplo[lo_] + (idx - s_lpre[lo_]);
                        const double4 p = a.src[s];
                        s_px[lane] = p.x;
                        s_py[lane] = p.y;
                        s_pz[lane] = p.z;
                        s_slot[lane] = s;
                        if (REV) {
                            const uint32_t j = a.orderbuf[s];
                            const double hj = (j < a.n_gas) ? a.h[j] : 0.0;
                            s_h2[lane] = (j < a.n_gas) ? hj * hj : -1.0; // -1: not gas
                        }
                    }
                    __syncthreads();
                    const int cnt = min(64, ptot - base);
                    for (int e = 0; e < cnt; ++e) {
                        const double ex = s_px[e] - qx, ey = s_py[e] - qy, ez = s_pz[e] - qz;
                        const double d2 = ex * ex + ey * ey + ez * ez;
                        const bool in = REV ? (d2 >= r2 && d2 < s_h2[e]) : (d2 < r2);
                        if (live && in) {
                            if (FILL) a.slots[out++] = (uint32_t)s_slot[e];
                            ++found;
                        }
                    }
                    __syncthreads();
                }
            }
            __threadfence_block();
            __syncthreads();
        }
        if (overflow) {
            if (lane == 0) atomicOr(a.status, 1);
            continue;
        }
        if (!FILL && live) a.count[q] = found;
    }
}

// The per-query lists of more than kFinishLds neighbours: a Shell sort in place.
__device__ void shell_sort_u32(uint32_t* s, long long n) {
    for (long long gap = n / 2; gap > 0; gap = (gap == 2) ? 1 : gap * 5 / 11) {
        for (long long i = gap; i < n; ++i) {
            const uint32_t v = s[i];
            long long j = i;
            while (j >= gap && s[j - gap] > v) {
                s[j] = s[j - gap];
                j -= gap;
            }
            s[j] = v;
        }
    }
}

// Each query's neighbours in slot (Morton) order, then as particle indices: one wavefront per
// query (a grid-stride loop over the queries). Up to 64 neighbours: a bitonic sort across the
// lanes' registers; up to kFinishLds: a bitonic sort in LDS, each lane taking every 64th element;
// longer lists (rare: a radius far above the local spacing): a Shell sort by one lane.
constexpr int kFinishLds = 1024;

__device__ __forceinline__ uint32_t bitonic_lanes(uint32_t v, int lane) {
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            const uint32_t o = (uint32_t)__shfl_xor((int)v, j);
            const bool up = ((lane & k) == 0);
            const bool lower = ((lane & j) == 0);
            // the lower lane of a pair keeps the smaller value in an ascending block
            v = (lower == up) ? min(v, o) : max(v, o);
        }
    }
    return v;
}

__global__ __launch_bounds__(64) void k_ngb_finish(int nq, const size_t* __restrict__ start,
                                                   uint32_t* __restrict__ slots,
                                                   const uint32_t* __restrict__ orderbuf,
                                                   uint32_t* __restrict__ found) {
    __shared__ uint32_t buf[kFinishLds];
    const int lane = threadIdx.x;
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const size_t b = start[q];
        const int n = (int)(start[q + 1] - b);
        uint32_t* s = slots + b;
        uint32_t* f = found + b;
        if (n <= 64) {
            uint32_t v = (lane < n) ? s[lane] : 0xFFFFFFFFu;
            v = bitonic_lanes(v, lane);
            if (lane < n) f[lane] = orderbuf[v];
        } else if (n <= kFinishLds) {
            int m = 128;
            while (m < n) m <<= 1;
            for (int i = lane; i < m; i += 64) buf[i] = (i < n) ? s[i] : 0xFFFFFFFFu;
            __syncthreads();
            for (int k = 2; k <= m; k <<= 1) {
                for (int j = k >> 1; j > 0; j >>= 1) {
                    // each pair (i, i ^ j) with i < i ^ j is handled by one lane
                    for (int t = lane; t < (m >> 1); t += 64) {
                        const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
                        const int l = i | j;
                        const uint32_t x = buf[i], y = buf[l];
                        const bool up = ((i & k) == 0);
                        if ((x > y) == up) {
                            buf[i] = y;
                            buf[l] = x;
                        }
                    }
                    __syncthreads();
                }
            }
            for (int i = lane; i < n; i += 64) f[i] = orderbuf[buf[i]];
            __syncthreads();
        } else if (lane == 0) {
            shell_sort_u32(s, n);
            for (int i = 0; i < n; ++i) f[i] = orderbuf[s[i]];
        }
    }
}



// Some lists too long for the LDS sort: every list sorted at once by a device radix sort of
// (query << 32 | slot) keys, which keeps each query's range of the CSR and does not slow down on
// long lists (lists that all fit the LDS sort are faster in k_ngb_finish; past kFinishLds it
// falls back to a one-lane Shell sort, hundreds of times slower on lists of 10^4-10^5). One wave
// per query writes its keys.
constexpr size_t kRadixFinishMax = size_t(1) << 28; // 24 B of keys and values per pair
__global__ __launch_bounds__(64) void k_ngb_keys(int nq, const size_t* __restrict__ start,
                                                 const uint32_t* __restrict__ slots,
                                                 uint64_t* __restrict__ keys,
                                                 uint32_t* __restrict__ vals) {
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const size_t b = start[q], e = start[q + 1];
        for (size_t i = b + threadIdx.x; i < e; i += 64) {
            keys[i] = ((uint64_t)q << 32) | slots[i];
            vals[i] = 0;
        }
    }
}
__global__ __launch_bounds__(256) void k_ngb_map(size_t n, const uint64_t* __restrict__ keys,
                                                 const uint32_t* __restrict__ orderbuf,
                                                 uint32_t* __restrict__ found) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        found[i] = orderbuf[(uint32_t)keys[i]];
}

__global__ void k_u32_to_size(int n, const uint32_t* __restrict__ in, size_t* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = in[i];
}

// The sum of the per-query counts in 64 bits (the 32-bit scan that makes the offsets would wrap
// past 2^32 - 1 pairs): a grid-stride sum per thread, a wave sum, one atomic per wave.
// The largest count too, which picks how the lists are sorted.
__global__ __launch_bounds__(256) void k_sum_counts(int n, const uint32_t* __restrict__ in,
                                                    unsigned long long* __restrict__ out,
                                                    uint32_t* __restrict__ max_out) {
    long long s = 0;
    uint32_t mx = 0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
        s += in[i];
        mx = max(mx, in[i]);
    }
    for (int d = 32; d >= 1; d >>= 1) {
        s += __shfl_xor(s, d);
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, d));
    }
    if ((threadIdx.x & 63) == 0) {
        if (s) atomicAdd(out, (unsigned long long)s);
        if (mx) atomicMax(max_out, mx);
    }
}

// h rounded up to float, so that the float bound stays one-sided (tree.cc:286-290).
__device__ __forceinline__ float h_as_bound(double h) {
    const float f = (float)h;
    return ((double)f < h) ? nextafterf(f, 1e30f) : f;
}

// init_hmax (tree.cc:332-338): each leaf takes the largest rounded-up h of its gas particles (an
// atomic max on the float's bits, which order like the values for h >= 0), then each level from
// the bottom takes the largest of its children's. The reference's per-particle climb with an
// early exit ends with the same subtree maxima.
__global__ __launch_bounds__(256) void k_hmax_leaf(size_t n_live, const uint32_t* __restrict__ orderbuf,
                                                   const int* __restrict__ slot_leaf,
                                                   const double* __restrict__ h, size_t n_hgas,
                                                   unsigned* __restrict__ hbits) {
    const size_t s = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= n_live) return;
    const uint32_t j = orderbuf[s];
    if (j >= n_hgas) return;
    const double hj = h[j];
    const int leaf = slot_leaf[s];
    if (!(hj > 0) || leaf < 0) return;
    atomicMax(&hbits[leaf], __float_as_uint(h_as_bound(hj)));
}

// raise_hmax (tree.cc:293-301) for the leaf of each listed particle: up the ancestors with an
// atomic max on the float bits, stopping where a node already holds as much.
__global__ __launch_bounds__(256) void k_raise_hmax(const uint32_t* __restrict__ idx, size_t n_idx,
                                                    size_t n, const int* __restrict__ leaf_of,
                                                    const int* __restrict__ parent,
                                                    const double* __restrict__ h,
                                                    unsigned* __restrict__ hbits,
                                                    int* __restrict__ status) {
    const size_t k = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n_idx) return;
    const uint32_t i = idx[k];
    if (i >= n) {
        atomicOr(status, 1);
        return;
    }
    const int leaf = leaf_of[i];
    const double hi = h[i];
    if (leaf < 0 || !(hi > 0)) return;
    const unsigned want = __float_as_uint(h_as_bound(hi));
    for (int no = leaf; no >= 0; no = parent[no])
        if (atomicMax(&hbits[no], want) >= want) break;
}

__global__ __launch_bounds__(256) void k_hmax_up(int lo, int hi, const WNode* __restrict__ wn,
                                                 float* __restrict__ hmax) {
    const int n = lo + blockIdx.x * 256 + threadIdx.x;
    if (n >= hi) return;
    const int f = wn[n].first;
    if (f < 0) return;
    float m = 0.f;
    for (int c = f; c < f + wn[n].nchild; ++c) m = fmaxf(m, hmax[c]);
    hmax[n] = m;
}

} // namespace gz

using namespace gz;

namespace {

gz_status ngb_pass(NgbArgs& a, bool fill, bool rev, int waves, hipStream_t st) {
    GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(a.queue), 2, st));
    const int grid = std::max(1, std::min(waves, a.ngroups));
    if (rev) {
        if (fill) k_ngb<true, true><<<grid, 64, 0, st>>>(a);
        else k_ngb<false, true><<<grid, 64, 0, st>>>(a);
    } else {
        if (fill) k_ngb<true, false><<<grid, 64, 0, st>>>(a);
        else k_ngb<false, false><<<grid, 64, 0, st>>>(a);
    }
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}
} // namespace

// Both searches.
static gz_status ngb_run(gz_ctx* ctx, const gz_tree* tree, const double* cx, const double* cy,
                         const double* cz, const double* radius, size_t n_q, double tau,
                         size_t* start, uint32_t* found, size_t cap, size_t* total,
                         const double* h, size_t n_gas, bool rev) {
    const char* who = rev ? "gz_tree_ngb_search_reverse" : "gz_tree_ngb_search";
    clear_error();
    if (!ctx || !tree || !total) return fail(GZ_ERR_INVALID, "%s: null argument", who);
    GZ_HIP(hipSetDevice(ctx->device));
    *total = 0;
    if (n_q == 0) return GZ_OK;
    if (!cx || !cy || !cz || !radius || !start)
        return fail(GZ_ERR_INVALID, "%s: null centre, radius or start array", who);
    if (n_q >= 0x7FFFFFFFull) return fail(GZ_ERR_INVALID, "%s: n_q too large", who);
    if (rev && n_gas > 0 && !h) return fail(GZ_ERR_INVALID, "%s: null h", who);
    if (rev && n_gas > tree->n)
        return fail(GZ_ERR_INVALID, "%s: n_gas=%zu exceeds the particle count %zu", who, n_gas,
                    tree->n);
    if (rev && tree->nnodes > 0 && !tree->hmax)
        return fail(GZ_ERR_INVALID, "%s: no node hmax (call gz_tree_init_hmax first)", who);
    hipStream_t st = ctx->stream;
    if (tree->n_live == 0 || tree->nnodes == 0) {
        GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(start), 2 * (n_q + 1), st));
        GZ_HIP(hipStreamSynchronize(st));
        return GZ_OK;
    }
    int* stack = nullptr;
    int stack_cap = 0;
    GZ_TRY(ctx_spill(ctx, &stack, &stack_cap));
    int waves = 0;
    {
        int dev = 0, ncu = 0;
        GZ_HIP(hipGetDevice(&dev));
        GZ_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        waves = std::max(1, ncu) * 16; // the spill arena holds this many slices
    }
    DevMem& s = ctx->scratch[10];
    const size_t cnt_bytes = ((n_q + 1) * sizeof(uint32_t) + 255) & ~size_t(255);
    GZ_TRY(s.reserve(cnt_bytes + 256 + scan_tmp_bytes(n_q + 1)));
    uint8_t* b = s.as<uint8_t>();
    uint32_t* count = reinterpret_cast<uint32_t*>(b);
    int* misc = reinterpret_cast<int*>(b + cnt_bytes);
    NgbArgs a{};
    a.wn = tree->wn;
    a.gcx = tree->gcx; a.gcy = tree->gcy; a.gcz = tree->gcz;
    a.size = tree->size;
    a.vmax = tree->vmax;
    a.xpad = tree->xpad;
    a.src = tree->src;
    a.orderbuf = tree->orderbuf;
    a.cx = cx; a.cy = cy; a.cz = cz; a.radius = radius;
    a.nq = (int)n_q;
    // 64 consecutive queries per wave share its traversal; with too few queries to occupy the
    // GPU that way (a small set is often scattered too), fewer per wave, down to one, so that
    // more waves walk at once
    a.qpw = 64;
    while (a.qpw > 1 && (n_q + a.qpw - 1) / a.qpw < (size_t)waves) a.qpw >>= 1;
    a.ngroups = (int)((n_q + a.qpw - 1) / a.qpw);
    a.root = tree->root;
    a.nclusters = tree->n_clusters;
    a.snode_eps = 1e-6 * tree->root_box[3];
    a.tau = tau;
    a.queue = misc;
    a.status = misc + 1;
    a.stack = stack;
    a.stack_cap = stack_cap;
    a.count = count;
    a.hmax = tree->hmax;
    a.h = h;
    a.n_gas = rev ? n_gas : 0;
    GZ_TRY(fill_zero_u32(count, n_q + 1, st));
    unsigned long long* total64 = reinterpret_cast<unsigned long long*>(misc + 4);
    uint32_t* max_count = reinterpret_cast<uint32_t*>(misc + 6);
    GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(total64), 3, st));
    GZ_TRY(ngb_pass(a, false, rev, waves, st));
    k_sum_counts<<<(unsigned)std::min<size_t>(1024, (n_q + 255) / 256), 256, 0, st>>>(
        (int)n_q, count, total64, max_count);
    GZ_HIP(hipGetLastError());
    // offsets: exclusive scan of the counts (count[n_q] = 0 gives the total at start[n_q])
    uint32_t tot32 = 0;
    {
        DevMem& o = ctx->scratch[11];
        GZ_TRY(o.reserve((n_q + 1) * sizeof(uint32_t)));
        GZ_TRY(exclusive_scan_u32(count, o.as<uint32_t>(), n_q + 1, nullptr,
                                  b + cnt_bytes + 256, st));
        k_u32_to_size<<<(unsigned)((n_q + 1 + 255) / 256), 256, 0, st>>>((int)(n_q + 1),
                                                                          o.as<uint32_t>(), start);
        GZ_HIP(hipGetLastError());
        GZ_HIP(hipMemcpyAsync(&tot32, o.as<uint32_t>() + n_q, sizeof tot32,
                              hipMemcpyDeviceToHost, st));
    }
    int h_status = 0;
    unsigned long long h_total = 0;
    uint32_t h_max = 0;
    GZ_HIP(hipMemcpyAsync(&h_max, max_count, sizeof h_max, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipMemcpyAsync(&h_status, a.status, sizeof h_status, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipMemcpyAsync(&h_total, total64, sizeof h_total, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (h_status & 1)
        return fail(GZ_ERR_INTERNAL, "%s: a search stack outgrew %d entries", who,
                    stack_cap);
    if (h_total > 0xFFFFFFFFull)
        return fail(GZ_ERR_UNSUPPORTED,
                    "%s: %llu neighbour pairs exceed the 2^32 - 1 one call can return; split the "
                    "queries", who, h_total);
    *total = tot32;
    if (!found || cap < (size_t)tot32 || tot32 == 0) return GZ_OK;
    // pass 2: the neighbours' slots, then sorted and mapped to particle indices
    DevMem& sl = ctx->scratch[11];
    GZ_TRY(sl.reserve((size_t)tot32 * sizeof(uint32_t)));
    a.start = start;
    a.slots = sl.as<uint32_t>();
    GZ_TRY(ngb_pass(a, true, rev, waves, st));
    if (h_max > (uint32_t)kFinishLds && (size_t)tot32 <= kRadixFinishMax) {
        DevMem& kb = ctx->scratch[15];
        const size_t nk = tot32;
        GZ_TRY(kb.reserve(nk * (2 * sizeof(uint64_t) + 2 * sizeof(uint32_t))));
        uint64_t* keys = kb.as<uint64_t>();
        uint64_t* keys_alt = keys + nk;
        uint32_t* vals = reinterpret_cast<uint32_t*>(keys_alt + nk);
        uint32_t* vals_alt = vals + nk;
        k_ngb_keys<<<(unsigned)std::min<size_t>(n_q, 65536), 64, 0, st>>>((int)n_q, start,
                                                                        a.slots, keys, vals);
        GZ_HIP(hipGetLastError());
        int qbits = 1;
        while (qbits < 31 && (size_t(1) << qbits) < n_q) ++qbits;
        GZ_TRY(radix_sort_pairs(keys, vals, keys_alt, vals_alt, nk, 32 + qbits, ctx->scratch[3],
                                st));
        k_ngb_map<<<(unsigned)std::min<size_t>((nk + 255) / 256, 65536), 256, 0, st>>>(
            nk, keys, tree->orderbuf, found);
        GZ_HIP(hipGetLastError());
    } else {
        k_ngb_finish<<<(unsigned)std::min<size_t>(n_q, 65536), 64, 0, st>>>(
            (int)n_q, start, a.slots, tree->orderbuf, found);
        GZ_HIP(hipGetLastError());
    }
    GZ_HIP(hipMemcpyAsync(&h_status, a.status, sizeof h_status, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (h_status & 1)
        return fail(GZ_ERR_INTERNAL, "%s: a search stack outgrew %d entries", who,
                    stack_cap);
    return GZ_OK;
}

extern "C" gz_status gz_tree_ngb_search(gz_ctx* ctx, const gz_tree* tree, const double* cx,
                                        const double* cy, const double* cz, const double* radius,
                                        size_t n_q, double tau, size_t* start, uint32_t* found,
                                        size_t cap, size_t* total) {
    return ngb_run(ctx, tree, cx, cy, cz, radius, n_q, tau, start, found, cap, total, nullptr, 0,
                   false);
}

extern "C" gz_status gz_tree_ngb_search_reverse(gz_ctx* ctx, const gz_tree* tree,
                                                const double* h, size_t n_gas, const double* cx,
                                                const double* cy, const double* cz,
                                                const double* h_q, size_t n_q, double tau,
                                                size_t* start, uint32_t* found, size_t cap,
                                                size_t* total) {
    return ngb_run(ctx, tree, cx, cy, cz, h_q, n_q, tau, start, found, cap, total, h, n_gas,
                   true);
}

extern "C" gz_status gz_tree_init_hmax(gz_ctx* ctx, gz_tree* tree, const double* h, size_t n_h,
                                       size_t n_gas) {
    clear_error();
    if (!ctx || !tree) return fail(GZ_ERR_INVALID, "gz_tree_init_hmax: null argument");
    const size_t n_hgas = std::min(n_h, n_gas);
    if (n_hgas > 0 && !h) return fail(GZ_ERR_INVALID, "gz_tree_init_hmax: null h");
    GZ_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int nn = tree->nnodes;
    if (nn <= 0) return GZ_OK;
    if (!tree->hmax) {
        void* p = nullptr;
        GZ_HIP(hipMalloc(&p, (size_t)nn * sizeof(float)));
        tree->hmax = static_cast<float*>(p);
    }
    GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(tree->hmax), (size_t)nn, st));
    if (n_hgas > 0 && tree->n_live > 0) {
        k_hmax_leaf<<<(unsigned)((tree->n_live + 255) / 256), 256, 0, st>>>(
            tree->n_live, tree->orderbuf, tree->slot_leaf, h, n_hgas,
            reinterpret_cast<unsigned*>(tree->hmax));
        GZ_HIP(hipGetLastError());
        // levels bottom-up: children of level L are in level L + 1
        const std::vector<int>& lb = tree->level_begin;
        for (int L = (int)lb.size() - 2; L >= 0; --L) {
            const int lo = lb[L], hi = lb[L + 1];
            if (hi <= lo) continue;
            k_hmax_up<<<(unsigned)((hi - lo + 255) / 256), 256, 0, st>>>(lo, hi, tree->wn,
                                                                         tree->hmax);
            GZ_HIP(hipGetLastError());
        }
    }
    GZ_HIP(hipStreamSynchronize(st));
    return GZ_OK;
}

extern "C" gz_status gz_tree_raise_hmax(gz_ctx* ctx, gz_tree* tree, const uint32_t* idx, size_t n,
                                        const double* h) {
    clear_error();
    if (!ctx || !tree) return fail(GZ_ERR_INVALID, "gz_tree_raise_hmax: null argument");
    if (n == 0) return GZ_OK;
    if (!idx || !h) return fail(GZ_ERR_INVALID, "gz_tree_raise_hmax: null array");
    if (tree->nnodes > 0 && !tree->hmax)
        return fail(GZ_ERR_INVALID, "gz_tree_raise_hmax: no node hmax (call gz_tree_init_hmax first)");
    if (tree->nnodes <= 0) return GZ_OK;
    GZ_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    DevMem& sm = ctx->scratch[8];
    GZ_TRY(sm.reserve(256));
    int* status = sm.as<int>();
    GZ_TRY(fill_zero_u32(reinterpret_cast<uint32_t*>(status), 1, st));
    k_raise_hmax<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(
        idx, n, tree->n, tree->leaf_of, tree->parent, h, reinterpret_cast<unsigned*>(tree->hmax),
        status);
    GZ_HIP(hipGetLastError());
    int hs = 0;
    GZ_HIP(hipMemcpyAsync(&hs, status, sizeof hs, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    if (hs & 1) return fail(GZ_ERR_INVALID, "gz_tree_raise_hmax: an index is outside the particles");
    return GZ_OK;
}
