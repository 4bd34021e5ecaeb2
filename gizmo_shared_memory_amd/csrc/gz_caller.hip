// The caller side of a force evaluation, on the device: active targets in Morton order, the
// relative-criterion aold, the scatter of accelerations into the per-particle array and the
// gravity part of the KDK kick. Follows compute_gravity (shmem/mfm.cc:698-751, 828-831) and the
// kick (mfm.cc:3749-3767).
#include <algorithm>
#include <cmath>
#include <mutex>

#include "gz_internal.h"

namespace gz {

// ---- fixed scratch for gz_order_targets_noalloc ------------------------------------------------
// The flags, their exclusive scan and the scan's block partials for up to kNoAllocSlots live
// particles, in a file-scope device array: every context on a GPU shares that GPU's copy, which
// keeps the last call's contents. Host code never passes the array's name to a kernel (on the host
// it is only a stand-in for the symbol): it uses the address hipGetSymbolAddress returns for the
// current device, resolved once per device when a context is created there.
constexpr size_t kNoAllocSlots = size_t(1) << 23;
constexpr size_t kNoAllocFlagBytes = ((kNoAllocSlots + 64) * sizeof(uint32_t) + 255) & ~size_t(255);
constexpr size_t kNoAllocScanBytes = size_t(1) << 20;
constexpr size_t kNoAllocBytes = 2 * kNoAllocFlagBytes + kNoAllocScanBytes;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

static std::mutex g_scratch_mu;
static uint8_t* g_scratch_ptr[kNoAllocMaxDevices] = {};

gz_status resolve_noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return GZ_OK; // the entry point refuses it
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return GZ_OK;
    void* p = nullptr;
    GZ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(GZ_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
    return GZ_OK;
}

static uint8_t* noalloc_scratch(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

constexpr int kT = 256;
static inline unsigned blocks(size_t n) { return (unsigned)((n + kT - 1) / kT); }

__global__ __launch_bounds__(kT) void k_mark_active(const uint32_t* __restrict__ active, size_t na,
                                                    const uint32_t* __restrict__ rank, size_t n,
                                                    uint32_t* __restrict__ flag) {
    const size_t k = (size_t)blockIdx.x * kT + threadIdx.x;
    if (k >= na) return;
    const uint32_t i = active[k];
    if (i >= n) return;
    const uint32_t r = rank[i];
    if (r != 0xFFFFFFFFu) flag[r] = 1u;
}

__global__ __launch_bounds__(kT) void k_compact_active(const uint32_t* __restrict__ orderbuf,
                                                       size_t nl, const uint32_t* __restrict__ flag,
                                                       const uint32_t* __restrict__ pos,
                                                       uint32_t* __restrict__ out) {
    const size_t r = (size_t)blockIdx.x * kT + threadIdx.x;
    if (r < nl && flag[r]) out[pos[r]] = orderbuf[r];
}

__global__ __launch_bounds__(kT) void k_aold(const double* __restrict__ a_prev,
                                             const uint32_t* __restrict__ targets, size_t nt,
                                             double tol, double G, double* __restrict__ aold) {
    const size_t k = (size_t)blockIdx.x * kT + threadIdx.x;
    if (k >= nt) return;
    const size_t i = targets[k];
    const double ax = a_prev[3 * i], ay = a_prev[3 * i + 1], az = a_prev[3 * i + 2];
    aold[k] = tol * sqrt(ax * ax + ay * ay + az * az) / G; // mfm.cc:750
}

__global__ __launch_bounds__(kT) void k_scatter(const uint32_t* __restrict__ targets, size_t nt,
                                                const double* __restrict__ ax,
                                                const double* __restrict__ ay,
                                                const double* __restrict__ az,
                                                double* __restrict__ a_grav) {
    const size_t k = (size_t)blockIdx.x * kT + threadIdx.x;
    if (k >= nt) return;
    const size_t i = targets[k];
    a_grav[3 * i] = ax[k];
    a_grav[3 * i + 1] = ay[k];
    a_grav[3 * i + 2] = az[k];
}

// Close the previous step's half kick, open the next one (mfm.cc:3716-3730, 3753-3766). Every
// array is indexed by particle, dt included (the reference's dt_of[i], mfm.cc:3753, :3764).
__global__ __launch_bounds__(kT) void k_kick(const uint32_t* __restrict__ active, size_t na,
                                             const double* __restrict__ a_grav,
                                             const double* __restrict__ a_extra,
                                             const double* __restrict__ dt,
                                             double* __restrict__ pending, double* __restrict__ vx,
                                             double* __restrict__ vy, double* __restrict__ vz) {
    const size_t k = (size_t)blockIdx.x * kT + threadIdx.x;
    if (k >= na) return;
    const size_t i = active[k];
    double a0 = a_grav[3 * i], a1 = a_grav[3 * i + 1], a2 = a_grav[3 * i + 2];
    if (a_extra) {
        a0 += a_extra[3 * i];
        a1 += a_extra[3 * i + 1];
        a2 += a_extra[3 * i + 2];
    }
    const double close_dt = pending[i];
    const double open_dt = 0.5 * dt[i];
    double v0 = vx[i], v1 = vy[i], v2 = vz[i];
    v0 += a0 * close_dt; v1 += a1 * close_dt; v2 += a2 * close_dt;
    v0 += a0 * open_dt;  v1 += a1 * open_dt;  v2 += a2 * open_dt;
    vx[i] = v0; vy[i] = v1; vz[i] = v2;
    pending[i] = open_dt;
}

// The drift (drift_particle_to, mfm.cc:1287-1295, gravity part, no periodic fold): x += v dt for
// every particle with m > 0; a tombstone stays parked where it is.
__global__ __launch_bounds__(kT) void k_drift(size_t n, double dt, const double* __restrict__ vx,
                                              const double* __restrict__ vy,
                                              const double* __restrict__ vz,
                                              const double* __restrict__ m,
                                              double* __restrict__ x, double* __restrict__ y,
                                              double* __restrict__ z) {
    const size_t i = (size_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n || !(m[i] > 0.0)) return;
    x[i] = x[i] + vx[i] * dt;
    y[i] = y[i] + vy[i] * dt;
    z[i] = z[i] + vz[i] * dt;
}

} // namespace gz

using gz::fail;

extern "C" {

// The subset case of target ordering: flag the active slots of the Morton order and compact
// orderbuf by the flags, the reference's filtering scan (mfm.cc:710-717); particles outside the
// tree are dropped there. b holds 2 * fb + scan_tmp_bytes(nl) bytes.
static gz_status order_subset(gz_ctx* ctx, const gz_tree* tree, const uint32_t* active,
                              size_t n_active, uint32_t* out_targets, size_t* n_out, uint8_t* b,
                              size_t fb) {
    hipStream_t st = ctx->stream;
    const size_t nl = tree->n_live;
    uint32_t* flag = reinterpret_cast<uint32_t*>(b);
    uint32_t* pos = reinterpret_cast<uint32_t*>(b + fb);
    uint32_t* total = pos + nl;
    void* tmp = b + 2 * fb;
    GZ_TRY(gz::fill_zero_u32(flag, nl, st));
    if (n_active)
        gz::k_mark_active<<<gz::blocks(n_active), gz::kT, 0, st>>>(active, n_active, tree->rank,
                                                                   tree->n, flag);
    GZ_HIP(hipGetLastError());
    GZ_TRY(gz::exclusive_scan_u32(flag, pos, nl, total, tmp, st));
    gz::k_compact_active<<<gz::blocks(nl), gz::kT, 0, st>>>(tree->orderbuf, nl, flag, pos,
                                                            out_targets);
    GZ_HIP(hipGetLastError());
    uint32_t h = 0;
    GZ_HIP(hipMemcpyAsync(&h, total, sizeof h, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));
    *n_out = h;
    return GZ_OK;
}

// Everything active: the Morton order is orderbuf (mfm.cc:700-702).
static gz_status order_all(gz_ctx* ctx, const gz_tree* tree, uint32_t* out_targets,
                           size_t* n_out) {
    const size_t nl = tree->n_live;
    GZ_HIP(hipMemcpyAsync(out_targets, tree->orderbuf, nl * sizeof(uint32_t),
                          hipMemcpyDeviceToDevice, ctx->stream));
    GZ_HIP(hipStreamSynchronize(ctx->stream));
    *n_out = nl;
    return GZ_OK;
}

gz_status gz_order_targets(gz_ctx* ctx, const gz_tree* tree, const uint32_t* active,
                           size_t n_active, uint32_t* out_targets, size_t* n_out) {
    gz::clear_error();
    if (!ctx || !tree || !n_out) return fail(GZ_ERR_INVALID, "gz_order_targets: null argument");
    GZ_HIP(hipSetDevice(ctx->device));
    *n_out = 0;
    const size_t nl = tree->n_live;
    if (nl == 0) return GZ_OK;
    if (!out_targets) return fail(GZ_ERR_INVALID, "gz_order_targets: null output");
    if (!active || n_active >= tree->n) return order_all(ctx, tree, out_targets, n_out);
    gz::DevMem& s = ctx->scratch[7];
    const size_t fb = ((nl + 64) * sizeof(uint32_t) + 255) & ~size_t(255);
    GZ_TRY(s.reserve(2 * fb + gz::scan_tmp_bytes(nl) + 256));
    return order_subset(ctx, tree, active, n_active, out_targets, n_out, s.as<uint8_t>(), fb);
}

gz_status gz_order_targets_noalloc(gz_ctx* ctx, const gz_tree* tree, const uint32_t* active,
                                   size_t n_active, uint32_t* out_targets, size_t* n_out) {
    gz::clear_error();
    if (!ctx || !tree || !n_out)
        return fail(GZ_ERR_INVALID, "gz_order_targets_noalloc: null argument");
    *n_out = 0;
    const size_t nl = tree->n_live;
    if (nl > gz::kNoAllocSlots)
        return fail(GZ_ERR_INVALID,
                    "gz_order_targets_noalloc: %zu particles in the tree exceed its fixed scratch "
                    "of %zu; use gz_order_targets",
                    nl, gz::kNoAllocSlots);
    GZ_HIP(hipSetDevice(ctx->device));
    if (nl == 0) return GZ_OK;
    if (!out_targets) return fail(GZ_ERR_INVALID, "gz_order_targets_noalloc: null output");
    if (!active || n_active >= tree->n) return order_all(ctx, tree, out_targets, n_out);
    uint8_t* b = gz::noalloc_scratch(ctx->device);
    if (!b)
        return fail(GZ_ERR_HIP, "gz_order_targets_noalloc: the fixed scratch of device %d was not "
                                "resolved when its context was created", ctx->device);
    const size_t fb = ((nl + 64) * sizeof(uint32_t) + 255) & ~size_t(255);
    if (2 * fb > 2 * gz::kNoAllocFlagBytes || gz::scan_tmp_bytes(nl) > gz::kNoAllocScanBytes)
        return fail(GZ_ERR_INTERNAL, "gz_order_targets_noalloc: scratch layout exceeded");
    return order_subset(ctx, tree, active, n_active, out_targets, n_out, b, fb);
}

gz_status gz_aold_from_accel(gz_ctx* ctx, const double* a_prev, const uint32_t* targets,
                             size_t n_t, double err_tol_force_acc, double G, double* aold) {
    gz::clear_error();
    if (!ctx) return fail(GZ_ERR_INVALID, "gz_aold_from_accel: null context");
    if (n_t && (!a_prev || !targets || !aold))
        return fail(GZ_ERR_INVALID, "gz_aold_from_accel: null array");
    if (!(G != 0.0)) return fail(GZ_ERR_INVALID, "gz_aold_from_accel: G must be non-zero");
    GZ_HIP(hipSetDevice(ctx->device));
    if (n_t)
        gz::k_aold<<<gz::blocks(n_t), gz::kT, 0, ctx->stream>>>(a_prev, targets, n_t,
                                                               err_tol_force_acc, G, aold);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status gz_scatter_accel(gz_ctx* ctx, const uint32_t* targets, size_t n_t, const double* ax,
                           const double* ay, const double* az, double* a_grav) {
    gz::clear_error();
    if (!ctx) return fail(GZ_ERR_INVALID, "gz_scatter_accel: null context");
    if (n_t && (!targets || !ax || !ay || !az || !a_grav))
        return fail(GZ_ERR_INVALID, "gz_scatter_accel: null array");
    GZ_HIP(hipSetDevice(ctx->device));
    if (n_t)
        gz::k_scatter<<<gz::blocks(n_t), gz::kT, 0, ctx->stream>>>(targets, n_t, ax, ay, az,
                                                                  a_grav);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status gz_drift(gz_ctx* ctx, size_t n, double dt, const double* vx, const double* vy,
                   const double* vz, const double* m, double* x, double* y, double* z) {
    gz::clear_error();
    if (!ctx) return fail(GZ_ERR_INVALID, "gz_drift: null context");
    if (n && (!vx || !vy || !vz || !m || !x || !y || !z))
        return fail(GZ_ERR_INVALID, "gz_drift: null array");
    if (!std::isfinite(dt)) return fail(GZ_ERR_INVALID, "gz_drift: dt is not finite");
    GZ_HIP(hipSetDevice(ctx->device));
    if (n)
        gz::k_drift<<<gz::blocks(n), gz::kT, 0, ctx->stream>>>(n, dt, vx, vy, vz, m, x, y, z);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status gz_kick(gz_ctx* ctx, const uint32_t* active, size_t n_active, const double* a_grav,
                  const double* a_extra, const double* dt, double* pending_half_kick, double* vx,
                  double* vy, double* vz) {
    gz::clear_error();
    if (!ctx) return fail(GZ_ERR_INVALID, "gz_kick: null context");
    if (n_active && (!active || !a_grav || !dt || !pending_half_kick || !vx || !vy || !vz))
        return fail(GZ_ERR_INVALID, "gz_kick: null array");
    GZ_HIP(hipSetDevice(ctx->device));
    if (n_active)
        gz::k_kick<<<gz::blocks(n_active), gz::kT, 0, ctx->stream>>>(
            active, n_active, a_grav, a_extra, dt, pending_half_kick, vx, vy, vz);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // extern "C"
