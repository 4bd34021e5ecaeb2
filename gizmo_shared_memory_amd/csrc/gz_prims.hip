// Device-wide primitives used by the tree build: an exclusive scan and a stable LSD radix sort of
// (64-bit key, 32-bit value) pairs. Both are memory-bound; tiles are 2048 elements per 256-thread
// workgroup so one launch has thousands of workgroups at the sizes the tree sees (>= 1M keys).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "gz_internal.h"
#include "gz_wave.h"

namespace gz {

// ---- errors ----------------------------------------------------------------------------------
static thread_local std::string g_last_error;

void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
}

gz_status fail(gz_status s, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

gz_status hip_fail(hipError_t e, const char* what, const char* file, int line) {
    set_error("%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    return (e == hipErrorOutOfMemory) ? GZ_ERR_NOMEM : GZ_ERR_HIP;
}

const char* last_error_cstr() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

gz_status DevMem::reserve(size_t want) {
    if (want <= bytes && p) return GZ_OK;
    if (p) {
        // work already queued may still read the old buffer
        GZ_HIP(hipDeviceSynchronize());
        GZ_HIP(hipFree(p));
        p = nullptr;
        bytes = 0;
    }
    const size_t b = std::max<size_t>(want, 256);
    GZ_HIP(hipMalloc(&p, b));
    bytes = b;
    return GZ_OK;
}

void DevMem::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}

// ---- wave / block scans ----------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    return wave_incl_scan_dpp(x);
}

// Exclusive scan across a 256-thread block (4 waves). s4 is a 4-entry LDS scratch.
__device__ __forceinline__ uint32_t block256_exclusive_scan(uint32_t v, uint32_t* s4,
                                                            uint32_t& total) {
    const uint32_t inc = wave_inclusive_scan(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) s4[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const uint32_t s = s4[k];
        off += (k < w) ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

constexpr int kScanThreads = 256;
constexpr int kScanIpt = 8;
constexpr int kScanTile = kScanThreads * kScanIpt;

__global__ __launch_bounds__(kScanThreads) void k_scan_partials(const uint32_t* __restrict__ in,
                                                                size_t n,
                                                                uint32_t* __restrict__ partial) {
    __shared__ uint32_t s4[4];
    const size_t base = (size_t)blockIdx.x * kScanTile;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) {
        const size_t i = base + (size_t)j * kScanThreads + threadIdx.x;
        if (i < n) sum += in[i];
    }
    uint32_t total;
    (void)block256_exclusive_scan(sum, s4, total);
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// One workgroup of 1024 threads scans the per-tile partial sums in place, chunk by chunk.
__global__ __launch_bounds__(1024) void k_scan_partials_scan(uint32_t* __restrict__ partial,
                                                             size_t nb,
                                                             uint32_t* __restrict__ total_out) {
    __shared__ uint32_t s16[16];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (size_t base = 0; base < nb; base += 1024) {
        const size_t i = base + threadIdx.x;
        const uint32_t v = (i < nb) ? partial[i] : 0u;
        const uint32_t inc = wave_inclusive_scan(v);
        if (lane == 63) s16[w] = inc;
        __syncthreads();
        uint32_t off = 0, tot = 0;
        for (int k = 0; k < 16; ++k) {
            const uint32_t s = s16[k];
            off += (k < w) ? s : 0u;
            tot += s;
        }
        const uint32_t carry = s_carry;
        if (i < nb) partial[i] = carry + off + inc - v;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + tot;
        __syncthreads();
    }
    if (threadIdx.x == 0 && total_out) *total_out = s_carry;
}

// starts (optional): for a scan of 0/1 flags, starts[out[i]] = i wherever in[i] is set (the
// first index of each run that the flags open).
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(const uint32_t* in, uint32_t* out,
                                                             size_t n,
                                                             const uint32_t* __restrict__ partial,
                                                             int* __restrict__ starts, int vec) {
    __shared__ uint32_t s4[4];
    // blocked arrangement: thread t owns elements base + t*kScanIpt .. + kScanIpt - 1, read and
    // written as two 16-byte vectors when the tile is whole and the arrays are aligned
    const size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanIpt;
    const bool whole = vec && base + kScanIpt <= n;
    uint32_t v[kScanIpt];
    if (whole) {
        const uint4 a = reinterpret_cast<const uint4*>(in + base)[0];
        const uint4 b = reinterpret_cast<const uint4*>(in + base)[1];
        v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
        v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    } else {
#pragma unroll
        for (int j = 0; j < kScanIpt; ++j) v[j] = (base + j < n) ? in[base + j] : 0u;
    }
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) sum += v[j];
    uint32_t total;
    uint32_t run = block256_exclusive_scan(sum, s4, total) + partial[blockIdx.x];
    uint32_t o[kScanIpt];
#pragma unroll
    for (int j = 0; j < kScanIpt; ++j) {
        o[j] = run;
        if (starts && v[j] && base + j < n) starts[run] = (int)(base + j);
        run += v[j];
    }
    if (whole) {
        reinterpret_cast<uint4*>(out + base)[0] = make_uint4(o[0], o[1], o[2], o[3]);
        reinterpret_cast<uint4*>(out + base)[1] = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int j = 0; j < kScanIpt; ++j)
            if (base + j < n) out[base + j] = o[j];
    }
}

size_t scan_tmp_bytes(size_t n) {
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    return (nb + 64) * sizeof(uint32_t);
}

gz_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_dev,
                             void* tmp, hipStream_t st, int* starts) {
    if (n == 0) {
        if (total_dev) GZ_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), st));
        return GZ_OK;
    }
    const size_t nb = (n + kScanTile - 1) / kScanTile;
    uint32_t* partial = static_cast<uint32_t*>(tmp);
    k_scan_partials<<<(unsigned)nb, kScanThreads, 0, st>>>(in, n, partial);
    k_scan_partials_scan<<<1, 1024, 0, st>>>(partial, nb, total_dev);
    const int vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
    k_scan_apply<<<(unsigned)nb, kScanThreads, 0, st>>>(in, out, n, partial, starts, vec);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

__global__ __launch_bounds__(256) void k_fill_zero_u32(uint32_t* __restrict__ a, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = 0u;
}

gz_status fill_zero_u32(uint32_t* a, size_t n, hipStream_t st) {
    if (n == 0) return GZ_OK;
    k_fill_zero_u32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(a, n);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

// ---- radix sort ------------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsIpt = 8;
constexpr int kRsTile = kRsThreads * kRsIpt;
constexpr int kRsRadix = 256;
constexpr int kRsMaxPasses = 8;

// Histograms of every 8-bit digit in one read of the keys, so passes whose digit is the same for
// every key can be skipped.
__global__ __launch_bounds__(kRsThreads) void k_rs_digit_hist(const uint64_t* __restrict__ keys,
                                                              size_t n, int npasses,
                                                              uint32_t* __restrict__ ghist) {
    __shared__ uint32_t h[kRsMaxPasses * kRsRadix];
    for (int k = threadIdx.x; k < kRsMaxPasses * kRsRadix; k += kRsThreads) h[k] = 0;
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * kRsThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kRsThreads) {
        const uint64_t k = keys[i];
        for (int p = 0; p < npasses; ++p) atomicAdd(&h[p * kRsRadix + ((k >> (8 * p)) & 255u)], 1u);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < npasses * kRsRadix; k += kRsThreads)
        if (h[k]) atomicAdd(&ghist[k], h[k]);
}

// Per-tile digit counts, digit-major: hist[d * nb + b].
__global__ __launch_bounds__(kRsThreads) void k_rs_count(const uint64_t* __restrict__ keys,
                                                         size_t n, int shift, uint32_t nb,
                                                         uint32_t* __restrict__ hist) {
    __shared__ uint32_t c[kRsRadix];
    c[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * kRsTile;
#pragma unroll
    for (int j = 0; j < kRsIpt; ++j) {
        const size_t i = base + (size_t)j * kRsThreads + threadIdx.x;
        if (i < n) atomicAdd(&c[(uint32_t)(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = c[threadIdx.x];
}

// Stable scatter. Within a tile, an element's rank among equal digits is the number of equal
// digits at smaller indices: earlier rounds (s_run), earlier waves of this round (s_wcnt) and
// earlier lanes of this wave (match by 8 ballots).
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, size_t n, int shift, uint32_t nb,
    const uint32_t* __restrict__ offs) {
    __shared__ uint32_t s_off[kRsRadix];
    __shared__ uint32_t s_run[kRsRadix];
    __shared__ uint32_t s_wcnt[4][kRsRadix];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    s_off[tid] = offs[(size_t)tid * nb + blockIdx.x];
    s_run[tid] = 0;
    const uint64_t below_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t base = (size_t)blockIdx.x * kRsTile;
    for (int j = 0; j < kRsIpt; ++j) {
        const size_t i = base + (size_t)j * kRsThreads + tid;
        const bool valid = i < n;
        uint64_t key = 0;
        uint32_t val = 0, d = 0;
        if (valid) {
            key = kin[i];
            val = vin[i];
            d = (uint32_t)(key >> shift) & 255u;
        }
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const bool set = (d >> bit) & 1u;
            const uint64_t b = __ballot(set);
            peers &= set ? b : ~b;
        }
        for (int k = lane; k < kRsRadix; k += kWave) s_wcnt[w][k] = 0;
        __syncthreads();
        const uint32_t below = (uint32_t)__popcll(peers & below_mask);
        if (valid && below == 0) s_wcnt[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = s_off[d] + s_run[d] + below;
            for (int k = 0; k < w; ++k) pos += s_wcnt[k][d];
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
        s_run[tid] += s_wcnt[0][tid] + s_wcnt[1][tid] + s_wcnt[2][tid] + s_wcnt[3][tid];
        __syncthreads();
    }
}

gz_status radix_sort_pairs(uint64_t*& keys, uint32_t*& vals, uint64_t*& keys_alt,
                           uint32_t*& vals_alt, size_t n, int bits, DevMem& scratch,
                           hipStream_t st) {
    if (n <= 1) return GZ_OK;
    if (n >= 0xFFFFFFFFull) return fail(GZ_ERR_INVALID, "radix sort: %zu keys exceed 2^32-1", n);
    const int npasses = std::min(kRsMaxPasses, (bits + 7) / 8);
    const uint32_t nb = (uint32_t)((n + kRsTile - 1) / kRsTile);
    const size_t hist_n = (size_t)kRsRadix * nb;
    const size_t ghist_bytes = kRsMaxPasses * kRsRadix * sizeof(uint32_t);
    const size_t need = ghist_bytes + hist_n * sizeof(uint32_t) + scan_tmp_bytes(hist_n) + 256;
    GZ_TRY(scratch.reserve(need));
    uint8_t* base = scratch.as<uint8_t>();
    uint32_t* ghist = reinterpret_cast<uint32_t*>(base);
    uint32_t* hist = reinterpret_cast<uint32_t*>(base + ghist_bytes);
    void* scan_tmp = base + ghist_bytes + hist_n * sizeof(uint32_t);

    GZ_HIP(hipMemsetAsync(ghist, 0, ghist_bytes, st));
    const unsigned hgrid = (unsigned)std::min<size_t>((n + kRsThreads - 1) / kRsThreads, 1024);
    k_rs_digit_hist<<<hgrid, kRsThreads, 0, st>>>(keys, n, npasses, ghist);
    GZ_HIP(hipGetLastError());
    std::vector<uint32_t> h(kRsMaxPasses * kRsRadix);
    GZ_HIP(hipMemcpyAsync(h.data(), ghist, ghist_bytes, hipMemcpyDeviceToHost, st));
    GZ_HIP(hipStreamSynchronize(st));

    for (int p = 0; p < npasses; ++p) {
        const uint32_t* hp = h.data() + p * kRsRadix;
        if (*std::max_element(hp, hp + kRsRadix) == (uint32_t)n) continue; // one digit: no-op
        const int shift = 8 * p;
        k_rs_count<<<nb, kRsThreads, 0, st>>>(keys, n, shift, nb, hist);
        GZ_HIP(hipGetLastError());
        GZ_TRY(exclusive_scan_u32(hist, hist, hist_n, nullptr, scan_tmp, st));
        k_rs_scatter<<<nb, kRsThreads, 0, st>>>(keys, vals, keys_alt, vals_alt, n, shift, nb,
                                                hist);
        GZ_HIP(hipGetLastError());
        std::swap(keys, keys_alt);
        std::swap(vals, vals_alt);
    }
    return GZ_OK;
}

} // namespace gz
