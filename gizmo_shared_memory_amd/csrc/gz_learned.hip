// The learned opening criterion: per-group features, a 16 -> 64 -> 64 -> 1 GELU network with bf16
// weights and FP32 accumulation (inference, backpropagation and Adam), and the per-group opening
// parameter with its fallback to the reference criterion.
//
// The network maps a group's features to log10 of the relative-criterion tolerance
// (ErrTolForceAcc, shmem/mfm.cc:743-751; open_node's third test, tree.h:134-135) that keeps the
// group's force error at the training target. The walk then opens node X for the group when
// M_X L_X^2 > r^4 * (alpha_g / alpha_ref) * aold_min, i.e. the reference test with the group's own
// tolerance. A group whose features fall outside the ranges seen in training, or whose predicted
// tolerance falls outside the trained label range, keeps the reference tolerance.
#include <cmath>

#include "gz_features.h"

namespace gz {

constexpr int kIn = 16, kH = 64;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH;
static_assert(kOffB3 + 1 == kMlpParams, "parameter layout");

__device__ __forceinline__ float bf16_to_f32(uint16_t b) {
    return __uint_as_float((uint32_t)b << 16);
}
__device__ __forceinline__ uint16_t f32_to_bf16(float f) { // round to nearest even
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float gelu(float x) {
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) +
           x * expf(-0.5f * x * x) * 0.39894228040143268f;
}

// ---- group features ------------------------------------------------------------------------------
// amag is |a_prev| per target slot (no-G units, the basis of aold). The sums over a group's
// targets are taken by LPG consecutive lanes per group (LPG = 16, 32 or 64: groups share a
// wavefront, so its lanes are busy and its load latencies overlap); lanes read consecutive
// targets, so the reads coalesce. The features and the network, arithmetic on a
// handful of scalars, run one thread per group.

struct FeatureInputs {
    int ngroups, nt_total;
    const uint32_t* ngroups_dev; // the exact count when ngroups is an upper bound, or NULL
    const int* gstart;
    const uint32_t* targets;
    const int* tslot;        // sorted slot per target (data in src), -1 if not in the tree
    const double4* src;
    const double* src_soft;
    const double *x, *y, *z, *m, *soft, *amag;
    const int* slot_leaf;
    const int* level_begin;
    const WNode* wn;
    int root;
    const uint32_t* cstart;
    int nc;
};

// Reductions over aligned runs of W lanes (W a power of two up to 64): min and max are exact, the
// sum is the same doubling tree in every run.
template <int W> __device__ __forceinline__ double sub_sum_d(double v) {
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}
template <int W> __device__ __forceinline__ double sub_min_d(double v) {
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) {
        const double o = __shfl_xor(v, d);
        v = (o < v) ? o : v;
    }
    return v;
}
template <int W> __device__ __forceinline__ double sub_max_d(double v) {
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) {
        const double o = __shfl_xor(v, d);
        v = (v < o) ? o : v;
    }
    return v;
}
template <int W> __device__ __forceinline__ int sub_sum_i(int v) {
#pragma unroll
    for (int d = W / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

template <int LPG>
__global__ __launch_bounds__(256) void k_group_stats(FeatureInputs in,
                                                     GroupStats* __restrict__ out) {
    static_assert(LPG == 16 || LPG == 32 || LPG == 64, "lanes per group");
    const int sl_lane = threadIdx.x & (LPG - 1);
    const int g = (int)(((size_t)blockIdx.x * 256 + threadIdx.x) / LPG);
    const int ngroups = in.ngroups_dev ? (int)*in.ngroups_dev : in.ngroups;
    // a whole run of LPG lanes shares g, so a run is either all in range or all out
    if (g >= ngroups) return;
    const int tlo = in.gstart[g];
    const int thi = (g + 1 < ngroups) ? in.gstart[g + 1] : in.nt_total;
    double msum = 0, amin = 1e300, amax = 0, a2 = 0, emax = 0;
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (int i = tlo + sl_lane; i < thi; i += LPG) {
        const int sl = in.tslot[i];
        double px, py, pz, pm, pe;
        if (sl >= 0) { // the tree's sorted copy: consecutive targets, nearby slots
            const double4 q = in.src[sl];
            px = q.x; py = q.y; pz = q.z; pm = q.w;
            pe = in.src_soft[sl];
        } else {
            const uint32_t p = in.targets[i];
            px = in.x[p]; py = in.y[p]; pz = in.z[p]; pm = in.m[p];
            pe = in.soft ? in.soft[p] : 0.0;
        }
        msum += pm;
        const double am = in.amag ? in.amag[i] : 0.0;
        amin = fmin(amin, am);
        amax = fmax(amax, am);
        a2 += am * am;
        emax = fmax(emax, pe);
        lo0 = fmin(lo0, px); hi0 = fmax(hi0, px);
        lo1 = fmin(lo1, py); hi1 = fmax(hi1, py);
        lo2 = fmin(lo2, pz); hi2 = fmax(hi2, pz);
    }
    GroupStats S;
    S.lo[0] = sub_min_d<LPG>(lo0); S.lo[1] = sub_min_d<LPG>(lo1); S.lo[2] = sub_min_d<LPG>(lo2);
    S.hi[0] = sub_max_d<LPG>(hi0); S.hi[1] = sub_max_d<LPG>(hi1); S.hi[2] = sub_max_d<LPG>(hi2);
    S.emax = sub_max_d<LPG>(emax);
    S.msum = sub_sum_d<LPG>(msum);
    S.amin = sub_min_d<LPG>(amin);
    S.amax = sub_max_d<LPG>(amax);
    S.a2 = sub_sum_d<LPG>(a2);
    S.nt = thi - tlo;
    S.root = root_of(in.targets[tlo], in.cstart, in.nc, in.root);
    const int leaf = (in.tslot[tlo] >= 0) ? in.slot_leaf[in.tslot[tlo]] : -1;
    S.leaf_len = (leaf >= 0) ? (double)in.wn[leaf].len : -1.0;
    // node ids are level-major: the leaf's level is the number of level starts at or below it
    int below = 0;
    if (leaf >= 0)
        for (int l = sl_lane; l < kMaxLevel; l += LPG) below += (in.level_begin[l + 1] <= leaf);
    S.depth = sub_sum_i<LPG>(below);
    if (sl_lane == 0) out[g] = S;
}

__device__ __forceinline__ void features_of(const GroupStats& S0, const WNode* __restrict__ wn,
                                            float* f) {
    GroupStats S = S0;
    const WNode Rn = wn[S.root];
    const RootInfo R{Rn.cx, Rn.cy, Rn.cz, (double)Rn.len, Rn.mass};
    if (S.leaf_len < 0) S.leaf_len = fmax(R.side, 1e-300); // a target outside the tree
    group_features16(S, R, f);
}

__global__ __launch_bounds__(256) void k_group_features(const GroupStats* __restrict__ stats,
                                                        int ngroups, const WNode* __restrict__ wn,
                                                        float* __restrict__ feats) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ngroups) return;
    float f[kIn];
    features_of(stats[g], wn, f);
    float* o = feats + (size_t)g * kIn;
#pragma unroll
    for (int k = 0; k < kIn; ++k) o[k] = f[k];
}

// ---- inference -------------------------------------------------------------------------------------
// Weight matrices from their bf16 copy, biases from the FP32 masters, and the input map
// x' = (x - mean) * scale (xnorm = mean[16], scale[16]).
__device__ __forceinline__ bool is_bias(int k) {
    return (k >= kOffB1 && k < kOffW2) || (k >= kOffB2 && k < kOffW3) || k == kOffB3;
}
__device__ __forceinline__ void stage_weights(const uint16_t* __restrict__ wbf,
                                             const float* __restrict__ theta,
                                             const float* __restrict__ xnorm, float* sW,
                                             float* sN) {
    for (int k = threadIdx.x; k < kMlpParams; k += blockDim.x)
        sW[k] = is_bias(k) ? theta[k] : bf16_to_f32(wbf[k]);
    for (int k = threadIdx.x; k < 2 * kIn; k += blockDim.x) sN[k] = xnorm[k];
    __syncthreads();
}

// The network on one sample, weights read from LDS (the same address in every lane: broadcasts).
__device__ __forceinline__ float mlp_forward_one(const float* raw, const float* sW,
                                                const float* sN) {
    float xin[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) xin[k] = (raw[k] - sN[k]) * sN[kIn + k];
    // h1 stays in registers: every index into it is a compile-time constant
    float h1[kH];
#pragma unroll
    for (int i = 0; i < kH; ++i) {
        float zz = sW[kOffB1 + i];
#pragma unroll
        for (int k = 0; k < kIn; ++k) zz = __builtin_fmaf(sW[kOffW1 + i * kIn + k], xin[k], zz);
        h1[i] = gelu(zz);
    }
    float yy = sW[kOffB3];
    for (int i = 0; i < kH; ++i) {
        float zz = sW[kOffB2 + i];
#pragma unroll
        for (int j = 0; j < kH; ++j) zz = __builtin_fmaf(sW[kOffW2 + i * kH + j], h1[j], zz);
        yy = __builtin_fmaf(sW[kOffW3 + i], gelu(zz), yy);
    }
    return yy;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, int B,
                                                     const uint16_t* __restrict__ wbf,
                                                     const float* __restrict__ theta,
                                                     const float* __restrict__ xnorm,
                                                     float* __restrict__ out) {
    __shared__ float sW[kMlpParams];
    __shared__ float sN[2 * kIn];
    stage_weights(wbf, theta, xnorm, sW, sN);
    const int b = blockIdx.x * 256 + threadIdx.x;
    if (b >= B) return;
    float raw[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) raw[k] = X[(size_t)b * kIn + k];
    out[b] = mlp_forward_one(raw, sW, sN);
}

// ---- the per-group tolerance -------------------------------------------------------------------
// scale[g] = alpha_g / alpha_ref = 10^(pred - margin - log10 alpha_ref), the factor the walk applies
// to the group's aold minimum, or 1 (the reference tolerance) when a feature lies outside the range
// seen in training or the prediction outside the trained label range. Fallbacks are counted.
struct ToleranceModel {
    const uint16_t* wbf;
    const float* theta;
    const float* xnorm;
    const float* frange; // lo[16], hi[16]
    float lab_lo, lab_hi, margin, log_alpha_ref;
};

// One wavefront per batch of 32 groups; the two hidden layers are matrix products on the matrix
// cores (v_mfma_f32_32x32x2f32: FP32 operands, the bf16-rounded weights exactly representable).
// Operand layouts of that instruction, for lane l: A[l % 32][l / 32] (32 x 2), B[l / 32][l % 32]
// (2 x 32), and result register v holds D[8 (v / 4) + 4 (l / 32) + v % 4][l % 32] (32 x 32).
// Layer 1 is X[32 x 16] W1^T in 8 steps of K = 2 per block of 32 hidden units; layer 2 is
// H1[32 x 64] W2^T in 32 steps, its A operand read back from LDS (the result layout is not the
// operand layout); biases start the accumulators. The output unit is a dot product per group,
// summed in hidden-unit order by the group's lane. Each wave keeps its B operands (80 floats per
// lane) in registers and walks batches with a grid stride.
typedef float v16f __attribute__((ext_vector_type(16)));
constexpr int kScaleBatch = 32;

__global__ __launch_bounds__(64) void k_group_scale(const GroupStats* __restrict__ stats,
                                                    int ngroups_ub,
                                                    const uint32_t* __restrict__ ngroups_dev,
                                                    const WNode* __restrict__ wn,
                                                    ToleranceModel md, float* __restrict__ scale,
                                                    int* __restrict__ n_fallback) {
    const int ngroups = ngroups_dev ? (int)*ngroups_dev : ngroups_ub;
    static_assert(kIn == 16 && kH == 64, "the tiling assumes 16 inputs and 64 hidden units");
    __shared__ float sH[kScaleBatch][kH + 1];
    __shared__ float sW3[kH];
    const int l = threadIdx.x;
    const int col = l & 31, half = l >> 5;
    auto wmat = [&](int k) { return bf16_to_f32(md.wbf[k]); };
    // B operands: B[kk][jj] = W[32 jb + jj][2 ks + kk]. A lane reads its weight rows whole, in
    // 16-byte pieces (rows start on 16-byte boundaries), and keeps every other bf16: the low
    // halves of the 32-bit words in the first half-wave, the high halves in the second.
    auto pick = [&](uint32_t w) { return __uint_as_float(half ? (w & 0xFFFF0000u) : (w << 16)); };
    float w1b[2][8], w2b[2][32];
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const uint4* r1 = reinterpret_cast<const uint4*>(md.wbf + kOffW1 + (32 * jb + col) * kIn);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint4 w = r1[q];
            w1b[jb][4 * q] = pick(w.x); w1b[jb][4 * q + 1] = pick(w.y);
            w1b[jb][4 * q + 2] = pick(w.z); w1b[jb][4 * q + 3] = pick(w.w);
        }
        const uint4* r2 = reinterpret_cast<const uint4*>(md.wbf + kOffW2 + (32 * jb + col) * kH);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const uint4 w = r2[q];
            w2b[jb][4 * q] = pick(w.x); w2b[jb][4 * q + 1] = pick(w.y);
            w2b[jb][4 * q + 2] = pick(w.z); w2b[jb][4 * q + 3] = pick(w.w);
        }
    }
    const float b1[2] = {md.theta[kOffB1 + col], md.theta[kOffB1 + 32 + col]};
    const float b2[2] = {md.theta[kOffB2 + col], md.theta[kOffB2 + 32 + col]};
    const float b3 = md.theta[kOffB3];
    sW3[l] = wmat(kOffW3 + l); // the output unit's weights, read by every group's lane

    const int nbatch = (ngroups + kScaleBatch - 1) / kScaleBatch;
    for (int bt = blockIdx.x; bt < nbatch; bt += gridDim.x) {
        // features of group col of the batch (both halves of the wave hold them)
        const int g = bt * kScaleBatch + col;
        const bool valid = g < ngroups;
        float xin[kIn];
        bool ok = valid;
        if (valid) {
            float f[kIn];
            features_of(stats[g], wn, f);
#pragma unroll
            for (int k = 0; k < kIn; ++k) {
                ok = ok && (f[k] >= md.frange[k]) && (f[k] <= md.frange[kIn + k]);
                xin[k] = (f[k] - md.xnorm[k]) * md.xnorm[kIn + k];
            }
        } else {
#pragma unroll
            for (int k = 0; k < kIn; ++k) xin[k] = 0.f;
        }
        // layer 1
        v16f acc[2];
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[jb][v] = b1[jb];
#pragma unroll
            for (int ks = 0; ks < 8; ++ks) {
                const float a = half ? xin[2 * ks + 1] : xin[2 * ks];
                acc[jb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w1b[jb][ks], acc[jb], 0, 0, 0);
            }
        }
        // GELU, then to LDS in row (group) order for the next product's A operand
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                sH[8 * (v >> 2) + 4 * half + (v & 3)][32 * jb + col] = gelu(acc[jb][v]);
        __syncthreads();
        // layer 2
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int v = 0; v < 16; ++v) acc[jb][v] = b2[jb];
#pragma unroll
        for (int ks = 0; ks < 32; ++ks) {
            const float a = sH[col][2 * ks + half];
            acc[0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2b[0][ks], acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, w2b[1][ks], acc[1], 0, 0, 0);
        }
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 2; ++jb)
#pragma unroll
            for (int v = 0; v < 16; ++v)
                sH[8 * (v >> 2) + 4 * half + (v & 3)][32 * jb + col] = gelu(acc[jb][v]);
        __syncthreads();
        // output unit: group col's dot product, in hidden-unit order
        bool fell_back = false;
        if (half == 0 && valid) {
            float yy = b3;
            for (int j = 0; j < kH; ++j) yy = __builtin_fmaf(sW3[j], sH[col][j], yy);
            float sc = 1.0f;
            ok = ok && (yy >= md.lab_lo) && (yy <= md.lab_hi);
            if (ok) sc = exp10f(yy - md.margin - md.log_alpha_ref);
            scale[g] = sc;
            fell_back = !ok;
        }
        const uint64_t fb = __ballot(fell_back);
        if (fb && l == 0) atomicAdd(n_fallback, (int)__popcll(fb));
        __syncthreads(); // sH is rewritten by the next batch
    }
}

// ---- backpropagation -------------------------------------------------------------------------------
// One wavefront per block of 64 samples. Gradients of the loss are summed over the block in LDS
// and added to grad[] with one atomic per parameter per block. The loss is the mean squared error
// L = sum_b (y_b - t_b)^2 / B (tau <= 0), or the pinball loss of the tau-quantile,
// L = sum_b (t_b - y_b) (tau - [t_b < y_b]) / B (0 < tau < 1), which makes y the tau-quantile of
// the label given the features.
constexpr int kGB = 64;
__global__ __launch_bounds__(64) void k_mlp_grad(const float* __restrict__ X,
                                                 const float* __restrict__ T, int B, float inv_b,
                                                 const uint16_t* __restrict__ wbf,
                                                 const float* __restrict__ theta,
                                                 const float* __restrict__ xnorm, float tau,
                                                 float* __restrict__ grad,
                                                 float* __restrict__ loss) {
    __shared__ float sW[kMlpParams];
    __shared__ float sN[2 * kIn];
    __shared__ float sH1[kGB][kH + 1]; // h1, later dz1
    __shared__ float sZ2[kGB][kH + 1]; // z2, later dz2
    __shared__ float sX[kGB][kIn + 1];
    __shared__ float sDy[kGB];
    stage_weights(wbf, theta, xnorm, sW, sN);
    const int tid = threadIdx.x;
    const int b = blockIdx.x * kGB + tid;
    const bool valid = b < B;
    // forward, one sample per lane
    float xin[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) {
        xin[k] = valid ? (X[(size_t)b * kIn + k] - sN[k]) * sN[kIn + k] : 0.f;
        sX[tid][k] = xin[k];
    }
    float h1[kH];
#pragma unroll 4
    for (int i = 0; i < kH; ++i) {
        float zz = sW[kOffB1 + i];
#pragma unroll
        for (int k = 0; k < kIn; ++k) zz = __builtin_fmaf(sW[kOffW1 + i * kIn + k], xin[k], zz);
        h1[i] = gelu(zz);
        sH1[tid][i] = h1[i];
    }
    float yy = sW[kOffB3];
    for (int i = 0; i < kH; ++i) {
        float zz = sW[kOffB2 + i];
#pragma unroll
        for (int j = 0; j < kH; ++j) zz = __builtin_fmaf(sW[kOffW2 + i * kH + j], h1[j], zz);
        sZ2[tid][i] = zz;
        yy = __builtin_fmaf(sW[kOffW3 + i], gelu(zz), yy);
    }
    const float r = valid ? (yy - T[b]) : 0.f;
    float dy, l2;
    if (tau > 0.f) {
        dy = valid ? ((r > 0.f ? 1.0f : 0.0f) - tau) * inv_b : 0.f;
        l2 = (r > 0.f) ? (1.0f - tau) * r : -tau * r;
    } else {
        dy = 2.0f * r * inv_b;
        l2 = r * r;
    }
    sDy[tid] = dy;
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) l2 += __shfl_xor(l2, d);
    if (tid == 0) atomicAdd(loss, l2 * inv_b);
    __syncthreads();
    // dW3, db3 (lane = hidden unit)
    {
        float g3 = 0.f, gb = 0.f;
        for (int s = 0; s < kGB; ++s) {
            g3 = __builtin_fmaf(sDy[s], gelu(sZ2[s][tid]), g3);
            gb += sDy[s];
        }
        atomicAdd(&grad[kOffW3 + tid], g3);
        if (tid == 0) atomicAdd(&grad[kOffB3], gb);
    }
    __syncthreads();
    // dz2 = dy * w3 * gelu'(z2), row per lane
    for (int i = 0; i < kH; ++i) sZ2[tid][i] = dy * sW[kOffW3 + i] * gelu_grad(sZ2[tid][i]);
    __syncthreads();
    // dW2[i][j] = sum_s dz2[s][i] h1[s][j], db2 (lane = i)
    {
        float acc[kH];
#pragma unroll
        for (int j = 0; j < kH; ++j) acc[j] = 0.f;
        float gb = 0.f;
        for (int s = 0; s < kGB; ++s) {
            const float dz = sZ2[s][tid];
            gb += dz;
#pragma unroll
            for (int j = 0; j < kH; ++j) acc[j] = __builtin_fmaf(dz, sH1[s][j], acc[j]);
        }
#pragma unroll
        for (int j = 0; j < kH; ++j) atomicAdd(&grad[kOffW2 + tid * kH + j], acc[j]);
        atomicAdd(&grad[kOffB2 + tid], gb);
    }
    __syncthreads();
    // dh1 = dz2 W2, dz1 = dh1 * gelu'(z1), row per lane (overwrites h1)
    {
        float dh[kH];
#pragma unroll
        for (int j = 0; j < kH; ++j) dh[j] = 0.f;
        for (int i = 0; i < kH; ++i) {
            const float dz = sZ2[tid][i];
#pragma unroll
            for (int j = 0; j < kH; ++j) dh[j] = __builtin_fmaf(dz, sW[kOffW2 + i * kH + j], dh[j]);
        }
        for (int j = 0; j < kH; ++j) {
            float zz = sW[kOffB1 + j];
#pragma unroll
            for (int k = 0; k < kIn; ++k) zz = __builtin_fmaf(sW[kOffW1 + j * kIn + k], xin[k], zz);
            sH1[tid][j] = dh[j] * gelu_grad(zz);
        }
    }
    __syncthreads();
    // dW1[i][k] = sum_s dz1[s][i] x[s][k], db1 (lane = i)
    {
        float acc[kIn];
#pragma unroll
        for (int k = 0; k < kIn; ++k) acc[k] = 0.f;
        float gb = 0.f;
        for (int s = 0; s < kGB; ++s) {
            const float dz = sH1[s][tid];
            gb += dz;
#pragma unroll
            for (int k = 0; k < kIn; ++k) acc[k] = __builtin_fmaf(dz, sX[s][k], acc[k]);
        }
#pragma unroll
        for (int k = 0; k < kIn; ++k) atomicAdd(&grad[kOffW1 + tid * kIn + k], acc[k]);
        atomicAdd(&grad[kOffB1 + tid], gb);
    }
}

// ---- Adam ----------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ theta, float* __restrict__ mom,
                                              float* __restrict__ vel,
                                              const float* __restrict__ grad, int n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2,
                                              uint16_t* __restrict__ wbf) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = grad[i];
    const float m = b1 * mom[i] + (1.f - b1) * g;
    const float v = b2 * vel[i] + (1.f - b2) * g * g;
    mom[i] = m;
    vel[i] = v;
    const float th = theta[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    theta[i] = th;
    wbf[i] = f32_to_bf16(th);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ theta, int n,
                                                 uint16_t* __restrict__ wbf) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) wbf[i] = f32_to_bf16(theta[i]);
}

// ---- host launchers ---------------------------------------------------------------------------------
static FeatureInputs feature_inputs(int ngroups, int nt, const int* gstart,
                                    const uint32_t* targets, const int* tslot,
                                    const gz_particles* p, const double* amag, const gz_tree* t,
                                    const uint32_t* ngroups_dev) {
    FeatureInputs in;
    in.ngroups = ngroups;
    in.ngroups_dev = ngroups_dev;
    in.nt_total = nt;
    in.gstart = gstart;
    in.targets = targets;
    in.tslot = tslot;
    in.src = t->src;
    in.src_soft = t->src_soft;
    in.x = p->x; in.y = p->y; in.z = p->z; in.m = p->m; in.soft = p->soft;
    in.amag = amag;
    in.slot_leaf = t->slot_leaf;
    in.level_begin = t->level_begin_dev;
    in.wn = t->wn;
    in.root = t->root;
    in.cstart = t->cluster_start_dev;
    in.nc = t->n_clusters;
    return in;
}

gz_status launch_group_features(int ngroups, int nt, const int* gstart, const uint32_t* targets,
                                const int* tslot, const gz_particles* p, const double* amag,
                                const gz_tree* t, GroupStats* stats, float* feats,
                                hipStream_t st, const uint32_t* ngroups_dev) {
    if (ngroups <= 0) return GZ_OK;
    if (feats && ngroups_dev)
        return fail(GZ_ERR_INTERNAL, "group features need the exact group count");
    // lanes per group from the mean group size: the kernel waits on chains of dependent loads,
    // so several groups per wavefront (a few passes each) beat one pass of a whole wavefront
    const FeatureInputs in =
        feature_inputs(ngroups, nt, gstart, targets, tslot, p, amag, t, ngroups_dev);
    const double mean = (double)nt / (double)ngroups;
    if (mean <= 64.0)
        k_group_stats<16><<<(unsigned)(((size_t)ngroups + 15) / 16), 256, 0, st>>>(in, stats);
    else if (mean <= 128.0)
        k_group_stats<32><<<(unsigned)(((size_t)ngroups + 7) / 8), 256, 0, st>>>(in, stats);
    else
        k_group_stats<64><<<(unsigned)(((size_t)ngroups + 3) / 4), 256, 0, st>>>(in, stats);
    GZ_HIP(hipGetLastError());
    if (feats) {
        k_group_features<<<(unsigned)((ngroups + 255) / 256), 256, 0, st>>>(stats, ngroups, t->wn,
                                                                            feats);
        GZ_HIP(hipGetLastError());
    }
    return GZ_OK;
}

gz_status launch_group_tolerance(int ngroups, int nt, const int* gstart, const uint32_t* targets,
                                 const int* tslot, const gz_particles* p, const double* aold,
                                 const gz_tree* t, const gz_model* M, GroupStats* stats,
                                 float* scale, int* n_fallback, hipStream_t st,
                                 const uint32_t* ngroups_dev) {
    if (ngroups <= 0) return GZ_OK;
    GZ_TRY(launch_group_features(ngroups, nt, gstart, targets, tslot, p, aold, t, stats, nullptr,
                                 st, ngroups_dev));
    ToleranceModel md{M->wbf, M->theta, M->xnorm, M->frange, M->lab_lo, M->lab_hi, M->margin,
                      M->log_alpha_ref};
    // the weight rows are read in 16-byte pieces
    static_assert((kOffW1 * 2) % 16 == 0 && (kOffW2 * 2) % 16 == 0 && (kIn * 2) % 16 == 0 &&
                      (kH * 2) % 16 == 0, "bf16 weight rows on 16-byte boundaries");
    if (reinterpret_cast<uintptr_t>(M->wbf) % 16 != 0)
        return fail(GZ_ERR_INTERNAL, "group tolerance: bf16 weights not 16-byte aligned");
    const int nbatch = (ngroups + kScaleBatch - 1) / kScaleBatch;
    k_group_scale<<<(unsigned)std::min(nbatch, 1024), 64, 0, st>>>(stats, ngroups, ngroups_dev,
                                                                  t->wn, md, scale, n_fallback);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_mlp_forward(const float* X, int B, const uint16_t* wbf, const float* theta,
                             const float* xnorm, float* out, hipStream_t st) {
    if (B <= 0) return GZ_OK;
    k_mlp_forward<<<(unsigned)((B + 255) / 256), 256, 0, st>>>(X, B, wbf, theta, xnorm, out);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_mlp_grad(const float* X, const float* T, int B, const uint16_t* wbf,
                          const float* theta, const float* xnorm, float tau, float* grad,
                          float* loss, hipStream_t st) {
    if (B <= 0) return GZ_OK;
    k_mlp_grad<<<(unsigned)((B + kGB - 1) / kGB), 64, 0, st>>>(X, T, B, 1.0f / (float)B, wbf,
                                                                theta, xnorm, tau, grad, loss);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_adam(float* theta, float* m, float* v, const float* grad, float lr, float b1,
                      float b2, float eps, int step, uint16_t* wbf, hipStream_t st) {
    const float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
    k_adam<<<(kMlpParams + 255) / 256, 256, 0, st>>>(theta, m, v, grad, kMlpParams, lr, b1, b2,
                                                      eps, bc1, bc2, wbf);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

gz_status launch_to_bf16(const float* theta, uint16_t* wbf, hipStream_t st) {
    k_to_bf16<<<(kMlpParams + 255) / 256, 256, 0, st>>>(theta, kMlpParams, wbf);
    GZ_HIP(hipGetLastError());
    return GZ_OK;
}

} // namespace gz
