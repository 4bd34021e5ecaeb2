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
// amag is |a_prev| per target slot (no-G units, the basis of aold).

__global__ __launch_bounds__(256) void k_group_features(
    int ngroups, int nt_total, const int* __restrict__ gstart,
    const uint32_t* __restrict__ targets, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, const double* __restrict__ m,
    const double* __restrict__ soft, const double* __restrict__ amag,
    const int* __restrict__ leaf_of, const int* __restrict__ level_begin,
    const WNode* __restrict__ wn, int root, const uint32_t* __restrict__ cstart, int nc,
    float* __restrict__ feats) {
    // one thread per group: groups are small, and many independent groups in flight hide the
    // latency of the target gathers better than one wavefront per group
    const int g = blockIdx.x * 256 + threadIdx.x;
    if (g >= ngroups) return;
    GroupInfo G;
    G.tlo = gstart[g];
    G.thi = (g + 1 < ngroups) ? gstart[g + 1] : nt_total;
    const int nt = G.thi - G.tlo;
    double msum = 0, amin = 1e300, amax = 0, a2 = 0, emax = 0;
    double lo0 = 1e300, lo1 = 1e300, lo2 = 1e300, hi0 = -1e300, hi1 = -1e300, hi2 = -1e300;
    for (int i = G.tlo; i < G.thi; ++i) {
        const uint32_t p = targets[i];
        msum += m[p];
        const double am = amag ? amag[i] : 0.0;
        amin = fmin(amin, am);
        amax = fmax(amax, am);
        a2 += am * am;
        if (soft) emax = fmax(emax, soft[p]);
        lo0 = fmin(lo0, x[p]); hi0 = fmax(hi0, x[p]);
        lo1 = fmin(lo1, y[p]); hi1 = fmax(hi1, y[p]);
        lo2 = fmin(lo2, z[p]); hi2 = fmax(hi2, z[p]);
    }
    GroupStats S;
    S.lo[0] = lo0; S.lo[1] = lo1; S.lo[2] = lo2;
    S.hi[0] = hi0; S.hi[1] = hi1; S.hi[2] = hi2;
    S.emax = emax;
    S.msum = msum;
    S.amin = amin;
    S.amax = amax;
    S.a2 = a2;
    S.nt = nt;
    const WNode Rn = wn[root_of(targets[G.tlo], cstart, nc, root)];
    const RootInfo R{Rn.cx, Rn.cy, Rn.cz, (double)Rn.len, Rn.mass};
    const int leaf = leaf_of[targets[G.tlo]];
    S.leaf_len = (leaf >= 0) ? (double)wn[leaf].len : fmax(R.side, 1e-300);
    S.depth = 0; // node ids are level-major: the level is where the id falls in the table
    if (leaf >= 0)
        while (S.depth < kMaxLevel && level_begin[S.depth + 1] <= leaf) ++S.depth;
    float f[kIn];
    group_features16(S, R, f);
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
    float xin[kIn];
#pragma unroll
    for (int k = 0; k < kIn; ++k) xin[k] = (X[(size_t)b * kIn + k] - sN[k]) * sN[kIn + k];
    float h1[kH];
#pragma unroll 4
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
    out[b] = yy;
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
gz_status launch_group_features(int ngroups, int nt, const int* gstart, const uint32_t* targets,
                                const gz_particles* p, const double* amag, const gz_tree* t,
                                float* feats, hipStream_t st) {
    if (ngroups <= 0) return GZ_OK;
    k_group_features<<<(unsigned)((ngroups + 255) / 256), 256, 0, st>>>(
        ngroups, nt, gstart, targets, p->x, p->y, p->z, p->m, p->soft, amag, t->leaf_of,
        t->level_begin_dev, t->wn, t->root, t->cluster_start_dev, t->n_clusters, feats);
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
