// Device functions shared by the feature kernel (training data), the tolerance kernel (inference)
// and the walk: the 16 features of a target group, and the tree a target belongs to. One copy of
// the features is what makes the features a group sees in training the ones it sees in the walk.
#pragma once

#include "gz_internal.h"

namespace gz {

// What the features need to know about a group: its box, the sums over its targets, and the tree
// cell of its first target.
struct GroupStats {
    double lo[3], hi[3];
    double emax;       // largest softening
    double msum;       // total mass
    double amin, amax; // min / max of |a_prev| (aold units)
    double a2;         // sum of squares of |a_prev|
    int nt;            // targets
    double leaf_len;   // side of the first target's leaf (< 0: the target is not in the tree)
    int depth;         // level of that leaf
    int root;          // the root node of the group's tree
};

// Minimum and maximum over the wavefront; exact, so every lane gets the same value whatever the
// order of the exchange.
__device__ __forceinline__ double wave_min_d(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const double o = __shfl_xor(v, d);
        v = (o < v) ? o : v;
    }
    return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
        const double o = __shfl_xor(v, d);
        v = (v < o) ? o : v;
    }
    return v;
}

struct RootInfo {
    double cx, cy, cz, side, mass;
};

// The tree a target belongs to: in a forest, cluster c's root is node c.
__device__ __forceinline__ int root_of(uint32_t particle, const uint32_t* __restrict__ cstart,
                                       int nc, int root) {
    if (nc <= 1 || !cstart) return root;
    int lo = 0, hi = nc;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (cstart[mid] <= particle) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ float flog10(double v) { return (float)log10(fmax(v, 1e-300)); }

// All features are ratios of the group's own scales (extent, leaf cell, softening, field, mass)
// and of its position in the system, so none depends on the units or the box.
__device__ __forceinline__ void group_features16(const GroupStats& S, const RootInfo& R,
                                                 float* f) {
    const double side = fmax(R.side, 1e-300);
    const double Mr = fmax(R.mass, 1e-300);
    const double ex = S.hi[0] - S.lo[0], ey = S.hi[1] - S.lo[1], ez = S.hi[2] - S.lo[2];
    const double ext = fmax(fmax(ex, ey), fmax(ez, 1e-12 * side));
    const double emn = fmax(fmin(ex, fmin(ey, ez)), 1e-12 * side);
    const double cx = 0.5 * (S.lo[0] + S.hi[0]) - R.cx, cy = 0.5 * (S.lo[1] + S.hi[1]) - R.cy,
                 cz = 0.5 * (S.lo[2] + S.hi[2]) - R.cz;
    const double dist = fmax(sqrt(cx * cx + cy * cy + cz * cz), 1e-12 * side);
    const double arms = sqrt(S.a2 / fmax((double)S.nt, 1.0));
    const double mg = fmax(S.msum, 1e-300);
    const double rho_g = mg / (ext * ext * ext), rho_c = Mr / (dist * dist * dist);
    const double amn = fmax(S.amin, 1e-300);
    f[0] = flog10((double)S.nt);
    f[1] = flog10(ext / S.leaf_len);
    f[2] = flog10(amn * ext * ext / mg);      // external field / self field
    f[3] = flog10(S.amax / amn);
    f[4] = flog10(S.emax / ext);
    f[5] = flog10(emn / ext);                 // aspect of the group box
    f[6] = flog10(arms / amn);
    f[7] = (float)S.depth / 20.0f;
    f[8] = flog10(ext / dist);                // angular size seen from the centre of mass
    f[9] = (float)(fabs(cz) / dist);
    f[10] = flog10(Mr / (dist * dist * amn)); // bulk monopole field / actual field
    f[11] = flog10(mg / Mr);
    f[12] = flog10(rho_g / rho_c);            // local over enclosed-scale density
    f[13] = flog10(S.leaf_len / dist);
    f[14] = flog10(S.emax / dist);
    f[15] = flog10(S.emax / S.leaf_len);
}

} // namespace gz
