// FP64 gravity arithmetic shared by the tree walks: the reference's spline kernels (tree.h:59-96)
// and opening test (tree.h:109-144) restated operand for operand, and a fast FP64 1/sqrt.
#pragma once

#include <hip/hip_runtime.h>

namespace gz {

// std::min / std::max of two numbers that are never NaN: one v_min_f64 / v_max_f64 (a compare
// and two selects otherwise); the two forms differ only in the sign of a zero result.
__device__ __forceinline__ double dmin_w(double a, double b) { return __builtin_fmin(a, b); }
__device__ __forceinline__ double dmax_w(double a, double b) { return __builtin_fmax(a, b); }

// ---- kernels of the pair force (tree.h:59-67), FP64 as the reference writes them --------------
__device__ __forceinline__ double spline_force_over_r_d(double r, double h) {
    const double h_inv = 1.0 / h;
    const double h_inv3 = h_inv * h_inv * h_inv;
    const double u = r * h_inv;
    if (u >= 1.0) return 1.0 / (r * r * r);
    if (u < 0.5) return h_inv3 * (10.666666666666667 + u * u * (32.0 * u - 38.4));
    return h_inv3 * (21.333333333333332 - 48.0 * u + 38.4 * u * u - 10.666666666666667 * u * u * u -
                     0.06666666666666667 / (u * u * u));
}

// 1/sqrt(x) for x > 0 in FP64 to about an ulp: the hardware estimate y (measured: within 2^-24
// relative) and one third-order step y (1 + e/2 + 3 e^2/8), e = 1 - x y^2, which takes the error
// from d to about 5 d^3 / 2 (measured: within 1.4e-16 relative over 2^-60..2^60, against 2.8e-16
// for two Newton steps, in five operations instead of eight).
__device__ __forceinline__ double rsqrt_d(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = __builtin_fma(-(x * y), y, 1.0);
    return __builtin_fma(y, e * __builtin_fma(e, 0.375, 0.5), y);
}

// spline_force_over_r from r^2, with r and 1/r^3 from rsqrt_d (a few ulp from the reference's
// sqrt and division) and the spline's 1/u^3 as (h/r)^3; the branches and polynomials are the
// reference's (tree.h:59-67).
__device__ __forceinline__ double spline_force_over_r_r2(double r2, double h) {
    const double inv = rsqrt_d(r2);
    const double r = r2 * inv;
    const double h_inv = 1.0 / h;
    const double u = r * h_inv;
    if (u >= 1.0) return inv * inv * inv;
    const double h_inv3 = h_inv * h_inv * h_inv;
    if (u < 0.5) return h_inv3 * (10.666666666666667 + u * u * (32.0 * u - 38.4));
    const double hu = h * inv; // 1/u
    return h_inv3 * (21.333333333333332 - 48.0 * u + 38.4 * u * u - 10.666666666666667 * u * u * u -
                     0.06666666666666667 * (hu * hu * hu));
}

// The pair terms of the per-target walks inside the kernel (r^2 < h^2), from r^2 > 0, inv = 1/r
// (rsqrt_d) and the kernel extent h: the reference's polynomials in u = r/h (tree.h:59-80), with
// 1/u as h/r. Beyond the kernel the terms are the Newtonian inv^3 (force over r) and -inv
// (potential); the walks classify a pair by r^2 against h^2, and since the two sides of r = h
// meet continuously, that moves a result by rounding only.
__device__ __forceinline__ double pair_force_over_r_inside(double r2, double inv, double h) {
    const double h_inv = 1.0 / h;
    const double u = (r2 * inv) * h_inv;
    const double h_inv3 = h_inv * h_inv * h_inv;
    if (u < 0.5) return h_inv3 * (10.666666666666667 + u * u * (32.0 * u - 38.4));
    if (u < 1.0) {
        const double hu = h * inv;
        return h_inv3 * (21.333333333333332 - 48.0 * u + 38.4 * u * u -
                         10.666666666666667 * u * u * u - 0.06666666666666667 * (hu * hu * hu));
    }
    return inv * inv * inv;
}

__device__ __forceinline__ double pair_potential_inside(double r2, double inv, double h) {
    const double h_inv = 1.0 / h;
    const double u = (r2 * inv) * h_inv;
    if (u < 0.5) return h_inv * (-2.8 + u * u * (5.333333333333333 + u * u * (6.4 * u - 9.6)));
    if (u < 1.0)
        return h_inv * (-3.2 + 0.06666666666666667 * (h * inv) +
                        u * u * (10.666666666666667 +
                                 u * (-16.0 + u * (9.6 - 2.1333333333333333 * u))));
    return -inv;
}

// spline_potential (tree.h:72-80), operand for operand: exactly -1/r beyond h.
__device__ __forceinline__ double spline_potential_d(double r, double h) {
    const double h_inv = 1.0 / h;
    const double u = r * h_inv;
    if (u >= 1.0) return -1.0 / r;
    if (u < 0.5) return h_inv * (-2.8 + u * u * (5.333333333333333 + u * u * (6.4 * u - 9.6)));
    return h_inv * (-3.2 + 0.06666666666666667 / u +
                    u * u * (10.666666666666667 + u * (-16.0 + u * (9.6 - 2.1333333333333333 * u))));
}

// grav_tidal_factor (tree.h:86-96), operand for operand: the second kernel derivative, 3/r^5
// beyond h.
__device__ __forceinline__ double tidal_factor_d(double r, double h) {
    const double h_inv = 1.0 / h;
    const double u = r * h_inv;
    if (u >= 1.0) return 3.0 / (r * r * r * r * r);
    const double h_inv5 = h_inv * h_inv * h_inv * h_inv * h_inv;
    double wk;
    if (u < 0.5) wk = 76.8 - 96.0 * u;
    else wk = -0.2 / (u * u * u * u * u) + 48.0 / u - 76.8 + 32.0 * u;
    return wk * h_inv5;
}

// open_node (tree.h:109-144), operand for operand. The reference returns at the first arm that
// holds; every arm is a side-effect-free predicate, so their OR is the same decision, and
// evaluating all of them keeps the wave on one path (no divergent branches per node).
__device__ __forceinline__ bool open_node_d(double r_sq, double len, double l_open, double mass,
                                            double node_soft, double soft_t, double aold,
                                            double theta_sq, double adx, double ady, double adz) {
    const double grow_t = soft_t + 0.6 * len, grow_n = node_soft + 0.6 * len;
    const bool soft_arm = (r_sq < grow_t * grow_t) | (r_sq < grow_n * grow_n);
    const bool theta_arm = (theta_sq > 0.0) & (l_open * l_open >= theta_sq * r_sq);
    const bool rel_arm = (aold > 0) ? (mass * len * len > r_sq * r_sq * aold) : (theta_sq <= 0.0);
    const double box = 0.6 * len;
    const bool inside = (adx < box) & (ady < box) & (adz < box);
    return soft_arm | theta_arm | rel_arm | inside;
}

} // namespace gz
