// B-spline support geometry and weights (orders 2..4) shared by the HIP
// kernels and the C++ field API (include/psum/): the reference's quadratic,
// cubic and quartic B-splines (simple_interpolation.hpp:105-184) on order+1
// nodes per dimension around the nearest node (odd support) or the corner
// node (even support), wrapped once (simple_grid.hpp:316-385).  FP64.
#pragma once

#include <hip/hip_runtime.h>

#include "psum_cic.hpp"

namespace psum_hip {
namespace bspline {

// (1.5-a)^4 and friends as repeated products (the reference's std::pow, whose
// last bit is implementation-defined, evaluated the same way on every side)
__host__ __device__ inline double pw3(double r) { return (r * r) * r; }
__host__ __device__ inline double pw4(double r) { return ((r * r) * r) * r; }

__host__ __device__ inline double cubic_b(double t)
{
    const double a = fabs(t);
    if (a <= 1.0) return 2.0 / 3.0 - a * a + a * a * a / 2.0;
    if (a <= 2.0) {
        const double r = 2.0 - a;
        return r * r * r / 6.0;
    }
    return 0.0;
}
__host__ __device__ inline double cubic_d(double t)
{
    const double a = fabs(t);
    if (a <= 1.0) return -2.0 * t + 1.5 * t * a;
    if (a <= 2.0) {
        const double r = 2.0 - a;
        return -r * r / 2.0 * (t < 0 ? -1.0 : 1.0);
    }
    return 0.0;
}
__host__ __device__ inline double quartic_b(double t)
{
    const double a = fabs(t);
    if (a <= 0.5) return 115.0 / 192.0 - a * a * (5.0 / 8.0) + a * a * a * a / 4.0;
    if (a <= 1.5) {
        const double v = a - 0.5;
        return 0.5 - (2.0 * v / 3.0 - v * v * v / 3.0 + v * v * v * v / 8.0) - pw4(1.5 - a) / 24.0;
    }
    if (a <= 2.5) return pw4(2.5 - a) / 24.0;
    return 0.0;
}
__host__ __device__ inline double quartic_d(double t)
{
    const double a = fabs(t);
    const double sign = t < 0 ? -1.0 : 1.0;
    if (a <= 0.5) return sign * (-a * 1.25 + a * a * a);
    if (a <= 1.5) {
        const double v = a - 0.5;
        return sign * (-(2.0 / 3.0 - v * v + v * v * v / 2.0) + pw3(1.5 - a) / 6.0);
    }
    if (a <= 2.5) return sign * (-pw3(2.5 - a) / 6.0);
    return 0.0;
}

// 1D weights c[S] (and derivatives dc[S] when DIFF) at normalised offset s
template <int ORDER, bool DIFF>
__host__ __device__ inline void weights_1d(double s, double* c, double* dc)
{
    if constexpr (ORDER == 2) {
        c[0] = (0.5 - s) * (0.5 - s) * 0.5;
        c[1] = 0.75 - s * s;
        c[2] = (0.5 + s) * (0.5 + s) * 0.5;
        if constexpr (DIFF) {
            dc[0] = s - 0.5;
            dc[1] = -2.0 * s;
            dc[2] = s + 0.5;
        }
    } else if constexpr (ORDER == 3) {
#pragma unroll
        for (int k = 0; k < 4; ++k) c[k] = cubic_b(s + (double)(1 - k));
        if constexpr (DIFF) {
#pragma unroll
            for (int k = 0; k < 4; ++k) dc[k] = cubic_d(s + (double)(1 - k));
        }
    } else {
#pragma unroll
        for (int k = 0; k < 5; ++k) c[k] = quartic_b(s + (double)(2 - k));
        if constexpr (DIFF) {
#pragma unroll
            for (int k = 0; k < 5; ++k) dc[k] = quartic_d(s + (double)(2 - k));
        }
    }
}

template <int ORDER>
struct Support {
    static constexpr int S = ORDER + 1;
    static constexpr bool nearest = (S % 2) == 1;
    static constexpr int shift = nearest ? -(S / 2) : -(S / 2 - 1);
};

// base node per dimension and the 1D weights; false when the support leaves
// the single-wrap range or x is not finite
template <int D, int ORDER, bool DIFF>
__host__ __device__ inline bool support_of(const DevGrid& g, const double* x, int* base,
                                           double (*c)[ORDER + 1], double (*dc)[ORDER + 1])
{
    using Sp = Support<ORDER>;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (!isfinite(x[d])) return false;
        const double t = (x[d] - g.lb[d]) * g.delta_r[d];
        const double f = Sp::nearest ? round(t) : floor(t);
        const int w = g.w[d];
        if (!(f + Sp::shift >= (double)(-w) && f + Sp::shift + (Sp::S - 1) <= (double)(2 * w - 1))) return false;
        base[d] = (int)f;
        const double bp = g.lb[d] + (double)base[d] * g.delta[d];
        const double s = (x[d] - bp) * g.delta_r[d];
        weights_1d<ORDER, DIFF>(s, c[d], DIFF ? dc[d] : nullptr);
        if constexpr (DIFF) {
#pragma unroll
            for (int k = 0; k < Sp::S; ++k) dc[d][k] *= g.delta_r[d];
        }
    }
    return true;
}

// node of support offset k in dimension d, wrapped once
__host__ __device__ inline int wrap_node(int n, int w) { return n < 0 ? n + w : (n >= w ? n - w : n); }
// any node coordinate of an unwrapped box back into [0, w)
__host__ __device__ inline int mod_node(int n, int w) { return ((n % w) + w) % w; }

}  // namespace bspline
}  // namespace psum_hip
