// CIC support geometry shared by the HIP kernels and the C++ particle/field
// API (include/psum/): the grid as kernels see it, cell index and fractional
// offset in FP64, support-node indices and CIC (gradient) weights.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "psum_hip.h"

namespace psum_hip {

// Grid as the kernels see it (by value in kernel arguments).
struct DevGrid {
    int dim;
    int w[3];            // nodes per dimension = cells + 1
    double lb[3], ub[3], delta[3], delta_r[3];
    int periodic[3];
    double plo[3], phi[3], plen[3];   // periodic particle domain [plo, phi), phi = plo + plen
    int origin0;         // every lb is 0 and every ub exact in float
};

inline DevGrid to_dev(const psum_grid& g)
{
    DevGrid d{};
    d.dim = g.dim;
    for (int i = 0; i < 3; ++i) {
        d.w[i] = i < g.dim ? g.cells[i] + 1 : 1;
        d.lb[i] = g.lb[i];
        d.ub[i] = g.ub[i];
        d.delta[i] = g.delta[i];
        d.delta_r[i] = g.delta_r[i];
        d.periodic[i] = g.periodic[i];
        d.plo[i] = g.plo[i];
        d.phi[i] = g.plo[i] + g.plen[i];
        d.plen[i] = g.plen[i];
    }
    d.origin0 = 1;
    for (int i = 0; i < g.dim; ++i)
        if (!(g.lb[i] == 0.0 && (double)(float)g.ub[i] == g.ub[i])) d.origin0 = 0;
    return d;
}

inline int64_t grid_nodes(const psum_grid& g)
{
    int64_t n = 1;
    for (int i = 0; i < g.dim; ++i) n *= (int64_t)(g.cells[i] + 1);
    return n;
}

inline bool grid_ok(const psum_grid* g)
{
    if (!g || g->dim < 1 || g->dim > 3) return false;
    int64_t n = 1;
    for (int i = 0; i < g->dim; ++i) {
        if (g->cells[i] < 1) return false;
        n *= (int64_t)(g->cells[i] + 1);
    }
    return n < (int64_t)0xFFFFFFFF;
}

// ------------------------------------------------- device side (and host where marked)

// Base node of x and the fractional offset, FP64 (simple_grid.hpp:89-95,
// 201-206, 327-328).  cw is the base node after one periodic wrap over the
// w nodes of each dimension (simple_grid.hpp:338-343).  The reference result
// is defined only while that single wrap stays in range: cN in [-w, 2w-2].
// UNIT: every delta and delta_r is exactly 1, so the products with them are
// exact and are left out (same results).
template <int D, bool UNIT = false>
__host__ __device__ __forceinline__ bool base_cell(const DevGrid& g, const double* x, int* cw, double* s)
{
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double t = UNIT ? x[d] - g.lb[d] : (x[d] - g.lb[d]) * g.delta_r[d];
        const double f = floor(t);
        const int w = g.w[d];
        if (!(f >= (double)(-w) && f <= (double)(2 * w - 2))) return false;
        const int ci = (int)f;
        const double base = UNIT ? g.lb[d] + (double)ci : g.lb[d] + (double)ci * g.delta[d];
        s[d] = UNIT ? x[d] - base : (x[d] - base) * g.delta_r[d];
        cw[d] = ci < 0 ? ci + w : (ci >= w ? ci - w : ci);
    }
    return true;
}

// |c| > 1e-10 evaluated in FP64 (field_interp.hpp:70, :100), for c of type
// T.  For float c the same test is c > t32 with t32 the largest float not
// above 1e-10 (1e-10 is not a float, so (double)c > 1e-10 <=> c >= the next
// float up <=> c > t32): one float compare instead of a widening.
template <typename T>
__device__ __forceinline__ bool coef_kept(T c)
{
    if constexpr (sizeof(T) == 4) {
        constexpr float t32 = 9.999999439624929e-11f;   // largest float <= 1e-10
        return fabsf(c) > t32;
    } else {
        return fabs(c) > 1e-10;
    }
}

// Round-to-nearest-even of t to int64 for |t| < 2^51 (llrint in the default
// rounding mode): t + 1.5*2^52 lands in [2^52, 2^53), whose ulp is 1, so the
// addition rounds t to an integer and the difference of the bit patterns is
// that integer.
__device__ __forceinline__ long long round_to_ll(double t)
{
    const double m = 6755399441055744.0;   // 1.5 * 2^52
    const double u = t + m;
    return (long long)(__double_as_longlong(u) - __double_as_longlong(m));
}

// acc += v on a global int64 accumulator.  With ovf non-null, an add whose
// two's-complement partial sum wraps (old value and v of one sign, the sum
// of the other) is counted in *ovf, read from the value the atomic returns.
// Counting every global add this way leaves no wrap unseen; a sum of mixed
// signs may be counted although its final value is back in range (the
// integer result is then still exact: the adds are modulo 2^64).
__device__ __forceinline__ bool add_wraps(unsigned long long old, unsigned long long v)
{
    const unsigned long long sum = old + v;
    return (long long)((old ^ sum) & (v ^ sum)) < 0;
}
__device__ __forceinline__ void acc_add(unsigned long long* acc, unsigned long long v, unsigned long long* ovf)
{
    if (ovf) {
        if (add_wraps(atomicAdd(acc, v), v)) atomicAdd(ovf, 1ull);
    } else {
        atomicAdd(acc, v);
    }
}

// Linear node index of support point i: bit j of i offsets dimension D-1-j;
// cw + 1 == w wraps to node 0.
template <int D>
__host__ __device__ __forceinline__ uint32_t support_index(const DevGrid& g, const int* cw, int i)
{
    uint32_t lin = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        int nd = cw[d] + ((i >> (D - 1 - d)) & 1);
        if (nd == g.w[d]) nd = 0;
        lin = d == 0 ? (uint32_t)nd : lin * (uint32_t)g.w[d] + (uint32_t)nd;
    }
    return lin;
}

// CIC weight of support point i: product of {1-s, s} factors taken from
// dimension D-1 down to 0 (simple_interpolation.hpp:36-54).
template <typename T, int D>
__host__ __device__ __forceinline__ T cic_coef(const T (*W)[2], int i)
{
    T c = W[D - 1][i & 1];
#pragma unroll
    for (int j = 1; j < D; ++j) c = c * W[D - 1 - j][(i >> j) & 1];
    return c;
}

// Gradient weight of support point i along dimension dd (interp_tensor_diff,
// simple_interpolation.hpp:61-90): factor G for dimension dd, W otherwise.
template <typename T, int D>
__host__ __device__ __forceinline__ T cic_grad_coef(const T (*W)[2], const T (*G)[2], int i, int dd)
{
    T c = (D - 1 == dd) ? G[D - 1][i & 1] : W[D - 1][i & 1];
#pragma unroll
    for (int j = 1; j < D; ++j) {
        const int dim = D - 1 - j;
        c = c * (dim == dd ? G[dim][(i >> j) & 1] : W[dim][(i >> j) & 1]);
    }
    return c;
}

}  // namespace psum_hip
