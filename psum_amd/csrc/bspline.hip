// Higher-order particle/field coupling: quadratic, cubic and quartic
// B-spline gather + push and deposit (the reference's interp_B_spline /
// add_back_B_spline, higher_order_interp.hpp:79-248, selected in TSI by
// interp_order, TSI/main.cpp:29,54,64).
//
// Support of a particle per dimension: S = order + 1 nodes starting at
// base + shift, where base is the nearest node (round) for odd S and the
// corner node (floor) for even S, shift = -S/2 (odd) or -(S/2 - 1) (even),
// each node wrapped once over the w = cells + 1 nodes (simple_grid.hpp:
// 316-385).  Weights are FP64 for both particle dtypes, as in the reference;
// support point i takes digit j of i (base S) as the offset of dimension
// D-1-j, and its weight is the product of the 1D factors for dimensions D-1
// down to 0.  |weight| <= 1e-10 is skipped (field_interp.hpp:70,100).
//
// Deposit: acc[node] += round((w * coef) * scale) in int64 (scale a power of
// two), exact and order independent, staged through an LDS tile over the
// block's support box (global integer atomics when the box exceeds the
// tile).  Gather: the FP64 sum of phi * gradient weight in support order,
// rounded once to the particle dtype; then kick, drift and the boundaries
// of the CIC push.
#include <climits>
#include <cmath>

#include "common.hpp"
#include "psum_bspline.hpp"

namespace psum_hip {
namespace {

using namespace bspline;

constexpr int kBsThreads = 256;
constexpr int kBsPPT = 4;                           // particles per thread
constexpr int kBsTile = kBsThreads * kBsPPT;        // = the CIC push tile (diag block layout)
constexpr int kBsTileNodes = 2048;

// ------------------------------------------------------------ gather + push

template <typename T, int D, int ORDER>
__global__ __launch_bounds__(kBsThreads) void k_gather_push_bs(
    DevGrid g, T* __restrict__ x0, T* __restrict__ x1, T* __restrict__ x2, T* __restrict__ v0,
    T* __restrict__ v1, T* __restrict__ v2, int64_t n, const T* __restrict__ phi, T qm, T dt, double mass,
    double* __restrict__ diag, unsigned long long* __restrict__ counters)
{
    constexpr int S = ORDER + 1;
    constexpr int NS = D == 2 ? S * S : S * S * S;
    __shared__ double wd[kBsThreads / 64][4];
    T* const xs[3] = {x0, x1, x2};
    T* const vs[3] = {v0, v1, v2};
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double ke = 0.0, pm[3] = {0.0, 0.0, 0.0};
    unsigned absorbed = 0;
    for (int j = 0; j < kBsPPT; ++j) {
        const int64_t p = (int64_t)blockIdx.x * kBsTile + (int64_t)j * kBsThreads + threadIdx.x;
        if (p >= n) break;
        T x[D], v[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            x[d] = xs[d][p];
            v[d] = vs[d][p];
        }
        if (!isfinite((double)x[0])) continue;
        double v2sum = 0.0;
#pragma unroll
        for (int d = 0; d < D; ++d) v2sum += (double)v[d] * (double)v[d];
        ke += 0.5 * mass * v2sum;
#pragma unroll
        for (int d = 0; d < D; ++d) pm[d] += mass * (double)v[d];

        double xd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xd[d] = (double)x[d];
        int base[D];
        double c[D][S], dc[D][S];
        T grad[D];
#pragma unroll
        for (int d = 0; d < D; ++d) grad[d] = (T)0;
        if (support_of<D, ORDER, true>(g, xd, base, c, dc)) {
            double r[D];
#pragma unroll
            for (int d = 0; d < D; ++d) r[d] = 0.0;
            for (int i = 0; i < NS; ++i) {
                int k[3] = {0, 0, 0};
                int t = i;
#pragma unroll
                for (int jj = 0; jj < D; ++jj) {
                    k[D - 1 - jj] = t % S;
                    t /= S;
                }
                uint32_t lin = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const int nd = wrap_node(base[d] + k[d] + Support<ORDER>::shift, g.w[d]);
                    lin = d == 0 ? (uint32_t)nd : lin * (uint32_t)g.w[d] + (uint32_t)nd;
                }
                const double ph = (double)phi[lin];
#pragma unroll
                for (int dg = 0; dg < D; ++dg) {
                    double gf = dg == D - 1 ? dc[D - 1][k[D - 1]] : c[D - 1][k[D - 1]];
#pragma unroll
                    for (int jj = 1; jj < D; ++jj) {
                        const int dd = D - 1 - jj;
                        gf = gf * (dd == dg ? dc[dd][k[dd]] : c[dd][k[dd]]);
                    }
                    if (fabs(gf) > 1e-10) r[dg] += ph * gf;
                }
            }
#pragma unroll
            for (int d = 0; d < D; ++d) grad[d] = (T)r[d];
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T e = -grad[d];
            v[d] = v[d] + (e * qm) * dt;
            x[d] = x[d] + v[d] * dt;
        }
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (g.periodic[d]) {
                const T lo = (T)g.plo[d], hi = (T)g.phi[d], L = (T)g.plen[d];
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    if (x[d] < lo) x[d] = x[d] + L;
                    else if (x[d] >= hi) x[d] = x[d] - L;
                }
            }
        bool alive = true;
#pragma unroll
        for (int d = 0; d < D; ++d)
            if (!g.periodic[d]) {
                const double xx = (double)x[d];
                if (!(xx >= g.lb[d] && xx < g.ub[d])) alive = false;
            }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            xs[d][p] = x[d];
            vs[d][p] = v[d];
        }
        if (!alive) {
            xs[0][p] = (T)NAN;
            ++absorbed;
        }
    }
    if (diag) {
        const double vals[4] = {ke, pm[0], pm[1], pm[2]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double s = wave_sum_full(vals[q]);
            if (lane == 0) wd[wave][q] = s;
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            double s = 0.0;
            for (int w = 0; w < kBsThreads / 64; ++w) s += wd[w][threadIdx.x];
            diag[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
        }
    }
    if (counters) {
        const int ab = wave_isum_full((int)absorbed);
        if (lane == 0 && ab) atomicAdd(&counters[2], (unsigned long long)ab);
    }
}

// ------------------------------------------------------------------ deposit

template <typename T, int D, int ORDER>
__global__ __launch_bounds__(kBsThreads) void k_deposit_bs(DevGrid g, const T* __restrict__ x0,
                                                           const T* __restrict__ x1, const T* __restrict__ x2,
                                                           int64_t n, double wsc,
                                                           unsigned long long* __restrict__ acc,
                                                           unsigned long long* __restrict__ counters)
{
    constexpr int S = ORDER + 1;
    constexpr int NS = D == 2 ? S * S : S * S * S;
    constexpr int shift = Support<ORDER>::shift;
    __shared__ unsigned long long tile[kBsTileNodes];
    __shared__ int bmin[3], bmax[3], bref[3];
    const T* const xs[3] = {x0, x1, x2};
    const int lane = threadIdx.x & 63;
    // wraps of the global int64 sums (an LDS entry: <= kBsTile terms below 2^50)
    unsigned long long* const ovf = counters ? counters + 3 : nullptr;
    if (threadIdx.x < 3) {
        bmin[threadIdx.x] = INT_MAX;
        bmax[threadIdx.x] = INT_MIN;
    }
    // the thread's particles: base nodes and 1D weights
    bool ok[kBsPPT];
    int base[kBsPPT][D];
    double c[kBsPPT][D][S];
    unsigned oor = 0;
    int lmin[3] = {INT_MAX, INT_MAX, INT_MAX}, lmax[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
    for (int j = 0; j < kBsPPT; ++j) {
        ok[j] = false;
        const int64_t p = (int64_t)blockIdx.x * kBsTile + (int64_t)j * kBsThreads + threadIdx.x;
        if (p >= n) continue;
        double xd[D];
#pragma unroll
        for (int d = 0; d < D; ++d) xd[d] = (double)xs[d][p];
        if (!isfinite(xd[0])) continue;
        if (!support_of<D, ORDER, false>(g, xd, base[j], c[j], nullptr)) {
            ++oor;
            continue;
        }
        ok[j] = true;
    }
    if (counters) {
        const int oo = wave_isum_full((int)oor);
        if (lane == 0 && oo) atomicAdd(&counters[0], (unsigned long long)oo);
    }
    // box over base nodes unwrapped around thread 0's first particle on
    // periodic axes (a block straddling the seam keeps a small box)
    if (threadIdx.x == 0) {
        int r[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
        for (int j = kBsPPT - 1; j >= 0; --j)
            if (ok[j])
#pragma unroll
                for (int d = 0; d < D; ++d) r[d] = base[j][d];
#pragma unroll
        for (int d = 0; d < D; ++d) bref[d] = r[d];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kBsPPT; ++j)
        if (ok[j])
#pragma unroll
            for (int d = 0; d < D; ++d) {
                if (g.periodic[d] && bref[d] != INT_MIN) {
                    const int w = g.w[d];
                    if (base[j][d] - bref[d] > w / 2) base[j][d] -= w;
                    else if (bref[d] - base[j][d] > w / 2) base[j][d] += w;
                }
                lmin[d] = min(lmin[d], base[j][d]);
                lmax[d] = max(lmax[d], base[j][d]);
            }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int mn = wave_min_full(lmin[d]), mx = wave_max_full(lmax[d]);
        if (lane == 0 && mn != INT_MAX) {
            atomicMin(&bmin[d], mn);
            atomicMax(&bmax[d], mx);
        }
    }
    __syncthreads();
    const bool any = bmin[0] != INT_MAX;
    int e[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
    int64_t tsize = any ? 1 : 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = bmin[d] + shift;                 // unwrapped node coordinates
        e[d] = bmax[d] - bmin[d] + S;
        tsize *= e[d];
    }
    const bool use_lds = any && tsize <= kBsTileNodes;
    if (use_lds)
        for (int i = threadIdx.x; i < (int)tsize; i += kBsThreads) tile[i] = 0ull;
    __syncthreads();

#pragma unroll
    for (int j = 0; j < kBsPPT; ++j) {
        if (!ok[j]) continue;
        for (int i = 0; i < NS; ++i) {
            int k[3] = {0, 0, 0};
            int t = i;
#pragma unroll
            for (int jj = 0; jj < D; ++jj) {
                k[D - 1 - jj] = t % S;
                t /= S;
            }
            double cf = c[j][D - 1][k[D - 1]];
#pragma unroll
            for (int jj = 1; jj < D; ++jj) cf = cf * c[j][D - 1 - jj][k[D - 1 - jj]];
            if (!(fabs(cf) > 1e-10)) continue;
            const long long term = round_to_ll(cf * wsc);
            if (use_lds) {
                int li = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) li = li * e[d] + (base[j][d] + k[d] + shift - lo[d]);
                atomicAdd(&tile[li], (unsigned long long)term);
            } else {
                uint32_t lin = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const int nd = mod_node(base[j][d] + k[d] + shift, g.w[d]);
                    lin = d == 0 ? (uint32_t)nd : lin * (uint32_t)g.w[d] + (uint32_t)nd;
                }
                acc_add(&acc[lin], (unsigned long long)term, ovf);
            }
        }
    }
    if (use_lds) {
        __syncthreads();
        for (int i = threadIdx.x; i < (int)tsize; i += kBsThreads) {
            const unsigned long long v = tile[i];
            if (v == 0ull) continue;
            int rem = i;
            int nd[3];
#pragma unroll
            for (int d = D - 1; d >= 0; --d) {
                nd[d] = mod_node(lo[d] + rem % e[d], g.w[d]);
                rem /= e[d];
            }
            uint32_t lin = (uint32_t)nd[0];
#pragma unroll
            for (int d = 1; d < D; ++d) lin = lin * (uint32_t)g.w[d] + (uint32_t)nd[d];
            acc_add(&acc[lin], v, ovf);
        }
    } else if (any && counters && threadIdx.x == 0) {
        atomicAdd(&counters[1], 1ull);
    }
}

unsigned bs_blocks(int64_t n) { return (unsigned)((n + kBsTile - 1) / kBsTile); }

template <typename T, int D>
psum_status launch_gather_push_bs(int order, const DevGrid& dg, void* const* pos, void* const* vel, int64_t n,
                                  const void* phi, double qm, double dt, double mass, double* diag,
                                  int64_t* counters, hipStream_t st)
{
    T* p[3] = {nullptr, nullptr, nullptr};
    T* v[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < D; ++d) {
        p[d] = static_cast<T*>(pos[d]);
        v[d] = static_cast<T*>(vel[d]);
    }
    auto* c = reinterpret_cast<unsigned long long*>(counters);
    const T* ph = static_cast<const T*>(phi);
#define PSUM_BS_PUSH(O)                                                                          \
    hipLaunchKernelGGL((k_gather_push_bs<T, D, O>), dim3(bs_blocks(n)), dim3(kBsThreads), 0, st, dg, \
                       p[0], p[1], p[2], v[0], v[1], v[2], n, ph, (T)qm, (T)dt, mass, diag, c)
    if (order == 2) PSUM_BS_PUSH(2);
    else if (order == 3) PSUM_BS_PUSH(3);
    else PSUM_BS_PUSH(4);
#undef PSUM_BS_PUSH
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

template <typename T, int D>
psum_status launch_deposit_bs(int order, const DevGrid& dg, void* const* pos, int64_t n, double wsc,
                              int64_t* acc, int64_t* counters, hipStream_t st)
{
    const T* p[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < D; ++d) p[d] = static_cast<const T*>(pos[d]);
    auto* a = reinterpret_cast<unsigned long long*>(acc);
    auto* c = reinterpret_cast<unsigned long long*>(counters);
#define PSUM_BS_DEP(O)                                                                            \
    hipLaunchKernelGGL((k_deposit_bs<T, D, O>), dim3(bs_blocks(n)), dim3(kBsThreads), 0, st, dg, p[0], \
                       p[1], p[2], n, wsc, a, c)
    if (order == 2) PSUM_BS_DEP(2);
    else if (order == 3) PSUM_BS_DEP(3);
    else PSUM_BS_DEP(4);
#undef PSUM_BS_DEP
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

bool scale_pow2(double s)
{
    int e = 0;
    return s > 0.0 && std::isfinite(s) && std::frexp(s, &e) == 0.5;
}

}  // namespace
}  // namespace psum_hip

extern "C" {

psum_status psum_hip_deposit_bspline(const psum_grid* g, int dtype, int order, void* const* pos, int64_t n,
                                     double w, double scale, int64_t* acc, int64_t* counters, void* stream)
{
    using namespace psum_hip;
    if (order == 1) return psum_hip_deposit_cic(g, dtype, pos, n, w, scale, acc, counters, stream);
    PSUM_ARG_CHECK(grid_ok(g) && g->dim >= 2, "psum_hip_deposit_bspline: invalid grid (2D or 3D)");
    PSUM_ARG_CHECK(order >= 2 && order <= 4, "psum_hip_deposit_bspline: order must be 1..4");
    PSUM_ARG_CHECK(acc && (n == 0 || pos), "psum_hip_deposit_bspline: null pointer");
    PSUM_ARG_CHECK(scale_pow2(scale), "psum_hip_deposit_bspline: scale must be a power of two");
    PSUM_ARG_CHECK(std::fabs(w) * scale <= 0x1p50, "psum_hip_deposit_bspline: |w| * scale must not exceed 2^50");
    if (n == 0) return PSUM_OK;
    const DevGrid dg = to_dev(*g);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const double wsc = w * scale;
    if (dtype == PSUM_F64)
        return g->dim == 2 ? launch_deposit_bs<double, 2>(order, dg, pos, n, wsc, acc, counters, st)
                           : launch_deposit_bs<double, 3>(order, dg, pos, n, wsc, acc, counters, st);
    if (dtype == PSUM_F32)
        return g->dim == 2 ? launch_deposit_bs<float, 2>(order, dg, pos, n, wsc, acc, counters, st)
                           : launch_deposit_bs<float, 3>(order, dg, pos, n, wsc, acc, counters, st);
    set_error("psum_hip_deposit_bspline: bad dtype");
    return PSUM_ERR_ARG;
}

psum_status psum_hip_gather_push_bspline(const psum_grid* g, int dtype, int order, void* const* pos,
                                         void* const* vel, int64_t n, const void* phi, double qm, double dt,
                                         double mass, double* diag, int64_t* counters, void* stream)
{
    using namespace psum_hip;
    if (order == 1)
        return psum_hip_gather_push(g, dtype, pos, vel, n, phi, qm, dt, mass, 0.0, 1.0, nullptr, diag, counters,
                                    stream);
    PSUM_ARG_CHECK(grid_ok(g) && g->dim >= 2, "psum_hip_gather_push_bspline: invalid grid (2D or 3D)");
    PSUM_ARG_CHECK(order >= 2 && order <= 4, "psum_hip_gather_push_bspline: order must be 1..4");
    PSUM_ARG_CHECK(n == 0 || (pos && vel && phi), "psum_hip_gather_push_bspline: null pointer");
    if (n == 0) return PSUM_OK;
    const DevGrid dg = to_dev(*g);
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (dtype == PSUM_F64)
        return g->dim == 2
                   ? launch_gather_push_bs<double, 2>(order, dg, pos, vel, n, phi, qm, dt, mass, diag, counters, st)
                   : launch_gather_push_bs<double, 3>(order, dg, pos, vel, n, phi, qm, dt, mass, diag, counters, st);
    if (dtype == PSUM_F32)
        return g->dim == 2
                   ? launch_gather_push_bs<float, 2>(order, dg, pos, vel, n, phi, qm, dt, mass, diag, counters, st)
                   : launch_gather_push_bs<float, 3>(order, dg, pos, vel, n, phi, qm, dt, mass, diag, counters, st);
    set_error("psum_hip_gather_push_bspline: bad dtype");
    return PSUM_ERR_ARG;
}

}  // extern "C"
