// Adjoint (reverse-mode) kernels of the electrostatic PIC step.
//
// Forward step n (the reference step, TSI/main.cpp:59-75):
//   rho = sum_p w_dep c_k(x_p) at idx_k(x_p)        deposit (field_interp.hpp:63-75)
//   phi = S rho                                     neutralise, pin, MG solve
//   g_e(x) = sum_k phi[idx_k] d_e c_k(x)            interp_diff (field_interp.hpp:93-113)
//   v' = v - qm dt g(x) ;  x' = x + dt v'            kick, drift (wrap has unit Jacobian)
// Reverse, given the adjoints (lx', lv') of (x', v'):
//   lvt  = lv' + dt lx'
//   lx   = lx' - qm dt H(x)^T lvt,    H_ef = sum_k phi[idx_k] d_e d_f c_k   (e != f)
//   lphi += -qm dt sum_e lvt_e d_e c_k(x)   at idx_k     (transpose of the gather)
//   lv   = lvt
//   lrho = S^T lphi  (self-adjoint MG solve, see stepper.hip)
//   lx  += w_dep g(x; lrho)                              (transpose of the deposit)
// Deposit and gather are mutual transposes: the adjoint of the deposit is a
// gradient gather of lrho, the adjoint of the gather a gradient-weight
// deposit.  The gradient-weight deposit accumulates in int64 fixed point
// (no float atomics) with a scale 2^(40-e) from the largest per-particle
// term bound, computed on the device by the previous adjoint gather.

#include <climits>
#include <cmath>

#include "common.hpp"

using namespace psum_hip;

namespace {

constexpr int kA = 256;

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

// CIC factors of x: base cell cw, W (weights), G (weight derivatives)
template <typename T, int D>
__device__ __forceinline__ bool cic_factors(const DevGrid& g, const T* x, int* cw, T (*W)[2],
                                            T (*G)[2])
{
    double xd[D], s[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xd[d] = (double)x[d];
    if (!base_cell<D>(g, xd, cw, s)) return false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T sd = (T)s[d];
        W[d][0] = (T)1 - sd;
        W[d][1] = sd;
        const T dr = (T)g.delta_r[d];
        G[d][0] = (T)-1 * dr;
        G[d][1] = (T)1 * dr;
    }
    return true;
}

// second derivative d_e d_f c_i (e != f): G for e and f, W otherwise
template <typename T, int D>
__device__ __forceinline__ T cic_hess_coef(const T (*W)[2], const T (*G)[2], int i, int e, int f)
{
    T c = (T)1;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const int dim = D - 1 - j;
        const int b = (i >> j) & 1;
        c = c * ((dim == e || dim == f) ? G[dim][b] : W[dim][b]);
    }
    return c;
}

__device__ __forceinline__ double wave_max_d(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Reverse of kick + drift + gather for one species (in place on lx, lv);
// gradient-weight deposit of lvt into acc with the device-side scale.
// T: precision of the positions (the forward state); A: precision of the
// adjoint state lx, lv, of phi and of the arithmetic (A = double with FP32
// positions keeps the reverse sweep's accumulations in FP64).
template <typename T, typename A, int D>
__global__ __launch_bounds__(kA) void k_adj_push(DevGrid g, const T* __restrict__ x0,
                                                 const T* __restrict__ x1, const T* __restrict__ x2,
                                                 A* __restrict__ lx0, A* __restrict__ lx1,
                                                 A* __restrict__ lx2, A* __restrict__ lv0,
                                                 A* __restrict__ lv1, A* __restrict__ lv2,
                                                 int64_t n, const A* __restrict__ phi, A qm, A dt,
                                                 const double* __restrict__ scale_dev,
                                                 unsigned long long* __restrict__ acc)
{
    const int64_t p = (int64_t)blockIdx.x * kA + threadIdx.x;
    if (p >= n) return;
    const T* xs[3] = {x0, x1, x2};
    A* lxs[3] = {lx0, lx1, lx2};
    A* lvs[3] = {lv0, lv1, lv2};
    A x[D], lx[D], lvt[D];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] = (A)xs[d][p];
    if (!isfinite((double)x[0])) return;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lx[d] = lxs[d][p];
        lvt[d] = lvs[d][p] + dt * lx[d];
    }
    int cw[D];
    A W[D][2], G[D][2];
    if (cic_factors<A, D>(g, x, cw, W, G)) {
        A ph[1 << D];
#pragma unroll
        for (int i = 0; i < (1 << D); ++i) ph[i] = phi[support_index<D>(g, cw, i)];
        // lx_f -= qm dt sum_e lvt_e H_ef
#pragma unroll
        for (int f = 0; f < D; ++f) {
            A acc_f = (A)0;
#pragma unroll
            for (int e = 0; e < D; ++e) {
                if (e == f) continue;
                A h = (A)0;
#pragma unroll
                for (int i = 0; i < (1 << D); ++i) h = h + ph[i] * cic_hess_coef<A, D>(W, G, i, e, f);
                acc_f = acc_f + lvt[e] * h;
            }
            lx[f] = lx[f] - (qm * dt) * acc_f;
        }
        // lphi[idx_i] += -qm dt sum_e lvt_e d_e c_i
        const double scale = scale_dev[0];
        const double mqdt = -(double)qm * (double)dt;
#pragma unroll
        for (int i = 0; i < (1 << D); ++i) {
            A t = (A)0;
#pragma unroll
            for (int e = 0; e < D; ++e) t = t + lvt[e] * cic_grad_coef<A, D>(W, G, i, e);
            const long long term = round_to_ll((mqdt * (double)t) * scale);
            if (term) atomicAdd(&acc[support_index<D>(g, cw, i)], (unsigned long long)term);
        }
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lxs[d][p] = lx[d];
        lvs[d][p] = lvt[d];
    }
}

// Reverse of the deposit for one species: lx += w_dep grad(lrho)(x).  Also
// writes per-block maxima of sum_e |lv_e + dt lx_e| (the bound the next
// reverse push needs for its fixed-point scale).
template <typename T, typename A, int D>
__global__ __launch_bounds__(kA) void k_adj_gather(DevGrid g, const T* __restrict__ x0,
                                                   const T* __restrict__ x1,
                                                   const T* __restrict__ x2, A* __restrict__ lx0,
                                                   A* __restrict__ lx1, A* __restrict__ lx2,
                                                   const A* __restrict__ lv0,
                                                   const A* __restrict__ lv1,
                                                   const A* __restrict__ lv2, int64_t n,
                                                   const A* __restrict__ lrho, A w_dep, A dt,
                                                   double* __restrict__ bmax)
{
    __shared__ double wm[kA / 64];
    const int64_t p = (int64_t)blockIdx.x * kA + threadIdx.x;
    const T* xs[3] = {x0, x1, x2};
    A* lxs[3] = {lx0, lx1, lx2};
    const A* lvs[3] = {lv0, lv1, lv2};
    double m = 0.0;
    if (p < n) {
        A x[D];
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = (A)xs[d][p];
        if (isfinite((double)x[0])) {
            int cw[D];
            A W[D][2], G[D][2];
            A lx[D];
#pragma unroll
            for (int d = 0; d < D; ++d) lx[d] = lxs[d][p];
            if (cic_factors<A, D>(g, x, cw, W, G)) {
                A lr[1 << D];
#pragma unroll
                for (int i = 0; i < (1 << D); ++i) lr[i] = lrho[support_index<D>(g, cw, i)];
#pragma unroll
                for (int f = 0; f < D; ++f) {
                    A s = (A)0;
#pragma unroll
                    for (int i = 0; i < (1 << D); ++i) s = s + lr[i] * cic_grad_coef<A, D>(W, G, i, f);
                    lx[f] = lx[f] + w_dep * s;
                }
#pragma unroll
                for (int d = 0; d < D; ++d) lxs[d][p] = lx[d];
            }
#pragma unroll
            for (int d = 0; d < D; ++d) m += fabs((double)lvs[d][p] + (double)dt * (double)lx[d]);
        }
    }
    m = wave_max_d(m);
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < kA / 64; ++w) r = fmax(r, wm[w]);
        bmax[blockIdx.x] = r;
    }
}

// max over nb block maxima, then scale = 2^(40 - e) with 2^(e-1) <= bound < 2^e,
// bound = max |qm dt| * max_d delta_r * max sum_e |lvt_e| (the largest term);
// scale_out[0] = scale, scale_out[1] = 1/scale
__global__ __launch_bounds__(kA) void k_adj_scale(const double* __restrict__ bmax, int64_t nb,
                                                  double factor, double* __restrict__ scale_out,
                                                  int accumulate)
{
    __shared__ double sh[kA];
    double m = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kA) m = fmax(m, bmax[b]);
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int w = kA / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double bound = sh[0] * factor;
        if (accumulate) bound = fmax(bound, scale_out[2]);
        scale_out[2] = bound;
        int e = 0;
        frexp(bound > 0.0 ? bound : 1.0, &e);
        const double s = ldexp(1.0, 40 - e);
        scale_out[0] = s;
        scale_out[1] = 1.0 / s;
    }
}

__global__ __launch_bounds__(kA) void k_fixed_to_double_dev(const long long* __restrict__ acc,
                                                            int64_t n, const double* __restrict__ sc,
                                                            double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kA + threadIdx.x;
    if (i < n) out[i] = (double)acc[i] * sc[1];
}

// dW/dphi_i = eps dV sum over existing edges (i, j) of dr_d^2 (phi_i - phi_j)
template <int D>
__global__ __launch_bounds__(kA) void k_field_energy_grad(DevGrid g, const double* __restrict__ phi,
                                                          int64_t nn, double eps_dv,
                                                          double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kA + threadIdx.x;
    if (i >= nn) return;
    int64_t idx[3] = {0, 0, 0};
    int64_t rem = i;
    for (int d = D - 1; d >= 0; --d) {
        idx[d] = rem % g.w[d];
        rem /= g.w[d];
    }
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double dr2 = g.delta_r[d] * g.delta_r[d];
        for (int o = -1; o <= 1; o += 2) {
            int64_t nb = idx[d] + o;
            if (nb < 0 || nb >= g.w[d]) {
                if (!g.periodic[d]) continue;
                nb = nb < 0 ? nb + g.w[d] : nb - g.w[d];
            }
            int64_t lin = 0;
#pragma unroll
            for (int q = 0; q < D; ++q) lin = lin * g.w[q] + (q == d ? nb : idx[q]);
            s += dr2 * (phi[i] - phi[lin]);
        }
    }
    out[i] = eps_dv * s;
}

// x_p = xi_p + sum_k a_k L/(2 pi k) sin(2 pi k xi_p / L), k = 1..nm
template <typename T>
__global__ __launch_bounds__(kA) void k_mode_displace(const T* __restrict__ xi, T* __restrict__ x,
                                                      int64_t n, const double* __restrict__ a, int nm,
                                                      double lo, double L)
{
    const int64_t p = (int64_t)blockIdx.x * kA + threadIdx.x;
    if (p >= n) return;
    const double th = 2.0 * M_PI * ((double)xi[p] - lo) / L;
    double s1, c1;
    sincos(th, &s1, &c1);
    double sk = s1, skm1 = 0.0, d = 0.0;
    for (int k = 1; k <= nm; ++k) {
        d += a[k - 1] * (L / (2.0 * M_PI * k)) * sk;
        const double sn = 2.0 * c1 * sk - skm1;     // sin((k+1) th)
        skm1 = sk;
        sk = sn;
    }
    double xv = (double)xi[p] + d;
    if (xv < lo) xv += L;
    else if (xv >= lo + L) xv -= L;
    x[p] = (T)xv;
}

// per-block partials of sum_p lx_p L/(2 pi k) sin(2 pi k xi_p / L), k = 1..nm (nm <= 64)
template <typename T>
__global__ __launch_bounds__(kA) void k_mode_project(const T* __restrict__ xi, const T* __restrict__ lx,
                                                     int64_t n, int nm, double lo, double L,
                                                     double* __restrict__ part)
{
    __shared__ double sh[kA / 64][64];
    const int64_t p = (int64_t)blockIdx.x * kA + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double sk = 0.0, skm1 = 0.0, c1 = 1.0, l = 0.0;
    if (p < n && isfinite((double)xi[p])) {
        const double th = 2.0 * M_PI * ((double)xi[p] - lo) / L;
        sincos(th, &sk, &c1);
        l = (double)lx[p];
    }
    for (int k = 1; k <= nm; ++k) {
        const double v = wave_sum(l * (L / (2.0 * M_PI * k)) * sk);
        if (lane == 0) sh[wave][k - 1] = v;
        const double sn = 2.0 * c1 * sk - skm1;
        skm1 = sk;
        sk = sn;
    }
    __syncthreads();
    if (threadIdx.x < nm) {
        double s = 0.0;
        for (int w = 0; w < kA / 64; ++w) s += sh[w][threadIdx.x];
        part[(int64_t)blockIdx.x * nm + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(kA) void k_mode_reduce(const double* __restrict__ part, int64_t nb,
                                                    int nm, double* __restrict__ out)
{
    // one thread per mode, fixed order over blocks
    const int k = threadIdx.x;
    if (k >= nm) return;
    double s = 0.0;
    for (int64_t b = 0; b < nb; ++b) s += part[b * nm + k];
    out[k] = s;
}

template <typename T, typename A, int D>
psum_status launch_adj_push(const psum_grid* g, void* const* x, void* const* lx, void* const* lv,
                            int64_t n, const void* phi, double qm, double dt,
                            const double* scale_dev, int64_t* acc, hipStream_t st)
{
    const DevGrid dg = to_dev(*g);
    auto P = [&](void* const* a, int d) { return d < D ? static_cast<T*>(a[d]) : nullptr; };
    auto Q = [&](void* const* a, int d) { return d < D ? static_cast<A*>(a[d]) : nullptr; };
    hipLaunchKernelGGL((k_adj_push<T, A, D>), dim3(nblk(n, kA)), dim3(kA), 0, st, dg, P(x, 0), P(x, 1),
                       P(x, 2), Q(lx, 0), Q(lx, 1), Q(lx, 2), Q(lv, 0), Q(lv, 1), Q(lv, 2), n,
                       static_cast<const A*>(phi), (A)qm, (A)dt, scale_dev,
                       reinterpret_cast<unsigned long long*>(acc));
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

template <typename T, typename A, int D>
psum_status launch_adj_gather(const psum_grid* g, void* const* x, void* const* lx, void* const* lv,
                              int64_t n, const void* lrho, double w_dep, double dt, double* bmax,
                              hipStream_t st)
{
    const DevGrid dg = to_dev(*g);
    auto P = [&](void* const* a, int d) { return d < D ? static_cast<T*>(a[d]) : nullptr; };
    auto Q = [&](void* const* a, int d) { return d < D ? static_cast<A*>(a[d]) : nullptr; };
    hipLaunchKernelGGL((k_adj_gather<T, A, D>), dim3(nblk(n, kA)), dim3(kA), 0, st, dg, P(x, 0),
                       P(x, 1), P(x, 2), Q(lx, 0), Q(lx, 1), Q(lx, 2), Q(lv, 0), Q(lv, 1), Q(lv, 2),
                       n, static_cast<const A*>(lrho), (A)w_dep, (A)dt, bmax);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // namespace

extern "C" {

psum_status psum_hip_adjoint_push(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                  void* const* lv, int64_t n, const void* phi, double qm, double dt,
                                  const double* scale_dev, int64_t* acc, void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g) && g->dim >= 2, "psum_hip_adjoint_push: bad grid");
    PSUM_ARG_CHECK(n == 0 || (x && lx && lv && phi && scale_dev && acc), "psum_hip_adjoint_push: null pointer");
    if (n == 0) return PSUM_OK;
    if (dtype == PSUM_F64)
        return g->dim == 2 ? launch_adj_push<double, double, 2>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream))
                           : launch_adj_push<double, double, 3>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream));
    if (dtype == PSUM_F32)
        return g->dim == 2 ? launch_adj_push<float, float, 2>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream))
                           : launch_adj_push<float, float, 3>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream));
    if (dtype == PSUM_F32_ADJ64)
        return g->dim == 2 ? launch_adj_push<float, double, 2>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream))
                           : launch_adj_push<float, double, 3>(g, x, lx, lv, n, phi, qm, dt, scale_dev, acc, S(stream));
    set_error("psum_hip_adjoint_push: bad dtype");
    return PSUM_ERR_ARG;
}

int64_t psum_hip_adjoint_blocks(int64_t n) { return (n + kA - 1) / kA; }

psum_status psum_hip_adjoint_gather(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                    void* const* lv, int64_t n, const void* lrho, double w_dep,
                                    double dt, double* bmax, void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g) && g->dim >= 2, "psum_hip_adjoint_gather: bad grid");
    PSUM_ARG_CHECK(n == 0 || (x && lx && lv && lrho && bmax), "psum_hip_adjoint_gather: null pointer");
    if (n == 0) return PSUM_OK;
    if (dtype == PSUM_F64)
        return g->dim == 2 ? launch_adj_gather<double, double, 2>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream))
                           : launch_adj_gather<double, double, 3>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream));
    if (dtype == PSUM_F32)
        return g->dim == 2 ? launch_adj_gather<float, float, 2>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream))
                           : launch_adj_gather<float, float, 3>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream));
    if (dtype == PSUM_F32_ADJ64)
        return g->dim == 2 ? launch_adj_gather<float, double, 2>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream))
                           : launch_adj_gather<float, double, 3>(g, x, lx, lv, n, lrho, w_dep, dt, bmax, S(stream));
    set_error("psum_hip_adjoint_gather: bad dtype");
    return PSUM_ERR_ARG;
}

psum_status psum_hip_adjoint_scale(const double* bmax, int64_t nb, double factor, double* scale_dev,
                                   int accumulate, void* stream)
{
    hipLaunchKernelGGL(k_adj_scale, dim3(1), dim3(kA), 0, S(stream), bmax, nb, factor, scale_dev,
                       accumulate);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_fixed_to_double_dev(const int64_t* acc, int64_t nn, const double* scale_dev,
                                         double* out, void* stream)
{
    if (nn == 0) return PSUM_OK;
    hipLaunchKernelGGL(k_fixed_to_double_dev, dim3(nblk(nn, kA)), dim3(kA), 0, S(stream),
                       reinterpret_cast<const long long*>(acc), nn, scale_dev, out);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_field_energy_grad(const psum_grid* g, const double* phi, double eps, double* out,
                                       void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g) && phi && out, "psum_hip_field_energy_grad: bad argument");
    const int64_t nn = grid_nodes(*g);
    double vol = 1.0;
    for (int d = 0; d < g->dim; ++d) vol *= g->delta[d];
    const DevGrid dg = to_dev(*g);
    if (g->dim == 2)
        hipLaunchKernelGGL(k_field_energy_grad<2>, dim3(nblk(nn, kA)), dim3(kA), 0, S(stream), dg, phi, nn, eps * vol, out);
    else if (g->dim == 3)
        hipLaunchKernelGGL(k_field_energy_grad<3>, dim3(nblk(nn, kA)), dim3(kA), 0, S(stream), dg, phi, nn, eps * vol, out);
    else
        hipLaunchKernelGGL(k_field_energy_grad<1>, dim3(nblk(nn, kA)), dim3(kA), 0, S(stream), dg, phi, nn, eps * vol, out);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_mode_displace(int dtype, const void* xi, void* x, int64_t n, const double* a_dev,
                                   int nm, double lo, double L, void* stream)
{
    PSUM_ARG_CHECK(nm >= 0 && nm <= 64, "psum_hip_mode_displace: 0..64 modes");
    if (n == 0) return PSUM_OK;
    if (dtype == PSUM_F64)
        hipLaunchKernelGGL(k_mode_displace<double>, dim3(nblk(n, kA)), dim3(kA), 0, S(stream),
                           static_cast<const double*>(xi), static_cast<double*>(x), n, a_dev, nm, lo, L);
    else
        hipLaunchKernelGGL(k_mode_displace<float>, dim3(nblk(n, kA)), dim3(kA), 0, S(stream),
                           static_cast<const float*>(xi), static_cast<float*>(x), n, a_dev, nm, lo, L);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

size_t psum_hip_mode_project_workspace(int64_t n, int nm)
{
    return (size_t)((n + kA - 1) / kA) * (size_t)(nm > 0 ? nm : 1) * sizeof(double);
}

psum_status psum_hip_mode_project(int dtype, const void* xi, const void* lx, int64_t n, int nm,
                                  double lo, double L, double* out_dev, void* work, void* stream)
{
    PSUM_ARG_CHECK(nm >= 1 && nm <= 64 && out_dev && work, "psum_hip_mode_project: bad argument");
    const int64_t nb = (n + kA - 1) / kA;
    if (nb == 0) return PSUM_OK;
    double* part = static_cast<double*>(work);
    if (dtype == PSUM_F64)
        hipLaunchKernelGGL(k_mode_project<double>, dim3((unsigned)nb), dim3(kA), 0, S(stream),
                           static_cast<const double*>(xi), static_cast<const double*>(lx), n, nm, lo, L, part);
    else
        hipLaunchKernelGGL(k_mode_project<float>, dim3((unsigned)nb), dim3(kA), 0, S(stream),
                           static_cast<const float*>(xi), static_cast<const float*>(lx), n, nm, lo, L, part);
    hipLaunchKernelGGL(k_mode_reduce, dim3(1), dim3(64), 0, S(stream), part, nb, nm, out_dev);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // extern "C"
