// Geometric-transfer Galerkin multigrid for the FD Poisson problem on
// MI355X, served through PSuM's field-solver plugin ABI.
//
// Algorithm (the reference's, restated; PSuM paths):
//   transfers  P = tensor product of 1-D linear interpolation on index axes
//              that keep every other node and the last one; R = P^T scaled
//              by fine/coarse lumped volumes
//              (field_solver/implements/multigrid/multigrid_utilities.hpp:28-104)
//   setup      A0^ = D0^-1 A, A(l+1)^ = D(l+1)^-1 (R A(l)^ P), relax = I - 2/3 A^,
//              dense inverse of the coarsest A^      (multigrid_core.hpp:49-150)
//   V-cycle    weighted Jacobi pre/post smoothing, residual, restrict + D^-1,
//              prolong-add, coarse dense solve        (multigrid_core.hpp:163-272)
//   backends   sycl_multigrid_gpu.cpp:22-230 (FP64) and
//              sycl_multigrid32f_gpu.cpp:23-266 (FP32 V-cycles inside an FP64
//              residual refinement loop; here the coarse solve stays FP64)
//
// MI355X form: the matrices never exist on the device.  Every level is a
// structured node grid (n0 x n1 [x n2], last dimension fastest) whose rows
// are 3^D-point stencils; rows with identical coefficients share a class,
// so a level is a uint16 class id per node plus a small coefficient table.
// Level 0 of a 5/7-point matrix uses the compact (axis-only) stencil kernel.
// Transfers are per-axis weight tables (<= 3 fine nodes per coarse node and
// <= 2 coarse nodes per fine node along each axis).

#include <algorithm>
#include <array>
#include <type_traits>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"
#include "mg_batch.hpp"
#include "psum_field_solver.h"

using namespace psum_hip;

namespace {

constexpr int kMgThreads = 256;

#define MG_HIP(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw std::runtime_error(std::string("hip_multigrid: ") + #expr + ": " +       \
                                     hipGetErrorString(e_));                                \
    } while (0)

template <typename F>
void parallel_for(int64_t n, F&& f)
{
    unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    const int64_t chunk = (n + nt - 1) / nt;
    host_threads(nt, [&](unsigned t) {
        const int64_t b = (int64_t)t * chunk, e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) f(i);
    });
}

// ------------------------------------------------------------ device data

struct DevLevel {
    int n[3];
    int wrap[3];
    int64_t N;
    const uint16_t* cls;
};

template <typename T>
struct DevTable {
    const T* coef;    // ncls * KS
    const T* dinv;    // ncls
};

template <typename T>
struct DevTransfer {
    const int* pc[3];     // fine f -> lower coarse index
    const T* pw[3];       // fine f -> weights (2 per f)
    const int* rf[3];     // coarse c -> first fine index
    const int* rn[3];     // coarse c -> number of fine nodes (1..3)
    const T* rw[3];       // coarse c -> weights (3 per c)
    int wrap[3];          // periodic axis: indices wrap modulo the axis length
};

__device__ __forceinline__ int wrap_index(int i, int n, int wrap)
{
    if (!wrap) return i;
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// node coordinates of linear index i (row-major, last axis fastest); 32-bit
// unsigned division when the level has fewer than 2^32 nodes (64-bit
// division is a long software sequence on the GPU)
template <int D>
__device__ __forceinline__ void node_coords(const DevLevel& L, int64_t i, int* c)
{
    if (L.N <= (int64_t)0xffffffffll) {
        uint32_t u = (uint32_t)i;
        if (D == 3) {
            const uint32_t q = u / (uint32_t)L.n[2];
            c[2] = (int)(u - q * (uint32_t)L.n[2]);
            u = q;
        }
        const uint32_t q = u / (uint32_t)L.n[1];
        c[1] = (int)(u - q * (uint32_t)L.n[1]);
        c[0] = (int)q;
        return;
    }
    if (D == 3) {
        c[2] = (int)(i % L.n[2]);
        i /= L.n[2];
    }
    c[1] = (int)(i % L.n[1]);
    c[0] = (int)(i / L.n[1]);
}

__device__ __forceinline__ int nb_index(int c, int o, int n, int wrap)
{
    int q = c + o;
    if (q < 0) q = wrap ? q + n : 0;
    else if (q >= n) q = wrap ? q - n : n - 1;
    return q;
}

// Workgroups are dealt to the eight XCDs round-robin (workgroup b runs on
// XCD b % 8).  xcd_block() renumbers them so that each XCD works through one
// contiguous eighth of the index range: on 3D levels the neighbour planes a
// stencil reads (n1*n2 nodes apart) then stay in the same XCD's L2.  Blocks
// of an incomplete last round keep their index, so the map is a bijection
// for any grid size.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ int64_t xcd_block()
{
    const unsigned g = gridDim.x, b = blockIdx.x;
    const unsigned full = g / kXcds * kXcds;
    if (b >= full) return b;
    return (int64_t)(b % kXcds) * (full / kXcds) + b / kXcds;
}

template <int D>
__device__ __forceinline__ int64_t level_block()
{
    return D == 3 ? xcd_block() : (int64_t)blockIdx.x;
}

// sum_k coef[k] * x[nb_k], k ascending over the 3^D offsets (o0 slowest),
// COMPACT: only the 2D+1 axis offsets (all other coefficients are zero).
template <typename T, int D, bool COMPACT>
__device__ __forceinline__ T stencil_apply(const DevLevel& L, const T* __restrict__ coef,
                                          const T* __restrict__ x, const int* c)
{
    T s = (T)0;
    if (D == 2) {
        const int n0 = L.n[0], n1 = L.n[1];
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                if (COMPACT && a != 0 && b != 0) continue;
                const int k = (a + 1) * 3 + (b + 1);
                const int q0 = nb_index(c[0], a, n0, L.wrap[0]);
                const int q1 = nb_index(c[1], b, n1, L.wrap[1]);
                s = s + coef[k] * x[(int64_t)q0 * n1 + q1];
            }
    } else {
        const int n0 = L.n[0], n1 = L.n[1], n2 = L.n[2];
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b)
#pragma unroll
                for (int e = -1; e <= 1; ++e) {
                    if (COMPACT && ((a != 0) + (b != 0) + (e != 0)) > 1) continue;
                    const int k = (a + 1) * 9 + (b + 1) * 3 + (e + 1);
                    const int q0 = nb_index(c[0], a, n0, L.wrap[0]);
                    const int q1 = nb_index(c[1], b, n1, L.wrap[1]);
                    const int q2 = nb_index(c[2], e, n2, L.wrap[2]);
                    s = s + coef[k] * x[((int64_t)q0 * n1 + q1) * n2 + q2];
                }
    }
    return s;
}

// x_out = relax x_in + b * (2/3)   (multiply_add(relax, x, b, buf, 2/3, 1));
// ACC: also xacc += x_out in FP64 (the mixed refinement's x += e, fused into
// the last level-0 sweep)
template <typename T, int D, bool COMPACT, bool ACC>
__device__ __forceinline__ void smooth_node(const DevLevel& L, const DevTable<T>& relax,
                                            const T* __restrict__ xin, const T* __restrict__ b,
                                            T* __restrict__ xout, T two3, double* __restrict__ xacc,
                                            int64_t i)
{
    int c[3];
    node_coords<D>(L, i, c);
    const T* coef = relax.coef + (int64_t)L.cls[i] * (D == 2 ? 9 : 27);
    const T t = stencil_apply<T, D, COMPACT>(L, coef, xin, c);
    const T v = t * (T)1 + b[i] * (two3 * (T)1);
    xout[i] = v;
    if (ACC) xacc[i] += (double)v;
}

template <typename T, int D, bool COMPACT, bool ACC = false>
__global__ __launch_bounds__(kMgThreads) void k_smooth(DevLevel L, DevTable<T> relax,
                                                       const T* __restrict__ xin,
                                                       const T* __restrict__ b,
                                                       T* __restrict__ xout, T two3,
                                                       double* __restrict__ xacc = nullptr)
{
    const int64_t i = level_block<D>() * kMgThreads + threadIdx.x;
    if (i >= L.N) return;
    smooth_node<T, D, COMPACT, ACC>(L, relax, xin, b, xout, two3, xacc, i);
}

// r = (A^ x)*(-1) + b*((-1)*(-1))   (multiply_add(A, x, b, buf, -1, -1))
template <typename T, int D, bool COMPACT>
__device__ __forceinline__ void residual_node(const DevLevel& L, const DevTable<T>& A, const T* __restrict__ x,
                                              const T* __restrict__ b, T* __restrict__ r, int64_t i)
{
    int c[3];
    node_coords<D>(L, i, c);
    const T* coef = A.coef + (int64_t)L.cls[i] * (D == 2 ? 9 : 27);
    const T t = stencil_apply<T, D, COMPACT>(L, coef, x, c);
    r[i] = t * (T)-1 + b[i] * (T)1;
}

template <typename T, int D, bool COMPACT>
__global__ __launch_bounds__(kMgThreads) void k_residual(DevLevel L, DevTable<T> A,
                                                         const T* __restrict__ x,
                                                         const T* __restrict__ b,
                                                         T* __restrict__ r)
{
    const int64_t i = level_block<D>() * kMgThreads + threadIdx.x;
    if (i >= L.N) return;
    residual_node<T, D, COMPACT>(L, A, x, b, r, i);
}

// ---- 3D levels in chunks of kLine consecutive nodes ------------------------
// On a 3D level nearly every node lies inside the box and carries the same
// stencil class.  A chunk of kLine nodes (linear indices kLine*t ..
// kLine*t + kLine-1) all on one grid line, strictly inside the box, all of
// that class, is updated by one thread with the class row in registers and
// vector loads of the line and its neighbour lines (four for a 7-point
// level, eight for a 27-point one): no coordinate division per node, no
// clamping, no per-node table reads.  The per-node terms and their order
// are stencil_apply's, so the result is bit-identical to k_smooth /
// k_residual.  The other nodes (box faces, line ends, other classes, the
// tail) are listed when the level is built and updated by the per-node
// kernels over that list.
constexpr int kLine = 4;

struct LineGeom {
    int n0, n1, n2;
    uint32_t nchunks;          // full chunks: N / kLine
    const uint16_t* cls;
    int cls_in;                // the class the chunk kernels handle
};

// the chunk of nodes i .. i+kLine-1 (i = kLine*t) is a chunk-kernel chunk
// (the host lists the others with line_chunk_fast_host, the same test)
__device__ __forceinline__ bool line_chunk_fast(const LineGeom& G, uint32_t i)
{
    const uint32_t q = i / (uint32_t)G.n2;
    const uint32_t z = i - q * (uint32_t)G.n2;
    const uint32_t x = q / (uint32_t)G.n1;
    const uint32_t y = q - x * (uint32_t)G.n1;
    if (z < 1u || z + kLine > (uint32_t)G.n2 - 1u || y < 1u || y + 2u > (uint32_t)G.n1 || x < 1u ||
        x + 2u > (uint32_t)G.n0)
        return false;
    unsigned long long c4;
    __builtin_memcpy(&c4, G.cls + i, sizeof c4);
    const unsigned long long in = (unsigned long long)(uint16_t)G.cls_in;
    return c4 == (in | in << 16 | in << 32 | in << 48);
}

template <typename T>
__device__ __forceinline__ void ld_line(const T* __restrict__ p, T* v)
{
    struct alignas(sizeof(T)) Q { T a[kLine]; } q;
    __builtin_memcpy(&q, p, sizeof q);
#pragma unroll
    for (int k = 0; k < kLine; ++k) v[k] = q.a[k];
}

template <typename T>
__device__ __forceinline__ void st_line(T* __restrict__ p, const T* v)
{
    struct alignas(sizeof(T)) Q { T a[kLine]; } q;
#pragma unroll
    for (int k = 0; k < kLine; ++k) q.a[k] = v[k];
    __builtin_memcpy(p, &q, sizeof q);
}

// stencil_apply<T, 3, COMPACT> at the chunk's nodes: the 7-point sum in
// offset order x-1, y-1, z-1, centre, z+1, y+1, x+1, or the 27-point sum
// with offsets (a, b, e) in x, y, z ascending, z fastest
template <typename T, bool COMPACT>
__device__ __forceinline__ void line_apply(const LineGeom& G, const T* __restrict__ row,
                                           const T* __restrict__ x, uint32_t i, T* s)
{
    const uint32_t sy = (uint32_t)G.n2, sx = (uint32_t)G.n1 * (uint32_t)G.n2;
    if constexpr (COMPACT) {
        const T cxm = row[4], cym = row[10], czm = row[12], cc = row[13], czp = row[14], cyp = row[16],
                cxp = row[22];
        T xm[kLine], ym[kLine], zz[kLine + 2], yp[kLine], xp[kLine];
        ld_line(x + (i - sx), xm);
        ld_line(x + (i - sy), ym);
        zz[0] = x[i - 1];
        ld_line(x + i, zz + 1);
        zz[kLine + 1] = x[i + kLine];
        ld_line(x + (i + sy), yp);
        ld_line(x + (i + sx), xp);
#pragma unroll
        for (int k = 0; k < kLine; ++k) {
            T t = (T)0;
            t = t + cxm * xm[k];
            t = t + cym * ym[k];
            t = t + czm * zz[k];
            t = t + cc * zz[k + 1];
            t = t + czp * zz[k + 2];
            t = t + cyp * yp[k];
            t = t + cxp * xp[k];
            s[k] = t;
        }
    } else {
#pragma unroll
        for (int k = 0; k < kLine; ++k) s[k] = (T)0;
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                const T* ln = x + (i + (uint32_t)((int)sx * a + (int)sy * b));
                T zz[kLine + 2];
                zz[0] = ln[-1];
                ld_line(ln, zz + 1);
                zz[kLine + 1] = ln[kLine];
#pragma unroll
                for (int e = -1; e <= 1; ++e) {
                    const T c = row[(a + 1) * 9 + (b + 1) * 3 + (e + 1)];
#pragma unroll
                    for (int k = 0; k < kLine; ++k) s[k] = s[k] + c * zz[k + 1 + e];
                }
            }
    }
}

// A launch takes nbc blocks of chunks and nblk(n) blocks of listed nodes
// (one per thread, with the per-node kernels' code).  Workgroup b runs on
// XCD x = b % kXcds as that XCD's j = b / kXcds-th: j < lb takes listed block
// x*lb + j, later ones chunk block x*cb + (j - lb) (lb, cb: the listed and
// chunk blocks per XCD, rounded up; indexes past the end exit).  Every XCD
// thus gets an eighth of the listed nodes, whose per-node path is slower,
// started first, and a contiguous eighth of the chunks.
struct LineList {
    const uint32_t* list;
    uint32_t n;
    uint32_t nbc;              // blocks of chunks
};

__host__ __device__ __forceinline__ uint32_t line_listed_blocks(const LineList& LL)
{
    return (LL.n + kMgThreads - 1) / kMgThreads;
}
__host__ __device__ __forceinline__ uint32_t line_grid(const LineList& LL)
{
    const uint32_t lb = (line_listed_blocks(LL) + kXcds - 1) / kXcds, cb = (LL.nbc + kXcds - 1) / kXcds;
    return kXcds * (lb + cb);
}
// this workgroup's block: listed (true) or chunk, and its index (-1: none)
__device__ __forceinline__ int64_t line_block(const LineList& LL, bool* listed)
{
    const uint32_t nbl = line_listed_blocks(LL);
    const uint32_t lb = (nbl + kXcds - 1) / kXcds, cb = (LL.nbc + kXcds - 1) / kXcds;
    const uint32_t x = blockIdx.x % kXcds, j = blockIdx.x / kXcds;
    if (j < lb) {
        *listed = true;
        const uint32_t k = x * lb + j;
        return k < nbl ? (int64_t)k : -1;
    }
    *listed = false;
    const uint32_t k = x * cb + (j - lb);
    return k < LL.nbc ? (int64_t)k : -1;
}

template <typename T, bool COMPACT, bool ACC>
__global__ __launch_bounds__(kMgThreads) void k_smooth_line(LineGeom G, LineList LL, DevLevel L, DevTable<T> relax,
                                                            const T* __restrict__ row,
                                                            const T* __restrict__ xin,
                                                            const T* __restrict__ b,
                                                            T* __restrict__ xout, T two3,
                                                            double* __restrict__ xacc)
{
    bool listed;
    const int64_t blk = line_block(LL, &listed);
    if (blk < 0) return;
    if (listed) {
        const int64_t t = blk * kMgThreads + threadIdx.x;
        if (t < (int64_t)LL.n) smooth_node<T, 3, COMPACT, ACC>(L, relax, xin, b, xout, two3, xacc, (int64_t)LL.list[t]);
        return;
    }
    const int64_t t = blk * kMgThreads + threadIdx.x;
    if (t >= (int64_t)G.nchunks) return;
    const uint32_t i = (uint32_t)t * kLine;
    if (!line_chunk_fast(G, i)) return;
    T s[kLine], bb[kLine], v[kLine];
    line_apply<T, COMPACT>(G, row, xin, i, s);
    ld_line(b + i, bb);
#pragma unroll
    for (int k = 0; k < kLine; ++k) v[k] = s[k] * (T)1 + bb[k] * (two3 * (T)1);
    st_line(xout + i, v);
    if (ACC) {
        double a[kLine];
        ld_line(xacc + i, a);
#pragma unroll
        for (int k = 0; k < kLine; ++k) a[k] += (double)v[k];
        st_line(xacc + i, a);
    }
}

template <typename T, bool COMPACT>
__global__ __launch_bounds__(kMgThreads) void k_residual_line(LineGeom G, LineList LL, DevLevel L, DevTable<T> A,
                                                              const T* __restrict__ row,
                                                              const T* __restrict__ x,
                                                              const T* __restrict__ b, T* __restrict__ r)
{
    bool listed;
    const int64_t blk = line_block(LL, &listed);
    if (blk < 0) return;
    if (listed) {
        const int64_t t = blk * kMgThreads + threadIdx.x;
        if (t < (int64_t)LL.n) residual_node<T, 3, COMPACT>(L, A, x, b, r, (int64_t)LL.list[t]);
        return;
    }
    const int64_t t = blk * kMgThreads + threadIdx.x;
    if (t >= (int64_t)G.nchunks) return;
    const uint32_t i = (uint32_t)t * kLine;
    if (!line_chunk_fast(G, i)) return;
    T s[kLine], bb[kLine], v[kLine];
    line_apply<T, COMPACT>(G, row, x, i, s);
    ld_line(b + i, bb);
#pragma unroll
    for (int k = 0; k < kLine; ++k) v[k] = s[k] * (T)-1 + bb[k] * (T)1;
    st_line(r + i, v);
}

// b_c = (R r) o D_c^-1 ; x_c = 0, or, when the coarse level is pre-smoothed,
// the first Jacobi sweep from x_c = 0 fused in: relax*0 + b_c*(2/3) is
// exactly b_c * (2/3) (first_sweep = 2/3; 0 means plain x_c = 0).
template <typename T, int D>
__global__ __launch_bounds__(kMgThreads) void k_restrict(DevLevel F, DevLevel Cl,
                                                         DevTransfer<T> tr, DevTable<T> Ac,
                                                         const T* __restrict__ r,
                                                         T* __restrict__ bc, T* __restrict__ xc,
                                                         T first_sweep)
{
    const int64_t i = level_block<D>() * kMgThreads + threadIdx.x;
    if (i >= Cl.N) return;
    int c[3];
    node_coords<D>(Cl, i, c);
    T s = (T)0;
    const int f0 = tr.rf[0][c[0]], m0 = tr.rn[0][c[0]];
    const int f1 = tr.rf[1][c[1]], m1 = tr.rn[1][c[1]];
    if (D == 2) {
        for (int a = 0; a < m0; ++a) {
            const T w0 = tr.rw[0][3 * c[0] + a];
            const int i0 = wrap_index(f0 + a, F.n[0], tr.wrap[0]);
            for (int b = 0; b < m1; ++b) {
                const T w = w0 * tr.rw[1][3 * c[1] + b];
                const int i1 = wrap_index(f1 + b, F.n[1], tr.wrap[1]);
                s = s + w * r[(int64_t)i0 * F.n[1] + i1];
            }
        }
    } else {
        const int f2 = tr.rf[2][c[2]], m2 = tr.rn[2][c[2]];
        for (int a = 0; a < m0; ++a) {
            const T w0 = tr.rw[0][3 * c[0] + a];
            const int i0 = wrap_index(f0 + a, F.n[0], tr.wrap[0]);
            for (int b = 0; b < m1; ++b) {
                const T w01 = w0 * tr.rw[1][3 * c[1] + b];
                const int i1 = wrap_index(f1 + b, F.n[1], tr.wrap[1]);
                const int64_t row = ((int64_t)i0 * F.n[1] + i1) * F.n[2];
                for (int e = 0; e < m2; ++e) {
                    const T w = w01 * tr.rw[2][3 * c[2] + e];
                    s = s + w * r[row + wrap_index(f2 + e, F.n[2], tr.wrap[2])];
                }
            }
        }
    }
    const T bv = s * Ac.dinv[Cl.cls[i]];
    bc[i] = bv;
    xc[i] = first_sweep != (T)0 ? bv * (first_sweep * (T)1) : (T)0;
}

// x_f = (P x_c)*1 + x_f*(1*1)
template <typename T, int D>
__global__ __launch_bounds__(kMgThreads) void k_prolong_add(DevLevel F, DevLevel Cl,
                                                            DevTransfer<T> tr,
                                                            const T* __restrict__ xc,
                                                            T* __restrict__ xf)
{
    const int64_t i = level_block<D>() * kMgThreads + threadIdx.x;
    if (i >= F.N) return;
    int c[3];
    node_coords<D>(F, i, c);
    T s = (T)0;
    const int p0 = tr.pc[0][c[0]], p1 = tr.pc[1][c[1]];
    const T* w0 = tr.pw[0] + 2 * c[0];
    const T* w1 = tr.pw[1] + 2 * c[1];
    if (D == 2) {
        for (int a = 0; a < 2; ++a) {
            if (w0[a] == (T)0) continue;
            const int i0 = wrap_index(p0 + a, Cl.n[0], tr.wrap[0]);
            for (int b = 0; b < 2; ++b) {
                if (w1[b] == (T)0) continue;
                const int i1 = wrap_index(p1 + b, Cl.n[1], tr.wrap[1]);
                s = s + (w0[a] * w1[b]) * xc[(int64_t)i0 * Cl.n[1] + i1];
            }
        }
    } else {
        const int p2 = tr.pc[2][c[2]];
        const T* w2 = tr.pw[2] + 2 * c[2];
        for (int a = 0; a < 2; ++a) {
            if (w0[a] == (T)0) continue;
            const int i0 = wrap_index(p0 + a, Cl.n[0], tr.wrap[0]);
            for (int b = 0; b < 2; ++b) {
                if (w1[b] == (T)0) continue;
                const T w01 = w0[a] * w1[b];
                const int i1 = wrap_index(p1 + b, Cl.n[1], tr.wrap[1]);
                for (int e = 0; e < 2; ++e) {
                    if (w2[e] == (T)0) continue;
                    const int i2 = wrap_index(p2 + e, Cl.n[2], tr.wrap[2]);
                    s = s + (w01 * w2[e]) * xc[((int64_t)i0 * Cl.n[1] + i1) * Cl.n[2] + i2];
                }
            }
        }
    }
    xf[i] = s * (T)1 + xf[i] * (T)1;
}

// ---- 3D transfers over planes ---------------------------------------------
// k_prolong_add / k_restrict on 3D levels with a two-dimensional launch:
// blockIdx.y is the output node's slowest coordinate, so its transfer table
// entries are uniform across the block, and one division splits the rest of
// the index.  Each node's terms and their order are the per-node kernels',
// so the bits are the same.
// index p + o on an axis of n nodes, wrapped on a periodic axis, else kept
// in range (the caller selects the term away when p + o is outside)
__device__ __forceinline__ uint32_t axis_index(int p, int n, int wrap)
{
    int q = wrap_index(p, n, wrap);
    q = q < 0 ? 0 : (q >= n ? n - 1 : q);
    return (uint32_t)q;
}

template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_prolong_add_plane(DevLevel F, DevLevel Cl, DevTransfer<T> tr,
                                                                  const T* __restrict__ xc, T* __restrict__ xf)
{
    const uint32_t j = blockIdx.x * kMgThreads + threadIdx.x;
    const uint32_t plane = (uint32_t)F.n[1] * (uint32_t)F.n[2];
    if (j >= plane) return;
    const int c0 = blockIdx.y;
    const int c1 = (int)(j / (uint32_t)F.n[2]);
    const int c2 = (int)(j - (uint32_t)c1 * (uint32_t)F.n[2]);
    const int64_t i = (int64_t)c0 * plane + j;
    const int p0 = tr.pc[0][c0], p1 = tr.pc[1][c1], p2 = tr.pc[2][c2];
    T w0[2], w1[2], w2[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        w0[a] = tr.pw[0][2 * c0 + a];
        w1[a] = tr.pw[1][2 * c1 + a];
        w2[a] = tr.pw[2][2 * c2 + a];
    }
    // all eight coarse values are loaded before any is used (addresses kept
    // in range); a term whose weight is zero is then skipped by a select, so
    // the sum is k_prolong_add's, term for term
    const uint32_t cs1 = (uint32_t)Cl.n[2], cs0 = (uint32_t)Cl.n[1] * (uint32_t)Cl.n[2];
    uint32_t q0[2], q1[2], q2[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        q0[a] = axis_index(p0 + a, Cl.n[0], tr.wrap[0]) * cs0;
        q1[a] = axis_index(p1 + a, Cl.n[1], tr.wrap[1]) * cs1;
        q2[a] = axis_index(p2 + a, Cl.n[2], tr.wrap[2]);
    }
    T v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = xc[q0[k >> 2] + q1[(k >> 1) & 1] + q2[k & 1]];
    const T xo = xf[i];
    T s = (T)0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const T w01 = w0[a] * w1[b];
#pragma unroll
            for (int e = 0; e < 2; ++e) {
                const bool use = w0[a] != (T)0 && w1[b] != (T)0 && w2[e] != (T)0;
                const T t = s + (w01 * w2[e]) * v[a * 4 + b * 2 + e];
                s = use ? t : s;
            }
        }
    xf[i] = s * (T)1 + xo * (T)1;
}

// k_restrict on 3D levels: the slowest coarse coordinate in blockIdx.y and
// all (up to 27) fine residuals loaded before any is used; a term outside
// the node's support (a >= m0, ...) is skipped by a select, so the sum is
// k_restrict's, term for term.
template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_restrict_plane(DevLevel F, DevLevel Cl, DevTransfer<T> tr,
                                                               DevTable<T> Ac, const T* __restrict__ r,
                                                               T* __restrict__ bc, T* __restrict__ xc,
                                                               T first_sweep)
{
    const uint32_t j = blockIdx.x * kMgThreads + threadIdx.x;
    const uint32_t plane = (uint32_t)Cl.n[1] * (uint32_t)Cl.n[2];
    if (j >= plane) return;
    const int c0 = blockIdx.y;
    const int c1 = (int)(j / (uint32_t)Cl.n[2]);
    const int c2 = (int)(j - (uint32_t)c1 * (uint32_t)Cl.n[2]);
    const int64_t i = (int64_t)c0 * plane + j;
    const int f0 = tr.rf[0][c0], m0 = tr.rn[0][c0];
    const int f1 = tr.rf[1][c1], m1 = tr.rn[1][c1];
    const int f2 = tr.rf[2][c2], m2 = tr.rn[2][c2];
    T w0[3], w1[3], w2[3];
    uint32_t g0[3], g1[3], g2[3];
    const uint32_t fs1 = (uint32_t)F.n[2], fs0 = (uint32_t)F.n[1] * (uint32_t)F.n[2];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w0[a] = tr.rw[0][3 * c0 + a];
        w1[a] = tr.rw[1][3 * c1 + a];
        w2[a] = tr.rw[2][3 * c2 + a];
        g0[a] = axis_index(f0 + a, F.n[0], tr.wrap[0]) * fs0;
        g1[a] = axis_index(f1 + a, F.n[1], tr.wrap[1]) * fs1;
        g2[a] = axis_index(f2 + a, F.n[2], tr.wrap[2]);
    }
    T v[27];
#pragma unroll
    for (int k = 0; k < 27; ++k) v[k] = r[g0[k / 9] + g1[(k / 3) % 3] + g2[k % 3]];
    const uint16_t ci = Cl.cls[i];
    T s = (T)0;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const T w01 = w0[a] * w1[b];
#pragma unroll
            for (int e = 0; e < 3; ++e) {
                const bool use = a < m0 && b < m1 && e < m2;
                const T t = s + (w01 * w2[e]) * v[a * 9 + b * 3 + e];
                s = use ? t : s;
            }
        }
    const T bv = s * Ac.dinv[ci];
    bc[i] = bv;
    xc[i] = first_sweep != (T)0 ? bv * (first_sweep * (T)1) : (T)0;
}

// ---- V-cycle tail in one workgroup -----------------------------------------
// The coarse end of the hierarchy (levels lt..Lc, at most kMaxTail) is
// latency-bound: a handful of nodes per level, three launches per level.
// k_vtail runs all of it in one 1024-thread workgroup.  It first stages in
// LDS every tail level's class ids, class tables and transfer tables, the
// iterate and right-hand side of level lt, and then keeps every level's two
// iterates and right-hand side there: pre-sweeps, residual, restriction down
// to the coarsest level, its dense solve (the inverse is read from global
// memory), then prolongation and post-sweeps back up to level lt, whose
// iterate it writes back.  Each phase is the per-node arithmetic of
// k_smooth, k_residual, k_restrict, k_coarse_solve and k_prolong_add (2D or
// 3D) in the same order, so the result is bit-identical to the
// launch-per-phase path.
constexpr int kMaxTail = 6;
constexpr int kTailThreads = 1024;
// coarsest nodes: one compute unit reads the dense inverse (n^2 doubles)
// from memory at a few tens of GB/s, so it must stay small (2D: 10^2, 3D:
// 5^3 nodes)
template <int D>
constexpr int tail_coarse() { return D == 2 ? 100 : 125; }
// one compute unit runs the tail: 64 lanes of VALU per clock, so a 2D level
// of 64^2 nodes already costs what its own multi-block launches cost; the
// tail takes 2D levels of at most 33^2 nodes and 3D levels of at most 9^3
// (a 17^3 level's 27-point sweep takes ~20 us on one compute unit against
// ~4.4 us as a launch of its own; levels 9^3 and 5^3 of the 3D box take
// 52 us per V-cycle in the tail, 54 us in their own launches)
template <int D>
constexpr int tail_level_nodes() { return D == 2 ? 1089 : 729; }
constexpr int kTailRows = (125 + kTailThreads / 64 - 1) / (kTailThreads / 64);   // rows per wave

template <typename T>
struct TailLevel {
    int n[3], wrap[3], N, compact, ncls;
    // byte offsets into the workgroup's LDS
    int ox, oy, ob, ocls, orx, oa, odinv;
    const uint16_t* cls;
    const T* rx;          // relax rows, ncls x 3^D
    const T* a;           // scaled operator rows, ncls x 3^D
    const T* dinv;        // ncls
};

template <typename T>
struct TailTransfer {
    int wrap[3];
    // byte offsets per axis: restriction (per coarse index) and prolongation
    // (per fine index) tables
    int orf[3], orn[3], orw[3], opc[3], opw[3];
    const int *rf[3], *rn[3], *pc[3];
    const T *rw[3], *pw[3];
};

template <typename T>
struct TailArgs {
    int nl;                         // tail levels; the last is the coarsest
    int pre, post;
    T two3;
    TailLevel<T> lev[kMaxTail];
    TailTransfer<T> tr[kMaxTail - 1];
    const T* x_in;                  // level lt iterate after the restriction
    const T* b_in;                  // level lt right-hand side
    T* x_out;
    const double* cinv;
};

template <typename U>
__device__ __forceinline__ void tail_copy(unsigned char* lds, int off, const U* __restrict__ g, int n)
{
    U* d = reinterpret_cast<U*>(lds + off);
    for (int i = threadIdx.x; i < n; i += kTailThreads) d[i] = g[i];
}

// stencil_apply<T, D, COMPACT> on a tail level: sum_k coef[k] * x[nb_k] over
// the 3^D offsets (first axis slowest), clamped or wrapped neighbours;
// COMPACT: the 2D+1 axis offsets
template <typename T, int D, bool COMPACT>
__device__ __forceinline__ T tail_stencil(const T* coef, const T* x, const int* c, const TailLevel<T>& t)
{
    T s = (T)0;
    if constexpr (D == 2) {
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b) {
                if (COMPACT && a != 0 && b != 0) continue;
                const int k = (a + 1) * 3 + (b + 1);
                const int q0 = nb_index(c[0], a, t.n[0], t.wrap[0]);
                const int q1 = nb_index(c[1], b, t.n[1], t.wrap[1]);
                s = s + coef[k] * x[q0 * t.n[1] + q1];
            }
    } else {
#pragma unroll
        for (int a = -1; a <= 1; ++a)
#pragma unroll
            for (int b = -1; b <= 1; ++b)
#pragma unroll
                for (int e = -1; e <= 1; ++e) {
                    if (COMPACT && ((a != 0) + (b != 0) + (e != 0)) > 1) continue;
                    const int k = (a + 1) * 9 + (b + 1) * 3 + (e + 1);
                    const int q0 = nb_index(c[0], a, t.n[0], t.wrap[0]);
                    const int q1 = nb_index(c[1], b, t.n[1], t.wrap[1]);
                    const int q2 = nb_index(c[2], e, t.n[2], t.wrap[2]);
                    s = s + coef[k] * x[(q0 * t.n[1] + q1) * t.n[2] + q2];
                }
    }
    return s;
}

template <int D, typename T>
__device__ __forceinline__ void tail_coords(const TailLevel<T>& t, int i, int* c)
{
    if constexpr (D == 3) {
        const int q = i / t.n[2];
        c[2] = i - q * t.n[2];
        c[1] = q % t.n[1];
        c[0] = q / t.n[1];
    } else {
        c[0] = i / t.n[1];
        c[1] = i - c[0] * t.n[1];
    }
}

// out = stencil(rows, x) * m + b * f for every node of level t
template <typename T, int D>
__device__ __forceinline__ void tail_apply(unsigned char* lds, const TailLevel<T>& t, int orows, const T* x,
                                           T* out, T m, T f)
{
    constexpr int K = D == 2 ? 9 : 27;
    const uint16_t* cls = reinterpret_cast<const uint16_t*>(lds + t.ocls);
    const T* rows = reinterpret_cast<const T*>(lds + orows);
    const T* b = reinterpret_cast<const T*>(lds + t.ob);
    for (int i = threadIdx.x; i < t.N; i += kTailThreads) {
        int c[3];
        tail_coords<D>(t, i, c);
        const T* coef = rows + (int)cls[i] * K;
        const T v = t.compact ? tail_stencil<T, D, true>(coef, x, c, t) : tail_stencil<T, D, false>(coef, x, c, t);
        out[i] = v * m + b[i] * f;
    }
}

template <typename T, int D>
__global__ __launch_bounds__(kTailThreads) void k_vtail(TailArgs<T> a)
{
    extern __shared__ __align__(16) unsigned char lds[];
    constexpr int K = D == 2 ? 9 : 27;
    const int nl = a.nl, Lc = a.nl - 1;
    // stage tables, level lt's iterate and right-hand side
    for (int l = 0; l < nl; ++l) {
        const TailLevel<T>& t = a.lev[l];
        tail_copy<uint16_t>(lds, t.ocls, t.cls, t.N);
        tail_copy<T>(lds, t.orx, t.rx, t.ncls * K);
        tail_copy<T>(lds, t.oa, t.a, t.ncls * K);
        tail_copy<T>(lds, t.odinv, t.dinv, t.ncls);
        if (l < Lc) {
            const TailTransfer<T>& r = a.tr[l];
            const TailLevel<T>& c = a.lev[l + 1];
#pragma unroll
            for (int d = 0; d < D; ++d) {
                tail_copy<int>(lds, r.orf[d], r.rf[d], c.n[d]);
                tail_copy<int>(lds, r.orn[d], r.rn[d], c.n[d]);
                tail_copy<T>(lds, r.orw[d], r.rw[d], 3 * c.n[d]);
                tail_copy<int>(lds, r.opc[d], r.pc[d], t.n[d]);
                tail_copy<T>(lds, r.opw[d], r.pw[d], 2 * t.n[d]);
            }
        }
    }
    tail_copy<T>(lds, a.lev[0].ox, a.x_in, a.lev[0].N);
    tail_copy<T>(lds, a.lev[0].ob, a.b_in, a.lev[0].N);
    __syncthreads();
    // bit l: which of level l's two LDS iterates holds its current iterate
    unsigned cur = 0;
    auto X = [&](int l) {
        return reinterpret_cast<T*>(lds + (((cur >> l) & 1u) ? a.lev[l].oy : a.lev[l].ox));
    };
    auto Y = [&](int l) {
        return reinterpret_cast<T*>(lds + (((cur >> l) & 1u) ? a.lev[l].ox : a.lev[l].oy));
    };
    for (int l = 0; l < Lc; ++l) {
        const TailLevel<T>& t = a.lev[l];
        // coarse levels received their first sweep from the restriction
        const int sweeps = a.pre > 0 ? a.pre - 1 : 0;
        for (int s = 0; s < sweeps; ++s) {
            tail_apply<T, D>(lds, t, t.orx, X(l), Y(l), (T)1, a.two3 * (T)1);
            __syncthreads();
            cur ^= 1u << l;
        }
        // r = (A x)*(-1) + b*1 into the free iterate
        T* r = Y(l);
        tail_apply<T, D>(lds, t, t.oa, X(l), r, (T)-1, (T)1);
        __syncthreads();
        // b_c = (R r) o D_c^-1 ; x_c = b_c * (2/3) (first sweep) or 0
        const TailLevel<T>& cl = a.lev[l + 1];
        const TailTransfer<T>& tr = a.tr[l];
        const int *rf[3], *rn[3];
        const T* rw[3];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            rf[d] = reinterpret_cast<const int*>(lds + tr.orf[d]);
            rn[d] = reinterpret_cast<const int*>(lds + tr.orn[d]);
            rw[d] = reinterpret_cast<const T*>(lds + tr.orw[d]);
        }
        const uint16_t* ccls = reinterpret_cast<const uint16_t*>(lds + cl.ocls);
        const T* cdinv = reinterpret_cast<const T*>(lds + cl.odinv);
        const T fs = (l + 1 < Lc && a.pre > 0) ? a.two3 : (T)0;
        T* bc = reinterpret_cast<T*>(lds + cl.ob);
        T* xc = reinterpret_cast<T*>(lds + cl.ox);
        for (int i = threadIdx.x; i < cl.N; i += kTailThreads) {
            int c[3];
            tail_coords<D>(cl, i, c);
            T sum = (T)0;
            const int f0 = rf[0][c[0]], m0 = rn[0][c[0]];
            const int f1 = rf[1][c[1]], m1 = rn[1][c[1]];
            if constexpr (D == 2) {
                for (int p = 0; p < m0; ++p) {
                    const T w0 = rw[0][3 * c[0] + p];
                    const int i0 = wrap_index(f0 + p, t.n[0], tr.wrap[0]);
                    for (int q = 0; q < m1; ++q) {
                        const T w = w0 * rw[1][3 * c[1] + q];
                        const int i1 = wrap_index(f1 + q, t.n[1], tr.wrap[1]);
                        sum = sum + w * r[i0 * t.n[1] + i1];
                    }
                }
            } else {
                const int f2 = rf[2][c[2]], m2 = rn[2][c[2]];
                for (int p = 0; p < m0; ++p) {
                    const T w0 = rw[0][3 * c[0] + p];
                    const int i0 = wrap_index(f0 + p, t.n[0], tr.wrap[0]);
                    for (int q = 0; q < m1; ++q) {
                        const T w01 = w0 * rw[1][3 * c[1] + q];
                        const int i1 = wrap_index(f1 + q, t.n[1], tr.wrap[1]);
                        const int row = (i0 * t.n[1] + i1) * t.n[2];
                        for (int e = 0; e < m2; ++e) {
                            const T w = w01 * rw[2][3 * c[2] + e];
                            sum = sum + w * r[row + wrap_index(f2 + e, t.n[2], tr.wrap[2])];
                        }
                    }
                }
            }
            const T bv = sum * cdinv[ccls[i]];
            bc[i] = bv;
            xc[i] = fs != (T)0 ? bv * (fs * (T)1) : (T)0;
        }
        cur &= ~(1u << (l + 1));
        __syncthreads();
    }
    {
        // x = Ainv b on the coarsest level: one wave per row, lane-strided
        // FP64 partials and a butterfly (k_coarse_solve)
        const int n = a.lev[Lc].N;
        const T* b = reinterpret_cast<const T*>(lds + a.lev[Lc].ob);
        T* x = reinterpret_cast<T*>(lds + a.lev[Lc].ox);
        // rows wave, wave + 16, ... (at most kTailRows per wave) are summed
        // together so that all their loads are in flight at once; each row
        // still adds its terms in ascending j
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        double sacc[kTailRows];
#pragma unroll
        for (int k = 0; k < kTailRows; ++k) sacc[k] = 0.0;
        for (int j = lane; j < n; j += 64) {
            const double bj = (double)b[j];
#pragma unroll
            for (int k = 0; k < kTailRows; ++k) {
                const int row = wave + k * (kTailThreads / 64);
                if (row < n) sacc[k] += a.cinv[(int64_t)row * n + j] * bj;
            }
        }
#pragma unroll
        for (int k = 0; k < kTailRows; ++k) {
            const int row = wave + k * (kTailThreads / 64);
            const double v = wave_sum(sacc[k]);
            if (row < n && lane == 0) x[row] = (T)v;
        }
        __syncthreads();
    }
    for (int l = Lc - 1; l >= 0; --l) {
        const TailLevel<T>& t = a.lev[l];
        const TailLevel<T>& cl = a.lev[l + 1];
        const TailTransfer<T>& tr = a.tr[l];
        const int* pc[3];
        const T* pw[3];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            pc[d] = reinterpret_cast<const int*>(lds + tr.opc[d]);
            pw[d] = reinterpret_cast<const T*>(lds + tr.opw[d]);
        }
        const T* xc = X(l + 1);
        T* xf = X(l);
        // x_f = (P x_c)*1 + x_f*(1*1)
        for (int i = threadIdx.x; i < t.N; i += kTailThreads) {
            int c[3];
            tail_coords<D>(t, i, c);
            T sum = (T)0;
            const int p0 = pc[0][c[0]], p1 = pc[1][c[1]];
            const T* w0 = pw[0] + 2 * c[0];
            const T* w1 = pw[1] + 2 * c[1];
            if constexpr (D == 2) {
                for (int p = 0; p < 2; ++p) {
                    if (w0[p] == (T)0) continue;
                    const int i0 = wrap_index(p0 + p, cl.n[0], tr.wrap[0]);
                    for (int q = 0; q < 2; ++q) {
                        if (w1[q] == (T)0) continue;
                        const int i1 = wrap_index(p1 + q, cl.n[1], tr.wrap[1]);
                        sum = sum + (w0[p] * w1[q]) * xc[i0 * cl.n[1] + i1];
                    }
                }
            } else {
                const int p2 = pc[2][c[2]];
                const T* w2 = pw[2] + 2 * c[2];
                for (int p = 0; p < 2; ++p) {
                    if (w0[p] == (T)0) continue;
                    const int i0 = wrap_index(p0 + p, cl.n[0], tr.wrap[0]);
                    for (int q = 0; q < 2; ++q) {
                        if (w1[q] == (T)0) continue;
                        const T w01 = w0[p] * w1[q];
                        const int i1 = wrap_index(p1 + q, cl.n[1], tr.wrap[1]);
                        for (int e = 0; e < 2; ++e) {
                            if (w2[e] == (T)0) continue;
                            const int i2 = wrap_index(p2 + e, cl.n[2], tr.wrap[2]);
                            sum = sum + (w01 * w2[e]) * xc[(i0 * cl.n[1] + i1) * cl.n[2] + i2];
                        }
                    }
                }
            }
            xf[i] = sum * (T)1 + xf[i] * (T)1;
        }
        __syncthreads();
        for (int s = 0; s < a.post; ++s) {
            tail_apply<T, D>(lds, t, t.orx, X(l), Y(l), (T)1, a.two3 * (T)1);
            __syncthreads();
            cur ^= 1u << l;
        }
    }
    const T* x0 = X(0);
    for (int i = threadIdx.x; i < a.lev[0].N; i += kTailThreads) a.x_out[i] = x0[i];
}

// ---- batched launches ------------------------------------------------------
// A batched launch runs one kernel body for several independent solver
// instances: instance i owns blocks first[i] .. first[i+1]-1 of the grid and
// reads its arguments from args[i].  The bodies below take the block index
// as a parameter, so a single-instance launch (block = blockIdx.x) and a
// batched one execute identical arithmetic.
__device__ __forceinline__ int batch_instance(const unsigned* __restrict__ first, int n)
{
    int lo = 0, hi = n - 1;
    const unsigned b = blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (first[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// ---- fused 2D kernels: several Jacobi sweeps of one tile in LDS ----------
// A block owns a tile of kTileRows x (kTileX - 2h) output nodes and loads it
// with a halo of h = `sweeps` nodes (kTileX = 64 columns: one wave per row,
// conflict-free LDS rows).  Sweep k updates the region shrunk by k, so after
// h sweeps the output nodes hold exactly what h launches of k_smooth produce
// (same operations in the same order).  Thread (tx, ty) owns column tx and
// rows ty, ty+4, ...; its right-hand side and class ids stay in registers,
// the iterate ping-pongs between two LDS planes, and class tables of up to
// kMaxClsLds classes are staged in LDS.  Halo nodes beyond a non-wrapped
// edge are clamped copies: the stencils of real nodes give them zero
// coefficients, as in k_smooth's clamped neighbour.  PROLONG adds P x_c to
// every loaded node first (k_prolong_add's arithmetic), fusing the
// coarse-grid correction with the post-sweeps; ACC adds the result into an
// FP64 iterate (the mixed refinement's x += e, fused into the last level-0
// sweep).
constexpr int kTileX = 64;
constexpr int kTileRows = 32;
constexpr int kMaxFuse = 3;
constexpr int kTileRmax = kTileRows + 2 * kMaxFuse;
constexpr int kTileThreads = 1024;
constexpr int kTileRowGroups = kTileThreads / kTileX;                       // 16
constexpr int kRowsPerThread = (kTileRmax + kTileRowGroups - 1) / kTileRowGroups;   // 3
constexpr int kMaxClsLds = 32;

// wrapped (periodic axis) or clamped node coordinate; g is at most a few
// nodes outside [0, n)
__device__ __forceinline__ int tile_coord(int g, int n, int wrap)
{
    if (wrap) {
        while (g < 0) g += n;
        while (g >= n) g -= n;
        return g;
    }
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

template <typename T>
struct SmoothTileArgs {
    DevLevel L;
    DevTable<T> relax;
    int ncls;
    const T* xin;
    const T* b;
    T* xout;
    T two3;
    int h;
    DevLevel Cl;
    DevTransfer<T> tr;
    const T* xc;
    double* xacc;
};

template <typename T, bool COMPACT, bool PROLONG, bool ACC>
__device__ __forceinline__ void smooth_tile_body(const SmoothTileArgs<T>& A, unsigned bid)
{
    const DevLevel& L = A.L;
    const DevTable<T>& relax = A.relax;
    const int ncls = A.ncls;
    const T* __restrict__ xin = A.xin;
    const T* __restrict__ b = A.b;
    T* __restrict__ xout = A.xout;
    const T two3 = A.two3;
    const int h = A.h;
    const DevLevel& Cl = A.Cl;
    const DevTransfer<T>& tr = A.tr;
    const T* __restrict__ xc = A.xc;
    double* __restrict__ xacc = A.xacc;
    __shared__ T sx[2][kTileRmax * kTileX];
    __shared__ T ctab[kMaxClsLds * 9];
    const int tx = threadIdx.x & (kTileX - 1), ty = threadIdx.x / kTileX;
    const int n0 = L.n[0], n1 = L.n[1];
    const int ostep = kTileX - 2 * h;
    const int ntc = (n1 + ostep - 1) / ostep;
    const int t0 = (int)(bid / ntc) * kTileRows, t1 = (int)(bid % ntc) * ostep;
    const int R = kTileRows + 2 * h;
    const bool lds_tab = ncls <= kMaxClsLds;
    if (lds_tab)
        for (int i = threadIdx.x; i < ncls * 9; i += kTileThreads) ctab[i] = relax.coef[i];
    const int g1 = tile_coord(t1 - h + tx, n1, L.wrap[1]);
    // column part of the prolongation, once per thread
    int q1[2] = {0, 0};
    T w1[2] = {(T)0, (T)0};
    if (PROLONG) {
        const int p1 = tr.pc[1][g1];
        w1[0] = tr.pw[1][2 * g1];
        w1[1] = tr.pw[1][2 * g1 + 1];
        q1[0] = wrap_index(p1, Cl.n[1], tr.wrap[1]);
        q1[1] = wrap_index(p1 + 1, Cl.n[1], tr.wrap[1]);
    }
    T breg[kRowsPerThread];
    int creg[kRowsPerThread];
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int i = ty + kTileRowGroups * k;
        breg[k] = (T)0;
        creg[k] = 0;
        if (i < R) {
            const int g0 = tile_coord(t0 - h + i, n0, L.wrap[0]);
            const int64_t gi = (int64_t)g0 * n1 + g1;
            T xv = xin[gi];
            if (PROLONG) {
                // k_prolong_add's sum: a over the row factor, b over the column factor
                const int p0 = tr.pc[0][g0];
                const T w0[2] = {tr.pw[0][2 * g0], tr.pw[0][2 * g0 + 1]};
                T s = (T)0;
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    if (w0[a] == (T)0) continue;
                    const int i0 = wrap_index(p0 + a, Cl.n[0], tr.wrap[0]);
#pragma unroll
                    for (int bb = 0; bb < 2; ++bb) {
                        if (w1[bb] == (T)0) continue;
                        s = s + (w0[a] * w1[bb]) * xc[(int64_t)i0 * Cl.n[1] + q1[bb]];
                    }
                }
                xv = s * (T)1 + xv * (T)1;
            }
            sx[0][i * kTileX + tx] = xv;
            breg[k] = b[gi];
            creg[k] = L.cls[gi];
        }
    }
    __syncthreads();
    int cur = 0;
    for (int sw = 1; sw <= h; ++sw) {
        const T* xs = sx[cur];
        T* xn = sx[cur ^ 1];
        const bool colok = tx >= sw && tx < kTileX - sw;
#pragma unroll
        for (int k = 0; k < kRowsPerThread; ++k) {
            const int i = ty + kTileRowGroups * k;
            if (!(colok && i >= sw && i < R - sw)) continue;
            T c[9];
            if (lds_tab) {
#pragma unroll
                for (int q = 0; q < 9; ++q) c[q] = ctab[creg[k] * 9 + q];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) c[q] = relax.coef[(int64_t)creg[k] * 9 + q];
            }
            const int li = i * kTileX + tx;
            T t = (T)0;
#pragma unroll
            for (int da = -1; da <= 1; ++da)
#pragma unroll
                for (int db = -1; db <= 1; ++db) {
                    if (COMPACT && da != 0 && db != 0) continue;
                    t = t + c[(da + 1) * 3 + (db + 1)] * xs[li + da * kTileX + db];
                }
            xn[li] = t * (T)1 + breg[k] * (two3 * (T)1);
        }
        __syncthreads();
        cur ^= 1;
    }
    const int g1o = t1 + tx - h;
    if (tx < h || tx >= kTileX - h || g1o >= n1) return;
#pragma unroll
    for (int k = 0; k < kRowsPerThread; ++k) {
        const int i = ty + kTileRowGroups * k;
        if (i < h || i >= h + kTileRows) continue;
        const int g0 = t0 + i - h;
        if (g0 >= n0) continue;
        const int64_t gi = (int64_t)g0 * n1 + g1o;
        const T v = sx[cur][i * kTileX + tx];
        xout[gi] = v;
        if (ACC) xacc[gi] += (double)v;
    }
}

template <typename T, bool COMPACT, bool PROLONG, bool ACC>
__global__ __launch_bounds__(kTileThreads) void k_smooth_tile(SmoothTileArgs<T> a)
{
    smooth_tile_body<T, COMPACT, PROLONG, ACC>(a, blockIdx.x);
}

template <typename T, bool COMPACT, bool PROLONG, bool ACC>
__global__ __launch_bounds__(kTileThreads) void k_smooth_tile_many(const SmoothTileArgs<T>* __restrict__ a,
                                                                   const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    smooth_tile_body<T, COMPACT, PROLONG, ACC>(a[i], blockIdx.x - first[i]);
}

// Tiled residual + restriction (2D): a block restricts a kRRRows x kRRCols
// coarse tile.  It loads x on the fine support of that tile plus a one-node
// halo into LDS, evaluates the fine residuals there (k_residual's
// arithmetic), then sums each coarse node's support (k_restrict's order).
// The host uses it only where every tile's fine support fits (at most
// kRRFineRows x kRRFineCols nodes; true for the halving and identity
// transfers).
constexpr int kRRRows = 16;
constexpr int kRRCols = 30;
constexpr int kRRFineRows = 34;
constexpr int kRRFineCols = 62;

template <typename T>
struct ResidRestrictArgs {
    DevLevel F, Cl;
    DevTransfer<T> tr;
    DevTable<T> Af;
    int ncls;
    DevTable<T> Ac;
    const T* x;
    const T* bf;
    T* bc;
    T* xc;
    T first_sweep;
};

template <typename T, bool COMPACT>
__device__ __forceinline__ void resid_restrict_tile_body(const ResidRestrictArgs<T>& A, unsigned bid)
{
    const DevLevel& F = A.F;
    const DevLevel& Cl = A.Cl;
    const DevTransfer<T>& tr = A.tr;
    const DevTable<T>& Af = A.Af;
    const int ncls = A.ncls;
    const DevTable<T>& Ac = A.Ac;
    const T* __restrict__ x = A.x;
    const T* __restrict__ bf = A.bf;
    T* __restrict__ bc = A.bc;
    T* __restrict__ xc = A.xc;
    const T first_sweep = A.first_sweep;
    __shared__ T sxx[(kRRFineRows + 2) * 64];
    __shared__ T sr[kRRFineRows * 64];
    __shared__ T ctab[kMaxClsLds * 9];
    const int ntc = (Cl.n[1] + kRRCols - 1) / kRRCols;
    const int c0lo = (int)(bid / ntc) * kRRRows, c1lo = (int)(bid % ntc) * kRRCols;
    const int c0hi = min(c0lo + kRRRows, Cl.n[0]), c1hi = min(c1lo + kRRCols, Cl.n[1]);
    const int f0lo = tr.rf[0][c0lo], f0hi = tr.rf[0][c0hi - 1] + tr.rn[0][c0hi - 1];
    const int f1lo = tr.rf[1][c1lo], f1hi = tr.rf[1][c1hi - 1] + tr.rn[1][c1hi - 1];
    const int FR = f0hi - f0lo, FC = f1hi - f1lo;
    const bool lds_tab = ncls <= kMaxClsLds;
    if (lds_tab)
        for (int i = threadIdx.x; i < ncls * 9; i += kTileThreads) ctab[i] = Af.coef[i];
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    const int n0 = F.n[0], n1 = F.n[1];
    if (tx < FC + 2) {
        const int g1 = tile_coord(f1lo - 1 + tx, n1, F.wrap[1]);
        for (int i = ty; i < FR + 2; i += kTileRowGroups) {
            const int g0 = tile_coord(f0lo - 1 + i, n0, F.wrap[0]);
            sxx[i * 64 + tx] = x[(int64_t)g0 * n1 + g1];
        }
    }
    __syncthreads();
    if (tx < FC) {
        const int g1 = tile_coord(f1lo + tx, n1, F.wrap[1]);
        for (int i = ty; i < FR; i += kTileRowGroups) {
            const int g0 = tile_coord(f0lo + i, n0, F.wrap[0]);
            const int64_t gi = (int64_t)g0 * n1 + g1;
            const int cl = F.cls[gi];
            T c[9];
            if (lds_tab) {
#pragma unroll
                for (int q = 0; q < 9; ++q) c[q] = ctab[cl * 9 + q];
            } else {
#pragma unroll
                for (int q = 0; q < 9; ++q) c[q] = Af.coef[(int64_t)cl * 9 + q];
            }
            const int li = (i + 1) * 64 + (tx + 1);
            T t = (T)0;
#pragma unroll
            for (int da = -1; da <= 1; ++da)
#pragma unroll
                for (int db = -1; db <= 1; ++db) {
                    if (COMPACT && da != 0 && db != 0) continue;
                    t = t + c[(da + 1) * 3 + (db + 1)] * sxx[li + da * 64 + db];
                }
            sr[i * 64 + tx] = t * (T)-1 + bf[gi] * (T)1;
        }
    }
    __syncthreads();
    if ((int)threadIdx.x >= kRRRows * kRRCols) return;
    const int c0 = c0lo + (int)threadIdx.x / kRRCols, c1 = c1lo + (int)threadIdx.x % kRRCols;
    if (c0 >= c0hi || c1 >= c1hi) return;
    const int f0 = tr.rf[0][c0], m0 = tr.rn[0][c0];
    const int f1 = tr.rf[1][c1], m1 = tr.rn[1][c1];
    T s = (T)0;
    for (int a = 0; a < m0; ++a) {
        const T w0 = tr.rw[0][3 * c0 + a];
        for (int b = 0; b < m1; ++b) {
            const T w = w0 * tr.rw[1][3 * c1 + b];
            s = s + w * sr[(f0 + a - f0lo) * 64 + (f1 + b - f1lo)];
        }
    }
    const int64_t ci = (int64_t)c0 * Cl.n[1] + c1;
    const T bv = s * Ac.dinv[Cl.cls[ci]];
    bc[ci] = bv;
    xc[ci] = first_sweep != (T)0 ? bv * (first_sweep * (T)1) : (T)0;
}

template <typename T, bool COMPACT>
__global__ __launch_bounds__(kTileThreads) void k_resid_restrict_tile(ResidRestrictArgs<T> a)
{
    resid_restrict_tile_body<T, COMPACT>(a, blockIdx.x);
}

template <typename T, bool COMPACT>
__global__ __launch_bounds__(kTileThreads) void k_resid_restrict_tile_many(
    const ResidRestrictArgs<T>* __restrict__ a, const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    resid_restrict_tile_body<T, COMPACT>(a[i], blockIdx.x - first[i]);
}

// k_residual fused into k_restrict (2D): the restriction evaluates the fine
// residual b - A^ x at each of its fine support nodes with k_residual's
// arithmetic instead of reading a stored residual.
template <typename T, bool COMPACT>
__device__ __forceinline__ void resid_restrict_body(const ResidRestrictArgs<T>& A, unsigned bid)
{
    const DevLevel& F = A.F;
    const DevLevel& Cl = A.Cl;
    const DevTransfer<T>& tr = A.tr;
    const DevTable<T>& Af = A.Af;
    const DevTable<T>& Ac = A.Ac;
    const T* __restrict__ x = A.x;
    const T* __restrict__ bf = A.bf;
    T* __restrict__ bc = A.bc;
    T* __restrict__ xc = A.xc;
    const T first_sweep = A.first_sweep;
    const int64_t i = (int64_t)bid * kMgThreads + threadIdx.x;
    if (i >= Cl.N) return;
    int c[3];
    node_coords<2>(Cl, i, c);
    T s = (T)0;
    const int f0 = tr.rf[0][c[0]], m0 = tr.rn[0][c[0]];
    const int f1 = tr.rf[1][c[1]], m1 = tr.rn[1][c[1]];
    for (int a = 0; a < m0; ++a) {
        const T w0 = tr.rw[0][3 * c[0] + a];
        const int i0 = wrap_index(f0 + a, F.n[0], tr.wrap[0]);
        for (int b = 0; b < m1; ++b) {
            const T w = w0 * tr.rw[1][3 * c[1] + b];
            const int i1 = wrap_index(f1 + b, F.n[1], tr.wrap[1]);
            const int64_t fi = (int64_t)i0 * F.n[1] + i1;
            const int fc[3] = {i0, i1, 0};
            const T t = stencil_apply<T, 2, COMPACT>(F, Af.coef + (int64_t)F.cls[fi] * 9, x, fc);
            const T r = t * (T)-1 + bf[fi] * (T)1;
            s = s + w * r;
        }
    }
    const T bv = s * Ac.dinv[Cl.cls[i]];
    bc[i] = bv;
    xc[i] = first_sweep != (T)0 ? bv * (first_sweep * (T)1) : (T)0;
}

template <typename T, bool COMPACT>
__global__ __launch_bounds__(kMgThreads) void k_resid_restrict(ResidRestrictArgs<T> a)
{
    resid_restrict_body<T, COMPACT>(a, blockIdx.x);
}

template <typename T, bool COMPACT>
__global__ __launch_bounds__(kMgThreads) void k_resid_restrict_many(const ResidRestrictArgs<T>* __restrict__ a,
                                                                    const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    resid_restrict_body<T, COMPACT>(a[i], blockIdx.x - first[i]);
}

// b0 = b o D0^-1  (hadamard_to)
template <typename T>
struct HadamardArgs {
    DevLevel L;
    DevTable<T> A;
    const T* b;
    T* out;
};

template <typename T>
__device__ __forceinline__ void hadamard_body(const HadamardArgs<T>& a, unsigned bid)
{
    const int64_t i = (int64_t)bid * kMgThreads + threadIdx.x;
    if (i < a.L.N) a.out[i] = a.b[i] * a.A.dinv[a.L.cls[i]];
}

template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_hadamard(HadamardArgs<T> a)
{
    hadamard_body<T>(a, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_hadamard_many(const HadamardArgs<T>* __restrict__ a,
                                                              const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    hadamard_body<T>(a[i], blockIdx.x - first[i]);
}

// x = Ainv b, one wave per row, fixed-order lane partials + butterfly.
template <typename T>
struct CoarseArgs {
    const double* inv;
    int n;
    const T* b;
    T* x;
};

template <typename T>
__device__ __forceinline__ void coarse_solve_body(const CoarseArgs<T>& a, unsigned bid)
{
    const int row = (int)(((int64_t)bid * kMgThreads + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    const int n = a.n;
    if (row >= n) return;
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += a.inv[(int64_t)row * n + j] * (double)a.b[j];
    s = wave_sum(s);
    if (lane == 0) a.x[row] = (T)s;
}

template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_coarse_solve(CoarseArgs<T> a)
{
    coarse_solve_body<T>(a, blockIdx.x);
}

template <typename T>
__global__ __launch_bounds__(kMgThreads) void k_coarse_solve_many(const CoarseArgs<T>* __restrict__ a,
                                                                  const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    coarse_solve_body<T>(a[i], blockIdx.x - first[i]);
}

// FP64 residual of the unscaled operator: r = b - A x, partial sums of r^2
// and of b^2 per block; when OUTF32 also emit (float)r o D0^-1 (the mixed
// backend's residual_to + hadamard_to).
// FP64 residual r = b - A x of the unscaled level-0 operator; partial sums
// of r^2 (skipping row `pin` of a pinned-periodic system, which the gauge
// shift satisfies exactly) and of bn^2 (the caller's right-hand side).
struct OuterArgs {
    DevLevel L;
    DevTable<double> K;
    const double* x;
    const double* b;
    const double* bn;
    int64_t pin;
    const float* dinv32;
    float* b0;
    double* part;
    float* e0;
    // 3D 7-point level 0: the dominant class and its row (K.coef +
    // cls_in * 27), read as uniform values; -1: every node reads its row
    int cls_in;
    const double* row_in;
};

template <int D, bool COMPACT, bool OUTF32, bool NORMS>
__device__ __forceinline__ void outer_residual_body(const OuterArgs& A, int64_t blk)
{
    const DevLevel& L = A.L;
    const DevTable<double>& K = A.K;
    const double* __restrict__ x = A.x;
    const double* __restrict__ b = A.b;
    const double* __restrict__ bn = A.bn;
    const int64_t pin = A.pin;
    const float* __restrict__ dinv32 = A.dinv32;
    float* __restrict__ b0 = A.b0;
    double* __restrict__ part = A.part;
    float* __restrict__ e0 = A.e0;
    __shared__ double ws[kMgThreads / 64][2];
    const int64_t i = blk * kMgThreads + threadIdx.x;
    double rr = 0.0, bb = 0.0;
    if (i < L.N) {
        int c[3];
        node_coords<D>(L, i, c);
        const uint16_t ci = L.cls[i];
        double ax;
        if (D == 3 && COMPACT && (int)ci == A.cls_in) {
            ax = stencil_apply<double, D, COMPACT>(L, A.row_in, x, c);
        } else {
            const double* coef = K.coef + (int64_t)ci * (D == 2 ? 9 : 27);
            ax = stencil_apply<double, D, COMPACT>(L, coef, x, c);
        }
        const double r = b[i] - ax;
        if (NORMS) {
            rr = i == pin ? 0.0 : r * r;
            bb = bn[i] * bn[i];
        }
        if (OUTF32) {
            const float bv = (float)r * dinv32[L.cls[i]];
            b0[i] = bv;
            // first Jacobi sweep of the correction from e = 0: relax*0 + b*(2/3)
            if (e0) e0[i] = bv * (2.0f / 3.0f * 1.0f);
        }
    }
    if (!NORMS) return;
    rr = wave_sum(rr);
    bb = wave_sum(bb);
    if ((threadIdx.x & 63) == 0) {
        ws[threadIdx.x >> 6][0] = rr;
        ws[threadIdx.x >> 6][1] = bb;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        double s = 0.0;
        for (int w = 0; w < kMgThreads / 64; ++w) s += ws[w][threadIdx.x];
        part[2 * blk + threadIdx.x] = s;
    }
}

template <int D, bool COMPACT, bool OUTF32, bool NORMS>
__global__ __launch_bounds__(kMgThreads) void k_outer_residual(OuterArgs a)
{
    outer_residual_body<D, COMPACT, OUTF32, NORMS>(a, level_block<D>());
}

template <int D, bool COMPACT, bool OUTF32, bool NORMS>
__global__ __launch_bounds__(kMgThreads) void k_outer_residual_many(const OuterArgs* __restrict__ a,
                                                                    const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    outer_residual_body<D, COMPACT, OUTF32, NORMS>(a[i], (int64_t)(blockIdx.x - first[i]));
}

// The outer residual without norms on a chunked 7-point 3D level: the
// FP32 right-hand side (and first sweep) of outer_residual_body at the
// chunk kernels' nodes (line_apply in FP64 over the unscaled class row),
// the listed nodes one per thread; same terms and order, same bits.
__device__ __forceinline__ void outer_node_f32(const OuterArgs& A, int64_t i)
{
    const DevLevel& L = A.L;
    int c[3];
    node_coords<3>(L, i, c);
    const double* coef = A.K.coef + (int64_t)L.cls[i] * 27;
    const double ax = stencil_apply<double, 3, true>(L, coef, A.x, c);
    const double r = A.b[i] - ax;
    const float bv = (float)r * A.dinv32[L.cls[i]];
    A.b0[i] = bv;
    if (A.e0) A.e0[i] = bv * (2.0f / 3.0f * 1.0f);
}

__global__ __launch_bounds__(kMgThreads) void k_outer_line(LineGeom G, LineList LL, OuterArgs A,
                                                           const double* __restrict__ row, float dinv_in)
{
    bool listed;
    const int64_t blk = line_block(LL, &listed);
    if (blk < 0) return;
    if (listed) {
        const int64_t t = blk * kMgThreads + threadIdx.x;
        if (t < (int64_t)LL.n) outer_node_f32(A, (int64_t)LL.list[t]);
        return;
    }
    const int64_t t = blk * kMgThreads + threadIdx.x;
    if (t >= (int64_t)G.nchunks) return;
    const uint32_t i = (uint32_t)t * kLine;
    if (!line_chunk_fast(G, i)) return;
    double ax[kLine], bb[kLine];
    line_apply<double, true>(G, row, A.x, i, ax);
    ld_line(A.b + i, bb);
    float bv[kLine], e[kLine];
#pragma unroll
    for (int k = 0; k < kLine; ++k) {
        const double r = bb[k] - ax[k];
        bv[k] = (float)r * dinv_in;
        e[k] = bv[k] * (2.0f / 3.0f * 1.0f);
    }
    st_line(A.b0 + i, bv);
    if (A.e0) st_line(A.e0 + i, e);
}

__global__ __launch_bounds__(kMgThreads) void k_reduce2(const double* __restrict__ part, int64_t nb,
                                                        double* __restrict__ out)
{
    __shared__ double sh[2][kMgThreads];
    double s0 = 0.0, s1 = 0.0;
    // thread t sums pairs 8t .. 8t+7, then 8(t + 256) .., in order; the eight
    // pairs of a round are loaded before they are added
    constexpr int kU = 8;
    for (int64_t b0 = (int64_t)threadIdx.x * kU; b0 < nb; b0 += (int64_t)kMgThreads * kU) {
        double v[2 * kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            const bool in = b0 + u < nb;
            v[2 * u] = in ? part[2 * (b0 + u)] : 0.0;
            v[2 * u + 1] = in ? part[2 * (b0 + u) + 1] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
            s0 += v[2 * u];
            s1 += v[2 * u + 1];
        }
    }
    sh[0][threadIdx.x] = s0;
    sh[1][threadIdx.x] = s1;
    __syncthreads();
    for (int w = kMgThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 2) out[threadIdx.x] = sh[threadIdx.x][0];
}

// per-block sums of x (part[b]), then a fixed-order single-block total
__global__ __launch_bounds__(kMgThreads) void k_psum(const double* __restrict__ x, int64_t n,
                                                     double* __restrict__ part)
{
    __shared__ double ws[kMgThreads / 64];
    const int64_t i = (int64_t)blockIdx.x * kMgThreads + threadIdx.x;
    double v = wave_sum(i < n ? x[i] : 0.0);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kMgThreads / 64; ++w) s += ws[w];
        part[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kMgThreads) void k_fsum(const double* __restrict__ part, int64_t nb,
                                                     double* __restrict__ out)
{
    __shared__ double sh[kMgThreads];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kMgThreads) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kMgThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// b'_p = b_p - sum(b): the consistent right-hand side of the singular system
__global__ void k_set_pin(double* __restrict__ bp, const double* __restrict__ b,
                          const double* __restrict__ sum, int64_t p)
{
    if (threadIdx.x == 0) bp[p] = b[p] - sum[0];
}

// gauge: shift = x_p - b_p ; x -= shift
__global__ void k_gauge_read(const double* __restrict__ x, const double* __restrict__ b, int64_t p,
                             double* __restrict__ shift)
{
    if (threadIdx.x == 0) shift[0] = x[p] - b[p];
}

__global__ __launch_bounds__(kMgThreads) void k_gauge_apply(double* __restrict__ x, int64_t n,
                                                            const double* __restrict__ shift)
{
    const int64_t i = (int64_t)blockIdx.x * kMgThreads + threadIdx.x;
    if (i < n) x[i] = x[i] - shift[0];
}

// Element-wise level-0 vector operations: x = 0 (FP32), x += e (FP64 +=
// FP32) and x = y (FP64 copy).
struct VecArgs {
    void* x;
    const void* y;
    int64_t n;
};

template <int OP>
__device__ __forceinline__ void vec_body(const VecArgs& a, unsigned bid)
{
    const int64_t i = (int64_t)bid * kMgThreads + threadIdx.x;
    if (i >= a.n) return;
    if (OP == 0) static_cast<float*>(a.x)[i] = 0.0f;
    else if (OP == 1) static_cast<double*>(a.x)[i] += (double)static_cast<const float*>(a.y)[i];
    else static_cast<double*>(a.x)[i] = static_cast<const double*>(a.y)[i];
}

template <int OP>
__global__ __launch_bounds__(kMgThreads) void k_vec(VecArgs a)
{
    vec_body<OP>(a, blockIdx.x);
}

template <int OP>
__global__ __launch_bounds__(kMgThreads) void k_vec_many(const VecArgs* __restrict__ a,
                                                        const unsigned* __restrict__ first, int n)
{
    const int i = batch_instance(first, n);
    vec_body<OP>(a[i], blockIdx.x - first[i]);
}

__global__ __launch_bounds__(kMgThreads) void k_zero_f32(float* __restrict__ x, int64_t n)
{
    vec_body<0>(VecArgs{x, nullptr, n}, blockIdx.x);
}

__global__ __launch_bounds__(kMgThreads) void k_addback(double* __restrict__ b,
                                                        const unsigned long long* __restrict__ idx,
                                                        const double* __restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kMgThreads + threadIdx.x;
    if (i < n) b[idx[i]] += v[i];
}

__global__ __launch_bounds__(kMgThreads) void k_replace(double* __restrict__ b,
                                                        const unsigned long long* __restrict__ idx,
                                                        const double* __restrict__ v, int64_t n,
                                                        int zero)
{
    const int64_t i = (int64_t)blockIdx.x * kMgThreads + threadIdx.x;
    if (i < n) b[idx[i]] = zero ? 0.0 : v[i];
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kMgThreads - 1) / kMgThreads); }

// ------------------------------------------------------------ host levels

struct Axis {
    std::vector<double> x;
};

Axis coarsen(const Axis& f)
{
    Axis c;
    for (size_t i = 0; i < f.x.size(); i += 2) c.x.push_back(f.x[i]);
    if (c.x.back() != f.x.back()) c.x.push_back(f.x.back());
    return c;
}

std::vector<double> volumes(const Axis& a)
{
    std::vector<double> w(a.x.size(), 1.0);
    if (a.x.size() == 1) return w;
    for (size_t i = 0; i < a.x.size(); ++i)
        w[i] = ((i ? a.x[i] - a.x[i - 1] : 0.0) + (i + 1 < a.x.size() ? a.x[i + 1] - a.x[i] : 0.0)) * 0.5;
    return w;
}

struct AxisTransfer {
    std::vector<int> pc;        // per fine f
    std::vector<double> pw;     // 2 per fine f (w_lo, w_hi); exact node: (1, 0)
    std::vector<int> rf, rn;    // per coarse c
    std::vector<double> rw;     // 3 per coarse c
    int wrap = 0;               // periodic axis: fine/coarse indices wrap
};

// Periodic axis of even length nf >= 4: coarse node c is fine node 2c; an odd
// fine node interpolates its two coarse neighbours (wrapping); volumes are
// uniform, so R = P^T / 2 = {1/4, 1/2, 1/4} around 2c.
AxisTransfer axis_transfer_periodic(int nf)
{
    AxisTransfer t;
    const int nc = nf / 2;
    t.wrap = 1;
    t.pc.resize(nf);
    t.pw.resize(2 * nf);
    for (int i = 0; i < nf; ++i) {
        if (i % 2 == 0) { t.pc[i] = i / 2; t.pw[2 * i] = 1.0; t.pw[2 * i + 1] = 0.0; }
        else { t.pc[i] = (i - 1) / 2; t.pw[2 * i] = 0.5; t.pw[2 * i + 1] = 0.5; }
    }
    t.rf.resize(nc);
    t.rn.assign(nc, 3);
    t.rw.resize(3 * nc);
    for (int c = 0; c < nc; ++c) {
        t.rf[c] = 2 * c - 1;
        t.rw[3 * c] = 0.25;
        t.rw[3 * c + 1] = 0.5;
        t.rw[3 * c + 2] = 0.25;
    }
    return t;
}

AxisTransfer axis_transfer_identity(int n, int wrap)
{
    AxisTransfer t;
    t.wrap = wrap;
    t.pc.resize(n);
    t.pw.assign(2 * n, 0.0);
    t.rf.resize(n);
    t.rn.assign(n, 1);
    t.rw.assign(3 * n, 0.0);
    for (int i = 0; i < n; ++i) {
        t.pc[i] = i;
        t.pw[2 * i] = 1.0;
        t.rf[i] = i;
        t.rw[3 * i] = 1.0;
    }
    return t;
}

AxisTransfer axis_transfer(const Axis& f, const Axis& c)
{
    AxisTransfer t;
    const int nf = (int)f.x.size(), nc = (int)c.x.size();
    t.pc.resize(nf);
    t.pw.resize(2 * nf);
    for (int i = 0; i < nf; ++i) {
        const double x = f.x[i];
        int hi = (int)(std::lower_bound(c.x.begin(), c.x.end(), x) - c.x.begin());
        if (hi >= nc) throw std::runtime_error("hip_multigrid: point outside coarse grid");
        if (c.x[hi] == x) {
            t.pc[i] = hi;
            t.pw[2 * i] = 1.0;
            t.pw[2 * i + 1] = 0.0;
        } else {
            if (hi == 0) throw std::runtime_error("hip_multigrid: point outside coarse grid");
            const double tt = (x - c.x[hi - 1]) / (c.x[hi] - c.x[hi - 1]);
            t.pc[i] = hi - 1;
            t.pw[2 * i] = 1.0 - tt;
            t.pw[2 * i + 1] = tt;
        }
    }
    const std::vector<double> vol = volumes(f);
    std::vector<double> cvol(nc, 0.0);
    for (int i = 0; i < nf; ++i) {
        cvol[t.pc[i]] += t.pw[2 * i] * vol[i];
        if (t.pw[2 * i + 1] != 0.0) cvol[t.pc[i] + 1] += t.pw[2 * i + 1] * vol[i];
    }
    t.rf.assign(nc, -1);
    t.rn.assign(nc, 0);
    t.rw.assign(3 * nc, 0.0);
    for (int i = 0; i < nf; ++i)
        for (int h = 0; h < 2; ++h) {
            const double p = t.pw[2 * i + h];
            if (p == 0.0) continue;
            const int cc = t.pc[i] + h;
            if (t.rf[cc] < 0) t.rf[cc] = i;
            const int k = i - t.rf[cc];
            if (k < 0 || k > 2) throw std::runtime_error("hip_multigrid: restriction support > 3");
            t.rw[3 * cc + k] = p * (vol[i] / cvol[cc]);
            t.rn[cc] = std::max(t.rn[cc], k + 1);
        }
    return t;
}

struct RowKey {
    const double* v;
    int ks;
};

struct HostLevel {
    int D = 2;
    int n[3] = {1, 1, 1};
    int wrap[3] = {0, 0, 0};
    int64_t N = 0;
    int KS = 9;
    bool compact = false;
    std::vector<uint16_t> cls;
    std::vector<double> K;     // unscaled class rows
    std::vector<double> dinv;  // per class
    std::vector<double> Ah;    // scaled rows
    std::vector<double> Rx;    // relax rows
};

// Dedupe rows (bitwise) into classes.  rows: N * KS doubles.
void classify(HostLevel& L, const std::vector<double>& rows)
{
    const int KS = L.KS;
    std::unordered_map<std::string, int> map;
    L.cls.resize(L.N);
    L.K.clear();
    for (int64_t i = 0; i < L.N; ++i) {
        std::string key(reinterpret_cast<const char*>(&rows[i * KS]), sizeof(double) * KS);
        auto it = map.find(key);
        int id;
        if (it == map.end()) {
            id = (int)map.size();
            if (id >= 65535) throw std::runtime_error("hip_multigrid: more than 65535 distinct stencil rows");
            map.emplace(std::move(key), id);
            L.K.insert(L.K.end(), rows.begin() + i * KS, rows.begin() + (i + 1) * KS);
        } else {
            id = it->second;
        }
        L.cls[i] = (uint16_t)id;
    }
    const int ncls = (int)map.size();
    const int center = (KS - 1) / 2;
    L.dinv.resize(ncls);
    L.Ah.resize((size_t)ncls * KS);
    L.Rx.resize((size_t)ncls * KS);
    L.compact = true;
    for (int c = 0; c < ncls; ++c) {
        const double diag = L.K[(size_t)c * KS + center];
        L.dinv[c] = 1 / diag;
        for (int k = 0; k < KS; ++k) {
            const double a = L.K[(size_t)c * KS + k] * L.dinv[c];
            L.Ah[(size_t)c * KS + k] = a;
            const double s = a * 2.0 / 3;
            L.Rx[(size_t)c * KS + k] = (k == center) ? 1.0 - s : 0.0 - s;
            if (a != 0.0) {
                int nz = 0;
                int kk = k;
                for (int d = 0; d < L.D; ++d) {
                    if (kk % 3 != 1) ++nz;
                    kk /= 3;
                }
                if (nz > 1) L.compact = false;
            }
        }
    }
}

// offset of column (cc) from row (rc) along an axis of n nodes; wrapped
// offsets flag the axis periodic.  Returns false if not a 3-point neighbour.
bool axis_offset(int rc, int cc, int n, int& off, int& wrapped)
{
    const int d = cc - rc;
    if (d >= -1 && d <= 1) {
        off = d;
        return true;
    }
    if (d == n - 1) {
        off = -1;
        wrapped = 1;
        return true;
    }
    if (d == -(n - 1)) {
        off = 1;
        wrapped = 1;
        return true;
    }
    return false;
}

inline int dims_n(const HostLevel& L, int d) { return L.n[d]; }

// Pinned-periodic detection on the level-0 rows (N x KS): every axis of
// size > 1 wrapped, exactly one identity row p, zero row sums elsewhere.
// Returns p, or -1.
int64_t detect_pinned_periodic(const HostLevel& L, const std::vector<double>& R)
{
    for (int d = 0; d < 3; ++d)
        if (L.n[d] > 1 && !L.wrap[d]) return -1;
    const int KS = L.KS, center = (KS - 1) / 2;
    int64_t pin = -1;
    for (int64_t i = 0; i < L.N; ++i) {
        const double* r = &R[(size_t)i * KS];
        bool ident = r[center] == 1.0;
        double s = 0.0, m = 0.0;
        for (int k = 0; k < KS; ++k) {
            if (k != center && r[k] != 0.0) ident = false;
            s += r[k];
            m += std::fabs(r[k]);
        }
        if (ident) {
            if (pin >= 0) return -1;
            pin = i;
        } else if (std::fabs(s) > 1e-12 * m) {
            return -1;
        }
    }
    return pin;
}

// Row p of the singular periodic operator from column p (the FD matrix is
// symmetric): A[p][p+o] = A[p+o][p]; the diagonal closes the zero row sum.
void restore_pinned_row(const HostLevel& L, std::vector<double>& R, int64_t p)
{
    const int KS = L.KS, center = (KS - 1) / 2, D = L.D;
    int c[3];
    int64_t t = p;
    c[2] = (int)(t % L.n[2]); t /= L.n[2]; c[1] = (int)(t % L.n[1]); c[0] = (int)(t / L.n[1]);
    double off = 0.0;
    for (int k = 0; k < KS; ++k) {
        if (k == center) continue;
        int o[3] = {0, 0, 0};
        if (D == 2) { o[0] = k / 3 - 1; o[1] = k % 3 - 1; }
        else { o[0] = k / 9 - 1; o[1] = (k / 3) % 3 - 1; o[2] = k % 3 - 1; }
        int g[3];
        for (int d = 0; d < 3; ++d) {
            g[d] = c[d] + o[d];
            if (g[d] < 0) g[d] += L.n[d];
            else if (g[d] >= L.n[d]) g[d] -= L.n[d];
        }
        const int64_t j = ((int64_t)g[0] * L.n[1] + g[1]) * L.n[2] + g[2];
        const double v = R[(size_t)j * KS + (KS - 1 - k)];
        R[(size_t)p * KS + k] = v;
        off += v;
    }
    R[(size_t)p * KS + center] = -off;
}

HostLevel level_from_coo(int I, int J, int K, unsigned long long n_row, unsigned long long nnz,
                         const unsigned long long* rows, const unsigned long long* cols,
                         const double* vals, bool periodic_aware, int64_t* pin_out,
                         std::vector<double>* pinned_rows)
{
    HostLevel L;
    L.D = K > 1 ? 3 : 2;
    L.n[0] = I;
    L.n[1] = J;
    L.n[2] = K;
    L.N = (int64_t)I * J * K;
    L.KS = L.D == 2 ? 9 : 27;
    if ((unsigned long long)L.N != n_row)
        throw std::runtime_error("hip_multigrid: I*J*K does not match the matrix size");
    std::vector<double> R((size_t)L.N * L.KS, 0.0);
    int wrapped[3] = {0, 0, 0};
    for (unsigned long long q = 0; q < nnz; ++q) {
        const unsigned long long r = rows[q], c = cols[q];
        if (r >= n_row || c >= n_row) throw std::runtime_error("hip_multigrid: index out of range");
        int rc[3], cc[3];
        unsigned long long a = r, b = c;
        rc[2] = (int)(a % K); a /= K; rc[1] = (int)(a % J); rc[0] = (int)(a / J);
        cc[2] = (int)(b % K); b /= K; cc[1] = (int)(b % J); cc[0] = (int)(b / J);
        int off[3] = {0, 0, 0};
        const int dims[3] = {I, J, K};
        for (int d = 0; d < 3; ++d)
            if (!axis_offset(rc[d], cc[d], dims[d], off[d], wrapped[d]))
                throw std::runtime_error("hip_multigrid: matrix is not a 3^D-point stencil on the I x J x K node grid");
        int k;
        if (L.D == 2) k = (off[0] + 1) * 3 + (off[1] + 1);
        else k = (off[0] + 1) * 9 + (off[1] + 1) * 3 + (off[2] + 1);
        R[(size_t)r * L.KS + k] += vals[q];
    }
    for (int d = 0; d < 3; ++d) L.wrap[d] = (dims_n(L, d) >= 3) ? wrapped[d] : 0;
    *pin_out = -1;
    if (periodic_aware) {
        *pin_out = detect_pinned_periodic(L, R);
        if (*pin_out >= 0) {
            if (pinned_rows) *pinned_rows = R;
            restore_pinned_row(L, R, *pin_out);
        }
    }
    classify(L, R);
    return L;
}

// Galerkin coarse level: rows of R A^ P with A^ the scaled fine operator.
HostLevel galerkin(const HostLevel& F, const AxisTransfer* tr, const int* nc)
{
    HostLevel C;
    C.D = F.D;
    C.KS = F.KS;
    for (int d = 0; d < 3; ++d) {
        C.n[d] = nc[d];
        C.wrap[d] = F.wrap[d];
    }
    C.N = (int64_t)nc[0] * nc[1] * nc[2];
    const int D = F.D, KS = F.KS;
    std::vector<double> rows((size_t)C.N * KS, 0.0);
    std::atomic<int> bad{0};
    parallel_for(C.N, [&](int64_t ci) {
        int c[3];
        int64_t t = ci;
        c[2] = (int)(t % nc[2]); t /= nc[2]; c[1] = (int)(t % nc[1]); c[0] = (int)(t / nc[1]);
        double* out = &rows[(size_t)ci * KS];
        const int m0 = tr[0].rn[c[0]], m1 = tr[1].rn[c[1]], m2 = D == 3 ? tr[2].rn[c[2]] : 1;
        for (int a = 0; a < m0; ++a)
            for (int b = 0; b < m1; ++b)
                for (int e = 0; e < m2; ++e) {
                    int f[3] = {tr[0].rf[c[0]] + a, tr[1].rf[c[1]] + b, D == 3 ? tr[2].rf[c[2]] + e : 0};
                    for (int d = 0; d < 3; ++d)
                        if (tr[d].wrap) f[d] = f[d] < 0 ? f[d] + F.n[d] : (f[d] >= F.n[d] ? f[d] - F.n[d] : f[d]);
                    double rw = tr[0].rw[3 * c[0] + a] * tr[1].rw[3 * c[1] + b];
                    if (D == 3) rw = rw * tr[2].rw[3 * c[2] + e];
                    const int64_t fi = ((int64_t)f[0] * F.n[1] + f[1]) * F.n[2] + f[2];
                    const double* arow = &F.Ah[(size_t)F.cls[fi] * KS];
                    for (int k = 0; k < KS; ++k) {
                        const double av = arow[k];
                        if (av == 0.0) continue;
                        int o[3] = {0, 0, 0};
                        if (D == 2) { o[0] = k / 3 - 1; o[1] = k % 3 - 1; }
                        else { o[0] = k / 9 - 1; o[1] = (k / 3) % 3 - 1; o[2] = k % 3 - 1; }
                        int g[3];
                        for (int d = 0; d < 3; ++d) {
                            g[d] = f[d] + o[d];
                            if (g[d] < 0) g[d] += F.n[d];
                            else if (g[d] >= F.n[d]) g[d] -= F.n[d];
                        }
                        const double ra = rw * av;
                        const int hm2 = D == 3 ? 2 : 1;
                        for (int h0 = 0; h0 < 2; ++h0) {
                            const double p0 = tr[0].pw[2 * g[0] + h0];
                            if (p0 == 0.0) continue;
                            for (int h1 = 0; h1 < 2; ++h1) {
                                const double p1 = tr[1].pw[2 * g[1] + h1];
                                if (p1 == 0.0) continue;
                                for (int h2 = 0; h2 < hm2; ++h2) {
                                    double p = p0 * p1;
                                    int cc[3] = {tr[0].pc[g[0]] + h0, tr[1].pc[g[1]] + h1, 0};
                                    if (D == 3) {
                                        const double p2 = tr[2].pw[2 * g[2] + h2];
                                        if (p2 == 0.0) continue;
                                        p = p * p2;
                                        cc[2] = tr[2].pc[g[2]] + h2;
                                    }
                                    for (int d = 0; d < 3; ++d)
                                        if (tr[d].wrap && cc[d] == nc[d]) cc[d] = 0;
                                    int oo[3] = {0, 0, 0};
                                    for (int d = 0; d < D; ++d) {
                                        int dd = cc[d] - c[d];
                                        if (C.wrap[d]) {
                                            if (dd > 1) dd -= nc[d];
                                            else if (dd < -1) dd += nc[d];
                                        }
                                        if (dd < -1 || dd > 1) { bad = 1; dd = 0; }
                                        oo[d] = dd;
                                    }
                                    const int kk = D == 2 ? (oo[0] + 1) * 3 + (oo[1] + 1)
                                                          : (oo[0] + 1) * 9 + (oo[1] + 1) * 3 + (oo[2] + 1);
                                    out[kk] += ra * p;
                                }
                            }
                        }
                    }
                }
    });
    if (bad) throw std::runtime_error("hip_multigrid: Galerkin operator wider than 3 nodes");
    classify(C, rows);
    return C;
}

void dense_inverse(const HostLevel& L, std::vector<double>& inv, bool pin0)
{
    const int64_t n = L.N;
    std::vector<double> m((size_t)(n * n), 0.0);
    inv.assign((size_t)(n * n), 0.0);
    const int D = L.D, KS = L.KS;
    for (int64_t i = 0; i < n; ++i) {
        int c[3];
        int64_t t = i;
        c[2] = (int)(t % L.n[2]); t /= L.n[2]; c[1] = (int)(t % L.n[1]); c[0] = (int)(t / L.n[1]);
        const double* row = &L.Ah[(size_t)L.cls[i] * KS];
        for (int k = 0; k < KS; ++k) {
            if (row[k] == 0.0) continue;
            int o[3] = {0, 0, 0};
            if (D == 2) { o[0] = k / 3 - 1; o[1] = k % 3 - 1; }
            else { o[0] = k / 9 - 1; o[1] = (k / 3) % 3 - 1; o[2] = k % 3 - 1; }
            int g[3];
            for (int d = 0; d < 3; ++d) {
                g[d] = c[d] + o[d];
                if (g[d] < 0) g[d] += L.n[d];
                else if (g[d] >= L.n[d]) g[d] -= L.n[d];
            }
            const int64_t j = ((int64_t)g[0] * L.n[1] + g[1]) * L.n[2] + g[2];
            m[(size_t)(i * n + j)] += row[k];
        }
        inv[(size_t)(i * n + i)] = 1.0;
    }
    if (pin0) {
        // singular coarsest operator: pin coarse node 0 (identity row)
        for (int64_t j = 0; j < n; ++j) m[(size_t)j] = 0.0;
        m[0] = 1.0;
    }
    for (int64_t col = 0; col < n; ++col) {
        int64_t p = col;
        for (int64_t r = col + 1; r < n; ++r)
            if (std::fabs(m[(size_t)(r * n + col)]) > std::fabs(m[(size_t)(p * n + col)])) p = r;
        if (m[(size_t)(p * n + col)] == 0.0) throw std::runtime_error("hip_multigrid: singular coarsest operator");
        if (p != col)
            for (int64_t j = 0; j < n; ++j) {
                std::swap(m[(size_t)(col * n + j)], m[(size_t)(p * n + j)]);
                std::swap(inv[(size_t)(col * n + j)], inv[(size_t)(p * n + j)]);
            }
        const double d = m[(size_t)(col * n + col)];
        for (int64_t j = 0; j < n; ++j) {
            m[(size_t)(col * n + j)] /= d;
            inv[(size_t)(col * n + j)] /= d;
        }
        parallel_for(n, [&](int64_t r) {
            if (r == col) return;
            const double f = m[(size_t)(r * n + col)];
            if (f == 0.0) return;
            for (int64_t j = 0; j < n; ++j) {
                m[(size_t)(r * n + j)] -= f * m[(size_t)(col * n + j)];
                inv[(size_t)(r * n + j)] -= f * inv[(size_t)(col * n + j)];
            }
        });
    }
    // with node 0 pinned its right-hand side entry is ignored (b_0 -> 0)
    if (pin0)
        for (int64_t i = 0; i < n; ++i) inv[(size_t)(i * n)] = 0.0;
}

// ------------------------------------------------------------ device buffers

template <typename T>
T* dev_copy(const std::vector<T>& v)
{
    T* p = nullptr;
    MG_HIP(hipMalloc(&p, std::max<size_t>(1, v.size()) * sizeof(T)));
    if (!v.empty()) MG_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return p;
}

template <typename T>
std::vector<T> cast_vec(const std::vector<double>& v)
{
    return std::vector<T>(v.begin(), v.end());
}

struct DevLevelBuf {
    uint16_t* cls = nullptr;
    // chunked 7-point 3D level: nodes the chunk kernels leave to the
    // per-node kernels, and the class the chunk kernels handle
    uint32_t* line_list = nullptr;
    uint32_t line_n = 0;
    int line_cls = -1;
    double *K = nullptr, *Ah = nullptr, *Rx = nullptr, *dinv = nullptr;
    float *Ah32 = nullptr, *Rx32 = nullptr, *dinv32 = nullptr;
    // vectors: FP64 x/b/t and FP32 x/b/t
    double *x = nullptr, *b = nullptr, *t = nullptr;
    float *xf = nullptr, *bf = nullptr, *tf = nullptr;
};

struct DevTransferBuf {
    int *pc[3] = {nullptr, nullptr, nullptr}, *rf[3] = {nullptr, nullptr, nullptr},
        *rn[3] = {nullptr, nullptr, nullptr};
    double *pw[3] = {nullptr, nullptr, nullptr}, *rw[3] = {nullptr, nullptr, nullptr};
    float *pw32[3] = {nullptr, nullptr, nullptr}, *rw32[3] = {nullptr, nullptr, nullptr};
};

}  // namespace

// ================================================================= solver

namespace psum_hip {

// k_vtail<T, D>'s dynamic LDS limit on the current device: raised once per
// device, precision and dimension to the device's per-workgroup maximum (the
// attribute is process-wide, so every solver shares the one largest
// value), and returned.  Called when a hierarchy is built, outside any
// stream capture.
template <typename T, int D>
int vtail_lds_limit()
{
    static std::mutex mu;
    static std::map<int, int> limit;
    int dev = 0;
    MG_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = limit.find(dev);
    if (it != limit.end()) return it->second;
    int lim = 0;
    MG_HIP(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    if (lim > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&k_vtail<T, D>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) {
        (void)hipGetLastError();
        lim = 65536;
    }
    limit[dev] = lim;
    return lim;
}

// kernel kinds of recorded ops (base + template variant bits)
constexpr uint16_t kKindSmooth = 0;      // + (FP64 << 3 | compact << 2 | prolong << 1 | acc)
constexpr uint16_t kKindRRTile = 16;     // + (FP64 << 1 | compact)
constexpr uint16_t kKindRR = 20;         // + (FP64 << 1 | compact)
constexpr uint16_t kKindCoarse = 24;     // + FP64
constexpr uint16_t kKindOuter = 26;      // + (norms << 2 | compact << 1 | emit FP32), 2D
constexpr uint16_t kKindVec = 34;        // + op (0 zero, 1 add correction, 2 copy)
constexpr uint16_t kKindHadamard = 37;   // FP64
constexpr uint16_t kKindCount = 38;

class HipMultigrid {
public:
    explicit HipMultigrid(bool mixed) : mixed_(mixed) {}
    ~HipMultigrid() { release(); }

    // options (sycl_multigrid_gpu.cpp:134-162, sycl_multigrid32f_gpu.cpp:162-196)
    void set_options(const char* options)
    {
        std::istringstream iss(options ? options : "");
        std::string key;
        bool k_provided = false;
        while (iss >> key) {
            if (key == "I") iss >> I_;
            else if (key == "J") iss >> J_;
            else if (key == "K") { iss >> K_; k_provided = true; }
            else if (key == "depth") iss >> depth_;
            else if (key == "pre_relax") iss >> pre_;
            else if (key == "post_relax") iss >> post_;
            else if (key == "max_cycles") iss >> max_cycles_;
            else if (key == "correction_interval") {
                int v = 0;
                if (!(iss >> v) || v < 1) throw std::invalid_argument("correction_interval must be positive");
                corr_ = v;
            } else if (key == "tol") iss >> tol_;
            else if (key == "periodic_mode") iss >> periodic_mode_;
            else if (key == "fuse") { iss >> fuse_; resolve_tail(); }
            else if (key == "tail") { iss >> tail_; resolve_tail(); }
            else if (key == "line") { iss >> line_; }
            else if (key == "graph") iss >> use_graph_;
            else if (key == "num_conv_mask" || key == "num_conv_mask_relax" ||
                     key == "num_conv_mask_restrict" || key == "num_conv_mask_interpolate" ||
                     key == "num_conv_mask_a") {
                int ignored;
                iss >> ignored;
            }
        }
        if (!k_provided) K_ = 1;
    }

    // Kernel-shape options that may change after set_matrix: "fuse 0|1"
    // (2D fused tile kernels), "tail 0|1" (the one-workgroup coarse end) and
    // "line 0|1" (chunked sweeps and residuals on 7-point 3D levels).
    // Other keys are rejected.
    void set_tuning(const char* options)
    {
        std::istringstream iss(options ? options : "");
        std::string key;
        while (iss >> key) {
            int v = 0;
            if (!(iss >> v)) throw std::invalid_argument("hip_multigrid: tuning option " + key + " needs a value");
            if (key == "tail") tail_ = v;
            else if (key == "fuse") fuse_ = v;
            else if (key == "line") line_ = v;
            else throw std::invalid_argument("hip_multigrid: unknown tuning option " + key);
        }
        rr_fit_.clear();
        resolve_tail();
        drop_graphs();   // captured cycles hold the previous kernel shapes
    }

    // Smoothing sweeps per level before and after the coarse correction
    // (pre_relax / post_relax), changed after set-up; captured cycles are
    // dropped.
    void set_smoothing(int pre, int post)
    {
        if (pre < 0 || post < 0 || pre > 64 || post > 64 || pre + post < 1)
            throw std::invalid_argument("hip_multigrid: smoothing sweeps need 0 <= pre, post <= 64, pre + post >= 1");
        pre_ = pre;
        post_ = post;
        rr_fit_.clear();
        resolve_tail();
        drop_graphs();
    }

    void set_stream(hipStream_t s) { stream_ = s; }

    void set_matrix(unsigned long long n_row, unsigned long long n_col, unsigned long long nnz,
                    const unsigned long long* rows, const unsigned long long* cols, const double* vals)
    {
        if (n_row != n_col) throw std::runtime_error("hip_multigrid: matrix must be square");
        if (I_ < 1 || J_ < 1 || K_ < 1 || depth_ < 1)
            throw std::invalid_argument("multigrid: positive grid dimensions and depth required");
        release();
        std::vector<HostLevel> L;
        int64_t pin = -1;
        L.push_back(level_from_coo(I_, J_, K_, n_row, nnz, rows, cols, vals, periodic_mode_ != 0,
                                   &pin, nullptr));
        pin_ = pin;
        std::vector<std::array<AxisTransfer, 3>> T;
        Axis ax[3];
        const int sz[3] = {I_, J_, K_};
        int per[3];
        for (int d = 0; d < 3; ++d) {
            for (int i = 0; i < sz[d]; ++i) ax[d].x.push_back(i);
            per[d] = periodic_mode_ != 0 && L[0].wrap[d];
        }
        for (int level = 0; level < depth_; ++level) {
            Axis co[3];
            int nc[3];
            bool same = true;
            std::array<AxisTransfer, 3> tr;
            for (int d = 0; d < 3; ++d) {
                const int nf = (int)ax[d].x.size();
                if (per[d]) {
                    // periodic axis: halve while even and >= 4, else keep
                    if (nf % 2 == 0 && nf >= 4) {
                        tr[d] = axis_transfer_periodic(nf);
                        for (int i = 0; i < nf; i += 2) co[d].x.push_back(ax[d].x[i]);
                    } else {
                        tr[d] = axis_transfer_identity(nf, 1);
                        co[d] = ax[d];
                    }
                } else {
                    co[d] = coarsen(ax[d]);
                }
                nc[d] = (int)co[d].x.size();
                if (co[d].x.size() != ax[d].x.size()) same = false;
            }
            if (same) break;   // endpoints cannot be coarsened further
            for (int d = 0; d < 3; ++d)
                if (!per[d]) tr[d] = axis_transfer(ax[d], co[d]);
            L.push_back(galerkin(L.back(), tr.data(), nc));
            T.push_back(tr);
            for (int d = 0; d < 3; ++d) ax[d] = co[d];
        }
        if (T.empty()) throw std::invalid_argument("multigrid: grid cannot be coarsened while retaining endpoints");
        if (L.back().N > 8192)
            throw std::runtime_error("hip_multigrid: coarse matrix too large (" + std::to_string(L.back().N) +
                                     " > 8192); increase depth");
        dense_inverse(L.back(), cinv_host_, pin_ >= 0);
        host_ = std::move(L);
        tr_host_ = std::move(T);
        rr_fit_.clear();
        upload();
        resolve_tail();
    }

    void set_source_replace(unsigned long long n, const unsigned long long* idx, const double* v)
    {
        // unique indices; the last value of a repeated index wins
        std::map<unsigned long long, double> m;
        for (unsigned long long i = 0; i < n; ++i) m[idx[i]] = v[i];
        upload_pairs(m, rep_idx_, rep_val_, rep_n_);
    }

    void set_source_addback(unsigned long long n, const unsigned long long* idx, const double* v)
    {
        // repeated indices accumulate (in argument order) before the device add
        std::map<unsigned long long, double> m;
        for (unsigned long long i = 0; i < n; ++i) m[idx[i]] += v[i];
        upload_pairs(m, add_idx_, add_val_, add_n_);
    }

    // solve(b, x): device pointers, b modified in place (addback, replace)
    // homogeneous: no addback and zero replace values (the linear part of the
    // solve, used by the adjoint)
    void solve(double* b, double* x, bool sync = true, bool homogeneous = false)
    {
        const double* bs = solve_begin(b, homogeneous);
        solve_core(b, bs, x);
        solve_end(b, x, sync);
    }

    // The three parts of solve(): begin (addback, replace, the consistent
    // right-hand side of a pinned periodic system; returns the right-hand
    // side the V-cycles see), core (the cycles) and end (gauge, sync).
    const double* solve_begin(double* b, bool homogeneous = false)
    {
        if (host_.empty()) throw std::runtime_error("hip_multigrid: solve before set_matrix");
        hipStream_t st = stream_;
        if (add_n_ && !homogeneous) hipLaunchKernelGGL(k_addback, dim3(nblk(add_n_)), dim3(kMgThreads), 0, st, b, add_idx_, add_val_, add_n_);
        if (rep_n_) hipLaunchKernelGGL(k_replace, dim3(nblk(rep_n_)), dim3(kMgThreads), 0, st, b, rep_idx_, rep_val_, rep_n_, homogeneous ? 1 : 0);
        MG_HIP(hipGetLastError());
        const int64_t N = N0();
        if (pin_ >= 0) {
            // consistent singular form: b' = b, b'_p = b_p - sum(b)
            MG_HIP(hipMemcpyAsync(bprime_, b, (size_t)N * sizeof(double), hipMemcpyDeviceToDevice, st));
            hipLaunchKernelGGL(k_psum, dim3(nblk(N)), dim3(kMgThreads), 0, st, b, N, part_);
            hipLaunchKernelGGL(k_fsum, dim3(1), dim3(kMgThreads), 0, st, part_, (int64_t)nblk(N), shift_);
            hipLaunchKernelGGL(k_set_pin, dim3(1), dim3(64), 0, st, bprime_, b, shift_, pin_);
            MG_HIP(hipGetLastError());
            return bprime_;
        }
        return b;
    }

    void solve_core(const double* b, const double* bs, double* x)
    {
        last_cycles_ = 0;
        last_rel_ = NAN;
        rec_iter_ = 1;                 // iteration 0 is the residual before the first cycle
        const bool use_tol = tol_ > 0.0;
        // With a tolerance, the first check comes after the number of cycles
        // the previous solve needed minus one (warm-started solves of a
        // time-stepping loop need similar counts); after that every cycle is
        // checked.  A check reads ||b - A x||^2 and ||b||^2 back to the host.
        const int first_check = std::max(1, std::min(max_cycles_, predicted_ - 1));
        if (!mixed_) {
            rec_phase(0, kStageOuter, 0);
            scale_rhs_f64(bs);
            if (!use_tol) {
                vcycles_f64(x, max_cycles_);
                last_cycles_ = max_cycles_;
            } else {
                double rr, bb;
                for (int c = 0; c < first_check; ++c) cycle_f64(bs, x);
                last_cycles_ = first_check;
                for (;;) {
                    outer_residual(bs, b, x, false, false, true, true, &rr, &bb);
                    last_rel_ = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                    if (last_rel_ <= tol_ || last_cycles_ >= max_cycles_) break;
                    cycle_f64(bs, x);
                    ++last_cycles_;
                }
            }
        } else {
            // refinement iteration: r = b - A x (FP64) -> FP32 V-cycles on A e = r -> x += e
            double rr, bb;
            const bool fm = fused_mixed();
            rec_phase(0, kStageOuter, 0);
            outer_residual(bs, b, x, true, fm, false, false, &rr, &bb);
            for (int cycle = 0; cycle < max_cycles_;) {
                if (use_tol && cycle >= first_check) {
                    read_norms(&rr, &bb);
                    last_rel_ = bb > 0 ? std::sqrt(rr / bb) : 0.0;
                    if (last_rel_ <= tol_) break;
                }
                if (use_tol) {
                    // the check before iteration cycle+1 reads this iteration's norms
                    iteration_f32(bs, b, x, cycle + 1 >= first_check);
                    cycle += 1;
                    last_cycles_ += 1;
                    continue;
                }
                // fixed count (the reference 32f backend): batches of
                // correction_interval V-cycles between FP64 residuals, no
                // host check (so the whole solve can be captured in a graph)
                const int batch = std::min(corr_, max_cycles_ - cycle);
                vcycles_f32(batch, fm, x);
                cycle += batch;
                last_cycles_ += batch;
                if (cycle < max_cycles_) {
                    rec_phase(rec_iter_ - 1, kStageOuter, 0);
                    outer_residual(bs, b, x, true, fm, false, false, &rr, &bb);
                }
            }
            if (use_tol) {
                if (last_cycles_ == 0) outer_residual(bs, b, x, false, false, true, false, &rr, &bb);
                read_norms(&rr, &bb);
                last_rel_ = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            }
        }
        if (use_tol) predicted_ = last_cycles_;
    }

    void solve_end(const double* b, double* x, bool sync)
    {
        hipStream_t st = stream_;
        if (pin_ >= 0) {
            hipLaunchKernelGGL(k_gauge_read, dim3(1), dim3(64), 0, st, x, b, pin_, shift_ + 1);
            hipLaunchKernelGGL(k_gauge_apply, dim3(nblk(N0())), dim3(kMgThreads), 0, st, x, N0(), shift_ + 1);
            MG_HIP(hipGetLastError());
        }
        if (sync) MG_HIP(hipStreamSynchronize(st));
    }

    // ---- batched fixed-cycle solves --------------------------------------
    // An op of the core of a fixed-cycle solve, recorded instead of launched:
    // the kernel body and variant (kind), its block count and its argument
    // struct.  Ops of different solvers with equal keys are independent and
    // run as one batched launch; keys grow along each solver's sequence.
    using RecOp = MgRecOp;

    // whether the core of a solve can be recorded: a fixed cycle count on a
    // 2D hierarchy through the fused tile kernels, without the one-workgroup
    // tail, with pre- and post-sweeps
    bool recordable() const
    {
        if (host_.empty() || tol_ > 0.0 || host_[0].D != 2 || !fuse_ || pre_ < 1 || post_ < 1) return false;
        for (const auto& h : host_)
            if (h.D != 2) return false;
        return tail_start<float>() < 0 && tail_start<double>() < 0;
    }

    // The ops of solve_core(b, bs, x), in order, without launching anything
    // (solve_begin must have produced bs; solve_end follows the ops).
    void record_core(const double* b, const double* bs, double* x, std::vector<RecOp>& ops)
    {
        if (!recordable()) throw std::logic_error("hip_multigrid: this solve cannot be batched");
        ops.clear();
        rec_ = &ops;
        try {
            solve_core(b, bs, x);
        } catch (...) {
            rec_ = nullptr;
            throw;
        }
        rec_ = nullptr;
    }

    double* pin_rhs() const { return pin_ >= 0 ? bprime_ : nullptr; }

    // fixed V-cycle count, one per refinement iteration, no host checks and
    // no graphs of its own: the solve can be captured into a caller's graph
    void set_fixed_cycles(int cycles)
    {
        if (cycles < 1) throw std::invalid_argument("hip_multigrid: fixed cycle count must be >= 1");
        tol_ = 0.0;
        max_cycles_ = cycles;
        corr_ = 1;
        use_graph_ = 0;
        for (auto& c : cap64_.e) drop_graph(c);
        for (auto& c : cap32_.e) drop_graph(c);
    }

    int last_cycles() const { return last_cycles_; }
    int predicted() const { return predicted_; }
    void set_predicted(int p) { predicted_ = p; }
    double last_rel() const { return last_rel_; }
    int64_t N0() const { return host_.empty() ? 0 : host_[0].N; }
    int levels() const { return (int)host_.size(); }

private:
    // ---- recording of batched solves (see record_core)
    static constexpr uint32_t kStageInit = 0, kStageDown = 1, kStageCoarse = 2, kStageUp = 3,
                              kStageCorr = 4, kStageOuter = 5;
    std::vector<RecOp>* rec_ = nullptr;
    uint32_t rec_iter_ = 0, rec_key_ = 0, rec_seq_ = 0;

    // key of the ops that follow: iteration, stage, level (ascending for the
    // downward pass, descending for the upward one), then emission order
    void rec_phase(uint32_t it, uint32_t stage, uint32_t lvl)
    {
        rec_key_ = (it << 20) | (stage << 16) | (lvl << 8);
        rec_seq_ = 0;
    }

    template <typename A>
    void emit(uint16_t kind, unsigned nb, const A& a)
    {
        RecOp op;
        op.key = rec_key_ | (rec_seq_ & 0xffu);
        ++rec_seq_;
        op.kind = kind;
        op.nblocks = nb;
        op.args.resize(sizeof(A));
        std::memcpy(op.args.data(), &a, sizeof(A));
        rec_->push_back(std::move(op));
    }

    template <typename T>
    static constexpr int tbit() { return std::is_same<T, double>::value ? 1 : 0; }

    template <typename T, bool C, bool P, bool A>
    void run_smooth_tile(unsigned nb, const SmoothTileArgs<T>& a)
    {
        if (rec_) { emit((uint16_t)(kKindSmooth + ((tbit<T>() << 3) | (C << 2) | (P << 1) | (A ? 1 : 0))), nb, a); return; }
        hipLaunchKernelGGL((k_smooth_tile<T, C, P, A>), dim3(nb), dim3(kTileThreads), 0, stream_, a);
    }

    template <typename T, bool C, bool TILE>
    void run_rr(unsigned nb, const ResidRestrictArgs<T>& a)
    {
        if (rec_) { emit((uint16_t)((TILE ? kKindRRTile : kKindRR) + ((tbit<T>() << 1) | (C ? 1 : 0))), nb, a); return; }
        if (TILE) hipLaunchKernelGGL((k_resid_restrict_tile<T, C>), dim3(nb), dim3(kTileThreads), 0, stream_, a);
        else hipLaunchKernelGGL((k_resid_restrict<T, C>), dim3(nb), dim3(kMgThreads), 0, stream_, a);
    }

    template <typename T>
    void run_coarse(unsigned nb, const CoarseArgs<T>& a)
    {
        if (rec_) { emit((uint16_t)(kKindCoarse + tbit<T>()), nb, a); return; }
        hipLaunchKernelGGL((k_coarse_solve<T>), dim3(nb), dim3(kMgThreads), 0, stream_, a);
    }

    template <int D, bool C, bool F32, bool NORMS>
    void run_outer(unsigned nb, const OuterArgs& a)
    {
        if (rec_) {
            if (D != 2) throw std::logic_error("hip_multigrid: only 2D solves are batched");
            emit((uint16_t)(kKindOuter + ((NORMS ? 4 : 0) | (C ? 2 : 0) | (F32 ? 1 : 0))), nb, a);
            return;
        }
        hipLaunchKernelGGL((k_outer_residual<D, C, F32, NORMS>), dim3(nb), dim3(kMgThreads), 0, stream_, a);
    }

    template <int OP>
    void run_vec(const VecArgs& a)
    {
        const unsigned nb = nblk(a.n);
        if (rec_) { emit((uint16_t)(kKindVec + OP), nb, a); return; }
        hipLaunchKernelGGL((k_vec<OP>), dim3(nb), dim3(kMgThreads), 0, stream_, a);
    }

    void run_hadamard(const HadamardArgs<double>& a)
    {
        const unsigned nb = nblk(a.L.N);
        if (rec_) { emit((uint16_t)kKindHadamard, nb, a); return; }
        hipLaunchKernelGGL((k_hadamard<double>), dim3(nb), dim3(kMgThreads), 0, stream_, a);
    }

    bool mixed_;
    int I_ = 0, J_ = 0, K_ = 0, depth_ = 3, pre_ = 2, post_ = 1, max_cycles_ = 10, corr_ = 4;
    int fuse_ = 1;   // 2D: fused multi-sweep tile smoothing and residual+restriction
    int line_ = 1;   // 3D 7-point levels: chunk kernels (k_smooth_line, k_residual_line)
    // the coarse end of the V-cycle in one workgroup (k_vtail); off unless
    // "tail 1" is set (measured: the levels' own launches are about as fast,
    // in 2D and in 3D)
    int tail_ = 0;
    int tail32_ = -2, tail64_ = -2;    // first tail level per precision (-1 none, -2 not yet chosen)
    size_t tail_bytes32_ = 0, tail_bytes64_ = 0;
    std::vector<int> rr_fit_;
    double tol_ = 0.0;
    int predicted_ = 1;                // cycles the previous solve needed
    int periodic_mode_ = 1;
    int64_t pin_ = -1;                 // pinned node of a pinned-periodic matrix
    double* bprime_ = nullptr;         // consistent singular right-hand side
    double* shift_ = nullptr;          // device scalar for the gauge shift / sums
    hipStream_t stream_ = nullptr;
    std::vector<HostLevel> host_;
    std::vector<std::array<AxisTransfer, 3>> tr_host_;
    std::vector<double> cinv_host_;
    std::vector<DevLevelBuf> dev_;
    std::vector<DevTransferBuf> dtr_;
    double* cinv_ = nullptr;
    double* part_ = nullptr;
    double* red_ = nullptr;
    double* red_host_ = nullptr;
    unsigned long long *rep_idx_ = nullptr, *add_idx_ = nullptr;
    double *rep_val_ = nullptr, *add_val_ = nullptr;
    int64_t rep_n_ = 0, add_n_ = 0, rep_cap_ = 0, add_cap_ = 0;
    int last_cycles_ = 0;
    double last_rel_ = NAN;

    void upload_pairs(const std::map<unsigned long long, double>& m, unsigned long long*& di,
                      double*& dv, int64_t& n)
    {
        std::vector<unsigned long long> hi;
        std::vector<double> hv;
        for (auto& kv : m) {
            if (!host_.empty() && kv.first >= (unsigned long long)host_[0].N)
                throw std::runtime_error("hip_multigrid: source index out of range");
            hi.push_back(kv.first);
            hv.push_back(kv.second);
        }
        if (di) (void)hipFree(di);
        if (dv) (void)hipFree(dv);
        di = nullptr;
        dv = nullptr;
        n = (int64_t)hi.size();
        if (n) {
            di = dev_copy(hi);
            dv = dev_copy(hv);
        }
    }

    void release()
    {
        for (auto& c : cap64_.e) drop_graph(c);
        for (auto& c : cap32_.e) drop_graph(c);
        for (auto& d : dev_) {
            for (void* p : {(void*)d.cls, (void*)d.line_list, (void*)d.K, (void*)d.Ah, (void*)d.Rx, (void*)d.dinv,
                            (void*)d.Ah32, (void*)d.Rx32, (void*)d.dinv32, (void*)d.t, (void*)d.bf,
                            (void*)d.xf, (void*)d.tf, (void*)d.b, (void*)d.x})
                if (p) (void)hipFree(p);
        }
        dev_.clear();
        for (auto& t : dtr_)
            for (int d = 0; d < 3; ++d)
                for (void* p : {(void*)t.pc[d], (void*)t.rf[d], (void*)t.rn[d], (void*)t.pw[d],
                                (void*)t.rw[d], (void*)t.pw32[d], (void*)t.rw32[d]})
                    if (p) (void)hipFree(p);
        dtr_.clear();
        if (cinv_) (void)hipFree(cinv_);
        if (part_) (void)hipFree(part_);
        if (red_) (void)hipFree(red_);
        if (red_host_) (void)hipHostFree(red_host_);
        if (shift_) (void)hipFree(shift_);
        if (bprime_) (void)hipFree(bprime_);
        shift_ = nullptr;
        bprime_ = nullptr;
        cinv_ = nullptr;
        part_ = nullptr;
        red_ = nullptr;
        red_host_ = nullptr;
    }

    void upload()
    {
        const int nl = (int)host_.size();
        dev_.resize(nl);
        for (int l = 0; l < nl; ++l) {
            const HostLevel& h = host_[l];
            DevLevelBuf& d = dev_[l];
            d.cls = dev_copy(h.cls);
            build_line_list(h, d);
            d.K = dev_copy(h.K);
            d.Ah = dev_copy(h.Ah);
            d.Rx = dev_copy(h.Rx);
            d.dinv = dev_copy(h.dinv);
            d.Ah32 = dev_copy(cast_vec<float>(h.Ah));
            d.Rx32 = dev_copy(cast_vec<float>(h.Rx));
            d.dinv32 = dev_copy(cast_vec<float>(h.dinv));
            const size_t N = (size_t)h.N;
            if (!mixed_) {
                if (l > 0) MG_HIP(hipMalloc(&d.x, N * sizeof(double)));
                if (l > 0) MG_HIP(hipMalloc(&d.b, N * sizeof(double)));
                else MG_HIP(hipMalloc(&d.b, N * sizeof(double)));
                MG_HIP(hipMalloc(&d.t, N * sizeof(double)));
            } else {
                MG_HIP(hipMalloc(&d.xf, N * sizeof(float)));
                MG_HIP(hipMalloc(&d.bf, N * sizeof(float)));
                MG_HIP(hipMalloc(&d.tf, N * sizeof(float)));
            }
        }
        dtr_.resize(tr_host_.size());
        for (size_t l = 0; l < tr_host_.size(); ++l)
            for (int d = 0; d < 3; ++d) {
                const AxisTransfer& t = tr_host_[l][d];
                dtr_[l].pc[d] = dev_copy(t.pc);
                dtr_[l].rf[d] = dev_copy(t.rf);
                dtr_[l].rn[d] = dev_copy(t.rn);
                dtr_[l].pw[d] = dev_copy(t.pw);
                dtr_[l].rw[d] = dev_copy(t.rw);
                dtr_[l].pw32[d] = dev_copy(cast_vec<float>(t.pw));
                dtr_[l].rw32[d] = dev_copy(cast_vec<float>(t.rw));
            }
        cinv_ = dev_copy(cinv_host_);
        const int64_t nb = (int64_t)nblk(host_[0].N);
        MG_HIP(hipMalloc(&part_, (size_t)nb * 2 * sizeof(double)));
        MG_HIP(hipMalloc(&red_, 2 * sizeof(double)));
        MG_HIP(hipMalloc(&shift_, 4 * sizeof(double)));
        if (pin_ >= 0) MG_HIP(hipMalloc(&bprime_, (size_t)host_[0].N * sizeof(double)));
        MG_HIP(hipHostMalloc(&red_host_, 2 * sizeof(double), hipHostMallocDefault));
    }

    // A 3D level of at least 2^16 nodes (and fewer than 2^32) runs its
    // smoothing sweeps and residuals in chunks of kLine nodes: the class that
    // covers most nodes goes to the chunk kernels, and every node of a chunk
    // they skip (by the same test as line_chunk_fast) is listed.
    static bool line_chunk_fast_host(const HostLevel& h, uint32_t i, int cls_in)
    {
        const uint32_t n0 = (uint32_t)h.n[0], n1 = (uint32_t)h.n[1], n2 = (uint32_t)h.n[2];
        const uint32_t q = i / n2, z = i - q * n2, x = q / n1, y = q - x * n1;
        if (z < 1u || z + kLine > n2 - 1u || y < 1u || y + 2u > n1 || x < 1u || x + 2u > n0) return false;
        for (int k = 0; k < kLine; ++k)
            if ((int)h.cls[i + k] != cls_in) return false;
        return true;
    }
    void build_line_list(const HostLevel& h, DevLevelBuf& d)
    {
        d.line_cls = -1;
        d.line_n = 0;
        if (h.D != 3 || h.N < (1 << 16) || h.N >= (int64_t)0xffffffffll) return;
        std::vector<int64_t> count(h.dinv.size(), 0);
        for (int64_t i = 0; i < h.N; ++i) ++count[h.cls[i]];
        const int cls_in = (int)(std::max_element(count.begin(), count.end()) - count.begin());
        std::vector<uint32_t> list;
        const uint32_t nchunks = (uint32_t)(h.N / kLine);
        for (uint32_t t = 0; t < nchunks; ++t)
            if (!line_chunk_fast_host(h, t * kLine, cls_in))
                for (int k = 0; k < kLine; ++k) list.push_back(t * kLine + k);
        for (int64_t i = (int64_t)nchunks * kLine; i < h.N; ++i) list.push_back((uint32_t)i);
        d.line_cls = cls_in;
        d.line_n = (uint32_t)list.size();
        if (!list.empty()) d.line_list = dev_copy(list);
    }
    // dinv of class k of level l as the FP32 table holds it
    float host_dinv32(int l, int k) const { return (float)host_[l].dinv[k]; }

    // 3D transfers writing level l over planes (k_prolong_add_plane,
    // k_restrict_plane): levels of 2^15 nodes or more
    bool planar(int l) const
    {
        return line_ && host_[l].D == 3 && host_[l].N >= (1 << 15) && host_[l].N < (int64_t)0xffffffffll &&
               (l == 0 || host_[l - 1].N < (int64_t)0xffffffffll);
    }

    LineList line_list(int l) const
    {
        LineList ll;
        ll.list = dev_[l].line_list;
        ll.n = dev_[l].line_n;
        ll.nbc = nblk(host_[l].N / kLine);
        return ll;
    }
    LineGeom line_geom(int l) const
    {
        LineGeom g;
        g.n0 = host_[l].n[0];
        g.n1 = host_[l].n[1];
        g.n2 = host_[l].n[2];
        g.nchunks = (uint32_t)(host_[l].N / kLine);
        g.cls = dev_[l].cls;
        g.cls_in = dev_[l].line_cls;
        return g;
    }

    DevLevel dl(int l) const
    {
        DevLevel d;
        for (int q = 0; q < 3; ++q) {
            d.n[q] = host_[l].n[q];
            d.wrap[q] = host_[l].wrap[q];
        }
        d.N = host_[l].N;
        d.cls = dev_[l].cls;
        return d;
    }

    template <typename T>
    DevTable<T> table(int l, bool relax) const
    {
        DevTable<T> t;
        if constexpr (std::is_same<T, double>::value) {
            t.coef = relax ? dev_[l].Rx : dev_[l].Ah;
            t.dinv = dev_[l].dinv;
        } else {
            t.coef = relax ? dev_[l].Rx32 : dev_[l].Ah32;
            t.dinv = dev_[l].dinv32;
        }
        return t;
    }

    template <typename T>
    DevTransfer<T> transfer(int l) const
    {
        DevTransfer<T> t;
        for (int d = 0; d < 3; ++d) {
            t.wrap[d] = tr_host_[l][d].wrap;
            t.pc[d] = dtr_[l].pc[d];
            t.rf[d] = dtr_[l].rf[d];
            t.rn[d] = dtr_[l].rn[d];
            if constexpr (std::is_same<T, double>::value) {
                t.pw[d] = dtr_[l].pw[d];
                t.rw[d] = dtr_[l].rw[d];
            } else {
                t.pw[d] = dtr_[l].pw32[d];
                t.rw[d] = dtr_[l].rw32[d];
            }
        }
        return t;
    }

    template <typename T>
    void smooth(int l, T*& x, T*& buf, const T* b, double* acc = nullptr)
    {
        const DevLevel L = dl(l);
        const T two3 = (T)(2.0 / 3);
        const unsigned nb = nblk(L.N);
#define SMOOTH(DIM, CMP) \
        if (acc) hipLaunchKernelGGL((k_smooth<T, DIM, CMP, true>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, true), x, b, buf, two3, acc); \
        else hipLaunchKernelGGL((k_smooth<T, DIM, CMP, false>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, true), x, b, buf, two3, nullptr)
        if (line_ && dev_[l].line_cls >= 0) {
            const LineGeom G = line_geom(l);
            const LineList LL = line_list(l);
            const DevTable<T> tb = table<T>(l, true);
            const T* row = tb.coef + (int64_t)G.cls_in * 27;
            const unsigned nb = line_grid(LL);
#define SLINE(CMP, ACC, A) hipLaunchKernelGGL((k_smooth_line<T, CMP, ACC>), dim3(nb), dim3(kMgThreads), 0, stream_, G, LL, L, tb, row, x, b, buf, two3, A)
            if (host_[l].compact) { if (acc) SLINE(true, true, acc); else SLINE(true, false, nullptr); }
            else { if (acc) SLINE(false, true, acc); else SLINE(false, false, nullptr); }
#undef SLINE
        } else if (host_[l].D == 2) {
            if (host_[l].compact) { SMOOTH(2, true); } else { SMOOTH(2, false); }
        } else {
            if (host_[l].compact) { SMOOTH(3, true); } else { SMOOTH(3, false); }
        }
#undef SMOOTH
        std::swap(x, buf);
    }

    template <typename T>
    void residual(int l, const T* x, const T* b, T* r)
    {
        const DevLevel L = dl(l);
        const unsigned nb = nblk(L.N);
        if (line_ && dev_[l].line_cls >= 0) {
            const LineGeom G = line_geom(l);
            const LineList LL = line_list(l);
            const DevTable<T> tb = table<T>(l, false);
            const T* row = tb.coef + (int64_t)G.cls_in * 27;
            if (host_[l].compact) hipLaunchKernelGGL((k_residual_line<T, true>), dim3(line_grid(LL)), dim3(kMgThreads), 0, stream_, G, LL, L, tb, row, x, b, r);
            else hipLaunchKernelGGL((k_residual_line<T, false>), dim3(line_grid(LL)), dim3(kMgThreads), 0, stream_, G, LL, L, tb, row, x, b, r);
            return;
        }
        if (host_[l].D == 2) {
            if (host_[l].compact) hipLaunchKernelGGL((k_residual<T, 2, true>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, false), x, b, r);
            else hipLaunchKernelGGL((k_residual<T, 2, false>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, false), x, b, r);
        } else {
            if (host_[l].compact) hipLaunchKernelGGL((k_residual<T, 3, true>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, false), x, b, r);
            else hipLaunchKernelGGL((k_residual<T, 3, false>), dim3(nb), dim3(kMgThreads), 0, stream_, L, table<T>(l, false), x, b, r);
        }
    }

    // `count` sweeps on level l: 2D levels in fused tile launches of up to
    // kMaxFuse sweeps (PROLONG: the first launch also applies x += P x_c)
    // acc (FP32 levels only): the last launch also adds its result into the
    // FP64 array acc; returns whether it did (only the fused 2D path can)
    template <typename T>
    bool smooth_n(int l, T*& x, T*& buf, const T* b, int count, const T* xc = nullptr,
                  double* acc = nullptr)
    {
        if (host_[l].D != 2 || !fuse_ || count == 0) {
            if (xc) prolong(l, xc, x);
            for (int s = 0; s < count; ++s) smooth<T>(l, x, buf, b, s == count - 1 ? acc : nullptr);
            return acc != nullptr && count > 0;
        }
        const DevLevel L = dl(l);
        const T two3 = (T)(2.0 / 3);
        const DevLevel C = xc ? dl(l + 1) : L;
        const DevTransfer<T> tr = xc ? transfer<T>(l) : DevTransfer<T>{};
        const bool cmp = host_[l].compact;
        const int ncls = (int)host_[l].dinv.size();
        bool accd = false;
        while (count > 0) {
            const int k = std::min(count, kMaxFuse);
            const int ostep = kTileX - 2 * k;
            const unsigned nb = (unsigned)(((L.n[0] + kTileRows - 1) / kTileRows) * ((L.n[1] + ostep - 1) / ostep));
            double* a = (k == count) ? acc : nullptr;
            const SmoothTileArgs<T> ta{L, table<T>(l, true), ncls, x, b, buf, two3, k, C, tr, xc, a};
#define TILE(CMP, PRO, ACC) run_smooth_tile<T, CMP, PRO, ACC>(nb, ta)
            if (a) {
                if (xc) { if (cmp) TILE(true, true, true); else TILE(false, true, true); }
                else { if (cmp) TILE(true, false, true); else TILE(false, false, true); }
                accd = true;
            } else {
                if (xc) { if (cmp) TILE(true, true, false); else TILE(false, true, false); }
                else { if (cmp) TILE(true, false, false); else TILE(false, false, false); }
            }
#undef TILE
            std::swap(x, buf);
            xc = nullptr;
            count -= k;
        }
        return accd;
    }

    // whether every coarse tile of level l+1 has its fine support within the
    // tiled residual + restriction kernel's LDS tile (checked once per level)
    bool rr_tile_fits(int l)
    {
        if ((int)rr_fit_.size() != (int)host_.size()) rr_fit_.assign(host_.size(), -1);
        if (rr_fit_[l] >= 0) return rr_fit_[l] == 1;
        const HostLevel& C = host_[l + 1];
        bool ok = true;
        const int lim[2] = {kRRFineRows, kRRFineCols}, step[2] = {kRRRows, kRRCols};
        for (int d = 0; d < 2 && ok; ++d) {
            const AxisTransfer& t = tr_host_[l][d];
            for (int lo = 0; lo < C.n[d] && ok; lo += step[d]) {
                const int hi = std::min(lo + step[d], C.n[d]) - 1;
                for (int c = lo; c < hi; ++c)
                    if (t.rf[c + 1] < t.rf[c] || t.rf[c + 1] + t.rn[c + 1] < t.rf[c] + t.rn[c]) ok = false;
                if (t.rf[hi] + t.rn[hi] - t.rf[lo] > lim[d]) ok = false;
            }
        }
        rr_fit_[l] = ok ? 1 : 0;
        return ok;
    }

    template <typename T>
    void prolong(int l, const T* xc, T* x)
    {
        const DevLevel F = dl(l), C = dl(l + 1);
        if (host_[l].D == 2)
            hipLaunchKernelGGL((k_prolong_add<T, 2>), dim3(nblk(F.N)), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), xc, x);
        else if (planar(l))
            hipLaunchKernelGGL((k_prolong_add_plane<T>), dim3(nblk((int64_t)F.n[1] * F.n[2]), F.n[0]), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), xc, x);
        else
            hipLaunchKernelGGL((k_prolong_add<T, 3>), dim3(nblk(F.N)), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), xc, x);
    }

    // First level of the one-workgroup tail for precision T, or -1: a 2D
    // fused or a 3D hierarchy whose levels lt..Lc (two or more, at most
    // kMaxTail, lt >= 1, at most tail_level_nodes<D> nodes each) keep their
    // iterates and right-hand sides in LDS and whose coarsest level has at
    // most tail_coarse<D> nodes (its dense solve reads the inverse from one
    // compute unit).
    // Chosen by resolve_tail() when the hierarchy is built or the option
    // changes, never while a stream is being captured.
    template <typename T>
    int tail_start() const
    {
        const int c = std::is_same<T, double>::value ? tail64_ : tail32_;
        if (c == -2) throw std::logic_error("hip_multigrid: V-cycle tail used before it was resolved");
        return c;
    }

    template <typename T>
    void resolve_tail_for()
    {
        int& c = std::is_same<T, double>::value ? tail64_ : tail32_;
        size_t& bytes = std::is_same<T, double>::value ? tail_bytes64_ : tail_bytes32_;
        c = -1;
        bytes = 0;
        const int nl = (int)host_.size(), Lc = nl - 1;
        if (nl < 3) return;
        const int D = host_[0].D;
        if (D != 2 && D != 3) return;
        for (const auto& h : host_)
            if (h.D != D) return;
        if (tail_ != 1 || (D == 2 && !fuse_)) return;
        const int64_t coarse = D == 2 ? tail_coarse<2>() : tail_coarse<3>();
        const int64_t maxn = D == 2 ? tail_level_nodes<2>() : tail_level_nodes<3>();
        if (host_[Lc].N > coarse) return;
        const int lim = D == 2 ? vtail_lds_limit<T, 2>() : vtail_lds_limit<T, 3>();
        for (int lt = 1; lt <= Lc - 1; ++lt) {
            if (Lc - lt + 1 > kMaxTail || host_[lt].N > maxn) continue;
            TailArgs<T> probe{};
            const size_t b = tail_layout<T>(lt, probe);
            if (b > (size_t)lim) continue;
            c = lt;
            bytes = b;
            break;
        }
    }

    // Both precisions' tails for the current hierarchy and options.
    void resolve_tail()
    {
        if (host_.empty()) {
            tail32_ = tail64_ = -2;
            return;
        }
        resolve_tail_for<float>();
        resolve_tail_for<double>();
    }

    // Fills the tail arguments for levels lt..Lc; returns the LDS bytes
    template <typename T>
    size_t tail_layout(int lt, TailArgs<T>& a) const
    {
        const int Lc = (int)host_.size() - 1;
        const int D = host_[lt].D;
        const int K = D == 2 ? 9 : 27;
        size_t off = 0;
        auto take = [&](size_t nbytes) {
            const size_t o = off;
            off = (off + nbytes + 15) & ~(size_t)15;
            return (int)o;
        };
        a.nl = Lc - lt + 1;
        for (int k = 0; k < a.nl; ++k) {
            const int l = lt + k;
            const HostLevel& h = host_[l];
            TailLevel<T>& t = a.lev[k];
            for (int d = 0; d < 3; ++d) {
                t.n[d] = d < D ? (int)h.n[d] : 1;
                t.wrap[d] = d < D ? h.wrap[d] : 0;
            }
            t.N = (int)h.N;
            t.compact = h.compact ? 1 : 0;
            t.ncls = (int)h.dinv.size();
            t.ox = take((size_t)h.N * sizeof(T));
            t.oy = l < Lc ? take((size_t)h.N * sizeof(T)) : t.ox;
            t.ob = take((size_t)h.N * sizeof(T));
            t.ocls = take((size_t)h.N * sizeof(uint16_t));
            t.orx = take((size_t)t.ncls * K * sizeof(T));
            t.oa = take((size_t)t.ncls * K * sizeof(T));
            t.odinv = take((size_t)t.ncls * sizeof(T));
            if (!dev_.empty()) {
                const DevTable<T> rx = table<T>(l, true), A = table<T>(l, false);
                t.cls = dev_[l].cls;
                t.rx = rx.coef;
                t.a = A.coef;
                t.dinv = A.dinv;
            }
        }
        for (int k = 0; k + 1 < a.nl; ++k) {
            const int l = lt + k;
            const HostLevel &f = host_[l], &c = host_[l + 1];
            TailTransfer<T>& r = a.tr[k];
            const DevTransfer<T> dt = dtr_.empty() ? DevTransfer<T>{} : transfer<T>(l);
            for (int d = 0; d < D; ++d) {
                r.wrap[d] = tr_host_[l][d].wrap;
                r.orf[d] = take((size_t)c.n[d] * sizeof(int));
                r.orn[d] = take((size_t)c.n[d] * sizeof(int));
                r.orw[d] = take((size_t)3 * c.n[d] * sizeof(T));
                r.opc[d] = take((size_t)f.n[d] * sizeof(int));
                r.opw[d] = take((size_t)2 * f.n[d] * sizeof(T));
                if (!dtr_.empty()) {
                    r.rf[d] = dt.rf[d];
                    r.rn[d] = dt.rn[d];
                    r.rw[d] = dt.rw[d];
                    r.pc[d] = dt.pc[d];
                    r.pw[d] = dt.pw[d];
                }
            }
        }
        return off;
    }

    template <typename T>
    void launch_tail(int lt, T* x, const T* b)
    {
        TailArgs<T> a{};
        const size_t bytes = tail_layout<T>(lt, a);
        a.pre = pre_;
        a.post = post_;
        a.two3 = (T)(2.0 / 3);
        a.x_in = x;
        a.b_in = b;
        a.x_out = x;
        a.cinv = cinv_;
        const int D = host_[lt].D;
        const int lim = D == 2 ? vtail_lds_limit<T, 2>() : vtail_lds_limit<T, 3>();
        const size_t chosen = std::is_same<T, double>::value ? tail_bytes64_ : tail_bytes32_;
        if (bytes != chosen || bytes > (size_t)lim)
            throw std::logic_error("hip_multigrid: V-cycle tail LDS size differs from the one resolved");
        if (D == 2) hipLaunchKernelGGL((k_vtail<T, 2>), dim3(1), dim3(kTailThreads), bytes, stream_, a);
        else hipLaunchKernelGGL((k_vtail<T, 3>), dim3(1), dim3(kTailThreads), bytes, stream_, a);
        MG_HIP(hipGetLastError());
    }

    // One or more V-cycles on level-0 arrays (b0 scaled, x0 warm start).
    // xs[l] / ts[l] are the ping-pong pair of level l, bs[l] its rhs.
    // first_swept: xs[0] already holds the first level-0 sweep from zero.
    // acc: the last cycle's final level-0 sweep adds its result into acc
    // (returns whether that happened).
    template <typename T>
    bool vcycles(std::vector<T*>& xs, std::vector<T*>& ts, std::vector<T*>& bs, int cycles,
                 bool first_swept = false, double* acc = nullptr)
    {
        const int nl = (int)host_.size();
        const int Lc = nl - 1;
        const T two3 = (T)(2.0 / 3);
        bool accd = false;
        const int lt = tail_start<T>();          // > 0: levels lt..Lc run in k_vtail
        const int ld = lt > 0 ? lt : Lc;
        for (int cyc = 0; cyc < cycles; ++cyc) {
            const uint32_t it = rec_iter_ + (uint32_t)cyc;
            for (int l = 0; l < ld; ++l) {
                rec_phase(it, kStageDown, (uint32_t)l);
                // coarse levels received their first sweep from the restriction
                const bool swept = (l > 0 || (cyc == 0 && first_swept)) && pre_ > 0;
                smooth_n<T>(l, xs[l], ts[l], bs[l], swept ? pre_ - 1 : pre_);
                const DevLevel F = dl(l), C = dl(l + 1);
                const T fs = (l + 1 < Lc && pre_ > 0) ? two3 : (T)0;
                if (host_[l].D == 2 && fuse_) {
                    if (rr_tile_fits(l)) {
                        const unsigned nbt = (unsigned)(((C.n[0] + kRRRows - 1) / kRRRows) * ((C.n[1] + kRRCols - 1) / kRRCols));
                        const int ncls = (int)host_[l].dinv.size();
                        const ResidRestrictArgs<T> ra{F, C, transfer<T>(l), table<T>(l, false), ncls, table<T>(l + 1, false), xs[l], bs[l], bs[l + 1], xs[l + 1], fs};
                        if (host_[l].compact) run_rr<T, true, true>(nbt, ra);
                        else run_rr<T, false, true>(nbt, ra);
                    } else {
                        const ResidRestrictArgs<T> ra{F, C, transfer<T>(l), table<T>(l, false), 0, table<T>(l + 1, false), xs[l], bs[l], bs[l + 1], xs[l + 1], fs};
                        if (host_[l].compact) run_rr<T, true, false>(nblk(C.N), ra);
                        else run_rr<T, false, false>(nblk(C.N), ra);
                    }
                    continue;
                }
                residual<T>(l, xs[l], bs[l], ts[l]);
                if (host_[l].D == 2)
                    hipLaunchKernelGGL((k_restrict<T, 2>), dim3(nblk(C.N)), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), table<T>(l + 1, false), ts[l], bs[l + 1], xs[l + 1], fs);
                else if (planar(l + 1))
                    hipLaunchKernelGGL((k_restrict_plane<T>), dim3(nblk((int64_t)C.n[1] * C.n[2]), C.n[0]), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), table<T>(l + 1, false), ts[l], bs[l + 1], xs[l + 1], fs);
                else
                    hipLaunchKernelGGL((k_restrict<T, 3>), dim3(nblk(C.N)), dim3(kMgThreads), 0, stream_, F, C, transfer<T>(l), table<T>(l + 1, false), ts[l], bs[l + 1], xs[l + 1], fs);
            }
            rec_phase(it, kStageCoarse, 0);
            if (lt > 0) {
                launch_tail<T>(lt, xs[lt], bs[lt]);
            } else {
                const int n = (int)host_[Lc].N;
                const unsigned nb = (unsigned)(((int64_t)n * 64 + kMgThreads - 1) / kMgThreads);
                run_coarse<T>(nb, CoarseArgs<T>{cinv_, n, bs[Lc], xs[Lc]});
            }
            for (int l = ld - 1; l >= 0; --l) {
                rec_phase(it, kStageUp, 255u - (uint32_t)l);
                if (post_ > 0) {
                    double* a = (l == 0 && cyc == cycles - 1) ? acc : nullptr;
                    accd = smooth_n<T>(l, xs[l], ts[l], bs[l], post_, xs[l + 1], a) || accd;
                } else {
                    prolong<T>(l, xs[l + 1], xs[l]);
                }
            }
        }
        rec_iter_ += (uint32_t)cycles;
        MG_HIP(hipGetLastError());
        return accd;
    }

    // b0 = b o D0^-1 into dev_[0].b (once per solve: the V-cycles read it)
    void scale_rhs_f64(const double* b)
    {
        run_hadamard(HadamardArgs<double>{dl(0), table<double>(0, false), b, dev_[0].b});
    }

    // FP64 V-cycles on x with the scaled right-hand side in dev_[0].b
    void vcycles_f64(double* x, int cycles)
    {
        const int nl = (int)host_.size();
        std::vector<double*> xs(nl), ts(nl), bs(nl);
        for (int l = 0; l < nl; ++l) {
            xs[l] = l == 0 ? x : dev_[l].x;
            ts[l] = dev_[l].t;
            bs[l] = dev_[l].b;
        }
        vcycles<double>(xs, ts, bs, cycles);
        rec_phase(rec_iter_ - 1, kStageCorr, 0);
        if (xs[0] != x) run_vec<2>(VecArgs{x, xs[0], N0()});
    }

    // The mixed iteration fuses work into its neighbours when the level-0
    // smoothing runs through the 2D tile kernels: the outer residual also
    // writes the first sweep of e from zero, and the last post-sweep adds e
    // into x.
    // the mixed solve's first level-0 sweep from e = 0 taken from the FP64
    // outer residual (2D with the fused tile kernels, 3D with the chunk
    // kernels; "fuse 0" / "line 0" run it as a sweep of its own)
    bool fused_mixed() const { return (host_[0].D == 2 ? fuse_ : line_) && pre_ > 0 && post_ > 0; }

    // FP32 V-cycles on dev_[0].bf (already scaled) for the correction e from
    // e = 0.  first_swept: dev_[0].xf already holds the first level-0 sweep.
    // x != nullptr: x += e is applied (fused when possible).  Otherwise e
    // ends in dev_[0].xf (a fixed buffer, so the sequence can be captured).
    void vcycles_f32(int cycles, bool first_swept, double* x)
    {
        const int nl = (int)host_.size();
        std::vector<float*> xs(nl), ts(nl), bs(nl);
        for (int l = 0; l < nl; ++l) {
            xs[l] = dev_[l].xf;
            ts[l] = dev_[l].tf;
            bs[l] = dev_[l].bf;
        }
        rec_phase(rec_iter_, kStageInit, 0);
        if (!first_swept)
            run_vec<0>(VecArgs{xs[0], nullptr, N0()});
        const bool accd = vcycles<float>(xs, ts, bs, cycles, first_swept, x);
        rec_phase(rec_iter_ - 1, kStageCorr, 0);
        if (x && !accd) {
            run_vec<1>(VecArgs{x, xs[0], N0()});
        } else if (!x && xs[0] != dev_[0].xf) {
            MG_HIP(hipMemcpyAsync(dev_[0].xf, xs[0], (size_t)N0() * sizeof(float), hipMemcpyDeviceToDevice, stream_));
        }
    }

    // One V-cycle (FP64) or one refinement iteration (mixed: FP32 V-cycle,
    // x += e, next residual), replayed from a captured hipGraph when the
    // solver runs on its own stream.  Graphs are keyed on (b, x); a few are
    // kept so that callers alternating between solves on different buffers
    // (a forward and an adjoint solve) do not re-capture each time.
    struct Captured {
        hipGraphExec_t exec = nullptr;
        const void* b = nullptr;
        const void* x = nullptr;
        int flag = 0;
    };
    struct GraphCache {
        Captured e[4];
        int next = 0;
    };
    GraphCache cap64_, cap32_;
    int use_graph_ = 1;

    void drop_graph(Captured& c)
    {
        if (c.exec) (void)hipGraphExecDestroy(c.exec);
        c = Captured{};
    }
    void drop_graphs()
    {
        for (auto* gc : {&cap64_, &cap32_})
            for (auto& e : gc->e) drop_graph(e);
    }

    template <typename F>
    void replay(GraphCache& gc, const void* b, const void* x, int flag, F&& body)
    {
        if (!use_graph_ || stream_ == nullptr) {
            body();
            return;
        }
        Captured* c = nullptr;
        for (auto& e : gc.e)
            if (e.exec && e.b == b && e.x == x && e.flag == flag) c = &e;
        if (!c) {
            c = &gc.e[gc.next];
            gc.next = (gc.next + 1) % 4;
            drop_graph(*c);
            MG_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
            CaptureGuard guard(stream_);
            body();
            hipGraph_t graph = nullptr;
            guard.release();
            MG_HIP(hipStreamEndCapture(stream_, &graph));
            MG_HIP(hipGraphInstantiate(&c->exec, graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
            c->b = b;
            c->x = x;
            c->flag = flag;
        }
        MG_HIP(hipGraphLaunch(c->exec, stream_));
    }

    void cycle_f64(const double* bs, double* x)
    {
        replay(cap64_, bs, x, 0, [&]() { vcycles_f64(x, 1); });
    }

    // one refinement iteration: FP32 V-cycle on A e = r, x += e, next
    // residual r = b - A x (FP64); norms: also reduce ||r||^2, ||b||^2 for a
    // host check
    void iteration_f32(const double* bs, const double* b, double* x, bool norms)
    {
        replay(cap32_, bs, x, norms ? 1 : 0, [&]() {
            const bool fm = fused_mixed();
            vcycles_f32(1, fm, x);
            outer_residual(bs, b, x, true, fm, norms, false, nullptr, nullptr);
        });
    }

    // ||b - A x||^2 and ||bn||^2 into red_ (device) and, asynchronously, into
    // pinned host memory; `wait` synchronises and returns them.
    // emit_f32: also the scaled FP32 residual into dev_[0].bf; first_sweep:
    // and the first sweep of e from zero into dev_[0].xf; norms: reduce the
    // norms and start their copy to the host.
    void outer_residual(const double* b, const double* bn, const double* x, bool emit_f32,
                        bool first_sweep, bool norms, bool wait, double* rr, double* bb)
    {
        const DevLevel L = dl(0);
        const unsigned nb = nblk(L.N);
        DevTable<double> K;
        K.coef = dev_[0].K;
        K.dinv = dev_[0].dinv;
        const bool c = host_[0].compact;
        float* b0 = emit_f32 ? dev_[0].bf : nullptr;
        const float* d32 = dev_[0].dinv32;
        float* e0 = (emit_f32 && first_sweep) ? dev_[0].xf : nullptr;
        const bool dom = line_ && host_[0].D == 3 && c && dev_[0].line_cls >= 0;
        const OuterArgs oa{L, K, x, b, bn, pin_, d32, b0, part_, e0, dom ? dev_[0].line_cls : -1,
                           dom ? K.coef + (int64_t)dev_[0].line_cls * 27 : nullptr};
        if (!norms && emit_f32 && !rec_ && line_ && dev_[0].line_cls >= 0 && host_[0].compact) {
            const LineGeom G = line_geom(0);
            const LineList LL = line_list(0);
            hipLaunchKernelGGL(k_outer_line, dim3(line_grid(LL)), dim3(kMgThreads), 0, stream_, G, LL, oa,
                               K.coef + (int64_t)G.cls_in * 27, host_dinv32(0, G.cls_in));
            MG_HIP(hipGetLastError());
            return;
        }
        // the squared norms only when they are read (norms); otherwise the
        // kernel skips bn and the reductions
#define OUTER(DIM, CMP, F32) \
        do { if (norms) run_outer<DIM, CMP, F32, true>(nb, oa); else run_outer<DIM, CMP, F32, false>(nb, oa); } while (0)
        if (host_[0].D == 2) {
            if (c) { if (emit_f32) OUTER(2, true, true); else OUTER(2, true, false); }
            else { if (emit_f32) OUTER(2, false, true); else OUTER(2, false, false); }
        } else {
            if (c) { if (emit_f32) OUTER(3, true, true); else OUTER(3, true, false); }
            else { if (emit_f32) OUTER(3, false, true); else OUTER(3, false, false); }
        }
#undef OUTER
        MG_HIP(hipGetLastError());
        if (!norms) return;
        hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kMgThreads), 0, stream_, part_, (int64_t)nb, red_);
        MG_HIP(hipGetLastError());
        MG_HIP(hipMemcpyAsync(red_host_, red_, 2 * sizeof(double), hipMemcpyDeviceToHost, stream_));
        if (wait) read_norms(rr, bb);
    }

    void read_norms(double* rr, double* bb)
    {
        MG_HIP(hipStreamSynchronize(stream_));
        *rr = red_host_[0];
        *bb = red_host_[1];
    }
};

// ---- batched launches of recorded solves -----------------------------------
// Ops of all solvers are ordered by (key, kind, solver); each run of equal
// (key, kind) becomes one launch of the kind's batched kernel over the
// concatenated blocks of its solvers.  Every solver's ops keep their own
// order (keys grow along each solver's sequence), and ops sharing a launch
// belong to different, independent solvers.
struct MgBatchGroup {
    uint16_t kind;
    unsigned blocks;
    int n;
    size_t args_off, first_off;
};

struct MgBatchPlan {
    std::vector<MgBatchGroup> groups;
    void* dev = nullptr;
    size_t bytes = 0;
    int64_t ops = 0;
    ~MgBatchPlan() { release(); }
    void release()
    {
        if (dev) (void)hipFree(dev);
        dev = nullptr;
        bytes = 0;
        groups.clear();
        ops = 0;
    }
};

void mg_batch_build(const std::vector<const std::vector<MgRecOp>*>& per, MgBatchPlan& plan)
{
    plan.release();
    struct E {
        uint32_t key;
        uint16_t kind;
        int inst;
        const MgRecOp* op;
    };
    std::vector<E> all;
    for (int i = 0; i < (int)per.size(); ++i)
        for (const auto& op : *per[i]) {
            if (op.kind >= kKindCount) throw std::logic_error("hip_multigrid: unknown recorded kind");
            all.push_back(E{op.key, op.kind, i, &op});
        }
    std::stable_sort(all.begin(), all.end(), [](const E& a, const E& b) {
        if (a.key != b.key) return a.key < b.key;
        if (a.kind != b.kind) return a.kind < b.kind;
        return a.inst < b.inst;
    });
    std::vector<unsigned char> host;
    auto align16 = [&]() { host.resize((host.size() + 15) & ~(size_t)15, 0); };
    for (size_t a = 0; a < all.size();) {
        size_t b = a;
        while (b < all.size() && all[b].key == all[a].key && all[b].kind == all[a].kind) ++b;
        MgBatchGroup g{};
        g.kind = all[a].kind;
        g.n = (int)(b - a);
        const size_t asz = all[a].op->args.size();
        align16();
        g.args_off = host.size();
        for (size_t k = a; k < b; ++k) {
            if (all[k].op->args.size() != asz) throw std::logic_error("hip_multigrid: argument size mismatch");
            host.insert(host.end(), all[k].op->args.begin(), all[k].op->args.end());
        }
        align16();
        g.first_off = host.size();
        uint64_t blocks = 0;
        for (size_t k = a; k < b; ++k) {
            const unsigned f = (unsigned)blocks;
            const unsigned char* q = reinterpret_cast<const unsigned char*>(&f);
            host.insert(host.end(), q, q + sizeof(unsigned));
            blocks += all[k].op->nblocks;
        }
        if (blocks > 0x7fffffffull) throw std::runtime_error("hip_multigrid: batched launch too large");
        g.blocks = (unsigned)blocks;
        if (g.blocks > 0) plan.groups.push_back(g);
        a = b;
    }
    plan.ops = (int64_t)all.size();
    plan.bytes = std::max<size_t>(host.size(), 16);
    MG_HIP(hipMalloc(&plan.dev, plan.bytes));
    if (!host.empty()) MG_HIP(hipMemcpy(plan.dev, host.data(), host.size(), hipMemcpyHostToDevice));
}

void mg_batch_launch(const MgBatchPlan& plan, hipStream_t st)
{
    const unsigned char* base = static_cast<const unsigned char*>(plan.dev);
    for (const auto& g : plan.groups) {
        const void* a = base + g.args_off;
        const unsigned* f = reinterpret_cast<const unsigned*>(base + g.first_off);
        const dim3 grid(g.blocks);
        const uint16_t k = g.kind;
        if (k < kKindRRTile) {
            const int v = k - kKindSmooth;
#define SMT(T, C, P, A)                                                                             \
    hipLaunchKernelGGL((k_smooth_tile_many<T, C, P, A>), grid, dim3(kTileThreads), 0, st,           \
                       static_cast<const SmoothTileArgs<T>*>(a), f, g.n)
#define SMT4(T, C)                                            \
    switch (v & 3) {                                          \
    case 0: SMT(T, C, false, false); break;                   \
    case 1: SMT(T, C, false, true); break;                    \
    case 2: SMT(T, C, true, false); break;                    \
    default: SMT(T, C, true, true); break;                    \
    }
            if (v & 8) { if (v & 4) { SMT4(double, true) } else { SMT4(double, false) } }
            else { if (v & 4) { SMT4(float, true) } else { SMT4(float, false) } }
#undef SMT4
#undef SMT
        } else if (k < kKindCoarse) {
            const bool tile = k < kKindRR;
            const int v = k - (tile ? kKindRRTile : kKindRR);
#define RRL(T, C)                                                                                   \
    if (tile) hipLaunchKernelGGL((k_resid_restrict_tile_many<T, C>), grid, dim3(kTileThreads), 0, st, \
                                 static_cast<const ResidRestrictArgs<T>*>(a), f, g.n);              \
    else hipLaunchKernelGGL((k_resid_restrict_many<T, C>), grid, dim3(kMgThreads), 0, st,            \
                            static_cast<const ResidRestrictArgs<T>*>(a), f, g.n)
            switch (v) {
            case 0: RRL(float, false); break;
            case 1: RRL(float, true); break;
            case 2: RRL(double, false); break;
            default: RRL(double, true); break;
            }
#undef RRL
        } else if (k < kKindOuter) {
            if (k - kKindCoarse)
                hipLaunchKernelGGL((k_coarse_solve_many<double>), grid, dim3(kMgThreads), 0, st,
                                   static_cast<const CoarseArgs<double>*>(a), f, g.n);
            else
                hipLaunchKernelGGL((k_coarse_solve_many<float>), grid, dim3(kMgThreads), 0, st,
                                   static_cast<const CoarseArgs<float>*>(a), f, g.n);
        } else if (k < kKindVec) {
            const OuterArgs* oa = static_cast<const OuterArgs*>(a);
#define OUT(C, F, N) hipLaunchKernelGGL((k_outer_residual_many<2, C, F, N>), grid, dim3(kMgThreads), 0, st, oa, f, g.n)
            switch (k - kKindOuter) {
            case 0: OUT(false, false, false); break;
            case 1: OUT(false, true, false); break;
            case 2: OUT(true, false, false); break;
            case 3: OUT(true, true, false); break;
            case 4: OUT(false, false, true); break;
            case 5: OUT(false, true, true); break;
            case 6: OUT(true, false, true); break;
            default: OUT(true, true, true); break;
            }
#undef OUT
        } else if (k < kKindHadamard) {
            const VecArgs* va = static_cast<const VecArgs*>(a);
            switch (k - kKindVec) {
            case 0: hipLaunchKernelGGL((k_vec_many<0>), grid, dim3(kMgThreads), 0, st, va, f, g.n); break;
            case 1: hipLaunchKernelGGL((k_vec_many<1>), grid, dim3(kMgThreads), 0, st, va, f, g.n); break;
            default: hipLaunchKernelGGL((k_vec_many<2>), grid, dim3(kMgThreads), 0, st, va, f, g.n); break;
            }
        } else {
            hipLaunchKernelGGL((k_hadamard_many<double>), grid, dim3(kMgThreads), 0, st,
                               static_cast<const HadamardArgs<double>*>(a), f, g.n);
        }
    }
    MG_HIP(hipGetLastError());
}

}  // namespace psum_hip

// ============================================================ plugin ABI

namespace {

std::map<std::string, func_table>& registry()
{
    static std::map<std::string, func_table> r;
    return r;
}

template <bool MIXED>
psum_field_solver_handle mg_init()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1)
        throw std::runtime_error(MIXED ? "hip_multigrid32f_gpu: no GPU device available"
                                       : "hip_multigrid_gpu: no GPU device available");
    return new HipMultigrid(MIXED);
}

void mg_set_matrix(psum_field_solver_handle h, unsigned long long n_row, unsigned long long n_col,
                   unsigned long long nnz, unsigned long long* rows, unsigned long long* cols,
                   double* vals)
{
    static_cast<HipMultigrid*>(h)->set_matrix(n_row, n_col, nnz, rows, cols, vals);
}

void mg_solve(psum_field_solver_handle h, double* b, double* x)
{
    static_cast<HipMultigrid*>(h)->solve(b, x, true);
}

void mg_set_options(psum_field_solver_handle h, const char* o)
{
    static_cast<HipMultigrid*>(h)->set_options(o);
}

void mg_set_replace(psum_field_solver_handle h, unsigned long long n, unsigned long long* i, double* v)
{
    static_cast<HipMultigrid*>(h)->set_source_replace(n, i, v);
}

void mg_set_addback(psum_field_solver_handle h, unsigned long long n, unsigned long long* i, double* v)
{
    static_cast<HipMultigrid*>(h)->set_source_addback(n, i, v);
}

void mg_free(psum_field_solver_handle h) { delete static_cast<HipMultigrid*>(h); }

const int registered = []() {
    func_table t64{mg_init<false>, mg_set_matrix, mg_solve, mg_set_options, mg_set_replace, mg_set_addback, mg_free};
    func_table t32{mg_init<true>, mg_set_matrix, mg_solve, mg_set_options, mg_set_replace, mg_set_addback, mg_free};
    registry()["hip_multigrid_gpu"] = t64;
    registry()["hip_multigrid32f_gpu"] = t32;
    return 0;
}();

template <typename F>
int guarded(F&& f)
{
    try {
        f();
        return PSUM_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return PSUM_ERR_ARG;
    } catch (const std::exception& e) {
        set_error(e.what());
        return PSUM_ERR_UNSUPPORTED;
    }
}

}  // namespace

namespace psum_hip {

// in-library access for the step engine
HipMultigrid* mg_create(bool mixed) { return new HipMultigrid(mixed); }
void mg_destroy(HipMultigrid* h) { delete h; }
void mg_set_options(HipMultigrid* h, const char* o) { h->set_options(o); }
void mg_set_tuning(HipMultigrid* h, const char* o) { h->set_tuning(o); }
void mg_set_smoothing(HipMultigrid* h, int pre, int post) { h->set_smoothing(pre, post); }
void mg_set_matrix(HipMultigrid* h, unsigned long long n, unsigned long long nnz,
                   const unsigned long long* r, const unsigned long long* c, const double* v)
{
    h->set_matrix(n, n, nnz, r, c, v);
}
void mg_set_replace(HipMultigrid* h, unsigned long long n, const unsigned long long* i, const double* v)
{
    h->set_source_replace(n, i, v);
}
void mg_set_stream(HipMultigrid* h, hipStream_t s) { h->set_stream(s); }
void mg_solve(HipMultigrid* h, double* b, double* x, bool sync) { h->solve(b, x, sync); }
void mg_solve_homogeneous(HipMultigrid* h, double* b, double* x) { h->solve(b, x, false, true); }
int mg_predicted(HipMultigrid* h) { return h->predicted(); }
void mg_set_fixed_cycles(HipMultigrid* h, int cycles) { h->set_fixed_cycles(cycles); }
void mg_set_predicted(HipMultigrid* h, int p) { h->set_predicted(p); }
int mg_last_cycles(HipMultigrid* h) { return h->last_cycles(); }
double mg_last_rel(HipMultigrid* h) { return h->last_rel(); }
const double* mg_solve_begin(HipMultigrid* h, double* b) { return h->solve_begin(b); }
void mg_solve_core(HipMultigrid* h, const double* b, const double* bs, double* x) { h->solve_core(b, bs, x); }
void mg_solve_end(HipMultigrid* h, const double* b, double* x) { h->solve_end(b, x, false); }
bool mg_recordable(HipMultigrid* h) { return h->recordable(); }
const double* mg_core_rhs(HipMultigrid* h, double* b)
{
    double* p = h->pin_rhs();
    return p ? p : b;
}
void mg_record_core(HipMultigrid* h, const double* b, const double* bs, double* x,
                    std::vector<MgRecOp>& ops)
{
    h->record_core(b, bs, x, ops);
}
MgBatchPlan* mg_batch_plan_create() { return new MgBatchPlan(); }
void mg_batch_plan_destroy(MgBatchPlan* p) { delete p; }
void mg_batch_plan_build(MgBatchPlan* p, const std::vector<const std::vector<MgRecOp>*>& per)
{
    mg_batch_build(per, *p);
}
void mg_batch_plan_launch(const MgBatchPlan* p, hipStream_t st) { mg_batch_launch(*p, st); }
int64_t mg_batch_plan_launches(const MgBatchPlan* p) { return (int64_t)p->groups.size(); }
int64_t mg_batch_plan_ops(const MgBatchPlan* p) { return p->ops; }

}  // namespace psum_hip

extern "C" {

void register_solver(const char* name, func_table table) { registry()[name] = table; }

func_table get_solver_funcs(const char* name)
{
    func_table ans{};
    ans.init = nullptr;
    auto& reg = registry();
    auto it = reg.find(name ? name : "");
    if (it != reg.end()) return it->second;
    return ans;
}

void* psum_solver_create(const char* kind)
{
    func_table t = get_solver_funcs(kind);
    if (!t.init) {
        set_error(std::string("psum_solver_create: unknown backend ") + (kind ? kind : "(null)"));
        return nullptr;
    }
    void* h = nullptr;
    if (guarded([&]() { h = t.init(); }) != PSUM_OK) return nullptr;
    return h;
}

int psum_solver_set_options(void* h, const char* options)
{
    if (!h) { set_error("psum_solver_set_options: null handle"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<HipMultigrid*>(h)->set_options(options); });
}

int psum_solver_set_matrix(void* h, unsigned long long n_row, unsigned long long n_col,
                           unsigned long long nnz, unsigned long long* rows,
                           unsigned long long* cols, double* vals)
{
    if (!h) { set_error("psum_solver_set_matrix: null handle"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<HipMultigrid*>(h)->set_matrix(n_row, n_col, nnz, rows, cols, vals); });
}

int psum_solver_set_source_replace(void* h, unsigned long long n, unsigned long long* idxs, double* vals)
{
    if (!h) { set_error("null handle"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<HipMultigrid*>(h)->set_source_replace(n, idxs, vals); });
}

int psum_solver_set_source_addback(void* h, unsigned long long n, unsigned long long* idxs, double* vals)
{
    if (!h) { set_error("null handle"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<HipMultigrid*>(h)->set_source_addback(n, idxs, vals); });
}

int psum_solver_solve(void* h, double* b_dev, double* x_dev)
{
    if (!h) { set_error("null handle"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<HipMultigrid*>(h)->solve(b_dev, x_dev, true); });
}

int psum_solver_stats(void* h, int* cycles, double* rel_res)
{
    if (!h) { set_error("null handle"); return PSUM_ERR_ARG; }
    auto* m = static_cast<HipMultigrid*>(h);
    if (cycles) *cycles = m->last_cycles();
    if (rel_res) *rel_res = m->last_rel();
    return PSUM_OK;
}

void psum_solver_free(void* h) { delete static_cast<HipMultigrid*>(h); }

}  // extern "C"
