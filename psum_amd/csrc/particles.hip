// Particle-side kernels of the MI355X PIC step engine (gfx950):
//   cell binning, stable LSD radix sort, permutation, stable compaction,
//   deterministic CIC deposit (LDS int64 tile), fused gather + push (+ next
//   deposit), diagnostics reductions.
//
// Reference behaviour restated here (PSuM paths):
//   cN / n2i / periodic support wrap  src/field/simple_grid.hpp:89-128,316-385
//   CIC weights and gradient weights  src/field/simple_interpolation.hpp:33-98
//   add_back / interp_diff            src/field/field_interp.hpp:63-113,
//                                     src/field/device_field.hpp:270-310
//   push, wrap, absorb                application/benchmark/TSI/main.cpp:59-75,
//                                     example/big_bang/stage_6.cpp:48-65,
//                                     src/particle_container/validator.hpp:18-26
//   compress                          src/particle_container/particle_group.hpp:223-238
//
// Floating point: built with -ffp-contract=off, so every a*b+c below is two
// rounded operations in the order written (the CPU oracle does the same).

#include <climits>
#include <cmath>
#include <mutex>

#include "common.hpp"

namespace psum_hip {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

}  // namespace psum_hip

using namespace psum_hip;

extern "C" const char* psum_hip_last_error(void) { return g_last_error.c_str(); }

namespace {

constexpr int kThreads = 256;
constexpr int64_t kReduceBlocks = 256;

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_blocks(int64_t n, int64_t per_block)
{
    return (unsigned)((n + per_block - 1) / per_block);
}

// ------------------------------------------------------------------ binning

template <typename T, int D, bool VEL>
__global__ __launch_bounds__(kThreads) void k_bin(DevGrid g, const T* __restrict__ x0,
                                                  const T* __restrict__ x1,
                                                  const T* __restrict__ x2, const T* __restrict__ v0,
                                                  const T* __restrict__ v1, const T* __restrict__ v2,
                                                  int64_t n, uint32_t* __restrict__ keys, int shift)
{
    const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (p >= n) return;
    double x[3];
    x[0] = (double)x0[p];
    if (D > 1) x[1] = (double)x1[p];
    if (D > 2) x[2] = (double)x2[p];
    uint32_t key = PSUM_KEY_INVALID;
    int cw[3];
    double s[3];
    if (isfinite(x[0]) && base_cell<D>(g, x, cw, s)) {
        // shift 0: the reference cell key n2i(cN); shift b: the key of the
        // 2^b-cell bin (sort granularity of the stepper, not a cell index)
        uint32_t lin = (uint32_t)(cw[0] >> shift);
#pragma unroll
        for (int d = 1; d < D; ++d)
            lin = lin * (uint32_t)((g.w[d] + (1 << shift) - 1) >> shift) + (uint32_t)(cw[d] >> shift);
        if constexpr (VEL) {
            const T* const vs[3] = {v0, v1, v2};
#pragma unroll
            for (int d = 0; d < D; ++d) lin = (lin << 1) | (vs[d][p] >= (T)0 ? 1u : 0u);
        }
        key = lin;
    }
    keys[p] = key;
}

// ------------------------------------------------------------------- scan
// Exclusive scan of uint32, 4096 items per block (256 threads x 16 items),
// block sums scanned recursively.

constexpr int kScanItems = 16;
constexpr int kScanTile = kThreads * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* total)
{
    __shared__ uint32_t wsum[kThreads / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(inc, off, 64);
        if (lane >= off) inc += t;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, tot = 0;
    for (int w = 0; w < kThreads / 64; ++w) {
        if (w < wave) wave_off += wsum[w];
        tot += wsum[w];
    }
    __syncthreads();
    *total = tot;
    return wave_off + inc - v;
}

__global__ __launch_bounds__(kThreads) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                         uint32_t* __restrict__ out,
                                                         uint32_t* __restrict__ bsum, int64_t n)
{
    __shared__ uint32_t buf[kScanTile];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kThreads) {
        const int64_t p = base + i;
        buf[i] = p < n ? in[p] : 0u;
    }
    __syncthreads();
    uint32_t loc[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        loc[j] = buf[threadIdx.x * kScanItems + j];
        sum += loc[j];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, &total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        buf[threadIdx.x * kScanItems + j] = run;
        run += loc[j];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kScanTile; i += kThreads) {
        const int64_t p = base + i;
        if (p < n) out[p] = buf[i];
    }
    if (threadIdx.x == 0 && bsum) bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kThreads) void k_scan_add(uint32_t* __restrict__ out,
                                                       const uint32_t* __restrict__ boff, int64_t n)
{
    const uint32_t add = boff[blockIdx.x];
    const int64_t base = (int64_t)blockIdx.x * kScanTile;
    for (int i = threadIdx.x; i < kScanTile; i += kThreads) {
        const int64_t p = base + i;
        if (p < n) out[p] += add;
    }
}

// ------------------------------------------------------------- radix sort
// Stable LSD radix sort, 8-bit digits.  A block owns a tile of 4096 keys;
// wave w owns the contiguous quarter [w*1024, (w+1)*1024) and visits it in
// 16 rows of 64 keys, so lane order within a row and row order within the
// wave are index order.  Within-wave stable ranks come from 8 ballots per
// key (lanes with equal digit), across waves from per-wave digit counts.

constexpr int kSortWaveItems = 16;
constexpr int kSortTile = kThreads * kSortWaveItems;

__device__ __forceinline__ uint32_t sort_digit(uint32_t key, uint32_t nkeys, int shift)
{
    const uint32_t k = key == PSUM_KEY_INVALID ? nkeys : key;
    return (k >> shift) & 255u;
}

__global__ __launch_bounds__(kThreads) void k_sort_hist(const uint32_t* __restrict__ keys,
                                                        int64_t n, uint32_t nkeys, int shift,
                                                        uint32_t* __restrict__ ghist, int nb)
{
    __shared__ uint32_t hist[256];
    hist[threadIdx.x] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * kSortTile + wave * (64 * kSortWaveItems);
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
        const int64_t p = base + j * 64 + lane;
        if (base + j * 64 >= n) break;   // no active key in this row (uniform)
        const bool act = p < n;
        const uint32_t d = act ? sort_digit(keys[p], nkeys, shift) : 0u;
        // a row whose keys share one digit (nearly sorted keys) counts them
        // with one add; the counts are the same
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        const uint32_t nact = (uint32_t)__popcll(__ballot(act));   // with every lane of the wave
        if (__all(!act || d == d0)) {
            if (lane == 0) atomicAdd(&hist[d0], nact);
        } else if (act) {
            atomicAdd(&hist[d], 1u);
        }
    }
    __syncthreads();
    ghist[(int64_t)threadIdx.x * nb + blockIdx.x] = hist[threadIdx.x];
}

__global__ __launch_bounds__(kThreads) void k_sort_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, int64_t n,
    uint32_t nkeys, int shift, const uint32_t* __restrict__ goff, int nb,
    uint32_t* __restrict__ kout, uint32_t* __restrict__ vout)
{
    __shared__ uint32_t off[256];
    __shared__ uint32_t whist[kThreads / 64][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    off[threadIdx.x] = goff[(int64_t)threadIdx.x * nb + blockIdx.x];
#pragma unroll
    for (int q = 0; q < 4; ++q) whist[wave][lane + 64 * q] = 0;
    __syncthreads();

    const int64_t base = (int64_t)blockIdx.x * kSortTile + wave * (64 * kSortWaveItems);
    const uint64_t lt = __lanemask_lt();
    uint32_t key[kSortWaveItems], val[kSortWaveItems], rank[kSortWaveItems];
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
        const int64_t p = base + j * 64 + lane;
        const bool act = p < n;
        key[j] = act ? kin[p] : 0u;
        val[j] = act ? (vin ? vin[p] : (uint32_t)p) : 0u;
        const uint32_t d = sort_digit(key[j], nkeys, shift);
        uint64_t peers = __ballot(act);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        if (act) {
            const uint32_t before = whist[wave][d];
            rank[j] = before + (uint32_t)__popcll(peers & lt);
            if ((peers & lt) == 0) whist[wave][d] = before + (uint32_t)__popcll(peers);
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kSortWaveItems; ++j) {
        const int64_t p = base + j * 64 + lane;
        if (p < n) {
            const uint32_t d = sort_digit(key[j], nkeys, shift);
            uint32_t dst = off[d] + rank[j];
            for (int w = 0; w < wave; ++w) dst += whist[w][d];
            kout[dst] = key[j];
            vout[dst] = val[j];
        }
    }
}

// 16-byte vector load / store of four consecutive values (two for double)
template <typename T>
__device__ __forceinline__ void load4(const T* __restrict__ p, T* v)
{
    if constexpr (sizeof(T) == 4) {
        const float4 q = *reinterpret_cast<const float4*>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        const double2 a = *reinterpret_cast<const double2*>(p);
        const double2 b = *reinterpret_cast<const double2*>(p + 2);
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
    }
}

template <typename T>
__device__ __forceinline__ void store4(T* __restrict__ p, const T* v)
{
    if constexpr (sizeof(T) == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        *reinterpret_cast<double2*>(p) = make_double2(v[0], v[1]);
        *reinterpret_cast<double2*>(p + 2) = make_double2(v[2], v[3]);
    }
}

// -------------------------------------------------------------- permute

struct PtrPack {
    const void* src[8];
    void* dst[8];
};

// four outputs per thread: perm read and the outputs written with 16-byte
// vector accesses (perm and every output 16-byte aligned, n a multiple of 4
// for all but the last thread, which goes one by one)
template <typename T>
__global__ __launch_bounds__(kThreads) void k_permute4(PtrPack pk, int ncomp,
                                                       const uint32_t* __restrict__ perm, int64_t n)
{
    const int64_t i0 = ((int64_t)blockIdx.x * kThreads + threadIdx.x) * 4;
    if (i0 >= n) return;
    if (i0 + 4 <= n) {
        const uint4 s = *reinterpret_cast<const uint4*>(perm + i0);
        for (int c = 0; c < ncomp; ++c) {
            const T* src = static_cast<const T*>(pk.src[c]);
            T v[4] = {src[s.x], src[s.y], src[s.z], src[s.w]};
            store4(static_cast<T*>(pk.dst[c]) + i0, v);
        }
    } else {
        for (int64_t i = i0; i < n; ++i) {
            const uint32_t s = perm[i];
            for (int c = 0; c < ncomp; ++c)
                static_cast<T*>(pk.dst[c])[i] = static_cast<const T*>(pk.src[c])[s];
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_permute(PtrPack pk, int ncomp,
                                                      const uint32_t* __restrict__ perm, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    for (int c = 0; c < ncomp; ++c)
        static_cast<T*>(pk.dst[c])[i] = static_cast<const T*>(pk.src[c])[s];
}

// ------------------------------------------------------------ compaction
// Block b counts the valid particles of [b*1024, (b+1)*1024); the counts are
// scanned; the block then writes its survivors in index order.

constexpr int kCompactRows = 4;
constexpr int kCompactTile = kThreads * kCompactRows;

template <typename T>
__global__ __launch_bounds__(kThreads) void k_compact_count(const T* __restrict__ x0, int64_t n,
                                                            uint32_t* __restrict__ bcount)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) bcount[gridDim.x] = 0;   // becomes the total
    const int64_t base = (int64_t)blockIdx.x * kCompactTile;
    uint32_t c = 0;
#pragma unroll
    for (int r = 0; r < kCompactRows; ++r) {
        const int64_t p = base + r * kThreads + threadIdx.x;
        if (p < n && isfinite((double)x0[p])) ++c;
    }
    uint32_t total;
    block_exclusive_scan(c, &total);
    if (threadIdx.x == 0) bcount[blockIdx.x] = total;
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_compact_scatter(PtrPack pk, int ncomp, int64_t n,
                                                              const uint32_t* __restrict__ boff)
{
    const int64_t base = (int64_t)blockIdx.x * kCompactTile;
    uint32_t run = boff[blockIdx.x];
    const T* x0 = static_cast<const T*>(pk.src[0]);
    for (int r = 0; r < kCompactRows; ++r) {
        const int64_t p = base + r * kThreads + threadIdx.x;
        const bool ok = p < n && isfinite((double)x0[p]);
        uint32_t total;
        const uint32_t pre = block_exclusive_scan(ok ? 1u : 0u, &total);
        if (ok) {
            const uint32_t dst = run + pre;
            for (int c = 0; c < ncomp; ++c)
                static_cast<T*>(pk.dst[c])[dst] = static_cast<const T*>(pk.src[c])[p];
        }
        run += total;
    }
}

// --------------------------------------------- gather + push + deposit
// One block = 256 threads x 4 particles; thread t owns the 4 consecutive
// particles base + 4t .. base + 4t + 3 and moves them with 16-byte vector
// loads and stores.  With DEPOSIT the block deposits the (new) positions of
// its particles into an LDS int64 tile spanning the bounding box of their
// support nodes, then flushes the non-zero entries to the global int64
// accumulator; particles whose support leaves a box clipped to the tile use
// global integer atomics directly.  In 2D and 3D each particle adds its
// terms on its own; in 1D each thread first sums the terms of equal-cell
// runs in registers (exact integer adds) and issues one atomic per run and
// support node, in a lane-rotated node order to spread same-address
// conflicts.  Integer sums make every path exact and order independent.

constexpr int kPPT = 4;
constexpr int kPushTile = kThreads * kPPT;
constexpr int kTileNodes = 2048;
// 3D blocks get a larger tile (32 KB): a block's box of drifted particles
// grows with the cube of the drift
template <int D>
constexpr int tile_nodes() { return D == 3 ? 2 * kTileNodes : kTileNodes; }
// per-axis extent of a clipped box: kClip^D <= tile_nodes<D>()
template <int D>
constexpr int kClip() { return D == 3 ? 16 : (D == 2 ? 45 : 2048); }
static_assert(16 * 16 * 16 <= 2 * kTileNodes && 45 * 45 <= kTileNodes, "clipped box exceeds the tile");

// Linear indices of the 2^D support nodes of cell cw, as support_index for
// i = 0 .. 2^D-1: the cell's index plus, per dimension whose bit is set, one
// stride, or minus cw * stride where cw + 1 wraps to node 0.
template <int D>
__device__ __forceinline__ void support_indices(const DevGrid& g, const int* cw, uint32_t* idx)
{
    uint32_t stride[3];
    stride[D - 1] = 1u;
#pragma unroll
    for (int d = D - 2; d >= 0; --d) stride[d] = stride[d + 1] * (uint32_t)g.w[d + 1];
    uint32_t base = 0u, step[3];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        base += (uint32_t)cw[d] * stride[d];
        step[d] = cw[d] + 1 == g.w[d] ? 0u - (uint32_t)cw[d] * stride[d] : stride[d];
    }
#pragma unroll
    for (int i = 0; i < (1 << D); ++i) {
        uint32_t o = base;
#pragma unroll
        for (int d = 0; d < D; ++d)
            if ((i >> (D - 1 - d)) & 1) o += step[d];
        idx[i] = o;
    }
}

// grad phi when every delta_r is 1.  Gradient weight i along dd is then
// cic_grad_coef's product with G[dd] = -1 (bit of dd clear) or +1 in the
// place of the dd factor; multiplying by +-1 is exact and rounding is
// symmetric in sign, so the weight is +-q, q the product of the other
// dimensions' W factors in the same order, and ph * weight = +-(ph * q).
// The terms, their skips (|q| > 1e-10) and their order are the general
// form's: equal bits.
template <typename T, int D>
__device__ __forceinline__ void unit_grad(const T (*W)[2], const T* ph, T* grad)
{
#pragma unroll
    for (int dd = 0; dd < D; ++dd) {
        T r = (T)0;
#pragma unroll
        for (int i = 0; i < (1 << D); ++i) {
            T q = (T)1;
            bool first = true;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                const int dim = D - 1 - j;
                if (dim == dd) continue;
                const T f = W[dim][(i >> j) & 1];
                q = first ? f : q * f;
                first = false;
            }
            if (!coef_kept<T>(q)) continue;
            const T t = ph[i] * q;
            r = ((i >> (D - 1 - dd)) & 1) ? r + t : r - t;
        }
        grad[dd] = r;
    }
}

// Grid kinds the push is compiled for: any spacing (0), unit spacing (1),
// and unit spacing from the origin 0 with upper bounds exact in float and
// FP32 positions (2, DevGrid::origin0).
constexpr int kGridAny = 0, kGridUnit = 1, kGridUnit0 = 2;

// Cell index and offsets of x as base_cell<D, UNIT> takes them in FP64,
// rounded to T.  kGridUnit0 takes them in FP32 with the same bits: x - lb
// is x, floor and the range test are exact, and x - ci is exact (Sterbenz
// for ci >= 1 and ci <= -2; ci = 0 gives x) except for ci = -1 and
// |x| < 2^-29, where both round 1 - |x| to 1.0f.
template <typename T, int D, int UNIT>
__device__ __forceinline__ bool cell_of(const DevGrid& g, const T* x, int* cw, T* s)
{
    if constexpr (UNIT == kGridUnit0 && sizeof(T) == 4) {
        // every axis in turn, one validity test at the end (an axis out of
        // range takes cell 0; the caller uses no output of an invalid x)
        bool ok = true;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const float f = floorf(x[d]);
            const int w = g.w[d];
            const bool in = f >= (float)(-w) && f <= (float)(2 * w - 2);
            ok = ok && in;
            const float fc = in ? f : 0.0f;
            const int ci = (int)fc;
            s[d] = x[d] - fc;
            cw[d] = ci < 0 ? ci + w : (ci >= w ? ci - w : ci);
        }
        return ok;
    }
    double xd[D], sd[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xd[d] = (double)x[d];
    if (!base_cell<D, (UNIT > kGridAny)>(g, xd, cw, sd)) return false;
#pragma unroll
    for (int d = 0; d < D; ++d) s[d] = (T)sd[d];
    return true;
}

// two consecutive values at the alignment of one
template <typename T>
using pair_t = T __attribute__((ext_vector_type(2), aligned(sizeof(T))));

template <typename T, int D, bool PUSH, int UNIT>
__device__ __forceinline__ bool push_one(const DevGrid& g, const T* __restrict__ phi, T qm, T dt,
                                         double mass, T* x, T* v, double& ke, double* pm)
{
    double v2 = 0.0;
#pragma unroll
    for (int d = 0; d < D; ++d) v2 += (double)v[d] * (double)v[d];
    ke += 0.5 * mass * v2;
#pragma unroll
    for (int d = 0; d < D; ++d) pm[d] += mass * (double)v[d];

    // gather grad(phi) at x (field_interp.hpp:96-111)
    T grad[D];
#pragma unroll
    for (int d = 0; d < D; ++d) grad[d] = (T)0;
    T s[D];
    int c[D];
    if (cell_of<T, D, UNIT>(g, x, c, s)) {
        T W[D][2], G[D][2];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const T sd = s[d];
            W[d][0] = (T)1 - sd;
            W[d][1] = sd;
            const T dr = (T)g.delta_r[d];
            G[d][0] = (T)-1 * dr;
            G[d][1] = (T)1 * dr;
        }
        uint32_t sidx[1 << D];
        support_indices<D>(g, c, sidx);
        T ph[1 << D];
        if (__all(c[D - 1] + 1 != g.w[D - 1])) {
            // support nodes i and i + 1 differ in the fastest axis only and
            // are neighbours in memory unless that axis wraps: paired loads
            // (a wave with a wrapping particle loads them one by one)
#pragma unroll
            for (int i = 0; i < (1 << D); i += 2) {
                const pair_t<T> q = *reinterpret_cast<const pair_t<T>*>(phi + sidx[i]);
                ph[i] = q.x;
                ph[i + 1] = q.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < (1 << D); ++i) ph[i] = phi[sidx[i]];
        }
        if constexpr (UNIT > kGridAny) {
            unit_grad<T, D>(W, ph, grad);
        } else {
#pragma unroll
            for (int dd = 0; dd < D; ++dd) {
                T r = (T)0;
#pragma unroll
                for (int i = 0; i < (1 << D); ++i) {
                    const T gc = cic_grad_coef<T, D>(W, G, i, dd);
                    if (coef_kept<T>(gc)) r = r + ph[i] * gc;
                }
                grad[dd] = r;
            }
        }
    }
    // kick then drift (TSI/main.cpp:65-68)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T e = -grad[d];
        v[d] = v[d] + (e * qm) * dt;
        x[d] = x[d] + v[d] * dt;
    }
    // boundaries: bounded periodic wrap, then absorbing walls
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (g.periodic[d]) {
            const T lo = (T)g.plo[d], hi = (T)g.phi[d], L = (T)g.plen[d];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                if (x[d] < lo) x[d] = x[d] + L;
                else if (x[d] >= hi) x[d] = x[d] - L;
            }
        }
    }
    bool alive = true;
#pragma unroll
    for (int d = 0; d < D; ++d)
        if (!g.periodic[d]) {
            if constexpr (UNIT == kGridUnit0 && sizeof(T) == 4) {
                // lb = 0 and ub exact in float: the same test in FP32
                if (!(x[d] >= 0.0f && x[d] < (float)g.ub[d])) alive = false;
            } else {
                const double xx = (double)x[d];
                if (!(xx >= g.lb[d] && xx < g.ub[d])) alive = false;
            }
        }
    if (!alive) x[0] = (T)NAN;
    return alive;
}

// 2D: at least 5 waves per SIMD (the 2D kernels fit it without spills and
// were measured faster); 3D keeps the compiler's register budget
template <int D>
constexpr int push_min_waves() { return D == 2 ? 5 : 1; }

template <typename T, int D, bool PUSH, bool DEPOSIT, int UNIT = kGridAny>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(push_min_waves<D>(), 8))) void k_push_deposit(
    DevGrid g, T* __restrict__ x0, T* __restrict__ x1, T* __restrict__ x2, T* __restrict__ v0,
    T* __restrict__ v1, T* __restrict__ v2, int64_t n, const T* __restrict__ phi, T qm, T dt,
    double mass, double w_dep, double scale, unsigned long long* __restrict__ acc,
    double* __restrict__ diag, unsigned long long* __restrict__ counters)
{
    __shared__ unsigned long long tile[DEPOSIT ? tile_nodes<D>() : 1];
    __shared__ int bmin[3], bmax[3], bref[3];
    __shared__ long long bsum[3];
    __shared__ int bcnt;
    __shared__ double wdiag[kThreads / 16][4];   // per 16-lane row
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // wraps of the global int64 sums; an LDS tile entry holds at most
    // kPushTile terms below 2^50 (|w_dep| * scale <= 2^50) and cannot wrap
    unsigned long long* const ovf = counters ? counters + 3 : nullptr;
    if (DEPOSIT && threadIdx.x < 3) {
        bmin[threadIdx.x] = INT_MAX;
        bmax[threadIdx.x] = INT_MIN;
        bsum[threadIdx.x] = 0;
        if (threadIdx.x == 0) bcnt = 0;
    }
    T* const xs[3] = {x0, x1, x2};
    T* const vs[3] = {v0, v1, v2};
    const int64_t p0 = (int64_t)blockIdx.x * kPushTile + (int64_t)threadIdx.x * kPPT;
    const bool full = p0 + kPPT <= n;

    // load the thread's kPPT particles (component-major registers)
    T X[D][kPPT], V[D][kPPT];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (full) {
#pragma unroll
            for (int q = 0; q < kPPT; q += 4) {
                load4(xs[d] + p0 + q, X[d] + q);
                if (PUSH) load4(vs[d] + p0 + q, V[d] + q);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kPPT; ++j) {
                X[d][j] = p0 + j < n ? xs[d][p0 + j] : (T)NAN;
                if (PUSH) V[d][j] = p0 + j < n ? vs[d][p0 + j] : (T)0;
            }
        }
    }

    bool ok[kPPT];
    int cw[kPPT][D];
    T sw[kPPT][D];
    double ke = 0.0, pm[3] = {0.0, 0.0, 0.0};
    unsigned oor = 0, absorbed = 0;
    bool moved = false;

#pragma unroll
    for (int j = 0; j < kPPT; ++j) {
        ok[j] = false;
        T x[D], v[D];
#pragma unroll
        for (int d = 0; d < D; ++d) {
            x[d] = X[d][j];
            if (PUSH) v[d] = V[d][j];
        }
        if (!isfinite((double)x[0])) continue;       // invalid (or beyond n)
        bool alive = true;
        if (PUSH) {
            alive = push_one<T, D, PUSH, UNIT>(g, phi, qm, dt, mass, x, v, ke, pm);
            if (!alive) ++absorbed;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                X[d][j] = x[d];
                V[d][j] = v[d];
            }
            moved = true;
        }
        if (DEPOSIT && alive) {
            T s[D];
            if (cell_of<T, D, UNIT>(g, x, cw[j], s)) {
                ok[j] = true;
#pragma unroll
                for (int d = 0; d < D; ++d) sw[j][d] = s[d];
            } else {
                ++oor;
            }
        }
    }
    if (PUSH && moved) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (full) {
#pragma unroll
                for (int q = 0; q < kPPT; q += 4) {
                    store4(xs[d] + p0 + q, X[d] + q);
                    store4(vs[d] + p0 + q, V[d] + q);
                }
            } else {
#pragma unroll
                for (int j = 0; j < kPPT; ++j)
                    if (p0 + j < n) {
                        xs[d][p0 + j] = X[d][j];
                        vs[d][p0 + j] = V[d][j];
                    }
            }
        }
    }

    if (PUSH && diag) {
        double vals[4] = {ke, pm[0], pm[1], pm[2]};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            // row totals (16 lanes each) go to LDS; thread q sums them in
            // row order below
            const double r = row_sum_full(vals[q]);
            if ((lane & 15) == 0) wdiag[wave * 4 + (lane >> 4)][q] = r;
        }
    }
    if (counters) {
        // both counts (at most kPPT per thread, 256 per wave) in one sum
        const int both = wave_isum_full((int)((absorbed << 16) | oor));
        const int ab = both >> 16, oo = both & 0xFFFF;
        if (lane == 0) {
            if (ab) atomicAdd(&counters[2], (unsigned long long)ab);
            if (oo) atomicAdd(&counters[0], (unsigned long long)oo);
        }
    }

    if (DEPOSIT) {
        // the tile box is taken over cells unwrapped around a reference cell
        // (thread 0's first particle) on periodic axes, so a block whose
        // particles straddle the periodic seam still gets a small box
        if (threadIdx.x == 0) {
            int r[3] = {-1, -1, -1};
#pragma unroll
            for (int j = kPPT - 1; j >= 0; --j)
                if (ok[j])
#pragma unroll
                    for (int d = 0; d < D; ++d) r[d] = cw[j][d];
#pragma unroll
            for (int d = 0; d < D; ++d) bref[d] = r[d];
        }
        __syncthreads();   // bmin/bmax initialised, bref set
        int ref[3];
#pragma unroll
        for (int d = 0; d < D; ++d) ref[d] = bref[d];
        int uw[kPPT][D];
        int lmin[3] = {INT_MAX, INT_MAX, INT_MAX}, lmax[3] = {INT_MIN, INT_MIN, INT_MIN};
#pragma unroll
        for (int j = 0; j < kPPT; ++j)
            if (ok[j]) {
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    int u = cw[j][d];
                    if (g.periodic[d] && ref[d] >= 0) {
                        const int w = g.w[d];
                        if (u - ref[d] > w / 2) u -= w;
                        else if (ref[d] - u > w / 2) u += w;
                    }
                    uw[j][d] = u;
                    lmin[d] = min(lmin[d], u);
                    lmax[d] = max(lmax[d], u);
                }
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
        int e[3] = {1, 1, 1}, lo[3] = {0, 0, 0};
        int64_t tsize = 1;
        const bool any = bmin[0] != INT_MAX;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            lo[d] = bmin[d];
            e[d] = bmax[d] - bmin[d] + 2;   // support adds one node
            tsize *= any ? (int64_t)e[d] : 0;
        }
        // A box beyond the tile is clipped to at most kClip<D> nodes per
        // axis (kClip^D <= tile nodes) around the block's mean cell: runs
        // whose 2^D support lies inside go to the tile, the others (the few
        // particles far from the block's centre) to global integer atomics.
        const bool clipped = any && tsize > tile_nodes<D>();
        if (clipped) {
            long long ls[3] = {0, 0, 0};
            int lc = 0;
#pragma unroll
            for (int j = 0; j < kPPT; ++j)
                if (ok[j]) {
                    ++lc;
#pragma unroll
                    for (int d = 0; d < D; ++d) ls[d] += uw[j][d];
                }
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const long long t = wave_lsum_full(ls[d]);
                if (lane == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&bsum[d]), (unsigned long long)t);
            }
            const int tc = wave_isum_full(lc);
            if (lane == 0) atomicAdd(&bcnt, tc);
            __syncthreads();
            tsize = 1;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const int c = min(e[d], kClip<D>());
                // mean cell (floor division of a possibly negative sum)
                const long long sm = bsum[d];
                const int mean = (int)(sm >= 0 ? sm / bcnt : -((-sm + bcnt - 1) / bcnt));
                int l = mean - c / 2 + 1;
                l = max(l, bmin[d]);
                l = min(l, bmax[d] + 2 - c);
                lo[d] = l;
                e[d] = c;
                tsize *= c;
            }
        }
        const bool use_lds = any;
        if (use_lds)
            for (int i = threadIdx.x; i < (int)tsize; i += kThreads) tile[i] = 0ull;
        __syncthreads();

        constexpr int NS = 1 << D;
        const double wsc = w_dep * scale;
        if constexpr (D >= 2) {
            // 2D and 3D: each particle adds its 2^D terms directly, with the
            // support offsets fixed at compile time.  In 3D particles sorted
            // on cell bins rarely share a cell with their neighbour; in 2D
            // they often do (100 per cell), but the run bookkeeping below
            // cost more instructions than its fewer LDS atomics saved.
            int st[3];
            st[D - 1] = 1;
#pragma unroll
            for (int d = D - 2; d >= 0; --d) st[d] = st[d + 1] * e[d + 1];
#pragma unroll
            for (int j = 0; j < kPPT; ++j) {
                if (!ok[j]) continue;
                T W[D][2];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    W[d][0] = (T)1 - sw[j][d];
                    W[d][1] = sw[j][d];
                }
                // a coefficient below the cut is replaced by 0, whose term
                // rounds to 0 (one select on the coefficient, not two on
                // the int64 term)
                long long t8[NS];
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    const T c0 = cic_coef<T, D>(W, i);
                    const T cf = coef_kept<T>(c0) ? c0 : (T)0;
                    t8[i] = round_to_ll((double)cf * wsc);
                }
                bool inside = use_lds;
#pragma unroll
                for (int d = 0; d < D; ++d) inside = inside && uw[j][d] >= lo[d] && uw[j][d] + 1 < lo[d] + e[d];
                if (inside) {
                    int li0 = 0;
#pragma unroll
                    for (int d = 0; d < D; ++d) li0 += (uw[j][d] - lo[d]) * st[d];
                    // a zero term (a coefficient below the cut) adds nothing:
                    // no branch around it
#pragma unroll
                    for (int i = 0; i < NS; ++i) {
                        int o = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) o += ((i >> (D - 1 - d)) & 1) * st[d];
                        atomicAdd(&tile[li0 + o], (unsigned long long)t8[i]);
                    }
                } else if (ovf) {
                    // all eight adds are issued before their returned values
                    // are checked (a zero term leaves old unchanged: no wrap)
                    unsigned long long old8[NS];
#pragma unroll
                    for (int i = 0; i < NS; ++i)
                        old8[i] = t8[i] ? atomicAdd(&acc[support_index<D>(g, cw[j], i)], (unsigned long long)t8[i])
                                        : 0ull;
                    bool wr = false;
#pragma unroll
                    for (int i = 0; i < NS; ++i) wr |= add_wraps(old8[i], (unsigned long long)t8[i]);
                    if (wr) atomicAdd(ovf, 1ull);
                } else {
#pragma unroll
                    for (int i = 0; i < NS; ++i)
                        if (t8[i]) atomicAdd(&acc[support_index<D>(g, cw[j], i)], (unsigned long long)t8[i]);
                }
            }
        } else {
            long long run[NS];
            int cur[D], curu[D];
            bool have = false;
            auto flush = [&]() {
                // the run's support (cells curu .. curu + 1 per axis) inside the box
                bool inside = use_lds;
#pragma unroll
                for (int d = 0; d < D; ++d) inside = inside && curu[d] >= lo[d] && curu[d] + 1 < lo[d] + e[d];
                if (inside) {
#pragma unroll
                    for (int q = 0; q < NS; ++q) {
                        const int i = (q + lane) & (NS - 1);
                        const long long t = run[i];
                        if (t == 0) continue;
                        int li = 0;
#pragma unroll
                        for (int d = 0; d < D; ++d) li = li * e[d] + (curu[d] + ((i >> (D - 1 - d)) & 1) - lo[d]);
                        atomicAdd(&tile[li], (unsigned long long)t);
                    }
                } else if (ovf) {
                    unsigned long long old[NS];
#pragma unroll
                    for (int i = 0; i < NS; ++i)
                        old[i] = run[i] ? atomicAdd(&acc[support_index<D>(g, cur, i)], (unsigned long long)run[i])
                                        : 0ull;
                    bool wr = false;
#pragma unroll
                    for (int i = 0; i < NS; ++i) wr |= add_wraps(old[i], (unsigned long long)run[i]);
                    if (wr) atomicAdd(ovf, 1ull);
                } else {
#pragma unroll
                    for (int i = 0; i < NS; ++i)
                        if (run[i]) atomicAdd(&acc[support_index<D>(g, cur, i)], (unsigned long long)run[i]);
                }
            };
#pragma unroll
            for (int j = 0; j < kPPT; ++j) {
                if (!ok[j]) continue;
                bool same = have;
#pragma unroll
                for (int d = 0; d < D; ++d) same = same && cur[d] == cw[j][d];
                if (!same) {
                    if (have) flush();
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        cur[d] = cw[j][d];
                        curu[d] = uw[j][d];
                    }
#pragma unroll
                    for (int q = 0; q < NS; ++q) run[q] = 0;
                    have = true;
                }
                T W[D][2];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    W[d][0] = (T)1 - sw[j][d];
                    W[d][1] = sw[j][d];
                }
#pragma unroll
                for (int i = 0; i < NS; ++i) {
                    // a coefficient below the cut adds a zero term.
                    // (w*cf)*2^S: scaling by a power of two is exact, so one
                    // product with w*2^S rounds identically
                    const T c0 = cic_coef<T, D>(W, i);
                    const T cf = coef_kept<T>(c0) ? c0 : (T)0;
                    run[i] += round_to_ll((double)cf * wsc);
                }
            }
            if (have) flush();
        }
        if (use_lds) {
            __syncthreads();
            // tile entry i -> its global node.  Entry i holds box offsets k
            // with i = (k0 * e1 + k1) * e2 + k2; the divisions by e[d] >= 2
            // are high products with ceil(2^32 / e[d]), exact for
            // i < 2^32 / e[d] (entries < 2^13).  A box that crosses no
            // periodic seam maps to base + sum k[d] * stride[d].
            uint32_t mdiv[3], gst[3];
            gst[D - 1] = 1u;
#pragma unroll
            for (int d = D - 2; d >= 0; --d) gst[d] = gst[d + 1] * (uint32_t)g.w[d + 1];
            bool seam = false;
            uint32_t gbase = 0u;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                mdiv[d] = 0xFFFFFFFFu / (uint32_t)e[d] + 1u;
                seam = seam || lo[d] < 0 || lo[d] + e[d] > g.w[d];
                gbase += (uint32_t)lo[d] * gst[d];
            }
            auto node_of = [&](int i) {
                uint32_t rem = (uint32_t)i, k[3];
#pragma unroll
                for (int d = D - 1; d >= 1; --d) {
                    const uint32_t q = __umulhi(rem, mdiv[d]);
                    k[d] = rem - q * (uint32_t)e[d];
                    rem = q;
                }
                k[0] = rem;
                if (!seam) {
                    uint32_t lin = gbase;
#pragma unroll
                    for (int d = 0; d < D; ++d) lin += k[d] * gst[d];
                    return lin;
                }
                int nd[3];
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    nd[d] = lo[d] + (int)k[d];
                    if (nd[d] < 0) nd[d] += g.w[d];
                    else if (nd[d] >= g.w[d]) nd[d] -= g.w[d];
                }
                uint32_t lin = (uint32_t)nd[0];
#pragma unroll
                for (int d = 1; d < D; ++d) lin = lin * (uint32_t)g.w[d] + (uint32_t)nd[d];
                return lin;
            };
            if (ovf) {
                // kFlushBatch entries per thread in flight before the
                // returned values are checked
                constexpr int kFlushBatch = 4;
                bool wr = false;
                for (int i0 = threadIdx.x; i0 < (int)tsize; i0 += kFlushBatch * kThreads) {
                    unsigned long long v[kFlushBatch], old[kFlushBatch];
#pragma unroll
                    for (int u = 0; u < kFlushBatch; ++u) {
                        const int i = i0 + u * kThreads;
                        v[u] = i < (int)tsize ? tile[i] : 0ull;
                    }
#pragma unroll
                    for (int u = 0; u < kFlushBatch; ++u)
                        old[u] = v[u] ? atomicAdd(&acc[node_of(i0 + u * kThreads)], v[u]) : 0ull;
#pragma unroll
                    for (int u = 0; u < kFlushBatch; ++u) wr |= add_wraps(old[u], v[u]);
                }
                if (wr) atomicAdd(ovf, 1ull);
            } else {
                for (int i = threadIdx.x; i < (int)tsize; i += kThreads) {
                    const unsigned long long v = tile[i];
                    if (v) atomicAdd(&acc[node_of(i)], v);
                }
            }
        }
        if (clipped && counters && threadIdx.x == 0) atomicAdd(&counters[1], 1ull);
    }

    if (PUSH && diag) {
        __syncthreads();
        if (threadIdx.x < 4) {
            double s = 0.0;
            for (int w = 0; w < kThreads / 16; ++w) s += wdiag[w][threadIdx.x];
            diag[(int64_t)blockIdx.x * 4 + threadIdx.x] = s;
        }
    }
}

// -------------------------------------------------------- small kernels

__global__ __launch_bounds__(kThreads) void k_fixed_to_double(const long long* __restrict__ acc,
                                                              int64_t n, double inv_scale,
                                                              double offset,
                                                              double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = (double)acc[i] * inv_scale + offset;
}

__global__ __launch_bounds__(kThreads) void k_f64_to_f32(const double* __restrict__ in,
                                                         float* __restrict__ out, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

// Fixed-order reduction of nb x 4 partials: pass 1 has block b sum the
// contiguous chunk [b*chunk, (b+1)*chunk) (lane-strided, then a tree), pass 2
// sums the per-chunk results the same way in one block.
__global__ __launch_bounds__(kThreads) void k_reduce4(const double* __restrict__ part, int64_t nb,
                                                      int64_t chunk, double* __restrict__ out)
{
    __shared__ double sh[4][kThreads];
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    const int64_t b0 = (int64_t)blockIdx.x * chunk;
    const int64_t b1 = b0 + chunk < nb ? b0 + chunk : nb;
    for (int64_t b = b0 + threadIdx.x; b < b1; b += kThreads)
#pragma unroll
        for (int q = 0; q < 4; ++q) s[q] += part[b * 4 + q];
#pragma unroll
    for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 4; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 4) out[(int64_t)blockIdx.x * 4 + threadIdx.x] = sh[threadIdx.x][0];
}

// Field energy partials: one value per block (written to part[b*4]).
template <int D>
__global__ __launch_bounds__(kThreads) void k_field_energy(DevGrid g, const double* __restrict__ phi,
                                                           int64_t nn, double* __restrict__ part)
{
    __shared__ double wsum[kThreads / 64];
    const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    double acc = 0.0;
    if (i < nn) {
        int64_t idx[3] = {0, 0, 0};
        if (nn <= (int64_t)0xffffffffll) {
            // 32-bit unsigned division (a 64-bit one is a long software sequence)
            uint32_t rem = (uint32_t)i;
            for (int d = D - 1; d >= 0; --d) {
                const uint32_t q = rem / (uint32_t)g.w[d];
                idx[d] = rem - q * (uint32_t)g.w[d];
                rem = q;
            }
        } else {
            int64_t rem = i;
            for (int d = D - 1; d >= 0; --d) {
                idx[d] = rem % g.w[d];
                rem /= g.w[d];
            }
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            int64_t nb = idx[d] + 1;
            if (nb == g.w[d]) {
                if (!g.periodic[d]) continue;
                nb = 0;
            }
            int64_t lin = 0;
#pragma unroll
            for (int q = 0; q < D; ++q) lin = lin * g.w[q] + (q == d ? nb : idx[q]);
            const double gd = (phi[lin] - phi[i]) * g.delta_r[d];
            acc += gd * gd;
        }
    }
    const double r = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = r;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += wsum[w];
        part[blockIdx.x * 4 + 0] = s;
        part[blockIdx.x * 4 + 1] = 0.0;
        part[blockIdx.x * 4 + 2] = 0.0;
        part[blockIdx.x * 4 + 3] = 0.0;
    }
}

// v[0] = (f * v[0]) * vol, the order of orc_field_energy
__global__ void k_scale_first(double* v, double f, double vol)
{
    if (threadIdx.x == 0) v[0] = (f * v[0]) * vol;
}

// ------------------------------------------------------ launch helpers

template <typename T>
psum_status launch_push(const psum_grid* g, void* const* pos, void* const* vel, int64_t n,
                        const void* phi, double qm, double dt, double mass, double w_dep,
                        double scale, int64_t* acc, double* diag, int64_t* counters, bool push,
                        hipStream_t st)
{
    const DevGrid dg = to_dev(*g);
    const unsigned nb = grid_blocks(n, kPushTile);
    if (nb == 0) return PSUM_OK;
    T* p[3] = {nullptr, nullptr, nullptr};
    T* v[3] = {nullptr, nullptr, nullptr};
    for (int d = 0; d < g->dim; ++d) {
        p[d] = static_cast<T*>(pos[d]);
        if (push) v[d] = static_cast<T*>(vel[d]);
    }
    auto* a = reinterpret_cast<unsigned long long*>(acc);
    auto* c = reinterpret_cast<unsigned long long*>(counters);
    const T* ph = static_cast<const T*>(phi);
    // unit node spacing: the gather's gradient weights are +-products of W
    // and the cell index needs no products with the spacing; from the origin
    // 0, FP32 positions take it in FP32
    bool unit = true;
    for (int d = 0; d < g->dim; ++d) unit = unit && g->delta_r[d] == 1.0 && g->delta[d] == 1.0;
    const bool unit0 = unit && dg.origin0 && sizeof(T) == 4;
#define PSUM_LAUNCH(DIM, PU, DE, UN)                                                          \
    hipLaunchKernelGGL((k_push_deposit<T, DIM, PU, DE, UN>), dim3(nb), dim3(kThreads), 0, st, \
                       dg, p[0], p[1], p[2], v[0], v[1], v[2], n, ph, (T)qm, (T)dt, mass,      \
                       w_dep, scale, a, diag, c)
#define PSUM_DISPATCH(DIM)                                                                    \
    if (push && acc) {                                                                        \
        if (unit0) PSUM_LAUNCH(DIM, true, true, kGridUnit0);                                  \
        else if (unit) PSUM_LAUNCH(DIM, true, true, kGridUnit);                               \
        else PSUM_LAUNCH(DIM, true, true, kGridAny);                                          \
    } else if (push) {                                                                        \
        if (unit0) PSUM_LAUNCH(DIM, true, false, kGridUnit0);                                 \
        else if (unit) PSUM_LAUNCH(DIM, true, false, kGridUnit);                              \
        else PSUM_LAUNCH(DIM, true, false, kGridAny);                                         \
    } else PSUM_LAUNCH(DIM, false, true, kGridAny);
    switch (g->dim) {
    case 1: PSUM_DISPATCH(1); break;
    case 2: PSUM_DISPATCH(2); break;
    default: PSUM_DISPATCH(3); break;
    }
#undef PSUM_DISPATCH
#undef PSUM_LAUNCH
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // namespace

// ====================================================================== C ABI

namespace psum_hip {

size_t scan_workspace(int64_t n)
{
    size_t bytes = 0;
    while (n > kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        bytes += ((size_t)n * sizeof(uint32_t) + 255) & ~(size_t)255;
    }
    return bytes + 256;
}

psum_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* work,
                               hipStream_t st)
{
    if (n <= 0) return PSUM_OK;
    const unsigned nb = grid_blocks(n, kScanTile);
    if (nb == 1) {
        hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(kThreads), 0, st, in, out, nullptr, n);
        PSUM_HIP_TRY(hipGetLastError());
        return PSUM_OK;
    }
    uint32_t* bsum = static_cast<uint32_t*>(work);
    void* rest = static_cast<char*>(work) + (((size_t)nb * sizeof(uint32_t) + 255) & ~(size_t)255);
    hipLaunchKernelGGL(k_scan_tiles, dim3(nb), dim3(kThreads), 0, st, in, out, bsum, n);
    PSUM_HIP_TRY(hipGetLastError());
    psum_status s = exclusive_scan_u32(bsum, bsum, nb, rest, st);
    if (s != PSUM_OK) return s;
    hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(kThreads), 0, st, out, bsum, n);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // namespace psum_hip

static bool pow2(double s)
{
    int e = 0;
    return s > 0.0 && std::isfinite(s) && std::frexp(s, &e) == 0.5;
}

extern "C" {

psum_status psum_hip_bin(const psum_grid* g, int dtype, void* const* pos, int64_t n,
                         uint32_t* keys, void* stream)
{
    return psum_hip::bin_keys(g, dtype, pos, n, keys, 0, stream);
}

}  // extern "C"

namespace psum_hip {

uint32_t bin_count(const psum_grid& g, int shift)
{
    uint32_t n = 1;
    for (int d = 0; d < g.dim; ++d) n *= (uint32_t)((g.cells[d] + 1 + (1 << shift) - 1) >> shift);
    return n;
}

psum_status bin_keys(const psum_grid* g, int dtype, void* const* pos, int64_t n, uint32_t* keys,
                     int shift, void* stream, void* const* vel)
{
    PSUM_ARG_CHECK(grid_ok(g), "psum_hip_bin: invalid grid");
    PSUM_ARG_CHECK(shift >= 0 && shift <= 8, "psum_hip_bin: bin shift out of range");
    PSUM_ARG_CHECK(n >= 0 && (n == 0 || (pos && keys)), "psum_hip_bin: null pointer");
    if (vel) PSUM_ARG_CHECK(((uint64_t)bin_count(*g, shift) << g->dim) < (uint64_t)PSUM_KEY_INVALID,
                            "psum_hip_bin: too many bins for velocity-refined keys");
    if (n == 0) return PSUM_OK;
    const DevGrid dg = to_dev(*g);
    const unsigned nb = grid_blocks(n, kThreads);
#define PSUM_BIN(T, DIM)                                                                         \
    do {                                                                                         \
        const T* xp[3] = {static_cast<const T*>(pos[0]), DIM > 1 ? static_cast<const T*>(pos[1]) : nullptr, \
                          DIM > 2 ? static_cast<const T*>(pos[2]) : nullptr};                     \
        if (vel)                                                                                 \
            hipLaunchKernelGGL((k_bin<T, DIM, true>), dim3(nb), dim3(kThreads), 0, S(stream), dg, xp[0], \
                               xp[1], xp[2], static_cast<const T*>(vel[0]),                      \
                               DIM > 1 ? static_cast<const T*>(vel[1]) : nullptr,                \
                               DIM > 2 ? static_cast<const T*>(vel[2]) : nullptr, n, keys, shift); \
        else                                                                                     \
            hipLaunchKernelGGL((k_bin<T, DIM, false>), dim3(nb), dim3(kThreads), 0, S(stream), dg, xp[0], \
                               xp[1], xp[2], nullptr, nullptr, nullptr, n, keys, shift);        \
    } while (0)
    if (dtype == PSUM_F64) {
        if (g->dim == 1) PSUM_BIN(double, 1); else if (g->dim == 2) PSUM_BIN(double, 2); else PSUM_BIN(double, 3);
    } else if (dtype == PSUM_F32) {
        if (g->dim == 1) PSUM_BIN(float, 1); else if (g->dim == 2) PSUM_BIN(float, 2); else PSUM_BIN(float, 3);
    } else {
        set_error("psum_hip_bin: bad dtype");
        return PSUM_ERR_ARG;
    }
#undef PSUM_BIN
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // namespace psum_hip

extern "C" {

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t psum_hip_sort_workspace(int64_t n)
{
    if (n <= 0) return 256;
    const int64_t nb = (n + kSortTile - 1) / kSortTile;
    const int64_t nh = nb * 256;
    return 2 * align256((size_t)n * 4) + align256((size_t)nh * 4) + scan_workspace(nh);
}

psum_status psum_hip_sort(const uint32_t* keys, int64_t n, uint32_t nkeys, uint32_t* keys_out,
                          uint32_t* perm, void* work, size_t work_bytes, void* stream)
{
    PSUM_ARG_CHECK(n >= 0 && n < (int64_t)0xFFFFFFFF, "psum_hip_sort: n out of range");
    if (n == 0) return PSUM_OK;
    PSUM_ARG_CHECK(keys && keys_out && perm && work, "psum_hip_sort: null pointer");
    PSUM_ARG_CHECK(work_bytes >= psum_hip_sort_workspace(n), "psum_hip_sort: workspace too small");
    PSUM_ARG_CHECK(nkeys >= 1 && nkeys < PSUM_KEY_INVALID, "psum_hip_sort: bad nkeys");
    hipStream_t st = S(stream);
    const int nb = (int)((n + kSortTile - 1) / kSortTile);
    char* w = static_cast<char*>(work);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(w);
    w += align256((size_t)n * 4);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(w);
    w += align256((size_t)n * 4);
    uint32_t* gh = reinterpret_cast<uint32_t*>(w);
    w += align256((size_t)nb * 256 * 4);
    void* scan_work = w;

    int bits = 32 - __builtin_clz(nkeys);   // INVALID maps to nkeys
    const int passes = (bits + 7) / 8;
    const uint32_t* kin = keys;
    const uint32_t* vin = nullptr;
    for (int p = 0; p < passes; ++p) {
        // the last pass lands in (keys_out, perm)
        const bool to_out = ((passes - 1 - p) % 2) == 0;
        uint32_t* kout = to_out ? keys_out : k1;
        uint32_t* vout = to_out ? perm : v1;
        hipLaunchKernelGGL(k_sort_hist, dim3(nb), dim3(kThreads), 0, st, kin, n, nkeys, 8 * p, gh,
                           nb);
        PSUM_HIP_TRY(hipGetLastError());
        psum_status s = exclusive_scan_u32(gh, gh, (int64_t)nb * 256, scan_work, st);
        if (s != PSUM_OK) return s;
        hipLaunchKernelGGL(k_sort_scatter, dim3(nb), dim3(kThreads), 0, st, kin, vin, n, nkeys,
                           8 * p, gh, nb, kout, vout);
        PSUM_HIP_TRY(hipGetLastError());
        kin = kout;
        vin = vout;
    }
    return PSUM_OK;
}

psum_status psum_hip_permute(int dtype, int ncomp, void* const* src, void* const* dst,
                             const uint32_t* perm, int64_t n, void* stream)
{
    PSUM_ARG_CHECK(ncomp >= 1 && ncomp <= 8, "psum_hip_permute: ncomp must be 1..8");
    if (n == 0) return PSUM_OK;
    PtrPack pk{};
    for (int c = 0; c < ncomp; ++c) {
        PSUM_ARG_CHECK(src[c] && dst[c] && src[c] != dst[c], "psum_hip_permute: bad arrays");
        pk.src[c] = src[c];
        pk.dst[c] = dst[c];
    }
    bool a16 = ((uintptr_t)perm & 15) == 0;
    for (int c = 0; c < ncomp; ++c) a16 = a16 && ((uintptr_t)dst[c] & 15) == 0;
    const unsigned nb = grid_blocks(n, kThreads), nb4 = grid_blocks((n + 3) / 4, kThreads);
    if (dtype == PSUM_F64) {
        if (a16) hipLaunchKernelGGL(k_permute4<double>, dim3(nb4), dim3(kThreads), 0, S(stream), pk, ncomp, perm, n);
        else hipLaunchKernelGGL(k_permute<double>, dim3(nb), dim3(kThreads), 0, S(stream), pk, ncomp, perm, n);
    } else if (dtype == PSUM_F32) {
        if (a16) hipLaunchKernelGGL(k_permute4<float>, dim3(nb4), dim3(kThreads), 0, S(stream), pk, ncomp, perm, n);
        else hipLaunchKernelGGL(k_permute<float>, dim3(nb), dim3(kThreads), 0, S(stream), pk, ncomp, perm, n);
    }
    else {
        set_error("psum_hip_permute: bad dtype");
        return PSUM_ERR_ARG;
    }
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

size_t psum_hip_compact_workspace(int64_t n)
{
    const int64_t nb = (n + kCompactTile - 1) / kCompactTile;
    return align256((size_t)(nb + 1) * 4) + scan_workspace(nb) + 256;
}

}  // extern "C"

// Body of psum_hip_compact / psum_hip_compact_noalloc: `work` holds at least
// psum_hip_compact_workspace(n) bytes of device memory.  Every word of it
// that is read is written first in the same call (block counts, then the
// scan's partial sums), so it needs no clearing between calls.
static psum_status compact_impl(int dtype, int dim, void* const* pos, void* const* vel, int64_t n,
                                void* const* out_pos, void* const* out_vel, int64_t* count_host,
                                void* work, void* stream)
{
    PSUM_ARG_CHECK(dim >= 1 && dim <= 3, "psum_hip_compact: bad dim");
    hipStream_t st = S(stream);
    if (n == 0) {
        if (count_host) *count_host = 0;
        return PSUM_OK;
    }
    const unsigned nb = grid_blocks(n, kCompactTile);
    uint32_t* bc = static_cast<uint32_t*>(work);
    void* scan_work = static_cast<char*>(work) + align256((size_t)(nb + 1) * 4);
    PtrPack pk{};
    int nc = 0;
    for (int d = 0; d < dim; ++d) { pk.src[nc] = pos[d]; pk.dst[nc] = out_pos[d]; ++nc; }
    for (int d = 0; d < dim; ++d) { pk.src[nc] = vel[d]; pk.dst[nc] = out_vel[d]; ++nc; }
    for (int c = 0; c < nc; ++c) PSUM_ARG_CHECK(pk.src[c] && pk.dst[c], "psum_hip_compact: null array");
    if (dtype == PSUM_F64) {
        hipLaunchKernelGGL(k_compact_count<double>, dim3(nb), dim3(kThreads), 0, st,
                           static_cast<const double*>(pos[0]), n, bc);
    } else if (dtype == PSUM_F32) {
        hipLaunchKernelGGL(k_compact_count<float>, dim3(nb), dim3(kThreads), 0, st,
                           static_cast<const float*>(pos[0]), n, bc);
    } else {
        set_error("psum_hip_compact: bad dtype");
        return PSUM_ERR_ARG;
    }
    PSUM_HIP_TRY(hipGetLastError());
    // scan nb+1 entries: entry nb becomes the total
    psum_status s = exclusive_scan_u32(bc, bc, (int64_t)nb + 1, scan_work, st);
    if (s != PSUM_OK) return s;
    if (dtype == PSUM_F64)
        hipLaunchKernelGGL(k_compact_scatter<double>, dim3(nb), dim3(kThreads), 0, st, pk, nc, n, bc);
    else
        hipLaunchKernelGGL(k_compact_scatter<float>, dim3(nb), dim3(kThreads), 0, st, pk, nc, n, bc);
    PSUM_HIP_TRY(hipGetLastError());
    if (count_host) {
        uint32_t total = 0;
        PSUM_HIP_TRY(hipMemcpyAsync(&total, bc + nb, 4, hipMemcpyDeviceToHost, st));
        PSUM_HIP_TRY(hipStreamSynchronize(st));
        *count_host = total;
    }
    return PSUM_OK;
}

// ------------------------------------------------ allocation-free compaction
// Device scratch of psum_hip_compact_noalloc: block counts and the scan's
// partial sums for up to kNoAllocMaxParticles particles.  Each GPU the
// library runs on holds its own copy until the process ends; every call on
// a GPU shares it, so calls that overlap on one GPU race.
constexpr int64_t kNoAllocMaxParticles = (int64_t)1 << 31;
constexpr size_t kNoAllocScratchBytes = ((size_t)8 << 20) + ((size_t)64 << 10);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side name is only a stand-in).
psum_status scratch_on_current_device(void** out)
{
    int dev = -1;
    PSUM_HIP_TRY(hipGetDevice(&dev));
    PSUM_ARG_CHECK(dev >= 0 && dev < kMaxDevices, "psum_hip_compact_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        PSUM_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PSUM_ARG_CHECK(p != nullptr, "psum_hip_compact_noalloc: no device address for the scratch array");
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return PSUM_OK;
}
}  // namespace

extern "C" {

psum_status psum_hip_compact(int dtype, int dim, void* const* pos, void* const* vel, int64_t n,
                             void* const* out_pos, void* const* out_vel, int64_t* count_host,
                             void* work, size_t work_bytes, void* stream)
{
    PSUM_ARG_CHECK(work_bytes >= psum_hip_compact_workspace(n), "psum_hip_compact: workspace too small");
    return compact_impl(dtype, dim, pos, vel, n, out_pos, out_vel, count_host, work, stream);
}

psum_status psum_hip_compact_noalloc(int dtype, int dim, void* const* pos, void* const* vel, int64_t n,
                                     void* const* out_pos, void* const* out_vel, int64_t* count_host,
                                     void* stream)
{
    PSUM_ARG_CHECK(n >= 0 && n <= kNoAllocMaxParticles,
                   "psum_hip_compact_noalloc: n exceeds the 2^31-particle limit of its scratch");
    PSUM_ARG_CHECK(psum_hip_compact_workspace(n) <= kNoAllocScratchBytes,
                   "psum_hip_compact_noalloc: scratch too small for n");
    void* work = nullptr;
    const psum_status s = scratch_on_current_device(&work);
    if (s != PSUM_OK) return s;
    return compact_impl(dtype, dim, pos, vel, n, out_pos, out_vel, count_host, work, stream);
}

psum_status psum_hip_deposit_cic(const psum_grid* g, int dtype, void* const* pos, int64_t n,
                                 double w, double scale, int64_t* acc, int64_t* counters,
                                 void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g), "psum_hip_deposit_cic: invalid grid");
    PSUM_ARG_CHECK(acc && (n == 0 || pos), "psum_hip_deposit_cic: null pointer");
    PSUM_ARG_CHECK(pow2(scale), "psum_hip_deposit_cic: scale must be a power of two");
    PSUM_ARG_CHECK(std::fabs(w) * scale <= 0x1p50, "psum_hip_deposit_cic: |w| * scale must not exceed 2^50");
    if (dtype == PSUM_F64)
        return launch_push<double>(g, pos, nullptr, n, nullptr, 0, 0, 0, w, scale, acc, nullptr,
                                   counters, false, S(stream));
    if (dtype == PSUM_F32)
        return launch_push<float>(g, pos, nullptr, n, nullptr, 0, 0, 0, w, scale, acc, nullptr,
                                  counters, false, S(stream));
    set_error("psum_hip_deposit_cic: bad dtype");
    return PSUM_ERR_ARG;
}

psum_status psum_hip_fixed_to_double(const int64_t* acc, int64_t nn, double inv_scale,
                                     double offset, double* rho, void* stream)
{
    if (nn == 0) return PSUM_OK;
    hipLaunchKernelGGL(k_fixed_to_double, dim3(grid_blocks(nn, kThreads)), dim3(kThreads), 0,
                       S(stream), reinterpret_cast<const long long*>(acc), nn, inv_scale, offset,
                       rho);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

int64_t psum_hip_push_blocks(int64_t n) { return (n + kPushTile - 1) / kPushTile; }

psum_status psum_hip_gather_push(const psum_grid* g, int dtype, void* const* pos,
                                 void* const* vel, int64_t n, const void* phi, double qm,
                                 double dt, double mass, double w_dep, double scale, int64_t* acc,
                                 double* diag, int64_t* counters, void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g), "psum_hip_gather_push: invalid grid");
    PSUM_ARG_CHECK(n == 0 || (pos && vel && phi), "psum_hip_gather_push: null pointer");
    PSUM_ARG_CHECK(!acc || pow2(scale), "psum_hip_gather_push: scale must be a power of two");
    PSUM_ARG_CHECK(!acc || std::fabs(w_dep) * scale <= 0x1p50,
                   "psum_hip_gather_push: |w_dep| * scale must not exceed 2^50");
    if (dtype == PSUM_F64)
        return launch_push<double>(g, pos, vel, n, phi, qm, dt, mass, w_dep, scale, acc, diag,
                                   counters, true, S(stream));
    if (dtype == PSUM_F32)
        return launch_push<float>(g, pos, vel, n, phi, qm, dt, mass, w_dep, scale, acc, diag,
                                  counters, true, S(stream));
    set_error("psum_hip_gather_push: bad dtype");
    return PSUM_ERR_ARG;
}

size_t psum_hip_reduce_workspace(void) { return (size_t)kReduceBlocks * 4 * sizeof(double); }

static psum_status reduce4(const double* part, int64_t nb, double* out, double* work,
                           hipStream_t st)
{
    if (nb <= 4 * kThreads) {
        hipLaunchKernelGGL(k_reduce4, dim3(1), dim3(kThreads), 0, st, part, nb, nb, out);
    } else {
        const int64_t chunk = (nb + kReduceBlocks - 1) / kReduceBlocks;
        const int64_t nr = (nb + chunk - 1) / chunk;
        hipLaunchKernelGGL(k_reduce4, dim3((unsigned)nr), dim3(kThreads), 0, st, part, nb, chunk, work);
        hipLaunchKernelGGL(k_reduce4, dim3(1), dim3(kThreads), 0, st, work, nr, nr, out);
    }
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_reduce_diag(const double* diag, int64_t nblocks, double* out, void* work,
                                 void* stream)
{
    PSUM_ARG_CHECK(diag && out, "psum_hip_reduce_diag: null pointer");
    PSUM_ARG_CHECK(work || nblocks <= 4 * kThreads, "psum_hip_reduce_diag: workspace required");
    return reduce4(diag, nblocks, out, static_cast<double*>(work), S(stream));
}

size_t psum_hip_field_energy_workspace(const psum_grid* g)
{
    if (!grid_ok(g)) return 0;
    const int64_t nb = (grid_nodes(*g) + kThreads - 1) / kThreads;
    return (size_t)nb * 4 * sizeof(double) + (size_t)kReduceBlocks * 4 * sizeof(double);
}

psum_status psum_hip_field_energy(const psum_grid* g, const double* phi, double eps, double* out,
                                  void* work, size_t work_bytes, void* stream)
{
    PSUM_ARG_CHECK(grid_ok(g) && phi && out && work, "psum_hip_field_energy: bad argument");
    PSUM_ARG_CHECK(work_bytes >= psum_hip_field_energy_workspace(g),
                   "psum_hip_field_energy: workspace too small");
    const int64_t nn = grid_nodes(*g);
    const int64_t nb = (nn + kThreads - 1) / kThreads;
    double* part = static_cast<double*>(work);
    const DevGrid dg = to_dev(*g);
    hipStream_t st = S(stream);
    if (g->dim == 1)
        hipLaunchKernelGGL(k_field_energy<1>, dim3(nb), dim3(kThreads), 0, st, dg, phi, nn, part);
    else if (g->dim == 2)
        hipLaunchKernelGGL(k_field_energy<2>, dim3(nb), dim3(kThreads), 0, st, dg, phi, nn, part);
    else
        hipLaunchKernelGGL(k_field_energy<3>, dim3(nb), dim3(kThreads), 0, st, dg, phi, nn, part);
    PSUM_HIP_TRY(hipGetLastError());
    psum_status rs = reduce4(part, nb, out, part + nb * 4, st);
    if (rs != PSUM_OK) return rs;
    double vol = 1.0;
    for (int d = 0; d < g->dim; ++d) vol *= g->delta[d];
    hipLaunchKernelGGL(k_scale_first, dim3(1), dim3(64), 0, st, out, 0.5 * eps, vol);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

psum_status psum_hip_f64_to_f32(const double* in, float* out, int64_t n, void* stream)
{
    if (n == 0) return PSUM_OK;
    hipLaunchKernelGGL(k_f64_to_f32, dim3(grid_blocks(n, kThreads)), dim3(kThreads), 0, S(stream),
                       in, out, n);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // extern "C"
