// Shared host/device helpers of the MI355X PIC step engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <exception>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "psum_hip.h"

namespace psum_hip {

void set_error(const std::string& msg);

// Sort keys of 2^shift-cell bins (shift 0: the cell key of psum_hip_bin)
// and the number of bins.  With vel != NULL the key is refined by the signs
// of the velocity components, key = bin * 2^D + sum_d [v_d >= 0] 2^(D-1-d),
// so that particles of one bin moving the same way stay adjacent.
psum_status bin_keys(const psum_grid* g, int dtype, void* const* pos, int64_t n, uint32_t* keys,
                     int shift, void* stream, void* const* vel = nullptr);
uint32_t bin_count(const psum_grid& g, int shift);

#define PSUM_HIP_TRY(expr)                                                              \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::psum_hip::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
            return PSUM_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

#define PSUM_ARG_CHECK(cond, msg)                                                       \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::psum_hip::set_error(msg);                                                 \
            return PSUM_ERR_ARG;                                                        \
        }                                                                               \
    } while (0)

// Runs body(t) for t = 0..nt-1 on host threads.  A thread that cannot be
// created (std::system_error, e.g. under a thread limit) has its share run
// on the calling thread instead; the first exception a body throws is
// rethrown after every started thread has been joined.
template <typename F>
void host_threads(unsigned nt, F&& body)
{
    std::vector<std::thread> pool;
    std::vector<std::exception_ptr> err(nt);
    std::vector<unsigned> inline_ids;
    pool.reserve(nt);
    inline_ids.reserve(nt);
    for (unsigned t = 0; t < nt; ++t) {
        auto run = [t, &body, &err]() {
            try {
                body(t);
            } catch (...) {
                err[t] = std::current_exception();
            }
        };
        try {
            pool.emplace_back(run);
        } catch (const std::system_error&) {
            inline_ids.push_back(t);
        }
    }
    for (unsigned t : inline_ids) {
        try {
            body(t);
        } catch (...) {
            err[t] = std::current_exception();
        }
    }
    for (auto& th : pool) th.join();
    for (auto& e : err)
        if (e) std::rethrow_exception(e);
}

// Ends a stream capture that is still open when the scope is left (an
// exception between Begin and EndCapture) and drops the partial graph, so
// the stream is usable again.  release() after a successful EndCapture.
struct CaptureGuard {
    hipStream_t s;
    bool open = true;
    explicit CaptureGuard(hipStream_t st) : s(st) {}
    void release() { open = false; }
    ~CaptureGuard()
    {
        if (!open) return;
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
    }
};

}  // namespace psum_hip

#include "psum_cic.hpp"

namespace psum_hip {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_isum(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ long long wave_lsum(long long v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Reductions over a full wave (all 64 lanes active) without the LDS
// permute: a butterfly inside each row of 16 lanes through DPP moves
// (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror;
// every source lane is in range), after which each lane of a row holds the
// row's total, then the four row totals read from lanes 0, 16, 32 and 48.
// The result is the same in every lane.  Sums of doubles are taken in this
// fixed order (deterministic, not the order of wave_sum).
namespace dpp {
constexpr int kXor1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int kXor2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int kHalfMirror = 0x141;  // lane i of 8 reads lane 7 - i
constexpr int kMirror = 0x140;      // lane i of 16 reads lane 15 - i

template <int CTRL>
__device__ __forceinline__ int mov(int v, int identity)
{
    return __builtin_amdgcn_update_dpp(identity, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ long long mov64(long long v)
{
    const int lo = mov<CTRL>((int)(unsigned long long)v, 0);
    const int hi = mov<CTRL>((int)((unsigned long long)v >> 32), 0);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ long long readlane64(long long v, int lane)
{
    const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)v, lane);
    const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)v >> 32), lane);
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
}  // namespace dpp

// the butterfly of wave_sum_full within each row of 16 lanes: every lane
// holds its row's total (all 64 lanes active)
__device__ __forceinline__ double row_sum_full(double v)
{
    v += __longlong_as_double(dpp::mov64<dpp::kXor1>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kXor2>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kHalfMirror>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kMirror>(__double_as_longlong(v)));
    return v;
}

__device__ __forceinline__ double wave_sum_full(double v)
{
    v += __longlong_as_double(dpp::mov64<dpp::kXor1>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kXor2>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kHalfMirror>(__double_as_longlong(v)));
    v += __longlong_as_double(dpp::mov64<dpp::kMirror>(__double_as_longlong(v)));
    const long long b = __double_as_longlong(v);
    const double r0 = __longlong_as_double(dpp::readlane64(b, 0));
    const double r1 = __longlong_as_double(dpp::readlane64(b, 16));
    const double r2 = __longlong_as_double(dpp::readlane64(b, 32));
    const double r3 = __longlong_as_double(dpp::readlane64(b, 48));
    return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ long long wave_lsum_full(long long v)
{
    v += dpp::mov64<dpp::kXor1>(v);
    v += dpp::mov64<dpp::kXor2>(v);
    v += dpp::mov64<dpp::kHalfMirror>(v);
    v += dpp::mov64<dpp::kMirror>(v);
    return dpp::readlane64(v, 0) + dpp::readlane64(v, 16) + dpp::readlane64(v, 32) + dpp::readlane64(v, 48);
}

template <typename Op>
__device__ __forceinline__ int wave_reduce_full(int v, int identity, Op op)
{
    v = op(v, dpp::mov<dpp::kXor1>(v, identity));
    v = op(v, dpp::mov<dpp::kXor2>(v, identity));
    v = op(v, dpp::mov<dpp::kHalfMirror>(v, identity));
    v = op(v, dpp::mov<dpp::kMirror>(v, identity));
    const int r0 = __builtin_amdgcn_readlane(v, 0), r1 = __builtin_amdgcn_readlane(v, 16);
    const int r2 = __builtin_amdgcn_readlane(v, 32), r3 = __builtin_amdgcn_readlane(v, 48);
    return op(op(r0, r1), op(r2, r3));
}
__device__ __forceinline__ int wave_isum_full(int v)
{
    return wave_reduce_full(v, 0, [](int a, int b) { return a + b; });
}
__device__ __forceinline__ int wave_min_full(int v)
{
    return wave_reduce_full(v, INT_MAX, [](int a, int b) { return min(a, b); });
}
__device__ __forceinline__ int wave_max_full(int v)
{
    return wave_reduce_full(v, INT_MIN, [](int a, int b) { return max(a, b); });
}

// Exclusive scan of n uint32 values (out may alias in).  Workspace bytes
// from scan_workspace(n).
size_t scan_workspace(int64_t n);
psum_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* work,
                               hipStream_t stream);

}  // namespace psum_hip
