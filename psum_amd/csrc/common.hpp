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

// Exclusive scan of n uint32 values (out may alias in).  Workspace bytes
// from scan_workspace(n).
size_t scan_workspace(int64_t n);
psum_status exclusive_scan_u32(const uint32_t* in, uint32_t* out, int64_t n, void* work,
                               hipStream_t stream);

}  // namespace psum_hip
