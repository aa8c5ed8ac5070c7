// device_vector<T>: device memory with a size and a capacity, and a device
// accessor with operator[] and an atomic push_back
// (src/particle_container/device_vector.hpp:77-287), on a HIP stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace psum {

// Stands where the reference's SYCL handler is: the argument of the
// functions given to for_each, from which accessors are taken.  An accessor
// that needs work once the kernel has run (a fixed-point deposit folded into
// its field) registers it with after(); the launching for_each calls
// finish() when the kernel has completed.
struct handler {
    hipStream_t stream = nullptr;
    std::vector<std::pair<const void*, std::function<void()>>> done;

    // f runs once after the kernel, at most once per key
    void after(const void* key, std::function<void()> f)
    {
        for (const auto& d : done)
            if (d.first == key) return;
        done.emplace_back(key, std::move(f));
    }
    void finish()
    {
        auto d = std::move(done);
        done.clear();
        for (auto& x : d) x.second();
    }
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

namespace particle_container {

template <typename T>
struct device_vector_acc {
    T* data_ = nullptr;
    unsigned long long* size_ = nullptr;
    size_t capacity_ = 0;
    __device__ T& operator[](size_t i) const { return data_[i]; }
    // appends v; returns false when the vector is full (the element is dropped)
    __device__ bool push_back(const T& v) const
    {
        const unsigned long long i = atomicAdd(size_, 1ull);
        if (i >= capacity_) return false;
        data_[i] = v;
        return true;
    }
};

template <typename T, typename F>
__global__ void k_vector_for_each(T* data, size_t n, F f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f(data[i]);
}

template <typename T>
class device_vector {
public:
    explicit device_vector(hipStream_t s = nullptr, size_t capacity = 0) : stream_(s) { reserve(capacity); }
    device_vector(hipStream_t s, const std::vector<T>& host) : stream_(s)
    {
        reserve(host.size());
        if (!host.empty())
            hip_check(hipMemcpyAsync(data_, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream_),
                      "device_vector copy");
        set_size(host.size());
    }
    device_vector(const device_vector&) = delete;
    device_vector& operator=(const device_vector&) = delete;
    device_vector(device_vector&& o) noexcept { swap(o); }
    device_vector& operator=(device_vector&& o) noexcept { swap(o); return *this; }
    ~device_vector()
    {
        if (data_) (void)hipFree(data_);
        if (size_dev_) (void)hipFree(size_dev_);
    }

    void swap(device_vector& o) noexcept
    {
        std::swap(data_, o.data_);
        std::swap(size_dev_, o.size_dev_);
        std::swap(capacity_, o.capacity_);
        std::swap(stream_, o.stream_);
    }

    device_vector_acc<T> get_access(handler&) const { return {data_, size_dev_, capacity_}; }
    device_vector_acc<T> get_access() const { return {data_, size_dev_, capacity_}; }

    // the device-side size, clamped to the capacity (push_back past it drops)
    size_t size() const
    {
        unsigned long long n = 0;
        if (size_dev_) {
            hip_check(hipMemcpyAsync(&n, size_dev_, sizeof n, hipMemcpyDeviceToHost, stream_), "size");
            hip_check(hipStreamSynchronize(stream_), "size");
        }
        return n < capacity_ ? (size_t)n : capacity_;
    }
    size_t capacity() const { return capacity_; }
    T* data() { return data_; }
    const T* data() const { return data_; }
    hipStream_t get_queue() const { return stream_; }

    void reserve(size_t cap)
    {
        if (!size_dev_) {
            hip_check(hipMalloc(&size_dev_, sizeof(unsigned long long)), "device_vector");
            set_size(0);
        }
        if (cap <= capacity_) return;
        const size_t n = capacity_ ? size() : 0;
        T* nd = nullptr;
        hip_check(hipMalloc(&nd, cap * sizeof(T)), "device_vector reserve");
        if (n) hip_check(hipMemcpyAsync(nd, data_, n * sizeof(T), hipMemcpyDeviceToDevice, stream_), "reserve");
        hip_check(hipStreamSynchronize(stream_), "reserve");
        if (data_) (void)hipFree(data_);
        data_ = nd;
        capacity_ = cap;
    }

    void resize(size_t n)
    {
        reserve(n);
        set_size(n);
    }

    void clear() { set_size(0); }

    // the accessor without a capacity check on push_back's index: the
    // caller guarantees the room (device_vector.hpp:207-209)
    device_vector_acc<T> get_access_without_overflow_check(handler& h) const { return get_access(h); }

    // throws when a push_back found the vector full (its element was
    // dropped) since the size was last set (device_vector.hpp:158)
    void if_overflow() const
    {
        unsigned long long n = 0;
        if (size_dev_) {
            hip_check(hipMemcpyAsync(&n, size_dev_, sizeof n, hipMemcpyDeviceToHost, stream_), "if_overflow");
            hip_check(hipStreamSynchronize(stream_), "if_overflow");
        }
        if (n > capacity_) throw std::overflow_error("device_vector: push_back beyond the capacity");
    }

    // the host vector's elements, size and all
    void copy(const std::vector<T>& host)
    {
        reserve(host.size());
        if (!host.empty())
            hip_check(hipMemcpyAsync(data_, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, stream_),
                      "device_vector copy");
        set_size(host.size());
    }

    // func(handler&) returns the per-element function f(T&), applied to
    // every element (device_vector.hpp:229-266)
    template <typename FuncType>
    void for_each(FuncType&& func)
    {
        const size_t n = size();
        if (n == 0) return;
        handler h{stream_};
        auto f = func(h);
        hipLaunchKernelGGL((k_vector_for_each<T, decltype(f)>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           stream_, data_, n, f);
        hip_check(hipGetLastError(), "for_each");
        hip_check(hipStreamSynchronize(stream_), "for_each");
        h.finish();
    }
    template <typename FuncType>
    void for_each(FuncType&& func) const
    {
        const_cast<device_vector*>(this)->for_each(std::forward<FuncType>(func));
    }

    std::vector<T> to_host() const
    {
        std::vector<T> h(size());
        if (!h.empty()) {
            hip_check(hipMemcpyAsync(h.data(), data_, h.size() * sizeof(T), hipMemcpyDeviceToHost, stream_), "to_host");
            hip_check(hipStreamSynchronize(stream_), "to_host");
        }
        return h;
    }

    void set_size(size_t n)
    {
        const unsigned long long v = n;
        hip_check(hipMemcpyAsync(size_dev_, &v, sizeof v, hipMemcpyHostToDevice, stream_), "set_size");
        hip_check(hipStreamSynchronize(stream_), "set_size");
    }

private:
    T* data_ = nullptr;
    unsigned long long* size_dev_ = nullptr;
    size_t capacity_ = 0;
    hipStream_t stream_ = nullptr;
};

}  // namespace particle_container
}  // namespace psum
