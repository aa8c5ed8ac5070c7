// Particle records: tagged_struct<tag_bind<Tag, Type>...> with get<Tag>(p)
// (src/tag/tagged_struct.hpp:13-61), and vec<D>, the fixed-size vector used
// for positions and velocities where the reference uses Eigen::Vector<double, D>.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <type_traits>

namespace psum {

template <int D, typename S = double>
struct vec {
    S v[D] = {};
    __host__ __device__ S& operator[](int i) { return v[i]; }
    __host__ __device__ const S& operator[](int i) const { return v[i]; }
    __host__ __device__ S& operator()(int i) { return v[i]; }
    __host__ __device__ const S& operator()(int i) const { return v[i]; }
    __host__ __device__ S& x() { return v[0]; }
    __host__ __device__ const S& x() const { return v[0]; }
    __host__ __device__ S& y() { static_assert(D > 1); return v[1]; }
    __host__ __device__ const S& y() const { static_assert(D > 1); return v[1]; }
    __host__ __device__ S& z() { static_assert(D > 2); return v[2]; }
    __host__ __device__ const S& z() const { static_assert(D > 2); return v[2]; }
    static constexpr int size() { return D; }
    __host__ __device__ static vec Zero() { return vec{}; }
    __host__ __device__ vec& operator+=(const vec& o) { for (int i = 0; i < D; ++i) v[i] += o.v[i]; return *this; }
    __host__ __device__ vec& operator-=(const vec& o) { for (int i = 0; i < D; ++i) v[i] -= o.v[i]; return *this; }
    __host__ __device__ vec& operator*=(S a) { for (int i = 0; i < D; ++i) v[i] *= a; return *this; }
    __host__ __device__ friend vec operator+(vec a, const vec& b) { return a += b; }
    __host__ __device__ friend vec operator-(vec a, const vec& b) { return a -= b; }
    __host__ __device__ friend vec operator*(vec a, S s) { return a *= s; }
    __host__ __device__ friend vec operator*(S s, vec a) { return a *= s; }
};

namespace tag {
namespace property {
// tag_name: the block-name suffix of a tag in checkpoint files
struct position {
    static constexpr const char* tag_name = "position";
};
struct velocity {
    static constexpr const char* tag_name = "velocity";
};
struct weight {
    static constexpr const char* tag_name = "weight";
};
// the particle's own random stream state, a uint32_t (property.hpp:82)
struct random_seed {
    static constexpr const char* tag_name = "random_seed";
};
}  // namespace property

template <typename Tag, typename Type>
struct tag_bind {
    using tag = Tag;
    using type = Type;
};

namespace detail {
// the last binding has no tail member, so a record is exactly its bound
// types in order, with no empty sub-object padding it out
template <typename... B>
struct record;
template <>
struct record<> {};
template <typename B0>
struct record<B0> {
    typename B0::type head{};
};
template <typename B0, typename B1, typename... B>
struct record<B0, B1, B...> {
    typename B0::type head{};
    record<B1, B...> tail;
};

template <typename Tag, typename R>
struct finder;
template <typename Tag, typename B0>
struct finder<Tag, record<B0>> {
    static_assert(std::is_same_v<Tag, typename B0::tag>, "tag not bound in this record");
    __host__ __device__ static auto& get(record<B0>& r) { return r.head; }
    __host__ __device__ static const auto& get(const record<B0>& r) { return r.head; }
};
template <typename Tag, typename B0, typename B1, typename... B>
struct finder<Tag, record<B0, B1, B...>> {
    __host__ __device__ static auto& get(record<B0, B1, B...>& r)
    {
        if constexpr (std::is_same_v<Tag, typename B0::tag>) return r.head;
        else return finder<Tag, record<B1, B...>>::get(r.tail);
    }
    __host__ __device__ static const auto& get(const record<B0, B1, B...>& r)
    {
        if constexpr (std::is_same_v<Tag, typename B0::tag>) return r.head;
        else return finder<Tag, record<B1, B...>>::get(r.tail);
    }
};
}  // namespace detail

// A plain aggregate of the bound types in declaration order (the reference's
// std::tuple order, tag/tagged_struct.hpp:20-27), so device copies are
// bitwise copies.
template <typename... Binds>
struct tagged_struct : detail::record<Binds...> {};

template <typename Tag, typename... B>
__host__ __device__ auto& get(tagged_struct<B...>& s)
{
    return detail::finder<Tag, detail::record<B...>>::get(s);
}

template <typename Tag, typename... B>
__host__ __device__ const auto& get(const tagged_struct<B...>& s)
{
    return detail::finder<Tag, detail::record<B...>>::get(s);
}

}  // namespace tag
}  // namespace psum
