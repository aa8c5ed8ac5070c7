// Particle boundaries of the reference's box geometries: the axis-aligned
// absorbing box that boundary_router + basic_material_set::absorb implement
// through a triangle mesh (boundary_router.hpp:30-54,
// basic_material_set.hpp:100-102), and the periodic wrap of the reference
// step loops (TSI/main.cpp:71-72).
#pragma once

#include <hip/hip_runtime.h>

#include "tag.hpp"

namespace psum {
namespace particle_boundary {

template <int D>
struct box_absorber {
    double lo[D], hi[D];
    // invalidates (Validator::make_invalid) a particle outside [lo, hi)
    template <typename Validator, typename Particle>
    __host__ __device__ void deal(Particle& p) const
    {
        const auto& x = tag::get<tag::property::position>(p);
        for (int d = 0; d < D; ++d)
            if (!(x[d] >= lo[d] && x[d] < hi[d])) {
                Validator::make_invalid(p);
                return;
            }
    }
};

template <int D>
struct periodic_wrap {
    double lo[D], len[D];
    template <typename Particle>
    __host__ __device__ void deal(Particle& p) const
    {
        auto& x = tag::get<tag::property::position>(p);
        for (int d = 0; d < D; ++d) {
            while (x[d] < lo[d]) x[d] += len[d];
            while (x[d] >= lo[d] + len[d]) x[d] -= len[d];
        }
    }
};

}  // namespace particle_boundary
}  // namespace psum
