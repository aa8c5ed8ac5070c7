// The reference's particle / field API on HIP (names and argument meaning of
// PSuM's psum.hpp prelude), compiled by hipcc.  The user step loop of the
// reference ports by replacing sycl::handler with psum::handler and
// Eigen::Vector<double, D> with psum::vec<D>; the hot loops can also go
// through the fused psum_hip_* kernels (psum_hip.h) via to_soa / from_soa.
#pragma once

#include "../psum_hip.h"
#include "boundary.hpp"
#include "boundary_creator.hpp"
#include "device_field.hpp"
#include "device_vector.hpp"
#include "grid.hpp"
#include "host_field.hpp"
#include "particle_boundary.hpp"
#include "particle_group.hpp"
#include "poisson_solver.hpp"
#include "poisson_solver_1d.hpp"
#include "serialization.hpp"
#include "tag.hpp"

namespace psum {
namespace prelude {
using psum::handler;
using psum::vec;
using psum::field::add_back;
using psum::field::add_back_B_spline;
using psum::field::interp_B_spline;
using psum::field::interp_diff_B_spline;
using psum::field::cell_field1D;
using psum::field::cell_field2D;
using psum::field::cell_field3D;
using psum::field::device_field;
using psum::field::grid1D;
using psum::field::host_cell_field1D;
using psum::field::host_cell_field2D;
using psum::field::host_cell_field3D;
using psum::field::host_field;
using psum::field::host_node_field1D;
using psum::field::host_node_field2D;
using psum::field::host_node_field3D;
using psum::field::grid2D;
using psum::field::grid3D;
using psum::field::interp;
using psum::field::interp_diff;
using psum::field::node_field1D;
using psum::field::node_field2D;
using psum::field::node_field3D;
using psum::field::simple_grid;
using psum::field::var_loc;
using psum::particle_boundary::basic_material_set;
using psum::particle_boundary::boundary_router;
using psum::particle_boundary::box_absorber;
using psum::particle_boundary::triangle_mesh_trigger;
using psum::particle_boundary::triangle_mesh_with_id;
namespace geometry = psum::particle_boundary::geometry;
namespace mesh_generator = psum::particle_boundary::mesh_generator;
namespace stl_loader = psum::particle_boundary::stl_loader;
using psum::particle_boundary::periodic_wrap;
using psum::particle_container::device_vector;
using psum::particle_container::particle_group;
using psum::particle_container::pos_x_nan_is_invalid;
using psum::tag::get;
using psum::tag::tag_bind;
using psum::tag::tagged_struct;
using psum::field::boundary_direction_1d;
using psum::field::boundary_direction_2d;
using psum::field::boundary_direction_3d;
using psum::field_solver::fixed_boundary_1d_list;
using psum::field_solver::fixed_boundary_2d_list;
using psum::field_solver::fixed_boundary_3d_list;
using psum::field_solver::mixed_boundary_1d;
using psum::field_solver::mixed_boundary_1d_list;
using psum::field_solver::mixed_boundary_2d;
using psum::field_solver::mixed_boundary_2d_list;
using psum::field_solver::mixed_boundary_3d;
using psum::field_solver::mixed_boundary_3d_list;
using psum::field_solver::Poisson_solver_1d;
using psum::field_solver::Poisson_solver_2d;
using psum::field_solver::Poisson_solver_3d;
namespace boundary_creator = psum::field_solver::boundary_creator;
using psum::serialization::mas_file;
namespace property = psum::tag::property;
}  // namespace prelude
}  // namespace psum
