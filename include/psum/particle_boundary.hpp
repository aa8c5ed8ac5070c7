// Triangle-mesh particle boundaries of the reference's particle_boundary
// module (include/psum/particle_boundary.hpp): triangle data and meshes with
// ids (geometry.hpp:18-30, triangle_mesh_trigger.hpp:18-28), the mesh
// generators (mesh_generator.hpp:16-337), the STL reader (stl_loader.hpp:
// 20-270), triangle_mesh_trigger with its device accessor
// (triangle_mesh_trigger.hpp:31-420), basic_material_set
// (basic_material_set.hpp:15-120) and boundary_router
// (boundary_router.hpp:20-90).
//
// The partition is built by psum_mesh_create (the same build as the
// psum_hip_mesh_* kernels use); the accessors run the device code of
// psum_mesh_geom.hpp, so a router called from a for_each kernel gives the
// results psum_hip_mesh_boundary gives for the same particles.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../psum_hip.h"
#include "../psum_mesh_geom.hpp"
#include "device_vector.hpp"
#include "tag.hpp"

namespace psum {
namespace particle_boundary {

namespace geometry {

constexpr double finite_small = psum_hip::mesh::kSmall;

// three vertices and the unit normal (filled in by the trigger's build)
struct triangle_data {
    double p1x, p1y, p1z;
    double p2x, p2y, p2z;
    double p3x, p3y, p3z;
    double nx = 0, ny = 0, nz = 0;
};

}  // namespace geometry

template <typename MeshId>
struct triangle_mesh_with_id {
    std::vector<geometry::triangle_data> triangles_;
    MeshId id_;
};

template <typename MeshId>
struct triangle_with_id {
    geometry::triangle_data data_;
    MeshId id_;
};

// ---------------------------------------------------------------- generators

namespace mesh_generator {

using geometry::triangle_data;

inline triangle_data make_triangle(const double* a, const double* b, const double* c)
{
    return triangle_data{a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
}

// 12 triangles, two per face: z0, z1, y0, y1, x0, x1
inline std::vector<triangle_data> axis_aligned_cube(double x0, double y0, double z0, double x1, double y1, double z1)
{
    const double v[8][3] = {{x0, y0, z0}, {x1, y0, z0}, {x0, y1, z0}, {x1, y1, z0},
                            {x0, y0, z1}, {x1, y0, z1}, {x0, y1, z1}, {x1, y1, z1}};
    const int faces[6][4] = {{0, 1, 3, 2}, {4, 5, 7, 6}, {0, 1, 5, 4}, {2, 3, 7, 6}, {0, 2, 6, 4}, {1, 3, 7, 5}};
    std::vector<triangle_data> out;
    for (const auto& f : faces) {
        out.push_back(make_triangle(v[f[0]], v[f[1]], v[f[2]]));
        out.push_back(make_triangle(v[f[0]], v[f[2]], v[f[3]]));
    }
    return out;
}

// a polyline in the xy plane extruded over z in [-w/2, w/2]: four
// triangles per segment
inline std::vector<triangle_data> extrude_xy_line_to_width(const std::vector<std::pair<double, double>>& xy_points,
                                                          double z_width)
{
    std::vector<triangle_data> out;
    const double za = -z_width / 2.0, zb = z_width / 2.0;
    for (size_t i = 0; i + 1 < xy_points.size(); ++i) {
        const double a0[3] = {xy_points[i].first, xy_points[i].second, za};
        const double a1[3] = {xy_points[i].first, xy_points[i].second, zb};
        const double b0[3] = {xy_points[i + 1].first, xy_points[i + 1].second, za};
        const double b1[3] = {xy_points[i + 1].first, xy_points[i + 1].second, zb};
        out.push_back(make_triangle(a0, b0, b1));
        out.push_back(make_triangle(a0, b1, a1));
        out.push_back(make_triangle(a0, b0, a1));
        out.push_back(make_triangle(b0, b1, a1));
    }
    return out;
}

// two triangles over the rectangle of the one fixed coordinate; none when
// no coordinate is fixed
inline std::vector<triangle_data> axis_aligned_plane(double x0, double y0, double z0, double x1, double y1, double z1)
{
    double v[4][3];
    if (x0 == x1) {
        const double q[4][3] = {{x0, y0, z0}, {x0, y1, z0}, {x0, y1, z1}, {x0, y0, z1}};
        std::memcpy(v, q, sizeof(v));
    } else if (y0 == y1) {
        const double q[4][3] = {{x0, y0, z0}, {x1, y0, z0}, {x1, y0, z1}, {x0, y0, z1}};
        std::memcpy(v, q, sizeof(v));
    } else if (z0 == z1) {
        const double q[4][3] = {{x0, y0, z0}, {x1, y0, z0}, {x1, y1, z0}, {x0, y1, z0}};
        std::memcpy(v, q, sizeof(v));
    } else {
        return {};
    }
    return {make_triangle(v[0], v[1], v[2]), make_triangle(v[0], v[2], v[3])};
}

// surface of revolution of a (z, r) polyline about the z axis
inline std::vector<triangle_data> revolve_zr_curve(const std::vector<std::pair<double, double>>& zr_points,
                                                  int azimuthal_segments)
{
    std::vector<triangle_data> out;
    if (zr_points.size() < 2 || azimuthal_segments < 1) return out;
    const double dth = 2.0 * 3.14159265358979323846 / azimuthal_segments;
    for (int i = 0; i < azimuthal_segments; ++i) {
        const double c0 = std::cos(i * dth), s0 = std::sin(i * dth);
        const double c1 = std::cos((i + 1) * dth), s1 = std::sin((i + 1) * dth);
        for (size_t j = 0; j + 1 < zr_points.size(); ++j) {
            const double z0 = zr_points[j].first, r0 = zr_points[j].second;
            const double z1 = zr_points[j + 1].first, r1 = zr_points[j + 1].second;
            const double p00[3] = {r0 * c0, r0 * s0, z0}, p01[3] = {r0 * c1, r0 * s1, z0};
            const double p10[3] = {r1 * c0, r1 * s0, z1}, p11[3] = {r1 * c1, r1 * s1, z1};
            out.push_back(make_triangle(p00, p10, p11));
            out.push_back(make_triangle(p00, p11, p01));
        }
    }
    return out;
}

// scale about the mean of the triangle centroids
inline void rescale_triangles(std::vector<triangle_data>& triangles, double scale)
{
    if (triangles.empty()) return;
    double c[3] = {0, 0, 0};
    for (const auto& t : triangles) {
        c[0] += (t.p1x + t.p2x + t.p3x) / 3.0;
        c[1] += (t.p1y + t.p2y + t.p3y) / 3.0;
        c[2] += (t.p1z + t.p2z + t.p3z) / 3.0;
    }
    for (double& x : c) x /= (double)triangles.size();
    for (auto& t : triangles) {
        double* p = &t.p1x;
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) p[3 * v + k] = (p[3 * v + k] - c[k]) * scale + c[k];
    }
}

}  // namespace mesh_generator

// ---------------------------------------------------------------- STL

namespace stl_loader {

using geometry::triangle_data;

// "solid" at the start and a facet or vertex keyword in the first 1000 bytes
inline bool stl_file_is_ascii(const std::string& filename)
{
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("stl_loader: cannot open " + filename);
    std::string head(1000, '\0');
    f.read(&head[0], 1000);
    head.resize((size_t)f.gcount());
    return head.compare(0, 5, "solid") == 0 &&
           (head.find("facet") != std::string::npos || head.find("vertex") != std::string::npos);
}

inline std::vector<triangle_data> load_stl_file_ascii(const std::string& filename)
{
    std::ifstream f(filename);
    if (!f) throw std::runtime_error("stl_loader: cannot open " + filename);
    std::vector<double> v;
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string key;
        double x, y, z;
        if (is >> key && key == "vertex" && (is >> x >> y >> z)) {
            v.push_back(x);
            v.push_back(y);
            v.push_back(z);
        }
    }
    if (v.size() % 9) throw std::runtime_error("stl_loader: " + filename + ": vertex count not a multiple of 3");
    std::vector<triangle_data> out(v.size() / 9);
    for (size_t i = 0; i < out.size(); ++i) out[i] = mesh_generator::make_triangle(&v[9 * i], &v[9 * i + 3], &v[9 * i + 6]);
    return out;
}

// 80-byte header, uint32 count, then 50-byte records: normal, 3 vertices
// (float32, little endian), attribute word; the file's normals are unused
inline std::vector<triangle_data> load_stl_file_binary(const std::string& filename)
{
    std::ifstream f(filename, std::ios::binary);
    if (!f) throw std::runtime_error("stl_loader: cannot open " + filename);
    char head[84];
    if (!f.read(head, 84)) throw std::runtime_error("stl_loader: " + filename + ": too short for a binary STL");
    uint32_t n;
    std::memcpy(&n, head + 80, 4);
    f.seekg(0, std::ios::end);
    const std::streamoff bytes = f.tellg();
    if (bytes < 84 + 50 * (std::streamoff)n)
        throw std::runtime_error("stl_loader: " + filename + ": truncated (" + std::to_string(n) + " triangles declared)");
    f.seekg(84, std::ios::beg);
    std::vector<triangle_data> out(n);
    char rec[50];
    for (uint32_t i = 0; i < n; ++i) {
        if (!f.read(rec, 50)) throw std::runtime_error("stl_loader: " + filename + ": truncated");
        float q[9];
        std::memcpy(q, rec + 12, sizeof(q));
        double d[9];
        for (int k = 0; k < 9; ++k) d[k] = q[k];
        out[i] = mesh_generator::make_triangle(d, d + 3, d + 6);
    }
    return out;
}

inline std::vector<triangle_data> load(const std::string& filename)
{
    return stl_file_is_ascii(filename) ? load_stl_file_ascii(filename) : load_stl_file_binary(filename);
}

// (min corner, max corner) of the vertices
inline std::pair<std::array<double, 3>, std::array<double, 3>> box_of_triangles(
    const std::vector<triangle_data>& triangles)
{
    std::array<double, 3> lo{std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                             std::numeric_limits<double>::infinity()};
    std::array<double, 3> hi{-lo[0], -lo[1], -lo[2]};
    for (const auto& t : triangles) {
        const double* p = &t.p1x;
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::fmin(lo[k], p[3 * v + k]);
                hi[k] = std::fmax(hi[k], p[3 * v + k]);
            }
    }
    return {lo, hi};
}

}  // namespace stl_loader

// ---------------------------------------------------------------- trigger

template <typename MeshId>
class triangle_mesh_trigger_acc {
public:
    explicit triangle_mesh_trigger_acc(const psum_hip::mesh::View& v) : v_(v) {}

    // first hit of the segment p1 -> p2: segment parameter k, the
    // triangle's unit normal and its mesh id
    __device__ bool detect(double p1x, double p1y, double p1z, double p2x, double p2y, double p2z, double& k,
                           double& nx, double& ny, double& nz, MeshId& id) const
    {
        const double p1[3] = {p1x, p1y, p1z}, p2[3] = {p2x, p2y, p2z};
        double n[3];
        int32_t mid;
        if (!psum_hip::mesh::detect(v_, p1, p2, k, n, mid)) return false;
        nx = n[0];
        ny = n[1];
        nz = n[2];
        id = static_cast<MeshId>(mid);
        return true;
    }

private:
    psum_hip::mesh::View v_;
};

template <typename MeshId>
class triangle_mesh_trigger {
public:
    using triangle_data = geometry::triangle_data;
    using acc_type = triangle_mesh_trigger_acc<MeshId>;

    // the partition is uploaded synchronously, so no stream is kept
    explicit triangle_mesh_trigger(hipStream_t = nullptr) {}
    triangle_mesh_trigger(const triangle_mesh_trigger&) = delete;
    triangle_mesh_trigger& operator=(const triangle_mesh_trigger&) = delete;
    ~triangle_mesh_trigger()
    {
        if (h_) psum_mesh_destroy(h_);
    }

    // partition of the box (x0, y0, z0)-(x1, y1, z1) into I x J x K cells;
    // zero-area triangles are dropped
    void set(const std::vector<triangle_with_id<MeshId>>& triangles, double x0, double y0, double z0, double x1,
             double y1, double z1, int I, int J, int K)
    {
        const double box[6] = {x0, y0, z0, x1, y1, z1};
        create_(triangles, box, I, J, K);
    }

    // partition of the meshes' bounding box widened by about 10% a side
    void set(const std::vector<triangle_mesh_with_id<MeshId>>& meshes, int I, int J, int K)
    {
        std::vector<triangle_with_id<MeshId>> tri;
        for (const auto& m : meshes)
            for (const auto& t : m.triangles_) tri.push_back({t, m.id_});
        if (tri.empty()) throw std::runtime_error("triangle_mesh_trigger::set: no triangles");
        create_(tri, nullptr, I, J, K);
    }

    // traversal: the end cells of a step only (fast, the default) or every
    // cell the step crosses (restrict); the C ABI handle follows
    void set_fast() { mode_(false); }
    void set_restrict() { mode_(true); }
    // the distance-field skip of the full traversal (psum_mesh_set_udf)
    void set_acceleration_UDF(bool enabled)
    {
        if (!h_ && !enabled) return;
        if (!h_) throw std::runtime_error("triangle_mesh_trigger::set_acceleration_UDF called before set");
        if (psum_mesh_set_udf(h_, enabled ? 1 : 0) != PSUM_OK) throw std::runtime_error(psum_hip_last_error());
    }

    acc_type get_access(handler&) const
    {
        if (!h_) throw std::runtime_error("triangle_mesh_trigger::get_access called before set");
        psum_mesh_view v;
        if (psum_mesh_get_view(h_, &v) != PSUM_OK) throw std::runtime_error(psum_hip_last_error());
        return acc_type(psum_hip::mesh::view_of(v));
    }

    // the C ABI handle (psum_hip_mesh_detect / psum_hip_mesh_boundary)
    void* handle() const { return h_; }
    bool restricted() const { return restrict_; }

    int get_I() const { return I_; }
    int get_J() const { return J_; }
    int get_K() const { return K_; }
    double get_x0() const { return box_[0]; }
    double get_y0() const { return box_[1]; }
    double get_z0() const { return box_[2]; }
    double get_x1() const { return box_[3]; }
    double get_y1() const { return box_[4]; }
    double get_z1() const { return box_[5]; }
    double span_x() const { return box_[3] - box_[0]; }
    double span_y() const { return box_[4] - box_[1]; }
    double span_z() const { return box_[5] - box_[2]; }

private:
    void create_(const std::vector<triangle_with_id<MeshId>>& tri, const double* box, int I, int J, int K)
    {
        if (h_) throw std::runtime_error("triangle_mesh_trigger::set called twice");
        std::vector<double> t9(tri.size() * 9);
        std::vector<int32_t> ids(tri.size());
        for (size_t i = 0; i < tri.size(); ++i) {
            std::memcpy(&t9[9 * i], &tri[i].data_.p1x, 9 * sizeof(double));
            ids[i] = static_cast<int32_t>(tri[i].id_);
        }
        h_ = psum_mesh_create(t9.data(), ids.data(), (int64_t)tri.size(), I, J, K, box);
        if (!h_) throw std::runtime_error(psum_hip_last_error());
        I_ = I;
        J_ = J;
        K_ = K;
        int64_t info[4];
        if (psum_mesh_info(h_, info, box_) != PSUM_OK) throw std::runtime_error(psum_hip_last_error());
        mode_(restrict_);
    }

    void mode_(bool restrict_mode)
    {
        restrict_ = restrict_mode;
        if (h_ && psum_mesh_set_fast(h_, restrict_ ? 0 : 1) != PSUM_OK) throw std::runtime_error(psum_hip_last_error());
    }

    void* h_ = nullptr;
    bool restrict_ = false;
    int I_ = 0, J_ = 0, K_ = 0;
    double box_[6] = {0, 0, 0, 0, 0, 0};
};

// ---------------------------------------------------------------- materials

// ids are materials; the particle carries position, velocity (vec<3>) and
// random_seed (uint32_t)
template <typename ParticleType, template <typename> typename Validator>
class basic_material_set {
public:
    enum class material_type : int32_t {
        diffuse = PSUM_MAT_DIFFUSE,
        specular = PSUM_MAT_SPECULAR,
        maxwell_reflect = PSUM_MAT_MAXWELL_REFLECT,
        absorb = PSUM_MAT_ABSORB
    };

    class acc_type {
    public:
        acc_type(double default_temperature, double default_mass)
            : temperature_(default_temperature), mass_(default_mass)
        {
        }

        // the material's action on a particle whose step crossed a triangle
        // at the hit point with unit normal n
        __device__ void action(material_type mat, ParticleType& p, double hit_x, double hit_y, double hit_z,
                               double nx, double ny, double nz) const
        {
            if (mat == material_type::absorb) {
                Validator<ParticleType>::make_invalid(p);
                return;
            }
            auto& x = tag::get<tag::property::position>(p);
            auto& v = tag::get<tag::property::velocity>(p);
            auto& seed = tag::get<tag::property::random_seed>(p);
            double pos[3], vel[3];
            for (int c = 0; c < 3; ++c) {
                pos[c] = (double)x[c];
                vel[c] = (double)v[c];
            }
            const double hit[3] = {hit_x, hit_y, hit_z}, n[3] = {nx, ny, nz};
            uint32_t s = (uint32_t)seed;
            psum_hip::mesh::material_action((int)mat, pos, vel, s, hit, n, temperature_, mass_);
            for (int c = 0; c < 3; ++c) {
                x[c] = pos[c];
                v[c] = vel[c];
            }
            seed = s;
        }

    private:
        double temperature_, mass_;
    };

    basic_material_set() = default;

    double get_default_temperature() const { return default_temperature_; }
    double get_default_mass() const { return default_mass_; }
    void set_default_temperature(double t) { default_temperature_ = t; }
    void set_default_mass(double m) { default_mass_ = m; }

    acc_type get_access(handler&) const { return acc_type(default_temperature_, default_mass_); }

private:
    double default_temperature_ = 273.15;
    double default_mass_ = 9.11e-31;
};

// ---------------------------------------------------------------- router

template <typename MaterialSet, typename ParticleType>
class boundary_router_acc {
    using material_type = typename MaterialSet::material_type;
    using trigger_type = triangle_mesh_trigger<material_type>;

public:
    boundary_router_acc(const typename trigger_type::acc_type& trigger, const typename MaterialSet::acc_type& mats)
        : trigger_(trigger), mats_(mats)
    {
    }

    // the step p1 -> p2 of particle p (now at p2): on a hit, the hit
    // triangle's material acts on p; returns whether it hit
    __device__ bool deal(double p1x, double p1y, double p1z, double p2x, double p2y, double p2z,
                         ParticleType& p) const
    {
        double k, nx, ny, nz;
        material_type mat;
        if (!trigger_.detect(p1x, p1y, p1z, p2x, p2y, p2z, k, nx, ny, nz, mat)) return false;
        mats_.action(mat, p, p1x + k * (p2x - p1x), p1y + k * (p2y - p1y), p1z + k * (p2z - p1z), nx, ny, nz);
        return true;
    }

private:
    typename trigger_type::acc_type trigger_;
    typename MaterialSet::acc_type mats_;
};

template <typename MaterialSet, typename ParticleType>
class boundary_router {
public:
    using material_type = typename MaterialSet::material_type;
    using trigger_type = triangle_mesh_trigger<material_type>;
    using acc_type = boundary_router_acc<MaterialSet, ParticleType>;

    explicit boundary_router(hipStream_t s = nullptr) : trigger_(s) {}

    // the meshes partitioned into I x J x K cells; full traversal of each step
    void set(const std::vector<triangle_mesh_with_id<material_type>>& meshes, int I, int J, int K)
    {
        trigger_.set(meshes, I, J, K);
        trigger_.set_restrict();
    }

    trigger_type& get_trigger() { return trigger_; }
    MaterialSet& get_material_set() { return material_set_; }

    acc_type get_access(handler& h) const { return acc_type(trigger_.get_access(h), material_set_.get_access(h)); }

private:
    trigger_type trigger_;
    MaterialSet material_set_;
};

}  // namespace particle_boundary
}  // namespace psum
