// Dirichlet node sets of the reference field solvers: fixed_boundary<D>
// (a set of nodes with a constant or position-dependent value) and
// fixed_boundary_list<D> (src/field_solver/fixed_boundary.hpp:18-210), with
// the boundary creators of src/field_solver/boundary_creator/Dirichlet.hpp
// (:58-137, :184-226).
#pragma once

#include <array>
#include <cmath>
#include <functional>
#include <initializer_list>
#include <map>
#include <optional>
#include <stdexcept>
#include <utility>
#include <vector>

#include "grid.hpp"

namespace psum {
namespace field {
enum class boundary_direction_2d { N, S, E, W, Count };
enum class boundary_direction_3d { Xpos, Xneg, Ypos, Yneg, Zpos, Zneg, Count };
}  // namespace field

namespace field_solver {

template <int Dim>
class fixed_boundary {
public:
    using Pos = std::array<double, Dim>;
    fixed_boundary() = default;
    void set_related_elements(const std::map<size_t, Pos>& e) { elements_ = e; }
    const std::map<size_t, Pos>& related_elements() const { return elements_; }
    fixed_boundary& operator=(double v)
    {
        value_ = v;
        func_ = nullptr;
        return *this;
    }
    fixed_boundary& operator=(const std::function<double(const Pos&)>& f)
    {
        func_ = f;
        value_.reset();
        return *this;
    }
    void erase_at(const std::vector<size_t>& idx)
    {
        for (size_t i : idx) elements_.erase(i);
    }
    // (node, value) pairs in node order; 0 when no value was assigned
    void evaluate(std::vector<std::pair<size_t, double>>& out) const
    {
        out.clear();
        for (const auto& e : elements_) out.emplace_back(e.first, func_ ? func_(e.second) : value_.value_or(0.0));
    }

private:
    std::map<size_t, Pos> elements_;
    std::optional<double> value_;
    std::function<double(const Pos&)> func_;
};

// A node claimed by several boundaries takes the value of the last one.
template <int Dim>
class fixed_boundary_list {
public:
    fixed_boundary_list() = default;
    fixed_boundary_list(std::initializer_list<fixed_boundary<Dim>> in) : b_(in) {}
    void push_back(const fixed_boundary<Dim>& b) { b_.push_back(b); }
    size_t size() const { return b_.size(); }
    fixed_boundary<Dim>& operator[](size_t i) { return b_[i]; }
    void evaluate(std::vector<unsigned long long>& idx, std::vector<double>& val) const
    {
        std::map<size_t, double> m;
        std::vector<std::pair<size_t, double>> one;
        for (const auto& b : b_) {
            b.evaluate(one);
            for (const auto& e : one) m[e.first] = e.second;
        }
        idx.clear();
        val.clear();
        for (const auto& e : m) {
            idx.push_back(e.first);
            val.push_back(e.second);
        }
    }
    std::vector<uint8_t> mask(size_t nodes) const
    {
        std::vector<uint8_t> f(nodes, 0);
        for (const auto& b : b_)
            for (const auto& e : b.related_elements()) f.at(e.first) = 1;
        return f;
    }

private:
    std::vector<fixed_boundary<Dim>> b_;
};

using fixed_boundary_2d = fixed_boundary<2>;
using fixed_boundary_3d = fixed_boundary<3>;
using fixed_boundary_2d_list = fixed_boundary_list<2>;
using fixed_boundary_3d_list = fixed_boundary_list<3>;

namespace boundary_creator {

namespace detail {
template <int Dim, typename Pred>
fixed_boundary<Dim> select(const field::simple_grid<Dim>& g, Pred&& keep)
{
    std::map<size_t, std::array<double, Dim>> e;
    const size_t n = g.contentSize(field::var_loc::nodeCentered);
    for (size_t i = 0; i < n; ++i) {
        const auto nd = g.i2n(i);
        const auto p = g.nodePosition(nd);
        if (keep(nd, p)) e[i] = p;
    }
    fixed_boundary<Dim> b;
    b.set_related_elements(e);
    return b;
}
}  // namespace detail

// The nodes of one edge of a 2D grid (S: first index of y, N: last, W:
// first of x, E: last), optionally only where the coordinate along the edge
// lies in [min, max] (Dirichlet.hpp:58-102).
inline fixed_boundary_2d Dirichlet_line(const field::grid2D& g, field::boundary_direction_2d dir,
                                        std::pair<double, double> min_max = {NAN, NAN})
{
    using D = field::boundary_direction_2d;
    const int I = g.numCells<0>(), J = g.numCells<1>();
    return detail::select<2>(g, [&](const std::array<int, 2>& n, const std::array<double, 2>& p) {
        bool on = (dir == D::S && n[1] == 0) || (dir == D::N && n[1] == J) || (dir == D::W && n[0] == 0) ||
                  (dir == D::E && n[0] == I);
        const double c = (dir == D::W || dir == D::E) ? p[1] : p[0];
        return on && !(c < min_max.first || c > min_max.second);
    });
}

// One face of a 3D grid (Dirichlet.hpp:137-182).
inline fixed_boundary_3d Dirichlet_plane(const field::grid3D& g, field::boundary_direction_3d dir)
{
    using D = field::boundary_direction_3d;
    const int n[3] = {g.numCells<0>(), g.numCells<1>(), g.numCells<2>()};
    return detail::select<3>(g, [&](const std::array<int, 3>& c, const std::array<double, 3>&) {
        switch (dir) {
        case D::Xneg: return c[0] == 0;
        case D::Xpos: return c[0] == n[0];
        case D::Yneg: return c[1] == 0;
        case D::Ypos: return c[1] == n[1];
        case D::Zneg: return c[2] == 0;
        case D::Zpos: return c[2] == n[2];
        default: return false;
        }
    });
}

// Nodes inside a closed box, with a geometric tolerance (Dirichlet.hpp:104-119, :184-202).
inline fixed_boundary_2d Dirichlet_box(const field::grid2D& g, double x0, double y0, double x1, double y1,
                                       double tol = 1e-10)
{
    return detail::select<2>(g, [&](const std::array<int, 2>&, const std::array<double, 2>& p) {
        return p[0] >= x0 - tol && p[0] <= x1 + tol && p[1] >= y0 - tol && p[1] <= y1 + tol;
    });
}
inline fixed_boundary_3d Dirichlet_box(const field::grid3D& g, double x0, double y0, double z0, double x1,
                                       double y1, double z1, double tol = 1e-10)
{
    return detail::select<3>(g, [&](const std::array<int, 3>&, const std::array<double, 3>& p) {
        return p[0] >= x0 - tol && p[0] <= x1 + tol && p[1] >= y0 - tol && p[1] <= y1 + tol &&
               p[2] >= z0 - tol && p[2] <= z1 + tol;
    });
}

// Nodes where func(position) holds (Dirichlet.hpp:121-135, :226-240).
inline fixed_boundary_2d Dirichlet_func(const field::grid2D& g, const std::function<bool(double, double)>& f)
{
    return detail::select<2>(g, [&](const std::array<int, 2>&, const std::array<double, 2>& p) { return f(p[0], p[1]); });
}
inline fixed_boundary_3d Dirichlet_func(const field::grid3D& g,
                                        const std::function<bool(double, double, double)>& f)
{
    return detail::select<3>(g, [&](const std::array<int, 3>&, const std::array<double, 3>& p) {
        return f(p[0], p[1], p[2]);
    });
}

// A single node (the pinned node of a periodic problem, TSI/main.cpp:105-113).
template <int Dim>
fixed_boundary<Dim> Dirichlet_node(const field::simple_grid<Dim>& g, size_t node)
{
    std::map<size_t, std::array<double, Dim>> e;
    e[node] = g.nodePosition(g.i2n(node));
    fixed_boundary<Dim> b;
    b.set_related_elements(e);
    return b;
}

}  // namespace boundary_creator
}  // namespace field_solver
}  // namespace psum
