// Boundary node sets of the reference field solvers: fixed_boundary<D>
// (Dirichlet: a set of nodes with a constant or position-dependent value)
// and fixed_boundary_list<D> (src/field_solver/fixed_boundary.hpp:18-210),
// mixed_boundary<D> (Robin/Neumann: value_coef * phi + grad_coef * dphi/dn =
// value on one side of each node) and mixed_boundary_list<D>
// (mixed_boundary.hpp:17-262), with the boundary creators of
// src/field_solver/boundary_creator/Dirichlet.hpp (:58-137, :184-226),
// Neumann.hpp (:46-132) and Robin.hpp (:56-168).
#pragma once

#include <array>
#include <cmath>
#include <functional>
#include <initializer_list>
#include <map>
#include <optional>
#include <set>
#include <type_traits>
#include <stdexcept>
#include <utility>
#include <vector>

#include "grid.hpp"

namespace psum {
namespace field {
enum class boundary_direction_1d { L, R, Count };
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

using fixed_boundary_1d = fixed_boundary<1>;
using fixed_boundary_1d_list = fixed_boundary_list<1>;
using fixed_boundary_2d = fixed_boundary<2>;
using fixed_boundary_3d = fixed_boundary<3>;
using fixed_boundary_2d_list = fixed_boundary_list<2>;
using fixed_boundary_3d_list = fixed_boundary_list<3>;

// One mixed boundary: nodes (with their positions), one side, a value (constant
// or a function of position) and the (value_coef, grad_coef) pair as a function
// of position.  Nodes are kept in index order.
template <int Dim>
class mixed_boundary {
public:
    using Pos = std::array<double, Dim>;
    using direction_type =
        std::conditional_t<Dim == 1, field::boundary_direction_1d,
                           std::conditional_t<Dim == 2, field::boundary_direction_2d, field::boundary_direction_3d>>;
    using Coeffs = std::pair<double, double>;
    struct mixed_boundary_node {
        size_t element;
        direction_type direction;
        Coeffs coeffs;
    };

    mixed_boundary() = default;
    void set_related_elements(const std::map<size_t, Pos>& e) { elements_ = e; }
    const std::map<size_t, Pos>& related_elements() const { return elements_; }
    mixed_boundary& operator=(double v)
    {
        value_ = v;
        func_ = nullptr;
        return *this;
    }
    mixed_boundary& operator=(const std::function<double(const Pos&)>& f)
    {
        func_ = f;
        value_.reset();
        return *this;
    }
    void set_coeff_func(const std::function<Coeffs(const Pos&)>& f) { coeff_ = f; }
    void set_direction(direction_type d) { dir_ = d; }
    direction_type get_direction() const
    {
        if (!dir_) throw std::runtime_error("mixed_boundary: direction is not set");
        return *dir_;
    }
    void erase_at(const std::vector<size_t>& idx)
    {
        for (size_t i : idx) elements_.erase(i);
    }
    std::vector<size_t> get_element_indexes() const
    {
        std::vector<size_t> out;
        for (const auto& e : elements_) out.push_back(e.first);
        return out;
    }
    // values in node order; a boundary without a value is an error
    void evaluate(std::vector<double>& out) const
    {
        if (!func_ && !value_) throw std::runtime_error("mixed_boundary::evaluate: no value is set");
        out.clear();
        for (const auto& e : elements_) out.push_back(func_ ? func_(e.second) : *value_);
    }
    Coeffs evaluate_coeffs_at(size_t idx, direction_type d) const
    {
        if (!coeff_) throw std::runtime_error("mixed_boundary: coefficient function is not set");
        auto it = elements_.find(idx);
        if (it == elements_.end()) throw std::runtime_error("mixed_boundary: element is not on this boundary");
        if (get_direction() != d) throw std::runtime_error("mixed_boundary: direction differs");
        return coeff_(it->second);
    }
    std::vector<mixed_boundary_node> get_boundary_nodes() const
    {
        std::vector<mixed_boundary_node> out;
        for (const auto& e : elements_)
            out.push_back({e.first, get_direction(), coeff_ ? coeff_(e.second) : Coeffs{0.0, 0.0}});
        return out;
    }

private:
    std::map<size_t, Pos> elements_;
    std::optional<double> value_;
    std::function<double(const Pos&)> func_;
    std::function<Coeffs(const Pos&)> coeff_;
    std::optional<direction_type> dir_;
};

// A (node, side) pair claimed by several boundaries belongs to the last one.
template <int Dim>
class mixed_boundary_list {
public:
    using direction_type = typename mixed_boundary<Dim>::direction_type;
    using mixed_boundary_node = typename mixed_boundary<Dim>::mixed_boundary_node;
    mixed_boundary_list() = default;
    mixed_boundary_list(std::initializer_list<mixed_boundary<Dim>> in) : b_(in)
    {
        std::set<std::pair<size_t, int>> taken;
        for (size_t k = b_.size(); k-- > 0;) {
            std::vector<size_t> dup;
            const int d = (int)b_[k].get_direction();
            for (size_t e : b_[k].get_element_indexes())
                if (!taken.insert({e, d}).second) dup.push_back(e);
            b_[k].erase_at(dup);
        }
    }
    size_t size() const { return b_.size(); }
    void erase_at(const std::vector<size_t>& idx)
    {
        for (auto& b : b_) b.erase_at(idx);
    }
    void evaluate(std::vector<double>& out) const
    {
        out.clear();
        std::vector<double> one;
        for (const auto& b : b_) {
            b.evaluate(one);
            out.insert(out.end(), one.begin(), one.end());
        }
    }
    size_t elements_size() const
    {
        size_t n = 0;
        for (const auto& b : b_) n += b.related_elements().size();
        return n;
    }
    bool element_exists(size_t idx, direction_type d) const
    {
        for (const auto& b : b_)
            if (b.get_direction() == d && b.related_elements().count(idx)) return true;
        return false;
    }
    // (value_coef, grad_coef) of node idx on side d (mixed_boundary.hpp:
    // 219-229)
    typename mixed_boundary<Dim>::Coeffs evaluate_coeffs_at(size_t idx, direction_type d) const
    {
        for (const auto& b : b_)
            if (b.get_direction() == d && b.related_elements().count(idx)) return b.evaluate_coeffs_at(idx, d);
        throw std::runtime_error("mixed_boundary_list: no condition at this node and side");
    }
    bool element_exists(size_t idx) const
    {
        for (const auto& b : b_)
            if (b.related_elements().count(idx)) return true;
        return false;
    }
    std::vector<mixed_boundary_node> get_boundary_nodes() const
    {
        std::vector<mixed_boundary_node> out;
        for (const auto& b : b_) {
            const auto one = b.get_boundary_nodes();
            out.insert(out.end(), one.begin(), one.end());
        }
        return out;
    }

private:
    std::vector<mixed_boundary<Dim>> b_;
};

using mixed_boundary_1d = mixed_boundary<1>;
using mixed_boundary_1d_list = mixed_boundary_list<1>;
using mixed_boundary_2d = mixed_boundary<2>;
using mixed_boundary_3d = mixed_boundary<3>;
using mixed_boundary_2d_list = mixed_boundary_list<2>;
using mixed_boundary_3d_list = mixed_boundary_list<3>;

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

// The end node of a 1D grid (L: node 0, R: the last), or the nodes within
// [x0, x1] widened by tolerance * span (Dirichlet.hpp:24-56)
inline fixed_boundary_1d Dirichlet_point(const field::grid1D& g, field::boundary_direction_1d dir)
{
    const int last = g.numCells<0>();
    return detail::select<1>(g, [&](const std::array<int, 1>& n, const std::array<double, 1>&) {
        return n[0] == (dir == field::boundary_direction_1d::L ? 0 : last);
    });
}
inline fixed_boundary_1d Dirichlet_range(const field::grid1D& g, double x0, double x1, double geo_tolerance = 1e-10)
{
    const double tol = geo_tolerance * g.span<0>();
    return detail::select<1>(g, [&](const std::array<int, 1>&, const std::array<double, 1>& p) {
        return p[0] >= x0 - tol && p[0] <= x1 + tol;
    });
}

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


namespace detail {
template <int Dim>
bool on_side(const std::array<int, Dim>& n, const int* cells, int axis, bool plus)
{
    return n[axis] == (plus ? cells[axis] : 0);
}
inline void side2(field::boundary_direction_2d d, int& axis, bool& plus)
{
    using D = field::boundary_direction_2d;
    axis = (d == D::N || d == D::S) ? 1 : 0;
    plus = d == D::N || d == D::E;
}
inline void side3(field::boundary_direction_3d d, int& axis, bool& plus)
{
    axis = (int)d / 2;
    plus = ((int)d % 2) == 0;
}
template <int Dim, typename Dir>
mixed_boundary<Dim> mixed_side(const field::simple_grid<Dim>& g, Dir dir, std::pair<double, double> min_max,
                               const std::function<std::pair<double, double>(const std::array<double, Dim>&)>& coeffs)
{
    int axis;
    bool plus;
    if constexpr (Dim == 2) side2(dir, axis, plus);
    else side3(dir, axis, plus);
    int cells[Dim];
    for (int d = 0; d < Dim; ++d) cells[d] = g.to_psum().cells[d];
    std::map<size_t, std::array<double, Dim>> e;
    const size_t n = g.contentSize(field::var_loc::nodeCentered);
    for (size_t i = 0; i < n; ++i) {
        const auto nd = g.i2n(i);
        if (!on_side<Dim>(nd, cells, axis, plus)) continue;
        const auto p = g.nodePosition(nd);
        // one coordinate along the side may be limited to [min, max]: the
        // other axis of a 2D edge; y for an X face, x for a Y or Z face
        const double c = Dim == 2 ? p[1 - axis] : p[axis == 0 ? 1 : 0];
        if (c < min_max.first || c > min_max.second) continue;
        e[i] = p;
    }
    mixed_boundary<Dim> b;
    b.set_related_elements(e);
    b.set_direction(dir);
    b.set_coeff_func(coeffs);
    return b;
}
}  // namespace detail

// Neumann sides: (value_coef, grad_coef) = (0, 1) (Neumann.hpp:46-132)
inline mixed_boundary_2d Neumann_line(const field::grid2D& g, field::boundary_direction_2d dir,
                                      std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<2>(g, dir, min_max, [](const std::array<double, 2>&) { return std::make_pair(0.0, 1.0); });
}
inline mixed_boundary_3d Neumann_plane(const field::grid3D& g, field::boundary_direction_3d dir,
                                       std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<3>(g, dir, min_max, [](const std::array<double, 3>&) { return std::make_pair(0.0, 1.0); });
}

// mixed condition at an end node of a 1D grid (Neumann.hpp:24-44,
// Robin.hpp:24-54)
namespace detail {
inline mixed_boundary_1d mixed_point(const field::grid1D& g, field::boundary_direction_1d dir,
                                     const std::function<std::pair<double, double>(const std::array<double, 1>&)>& c)
{
    const size_t idx = dir == field::boundary_direction_1d::L ? 0 : (size_t)g.numCells<0>();
    std::map<size_t, std::array<double, 1>> e;
    e[idx] = g.nodePosition(g.i2n(idx));
    mixed_boundary_1d b;
    b.set_related_elements(e);
    b.set_coeff_func(c);
    b.set_direction(dir);
    return b;
}
}  // namespace detail
inline mixed_boundary_1d Neumann_point(const field::grid1D& g, field::boundary_direction_1d dir)
{
    return detail::mixed_point(g, dir, [](const std::array<double, 1>&) { return std::make_pair(0.0, 1.0); });
}
inline mixed_boundary_1d Robin_point(const field::grid1D& g, field::boundary_direction_1d dir,
                                     const std::function<std::pair<double, double>(double)>& coef)
{
    return detail::mixed_point(g, dir, [coef](const std::array<double, 1>& p) { return coef(p[0]); });
}
inline mixed_boundary_1d Robin_point(const field::grid1D& g, field::boundary_direction_1d dir,
                                     std::pair<double, double> coef)
{
    return detail::mixed_point(g, dir, [coef](const std::array<double, 1>&) { return coef; });
}

// Robin sides with (value_coef, grad_coef) constant or a function of position
// (Robin.hpp:56-168)
inline mixed_boundary_2d Robin_Line(const field::grid2D& g, field::boundary_direction_2d dir,
                                    const std::function<std::pair<double, double>(double, double)>& coef,
                                    std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<2>(g, dir, min_max, [coef](const std::array<double, 2>& p) { return coef(p[0], p[1]); });
}
inline mixed_boundary_2d Robin_Line(const field::grid2D& g, field::boundary_direction_2d dir,
                                    std::pair<double, double> coef, std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<2>(g, dir, min_max, [coef](const std::array<double, 2>&) { return coef; });
}
inline mixed_boundary_3d Robin_plane(const field::grid3D& g, field::boundary_direction_3d dir,
                                     const std::function<std::pair<double, double>(double, double, double)>& coef,
                                     std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<3>(g, dir, min_max,
                                 [coef](const std::array<double, 3>& p) { return coef(p[0], p[1], p[2]); });
}
inline mixed_boundary_3d Robin_plane(const field::grid3D& g, field::boundary_direction_3d dir,
                                     std::pair<double, double> coef, std::pair<double, double> min_max = {NAN, NAN})
{
    return detail::mixed_side<3>(g, dir, min_max, [coef](const std::array<double, 3>&) { return coef; });
}

}  // namespace boundary_creator
}  // namespace field_solver
}  // namespace psum
