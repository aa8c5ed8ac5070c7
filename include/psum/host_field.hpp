// host_field<D, Loc, Scalar>: a node- or cell-centred field on a simple_grid
// in host memory (src/field/host_field.hpp:36-235), with for_each over
// (index, value) or (index, value, position), the host interp /
// interp_diff / add_back (host_field.hpp:237-305) and its MAS blocks
// (host_field.hpp:310-337).  interp and interp_diff run the same
// arithmetic as the device accessors, so a host and a device field with the
// same content give the same bits.
#pragma once

#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <optional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "device_field.hpp"
#include "grid.hpp"

namespace psum {
namespace field {

template <int D, var_loc Loc, typename Scalar, int QuantitySize = 1>
class host_field {
    static_assert(QuantitySize == 1, "host_field: vector-valued fields are not part of this build");

public:
    static constexpr int Dimension = D;
    static constexpr var_loc Location = Loc;
    using Value = Scalar;
    using Grid = simple_grid<D>;
    using Position = vec<D>;   // x(), y(), z() and [d]
    using acc_type = device_field_acc<D, Loc, Scalar>;

    host_field() = default;
    explicit host_field(const Grid& g) : grid_(g), content_(g.contentSize(Loc), Scalar(0)) {}
    host_field(std::vector<Scalar>&& content, const Grid& g) : grid_(g), content_(std::move(content))
    {
        if (content_.size() != g.contentSize(Loc))
            throw std::runtime_error("host_field: content size is not the grid's");
    }
    // a host copy of a device field
    template <int Q>
    explicit host_field(const device_field<D, Loc, Scalar, Q>& f) : grid_(f.getGrid()), content_(f.to_host())
    {
    }

    bool has_grid() const { return grid_.has_value(); }
    size_t size() const { return grid().contentSize(Loc); }
    Scalar* data() { return (void)grid(), content_.data(); }
    const Scalar* data() const { return (void)grid(), content_.data(); }
    const Grid& getGrid() const { return grid(); }
    var_loc getLocation() const { return Loc; }
    std::vector<Scalar> getContent() const { return (void)grid(), content_; }

    // func(i, value) or func(i, value, position of element i)
    template <typename F>
    void for_each(F&& func)
    {
        const size_t n = size();
        for (size_t i = 0; i < n; ++i) {
            if constexpr (std::is_invocable_v<F&, size_t, Scalar&, const Position&>)
                func(i, content_[i], position(i));
            else
                func(i, content_[i]);
        }
    }

    void setConstant(const Scalar& v)
    {
        (void)grid();
        for (auto& x : content_) x = v;
    }
    void setZero() { setConstant(Scalar(0)); }
    // content of b (same size)
    host_field& operator<<(const host_field& b)
    {
        if (b.size() != size()) throw std::runtime_error("host_field: operator<< of another size");
        content_ = b.content_;
        return *this;
    }
    void copy(const std::vector<Scalar>& v)
    {
        if (v.size() != size()) throw std::runtime_error("host_field::copy: size mismatch");
        content_ = v;
    }
    Scalar& operator()(size_t i) { return content_[i]; }
    const Scalar& operator()(size_t i) const { return content_[i]; }

    // node position, or cell centre
    Position position(size_t i) const { return detail::element_position<D, Loc>(grid(), i); }

    // Tecplot ASCII file of the field (device_field.hpp:203-226): variables
    // x[,y[,z]],varName and the grid's zone; scientific, 4 digits
    void plot(const std::string& filename, std::string varName = "_not_defined_",
              double time = std::nan("")) const
    {
        std::ofstream fp(filename, std::ios::out);
        if (!fp) throw std::runtime_error(filename + " cannot be opened");
        fp << std::scientific << std::setprecision(4);
        if (varName == "_not_defined_") varName = "v1";
        fp << "variables=" << (D == 1 ? "x," : (D == 2 ? "x,y," : "x,y,z,")) << varName;
        const auto h = getContent();
        getGrid().plot(fp, h.data(), time, Loc);
    }

    // accessor over the host content (interp / interp_diff take it)
    acc_type acc() const
    {
        acc_type a;
        a.content = const_cast<Scalar*>(data());
        a.grid = psum_hip::to_dev(grid().to_psum());
        return a;
    }

private:
    const Grid& grid() const
    {
        if (!grid_) throw std::runtime_error("host_field has no grid");
        return *grid_;
    }

    std::optional<Grid> grid_;
    std::vector<Scalar> content_;

public:
    // grid and content together (the MAS load)
    void reset(const Grid& g, std::vector<Scalar>&& c)
    {
        if (c.size() != g.contentSize(Loc)) throw std::runtime_error("host_field: content size is not the grid's");
        grid_ = g;
        content_ = std::move(c);
    }
};

template <typename Scalar, int QS = 1> using host_node_field1D = host_field<1, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using host_node_field2D = host_field<2, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using host_node_field3D = host_field<3, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using host_cell_field1D = host_field<1, var_loc::cellCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using host_cell_field2D = host_field<2, var_loc::cellCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using host_cell_field3D = host_field<3, var_loc::cellCentered, Scalar, QS>;

template <typename Pos, int D, typename Scalar>
Scalar interp(const Pos& pos, const host_field<D, var_loc::nodeCentered, Scalar>& f)
{
    return interp(pos, f.acc());
}

template <typename Pos, int D, typename Scalar>
std::array<Scalar, D> interp_diff(const Pos& pos, const host_field<D, var_loc::nodeCentered, Scalar>& f)
{
    return interp_diff(pos, f.acc());
}

// f[idx_i] += w c_i, skipping |c_i| <= 1e-10, in support order
template <typename Pos, int D, typename Scalar>
void add_back(const Pos& pos, const Scalar& w, host_field<D, var_loc::nodeCentered, Scalar>& f)
{
    const auto a = f.acc();
    int cw[D];
    double W[D][2], G[D][2];
    if (!detail::support<D>(a.grid, pos, cw, W, G)) return;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (std::fabs(c) > 1e-10) a.content[psum_hip::support_index<D>(a.grid, cw, i)] += w * c;
    }
}

}  // namespace field
}  // namespace psum
