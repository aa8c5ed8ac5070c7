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
    static_assert(QuantitySize >= 1, "host_field: QuantitySize >= 1");

public:
    static constexpr int Dimension = D;
    static constexpr var_loc Location = Loc;
    static constexpr int QS = QuantitySize;
    using ScalarType = Scalar;
    using Value = field_value_t<Scalar, QuantitySize>;
    using Grid = simple_grid<D>;
    using Position = vec<D>;   // x(), y(), z() and [d]
    using acc_type = device_field_acc<D, Loc, Scalar, QuantitySize>;

    host_field() = default;
    explicit host_field(const Grid& g) : grid_(g), content_(g.contentSize(Loc), Value{}) {}
    host_field(std::vector<Value>&& content, const Grid& g) : grid_(g), content_(std::move(content))
    {
        if (content_.size() != g.contentSize(Loc))
            throw std::runtime_error("host_field: content size is not the grid's");
    }
    // a host copy of a device field
    explicit host_field(const device_field<D, Loc, Scalar, QuantitySize>& f)
        : grid_(f.getGrid()), content_(f.to_host())
    {
    }

    bool has_grid() const { return grid_.has_value(); }
    size_t size() const { return grid().contentSize(Loc); }
    Value* data() { return (void)grid(), content_.data(); }
    const Value* data() const { return (void)grid(), content_.data(); }
    const Grid& getGrid() const { return grid(); }
    var_loc getLocation() const { return Loc; }
    std::vector<Value> getContent() const { return (void)grid(), content_; }

    // func(i, value) or func(i, value, position of element i)
    template <typename F>
    void for_each(F&& func)
    {
        const size_t n = size();
        for (size_t i = 0; i < n; ++i) {
            if constexpr (std::is_invocable_v<F&, size_t, Value&, const Position&>)
                func(i, content_[i], position(i));
            else
                func(i, content_[i]);
        }
    }

    void setConstant(const Value& v)
    {
        (void)grid();
        for (auto& x : content_) x = v;
    }
    void setZero() { setConstant(Value{}); }
    // content of b (same size)
    host_field& operator<<(const host_field& b)
    {
        if (b.size() != size()) throw std::runtime_error("host_field: operator<< of another size");
        content_ = b.content_;
        return *this;
    }
    void copy(const std::vector<Value>& v)
    {
        if (v.size() != size()) throw std::runtime_error("host_field::copy: size mismatch");
        content_ = v;
    }
    Value& operator()(size_t i) { return content_[i]; }
    const Value& operator()(size_t i) const { return content_[i]; }
    // component col of element row (QuantitySize > 1)
    Scalar& operator()(size_t row, size_t col)
    {
        static_assert(QuantitySize > 1, "operator()(row, col) needs QuantitySize > 1");
        return content_[row][(int)col];
    }

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
        if (varName == "_not_defined_") {
            varName = "v1";
            for (int q = 1; q < QuantitySize; ++q) varName += ",v" + std::to_string(q + 1);
        }
        fp << "variables=" << (D == 1 ? "x," : (D == 2 ? "x,y," : "x,y,z,")) << varName;
        const auto h = getContent();
        getGrid().plot(fp, reinterpret_cast<const Scalar*>(h.data()), time, Loc, QuantitySize);
    }

    // accessor over the host content (interp / interp_diff take it)
    acc_type acc() const
    {
        acc_type a;
        a.content = const_cast<Value*>(data());
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
    std::vector<Value> content_;

public:
    // grid and content together (the MAS load)
    void reset(const Grid& g, std::vector<Value>&& c)
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

template <typename Pos, int D, typename Scalar, int QS>
field_value_t<Scalar, QS> interp(const Pos& pos, const host_field<D, var_loc::nodeCentered, Scalar, QS>& f)
{
    return interp(pos, f.acc());
}

template <typename Pos, int D, typename Scalar, int QS>
std::array<field_value_t<Scalar, QS>, D> interp_diff(const Pos& pos,
                                                     const host_field<D, var_loc::nodeCentered, Scalar, QS>& f)
{
    return interp_diff(pos, f.acc());
}

// f[idx_i] += w c_i, skipping |c_i| <= 1e-10, in support order
template <typename Pos, int D, typename Scalar, int QS>
void add_back(const Pos& pos, const field_value_t<Scalar, QS>& w, host_field<D, var_loc::nodeCentered, Scalar, QS>& f)
{
    const auto a = f.acc();
    int cw[D];
    double W[D][2], G[D][2];
    if (!detail::support<D>(a.grid, pos, cw, W, G)) return;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (std::fabs(c) > 1e-10) {
            auto& v = a.content[psum_hip::support_index<D>(a.grid, cw, i)];
            for (int q = 0; q < QS; ++q) detail::comp<QS>(v, q) += detail::comp<QS>(w, q) * c;
        }
    }
}

// the element nearest to pos (field_interp.hpp:38-45; host_field.hpp
// interp_nearest); zero outside the grid
template <typename Pos, int D, var_loc Loc, typename Scalar, int QS>
field_value_t<Scalar, QS> interp_nearest(const Pos& pos, const host_field<D, Loc, Scalar, QS>& f)
{
    return interp_nearest(pos, f.acc());
}

// the element nearest to pos += w (field_interp.hpp:24-31); nothing outside
template <typename Pos, int D, var_loc Loc, typename Scalar, int QS>
void add_back_nearest(const Pos& pos, const field_value_t<Scalar, QS>& w, host_field<D, Loc, Scalar, QS>& f)
{
    const auto a = f.acc();
    size_t k = 0;
    if (!detail::nearest_element<D, Loc>(a.grid, pos, &k)) return;
    for (int q = 0; q < QS; ++q) detail::comp<QS>(a.content[k], q) += detail::comp<QS>(w, q);
}

}  // namespace field
}  // namespace psum
