// device_field<D, Loc, Scalar>: a node- or cell-centred field on a
// simple_grid in device memory (src/field/device_field.hpp:120-240), with
// the device accessor that interp / interp_diff / add_back take
// (device_field.hpp:245-321, field_interp.hpp:63-113).
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <string>
#include <type_traits>
#include <vector>

#include "../psum_bspline.hpp"
#include "../psum_cic.hpp"
#include "device_vector.hpp"
#include "grid.hpp"
#include "tag.hpp"

namespace psum {
namespace field {

template <int D, var_loc Loc, typename Scalar>
struct device_field_acc {
    static constexpr int Dimension = D;
    static constexpr var_loc Location = Loc;
    using Value = Scalar;
    Scalar* content = nullptr;
    psum_hip::DevGrid grid{};
    __host__ __device__ Scalar& operator()(size_t i) const { return content[i]; }
};

namespace detail {
// position of element i (the node, or the cell centre) as a vec<D>, which
// has the x() / y() / z() of the reference's position rows
template <int D, var_loc Loc>
__host__ __device__ vec<D> element_position(const simple_grid<D>& g, size_t i)
{
    const std::array<double, D> a = Loc == var_loc::nodeCentered ? g.nodePosition(g.i2n(i)) : g.cellCenter(i);
    vec<D> p;
    for (int d = 0; d < D; ++d) p[d] = a[d];
    return p;
}

template <int D, var_loc Loc, typename Scalar, typename F>
__global__ void k_field_for_each(Scalar* v, size_t n, simple_grid<D> g, F f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (std::is_invocable_v<F&, size_t, Scalar&, const vec<D>&>)
        f(i, v[i], element_position<D, Loc>(g, i));
    else
        f(i, v[i]);
}

template <typename Scalar>
__global__ void k_field_fill(Scalar* v, size_t n, Scalar x)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = x;
}
}  // namespace detail

template <int D, var_loc Loc, typename Scalar, int QuantitySize = 1>
class device_field {
    static_assert(QuantitySize == 1, "device_field: vector-valued fields are not part of this build");

public:
    using acc_type = device_field_acc<D, Loc, Scalar>;
    using Grid = simple_grid<D>;

    device_field(hipStream_t s, const Grid& g) : grid_(g), stream_(s)
    {
        const size_t n = g.contentSize(Loc);
        hip_check(hipMalloc(&data_, n * sizeof(Scalar)), "device_field");
        setZero();
    }
    explicit device_field(const Grid& g) : device_field(nullptr, g) {}
    device_field(const device_field&) = delete;
    device_field& operator=(const device_field&) = delete;
    ~device_field()
    {
        if (data_) (void)hipFree(data_);
    }

    acc_type get_access(handler&) const { return acc(); }
    acc_type acc() const
    {
        acc_type a;
        a.content = data_;
        a.grid = psum_hip::to_dev(grid_.to_psum());
        return a;
    }
    size_t size() const { return grid_.contentSize(Loc); }
    Scalar* data() { return data_; }
    const Scalar* data() const { return data_; }
    const Grid& getGrid() const { return grid_; }
    var_loc getLocation() const { return Loc; }
    hipStream_t getQueue() const { return stream_; }

    void setZero()
    {
        hip_check(hipMemsetAsync(data_, 0, size() * sizeof(Scalar), stream_), "setZero");
        hip_check(hipStreamSynchronize(stream_), "setZero");
    }
    bool has_grid() const { return true; }
    void setConstant(const Scalar& x)
    {
        const size_t n = size();
        hipLaunchKernelGGL(detail::k_field_fill<Scalar>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_,
                           data_, n, x);
        hip_check(hipGetLastError(), "setConstant");
        hip_check(hipStreamSynchronize(stream_), "setConstant");
    }
    // content of b (same size)
    device_field& operator<<(const device_field& b)
    {
        if (b.size() != size()) throw std::invalid_argument("device_field: operator<< of another size");
        hip_check(hipMemcpyAsync(data_, b.data_, size() * sizeof(Scalar), hipMemcpyDeviceToDevice, stream_), "<<");
        hip_check(hipStreamSynchronize(stream_), "<<");
        return *this;
    }
    // func(handler&) returns f(i, value) or f(i, value, position of element
    // i), applied to every element on the device (device_field.hpp:138-163)
    template <typename FuncType>
    void for_each(FuncType&& func)
    {
        const size_t n = size();
        handler h{stream_};
        auto f = func(h);
        hipLaunchKernelGGL((detail::k_field_for_each<D, Loc, Scalar, decltype(f)>), dim3((unsigned)((n + 255) / 256)),
                           dim3(256), 0, stream_, data_, n, grid_, f);
        hip_check(hipGetLastError(), "for_each");
        hip_check(hipStreamSynchronize(stream_), "for_each");
    }
    void copy(const std::vector<Scalar>& host)
    {
        if (host.size() != size()) throw std::invalid_argument("device_field::copy: size mismatch");
        hip_check(hipMemcpyAsync(data_, host.data(), size() * sizeof(Scalar), hipMemcpyHostToDevice, stream_), "copy");
        hip_check(hipStreamSynchronize(stream_), "copy");
    }
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
        const auto h = to_host();
        getGrid().plot(fp, h.data(), time, Loc);
    }

    std::vector<Scalar> to_host() const
    {
        std::vector<Scalar> h(size());
        hip_check(hipMemcpyAsync(h.data(), data_, size() * sizeof(Scalar), hipMemcpyDeviceToHost, stream_), "to_host");
        hip_check(hipStreamSynchronize(stream_), "to_host");
        return h;
    }

private:
    Grid grid_;
    hipStream_t stream_ = nullptr;
    Scalar* data_ = nullptr;
};

template <typename Scalar, int QS = 1> using node_field1D = device_field<1, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using node_field2D = device_field<2, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using node_field3D = device_field<3, var_loc::nodeCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using cell_field1D = device_field<1, var_loc::cellCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using cell_field2D = device_field<2, var_loc::cellCentered, Scalar, QS>;
template <typename Scalar, int QS = 1> using cell_field3D = device_field<3, var_loc::cellCentered, Scalar, QS>;

namespace detail {
// CIC support of pos on a node field: indices and per-dimension factors
template <int D, typename Pos>
__host__ __device__ bool support(const psum_hip::DevGrid& g, const Pos& pos, int* cw, double (*W)[2], double (*G)[2])
{
    double x[D], s[D];
    for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
    if (!psum_hip::base_cell<D>(g, x, cw, s)) return false;
    for (int d = 0; d < D; ++d) {
        W[d][0] = 1.0 - s[d];
        W[d][1] = s[d];
        G[d][0] = -1.0 * g.delta_r[d];
        G[d][1] = 1.0 * g.delta_r[d];
    }
    return true;
}
}  // namespace detail

// phi(x): sum_i phi[idx_i] c_i, skipping |c_i| <= 1e-10 (field_interp.hpp:77-91)
template <typename Pos, int D, typename Scalar>
__host__ __device__ Scalar interp(const Pos& pos, const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    Scalar r = (Scalar)0;
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return r;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (fabs(c) > 1e-10) r = r + acc.content[psum_hip::support_index<D>(acc.grid, cw, i)] * c;
    }
    return r;
}

// grad phi(x) of the multilinear interpolant (field_interp.hpp:93-113)
template <typename Pos, int D, typename Scalar>
__host__ __device__ std::array<Scalar, D> interp_diff(const Pos& pos,
                                             const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    std::array<Scalar, D> g{};
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return g;
    Scalar ph[1 << D];
    for (int i = 0; i < (1 << D); ++i) ph[i] = acc.content[psum_hip::support_index<D>(acc.grid, cw, i)];
    for (int d = 0; d < D; ++d) {
        Scalar r = (Scalar)0;
        for (int i = 0; i < (1 << D); ++i) {
            const double c = psum_hip::cic_grad_coef<double, D>(W, G, i, d);
            if (fabs(c) > 1e-10) r = r + ph[i] * c;
        }
        g[d] = r;
    }
    return g;
}

// rho[idx_i] += w c_i with FP64 atomics, skipping |c_i| <= 1e-10
// (field_interp.hpp:63-75, atomic_add.hpp:34-48).  Summation order is that of
// the atomics; psum_hip_deposit_cic is the deterministic fixed-point form.
template <typename Pos, int D, typename Scalar>
__device__ void add_back(const Pos& pos, const Scalar& w,
                         const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (fabs(c) > 1e-10) atomicAdd(&acc.content[psum_hip::support_index<D>(acc.grid, cw, i)], w * c);
    }
}

// ---- B-spline facade (higher_order_interp.hpp:196-248): Order 1 is the
// CIC functions above; 2..4 the quadratic / cubic / quartic B-splines, with
// FP64 weights, support order+1 nodes per dimension, and the same 1e-10 cut.
namespace detail {
template <int D, int O, typename F>
__device__ void bspline_points(const psum_hip::DevGrid& g, const int* base, F&& f)
{
    using psum_hip::bspline::Support;
    constexpr int S = O + 1;
    constexpr int NS = D == 1 ? S : (D == 2 ? S * S : S * S * S);
    for (int i = 0; i < NS; ++i) {
        int k[3] = {0, 0, 0};
        int t = i;
        for (int j = 0; j < D; ++j) {
            k[D - 1 - j] = t % S;
            t /= S;
        }
        uint32_t lin = 0;
        for (int d = 0; d < D; ++d) {
            const int nd = psum_hip::bspline::wrap_node(base[d] + k[d] + Support<O>::shift, g.w[d]);
            lin = d == 0 ? (uint32_t)nd : lin * (uint32_t)g.w[d] + (uint32_t)nd;
        }
        f(k, lin);
    }
}
}  // namespace detail

template <int Order, typename Pos, int D, typename Scalar>
__device__ Scalar interp_B_spline(const Pos& pos, const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    if constexpr (Order == 1) {
        return interp(pos, acc);
    } else {
        double x[D], c[D][Order + 1];
        int base[D];
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
        double r = 0.0;
        if (!psum_hip::bspline::support_of<D, Order, false>(acc.grid, x, base, c, nullptr)) return (Scalar)r;
        detail::bspline_points<D, Order>(acc.grid, base, [&](const int* k, uint32_t lin) {
            double cf = c[D - 1][k[D - 1]];
            for (int j = 1; j < D; ++j) cf = cf * c[D - 1 - j][k[D - 1 - j]];
            if (fabs(cf) > 1e-10) r += (double)acc.content[lin] * cf;
        });
        return (Scalar)r;
    }
}

template <int Order, typename Pos, int D, typename Scalar>
__device__ std::array<Scalar, D> interp_diff_B_spline(const Pos& pos,
                                                      const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    if constexpr (Order == 1) {
        return interp_diff(pos, acc);
    } else {
        double x[D], c[D][Order + 1], dc[D][Order + 1], r[D] = {};
        int base[D];
        std::array<Scalar, D> out{};
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
        if (!psum_hip::bspline::support_of<D, Order, true>(acc.grid, x, base, c, dc)) return out;
        detail::bspline_points<D, Order>(acc.grid, base, [&](const int* k, uint32_t lin) {
            const double ph = (double)acc.content[lin];
            for (int dg = 0; dg < D; ++dg) {
                double gf = dg == D - 1 ? dc[D - 1][k[D - 1]] : c[D - 1][k[D - 1]];
                for (int j = 1; j < D; ++j) {
                    const int dd = D - 1 - j;
                    gf = gf * (dd == dg ? dc[dd][k[dd]] : c[dd][k[dd]]);
                }
                if (fabs(gf) > 1e-10) r[dg] += ph * gf;
            }
        });
        for (int d = 0; d < D; ++d) out[d] = (Scalar)r[d];
        return out;
    }
}

// FP64 atomics like add_back; psum_hip_deposit_bspline is the fixed-point form
template <int Order, typename Pos, int D, typename Scalar>
__device__ void add_back_B_spline(const Pos& pos, const Scalar& w,
                                  const device_field_acc<D, var_loc::nodeCentered, Scalar>& acc)
{
    if constexpr (Order == 1) {
        add_back(pos, w, acc);
    } else {
        double x[D], c[D][Order + 1];
        int base[D];
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
        if (!psum_hip::bspline::support_of<D, Order, false>(acc.grid, x, base, c, nullptr)) return;
        detail::bspline_points<D, Order>(acc.grid, base, [&](const int* k, uint32_t lin) {
            double cf = c[D - 1][k[D - 1]];
            for (int j = 1; j < D; ++j) cf = cf * c[D - 1 - j][k[D - 1 - j]];
            if (fabs(cf) > 1e-10) atomicAdd(&acc.content[lin], w * cf);
        });
    }
}

}  // namespace field
}  // namespace psum
