// device_field<D, Loc, Scalar, QuantitySize>: a node- or cell-centred field
// on a simple_grid in device memory (src/field/device_field.hpp:75-240), with
// the device accessor that interp / interp_diff / add_back take
// (device_field.hpp:36-71, :245-321, field_interp.hpp:24-113).  An element
// is a Scalar, or for QuantitySize > 1 a row of QuantitySize scalars
// (vec<QuantitySize, Scalar>, where the reference has
// Eigen::RowVector<Scalar, QuantitySize>).
//
// Deterministic add_back: after use_fixed_point(w_max), add_back through the
// field's accessor adds each term llrint((w*c)*2^S) to an int64 array beside
// the content instead of adding w*c with floating atomics.  Integer sums do
// not depend on the order of the adds, so the deposit is bit-reproducible
// and equals psum_hip_deposit_cic with the same scale.  The handler of the
// launching for_each folds the sums into the content when the kernel ends;
// a term of magnitude >= 2^62 or an add whose int64 partial sum wraps is
// counted, and the fold then throws std::overflow_error.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <stdexcept>
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

// element type of a field: the scalar, or a row of QuantitySize scalars
template <typename Scalar, int QS>
struct field_value {
    using type = vec<QS, Scalar>;
};
template <typename Scalar>
struct field_value<Scalar, 1> {
    using type = Scalar;
};
template <typename Scalar, int QS>
using field_value_t = typename field_value<Scalar, QS>::type;

namespace detail {
// component q of an element
template <int QS, typename V>
__host__ __device__ inline auto& comp(V& v, int q)
{
    if constexpr (QS == 1) {
        (void)q;
        return v;
    } else {
        return v[q];
    }
}

// The int64 sums of a deterministic add_back.  acc holds QuantitySize
// entries per element; state[0] is set when a term was added, state[1]
// counts terms and adds that left the int64 range.
struct fixed_point_sink {
    unsigned long long* acc = nullptr;
    unsigned long long* state = nullptr;
    double scale = 0.0;   // 2^S
};

// acc[k] += llrint(t) for t = (w*c)*2^S.  A wrap of the two's-complement
// partial sum is seen from the value the atomic returns: old and term of
// one sign and a sum of the other.
__device__ inline void fixed_add(const fixed_point_sink& s, size_t k, double t)
{
    const unsigned long long active = __ballot(1);
    if (__lane_id() == __ffsll((long long)active) - 1) s.state[0] = 1ull;
    if (!(fabs(t) < 0x1p62)) {
        atomicAdd(&s.state[1], 1ull);
        return;
    }
    const unsigned long long q = (unsigned long long)__double2ll_rn(t);
    const unsigned long long old = atomicAdd(&s.acc[k], q);
    const unsigned long long sum = old + q;
    if ((long long)((old ^ sum) & (q ^ sum)) < 0) atomicAdd(&s.state[1], 1ull);
}
}  // namespace detail

template <int D, var_loc Loc, typename Scalar, int QS = 1>
struct device_field_acc {
    static constexpr int Dimension = D;
    static constexpr var_loc Location = Loc;
    static constexpr int QuantitySize = QS;
    using Value = field_value_t<Scalar, QS>;
    using ScalarType = Scalar;
    Value* content = nullptr;
    psum_hip::DevGrid grid{};
    detail::fixed_point_sink fx{};   // fx.acc == nullptr: floating atomics
    __host__ __device__ Value& operator()(size_t i) const { return content[i]; }
    // component col of element row (QuantitySize > 1, device_field.hpp:54-62)
    __host__ __device__ Scalar& operator()(size_t row, size_t col) const
    {
        static_assert(QS > 1, "operator()(row, col) needs QuantitySize > 1");
        return content[row][(int)col];
    }
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

template <int D, var_loc Loc, typename Value, typename F>
__global__ void k_field_for_each(Value* v, size_t n, simple_grid<D> g, F f)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if constexpr (std::is_invocable_v<F&, size_t, Value&, const vec<D>&>)
        f(i, v[i], element_position<D, Loc>(g, i));
    else
        f(i, v[i]);
}

template <typename Value>
__global__ void k_field_fill(Value* v, size_t n, Value x)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = x;
}

// content += acc * 2^-S (FP64, rounded once to the scalar type), acc = 0.
// With content zero the result is bitwise psum_hip_fixed_to_double's.
template <typename Scalar>
__global__ void k_fixed_fold(Scalar* content, unsigned long long* acc, size_t n, double inv_scale)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long q = (long long)acc[i];
    if (q == 0) return;
    content[i] = (Scalar)((double)content[i] + (double)q * inv_scale);
    acc[i] = 0ull;
}
}  // namespace detail

template <int D, var_loc Loc, typename Scalar, int QuantitySize = 1>
class device_field {
    static_assert(QuantitySize >= 1, "device_field: QuantitySize >= 1");

public:
    static constexpr int Dimension = D;
    static constexpr var_loc Location = Loc;
    static constexpr int QS = QuantitySize;
    using ScalarType = Scalar;
    using Value = field_value_t<Scalar, QuantitySize>;
    using Position = vec<D>;
    using acc_type = device_field_acc<D, Loc, Scalar, QuantitySize>;
    using Grid = simple_grid<D>;

    device_field(hipStream_t s, const Grid& g) : grid_(g), stream_(s)
    {
        hip_check(hipMalloc(&data_, size() * sizeof(Value)), "device_field");
        setZero();
    }
    explicit device_field(const Grid& g) : device_field(nullptr, g) {}
    device_field(const device_field&) = delete;
    device_field& operator=(const device_field&) = delete;
    ~device_field()
    {
        if (data_) (void)hipFree(data_);
        if (fx_.acc) (void)hipFree(fx_.acc);
        if (fx_.state) (void)hipFree(fx_.state);
    }

    // the accessor of a kernel launched by a for_each given h; a
    // deterministic add_back through it is folded in when that kernel ends
    acc_type get_access(handler& h)
    {
        if (fx_.acc) h.after(this, [this]() { fold_fixed_point(); });
        return acc();
    }
    acc_type get_access(handler&) const { return acc(); }
    // the accessor for a kernel the caller launches; with use_fixed_point
    // on, call fold_fixed_point() after that kernel
    acc_type acc() const
    {
        acc_type a;
        a.content = data_;
        a.grid = psum_hip::to_dev(grid_.to_psum());
        a.fx = fx_;
        return a;
    }
    size_t size() const { return grid_.contentSize(Loc); }
    Value* data() { return data_; }
    const Value* data() const { return data_; }
    const Grid& getGrid() const { return grid_; }
    var_loc getLocation() const { return Loc; }
    hipStream_t getQueue() const { return stream_; }
    bool has_grid() const { return true; }

    // ---- deterministic add_back (this build's addition; the reference adds
    // with floating atomics, field_interp.hpp:63-75, atomic_add.hpp:34-48)
    //
    // S = 40 - e with 2^(e-1) <= w_max < 2^e, so one term of a weight
    // |w| <= w_max is below 2^40 in magnitude and an element holds about 2^23
    // such terms before its int64 sum would leave the range.  Returns 2^S.
    double use_fixed_point(double w_max)
    {
        if (!(w_max > 0.0) || !std::isfinite(w_max))
            throw std::invalid_argument("device_field::use_fixed_point: w_max must be positive and finite");
        int e = 0;
        std::frexp(w_max, &e);
        const size_t n = size() * (size_t)QuantitySize;
        if (!fx_.acc) {
            hip_check(hipMalloc(&fx_.acc, n * sizeof(unsigned long long)), "use_fixed_point");
            hip_check(hipMalloc(&fx_.state, 2 * sizeof(unsigned long long)), "use_fixed_point");
        }
        hip_check(hipMemsetAsync(fx_.acc, 0, n * sizeof(unsigned long long), stream_), "use_fixed_point");
        hip_check(hipMemsetAsync(fx_.state, 0, 2 * sizeof(unsigned long long), stream_), "use_fixed_point");
        hip_check(hipStreamSynchronize(stream_), "use_fixed_point");
        fx_.scale = std::ldexp(1.0, 40 - e);
        return fx_.scale;
    }
    // back to floating atomics (pending sums are folded first)
    void use_atomics()
    {
        if (!fx_.acc) return;
        fold_fixed_point();
        (void)hipFree(fx_.acc);
        (void)hipFree(fx_.state);
        fx_ = {};
    }
    // 2^S, or 0 when add_back uses floating atomics
    double fixed_point_scale() const { return fx_.acc ? fx_.scale : 0.0; }

    // Adds the pending int64 sums to the content and clears them; throws
    // std::overflow_error (after the fold) when terms or sums left the
    // int64 range since the last fold.
    void fold_fixed_point()
    {
        if (!fx_.acc) return;
        unsigned long long st[2] = {0, 0};
        hip_check(hipMemcpyAsync(st, fx_.state, sizeof st, hipMemcpyDeviceToHost, stream_), "fold");
        hip_check(hipStreamSynchronize(stream_), "fold");
        if (!st[0] && !st[1]) return;
        const size_t n = size() * (size_t)QuantitySize;
        hipLaunchKernelGGL(detail::k_fixed_fold<Scalar>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_,
                           reinterpret_cast<Scalar*>(data_), fx_.acc, n, 1.0 / fx_.scale);
        hip_check(hipGetLastError(), "fold");
        hip_check(hipMemsetAsync(fx_.state, 0, 2 * sizeof(unsigned long long), stream_), "fold");
        hip_check(hipStreamSynchronize(stream_), "fold");
        if (st[1])
            throw std::overflow_error("device_field: " + std::to_string(st[1]) +
                                      " fixed-point add_back terms or sums left the int64 range "
                                      "(w_max of use_fixed_point too small, or too many terms on one element)");
    }

    void setZero()
    {
        hip_check(hipMemsetAsync(data_, 0, size() * sizeof(Value), stream_), "setZero");
        hip_check(hipStreamSynchronize(stream_), "setZero");
    }
    void setConstant(const Value& x)
    {
        const size_t n = size();
        hipLaunchKernelGGL(detail::k_field_fill<Value>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_,
                           data_, n, x);
        hip_check(hipGetLastError(), "setConstant");
        hip_check(hipStreamSynchronize(stream_), "setConstant");
    }
    // content of b (same size)
    device_field& operator<<(const device_field& b)
    {
        if (b.size() != size()) throw std::invalid_argument("device_field: operator<< of another size");
        hip_check(hipMemcpyAsync(data_, b.data_, size() * sizeof(Value), hipMemcpyDeviceToDevice, stream_), "<<");
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
        hipLaunchKernelGGL((detail::k_field_for_each<D, Loc, Value, decltype(f)>),
                           dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream_, data_, n, grid_, f);
        hip_check(hipGetLastError(), "for_each");
        hip_check(hipStreamSynchronize(stream_), "for_each");
        h.finish();
    }
    void copy(const std::vector<Value>& host)
    {
        if (host.size() != size()) throw std::invalid_argument("device_field::copy: size mismatch");
        hip_check(hipMemcpyAsync(data_, host.data(), size() * sizeof(Value), hipMemcpyHostToDevice, stream_), "copy");
        hip_check(hipStreamSynchronize(stream_), "copy");
    }
    // Tecplot ASCII file of the field (device_field.hpp:203-226): variables
    // x[,y[,z]],varName (default v1[,v2,...] per component) and the grid's
    // zone; scientific, 4 digits
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
        const auto h = to_host();
        getGrid().plot(fp, reinterpret_cast<const Scalar*>(h.data()), time, Loc, QuantitySize);
    }

    std::vector<Value> to_host() const
    {
        std::vector<Value> h(size());
        hip_check(hipMemcpyAsync(h.data(), data_, size() * sizeof(Value), hipMemcpyDeviceToHost, stream_), "to_host");
        hip_check(hipStreamSynchronize(stream_), "to_host");
        return h;
    }

private:
    Grid grid_;
    hipStream_t stream_ = nullptr;
    Value* data_ = nullptr;
    detail::fixed_point_sink fx_{};
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

// element k (linear) += w c: a fixed-point term when the accessor has a
// sink, else floating atomics per component (atomic_add.hpp:34-48)
template <int QS, typename Acc, typename Value>
__device__ void deposit_term(const Acc& acc, size_t k, const Value& w, double c)
{
    if (acc.fx.acc) {
        for (int q = 0; q < QS; ++q)
            fixed_add(acc.fx, k * QS + q, ((double)comp<QS>(w, q) * c) * acc.fx.scale);
    } else {
        for (int q = 0; q < QS; ++q) atomicAdd(&comp<QS>(acc.content[k], q), comp<QS>(w, q) * c);
    }
}

// the element nearest to pos: round((x - lb) * delta_r) for nodes,
// floor for cells (simple_grid.hpp:80-95, :296-305); false outside the grid
template <int D, var_loc Loc, typename Pos>
__host__ __device__ bool nearest_element(const psum_hip::DevGrid& g, const Pos& pos, size_t* k)
{
    size_t lin = 0;
    for (int d = 0; d < D; ++d) {
        const double t = ((double)pos[d] - g.lb[d]) * g.delta_r[d];
        const double f = Loc == var_loc::nodeCentered ? round(t) : floor(t);
        const int ext = Loc == var_loc::nodeCentered ? g.w[d] : g.w[d] - 1;
        if (!(f >= 0.0 && f < (double)ext)) return false;
        lin = lin * (size_t)ext + (size_t)f;
    }
    *k = lin;
    return true;
}
}  // namespace detail

// phi(x): sum_i phi[idx_i] c_i, skipping |c_i| <= 1e-10 (field_interp.hpp:77-91);
// per component for QuantitySize > 1
template <typename Pos, int D, typename Scalar, int QS>
__host__ __device__ field_value_t<Scalar, QS> interp(const Pos& pos,
                                                     const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    field_value_t<Scalar, QS> r{};
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return r;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (fabs(c) > 1e-10) {
            auto& v = acc.content[psum_hip::support_index<D>(acc.grid, cw, i)];
            for (int q = 0; q < QS; ++q) detail::comp<QS>(r, q) = detail::comp<QS>(r, q) + detail::comp<QS>(v, q) * c;
        }
    }
    return r;
}

// grad phi(x) of the multilinear interpolant (field_interp.hpp:93-113): one
// element value per dimension
template <typename Pos, int D, typename Scalar, int QS>
__host__ __device__ std::array<field_value_t<Scalar, QS>, D> interp_diff(
    const Pos& pos, const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    std::array<field_value_t<Scalar, QS>, D> g{};
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return g;
    for (int q = 0; q < QS; ++q) {
        Scalar ph[1 << D];
        for (int i = 0; i < (1 << D); ++i)
            ph[i] = detail::comp<QS>(acc.content[psum_hip::support_index<D>(acc.grid, cw, i)], q);
        for (int d = 0; d < D; ++d) {
            Scalar r = (Scalar)0;
            for (int i = 0; i < (1 << D); ++i) {
                const double c = psum_hip::cic_grad_coef<double, D>(W, G, i, d);
                if (fabs(c) > 1e-10) r = r + ph[i] * c;
            }
            detail::comp<QS>(g[d], q) = r;
        }
    }
    return g;
}

// rho[idx_i] += w c_i, skipping |c_i| <= 1e-10 (field_interp.hpp:63-75):
// exact int64 terms after use_fixed_point, else floating atomics whose
// summation order is that of the threads
template <typename Pos, int D, typename Scalar, int QS>
__device__ void add_back(const Pos& pos, const field_value_t<Scalar, QS>& w,
                         const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
{
    int cw[D];
    double W[D][2], G[D][2];
    if (!detail::support<D>(acc.grid, pos, cw, W, G)) return;
    for (int i = 0; i < (1 << D); ++i) {
        const double c = psum_hip::cic_coef<double, D>(W, i);
        if (fabs(c) > 1e-10) detail::deposit_term<QS>(acc, psum_hip::support_index<D>(acc.grid, cw, i), w, c);
    }
}

// the value of the element nearest to pos (field_interp.hpp:38-45); zero
// outside the grid, where the reference indexes out of range
template <typename Pos, int D, var_loc Loc, typename Scalar, int QS>
__host__ __device__ field_value_t<Scalar, QS> interp_nearest(const Pos& pos,
                                                             const device_field_acc<D, Loc, Scalar, QS>& acc)
{
    size_t k = 0;
    if (!detail::nearest_element<D, Loc>(acc.grid, pos, &k)) return field_value_t<Scalar, QS>{};
    return acc.content[k];
}

// the element nearest to pos += w (field_interp.hpp:24-31); nothing outside
// the grid
template <typename Pos, int D, var_loc Loc, typename Scalar, int QS>
__device__ void add_back_nearest(const Pos& pos, const field_value_t<Scalar, QS>& w,
                                 const device_field_acc<D, Loc, Scalar, QS>& acc)
{
    size_t k = 0;
    if (detail::nearest_element<D, Loc>(acc.grid, pos, &k)) detail::deposit_term<QS>(acc, k, w, 1.0);
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

template <int Order, typename Pos, int D, typename Scalar, int QS>
__device__ field_value_t<Scalar, QS> interp_B_spline(const Pos& pos,
                                                     const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
{
    if constexpr (Order == 1) {
        return interp(pos, acc);
    } else {
        double x[D], c[D][Order + 1];
        int base[D];
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
        double r[QS] = {};
        field_value_t<Scalar, QS> out{};
        if (!psum_hip::bspline::support_of<D, Order, false>(acc.grid, x, base, c, nullptr)) return out;
        detail::bspline_points<D, Order>(acc.grid, base, [&](const int* k, uint32_t lin) {
            double cf = c[D - 1][k[D - 1]];
            for (int j = 1; j < D; ++j) cf = cf * c[D - 1 - j][k[D - 1 - j]];
            if (fabs(cf) > 1e-10)
                for (int q = 0; q < QS; ++q) r[q] += (double)detail::comp<QS>(acc.content[lin], q) * cf;
        });
        for (int q = 0; q < QS; ++q) detail::comp<QS>(out, q) = (Scalar)r[q];
        return out;
    }
}

template <int Order, typename Pos, int D, typename Scalar, int QS>
__device__ std::array<field_value_t<Scalar, QS>, D> interp_diff_B_spline(
    const Pos& pos, const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
{
    if constexpr (Order == 1) {
        return interp_diff(pos, acc);
    } else {
        double x[D], c[D][Order + 1], dc[D][Order + 1], r[D][QS] = {};
        int base[D];
        std::array<field_value_t<Scalar, QS>, D> out{};
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d];
        if (!psum_hip::bspline::support_of<D, Order, true>(acc.grid, x, base, c, dc)) return out;
        detail::bspline_points<D, Order>(acc.grid, base, [&](const int* k, uint32_t lin) {
            for (int dg = 0; dg < D; ++dg) {
                double gf = dg == D - 1 ? dc[D - 1][k[D - 1]] : c[D - 1][k[D - 1]];
                for (int j = 1; j < D; ++j) {
                    const int dd = D - 1 - j;
                    gf = gf * (dd == dg ? dc[dd][k[dd]] : c[dd][k[dd]]);
                }
                if (fabs(gf) > 1e-10)
                    for (int q = 0; q < QS; ++q) r[dg][q] += (double)detail::comp<QS>(acc.content[lin], q) * gf;
            }
        });
        for (int d = 0; d < D; ++d)
            for (int q = 0; q < QS; ++q) detail::comp<QS>(out[d], q) = (Scalar)r[d][q];
        return out;
    }
}

// fixed-point terms after use_fixed_point (as psum_hip_deposit_bspline),
// else floating atomics like add_back
template <int Order, typename Pos, int D, typename Scalar, int QS>
__device__ void add_back_B_spline(const Pos& pos, const field_value_t<Scalar, QS>& w,
                                  const device_field_acc<D, var_loc::nodeCentered, Scalar, QS>& acc)
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
            if (fabs(cf) > 1e-10) detail::deposit_term<QS>(acc, lin, w, cf);
        });
    }
}

}  // namespace field
}  // namespace psum
