// simple_grid<D>: the reference's Cartesian node/cell grid
// (src/field/simple_grid.hpp:38-51 ctor, :80-128 indices, :140-198 accessors),
// host and device callable, plus its conversion to the C ABI's psum_grid.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cmath>
#include <cstddef>
#include <ostream>
#include <stdexcept>
#include <vector>

#include "../psum_hip.h"

namespace psum {
namespace field {

enum class var_loc { nodeCentered, cellCentered };

template <int Dim>
class simple_grid {
public:
    static constexpr int Dimension = Dim;
    using node_index = std::array<int, Dim>;
    using cell_index = std::array<int, Dim>;

    simple_grid() : simple_grid(std::vector<double>(Dim, 0.0), std::vector<double>(Dim, 1.0),
                                std::vector<int>(Dim, 1)) {}

    // lower / upper bounds and cell counts per dimension; delta = (ub - lb) / n,
    // delta_r = 1 / delta (simple_grid.hpp:38-51)
    simple_grid(const std::vector<double>& lower, const std::vector<double>& upper,
                const std::vector<int>& cells)
    {
        if ((int)lower.size() != Dim || (int)upper.size() != Dim || (int)cells.size() != Dim)
            throw std::invalid_argument("simple_grid: bounds and cell counts need Dim entries");
        for (int d = 0; d < Dim; ++d) {
            if (cells[d] < 1 || !(upper[d] > lower[d]))
                throw std::invalid_argument("simple_grid: need cells >= 1 and upper > lower");
            lb_[d] = lower[d];
            ub_[d] = upper[d];
            n_[d] = cells[d];
            delta_[d] = (upper[d] - lower[d]) / cells[d];
            delta_r_[d] = 1.0 / delta_[d];
        }
    }

    template <int D> __host__ __device__ double del() const { return delta_[D]; }
    template <int D> __host__ __device__ double del_r() const { return delta_r_[D]; }
    template <int D> __host__ __device__ double lowerBound() const { return lb_[D]; }
    template <int D> __host__ __device__ double upperBound() const { return ub_[D]; }
    template <int D> __host__ __device__ int numCells() const { return n_[D]; }
    template <int D> __host__ __device__ double span() const { return ub_[D] - lb_[D]; }
    template <int D> __host__ __device__ double middle() const { return 0.5 * (lb_[D] + ub_[D]); }
    __host__ __device__ const double* get_deltas() const { return delta_; }
    __host__ __device__ const double* get_deltas_r() const { return delta_r_; }

    // cN: floor((x - lb) * delta_r) per dimension, in FP64 (simple_grid.hpp:89-95)
    template <typename Pos>
    __host__ __device__ cell_index cN(const Pos& p) const
    {
        cell_index c{};
        for (int d = 0; d < Dim; ++d) c[d] = (int)floor((p[d] - lb_[d]) * delta_r_[d]);
        return c;
    }

    // nearest node: round((x - lb) * delta_r) (simple_grid.hpp:80-86)
    template <typename Pos>
    __host__ __device__ node_index nN(const Pos& p) const
    {
        node_index c{};
        for (int d = 0; d < Dim; ++d) c[d] = (int)round((p[d] - lb_[d]) * delta_r_[d]);
        return c;
    }

    // linear node index, last dimension fastest (simple_grid.hpp:107-120)
    __host__ __device__ size_t n2i(const node_index& n) const
    {
        size_t i = (size_t)n[0];
        for (int d = 1; d < Dim; ++d) i = i * (size_t)(n_[d] + 1) + (size_t)n[d];
        return i;
    }
    __host__ __device__ node_index i2n(size_t i) const
    {
        node_index n{};
        for (int d = Dim - 1; d >= 0; --d) {
            n[d] = (int)(i % (size_t)(n_[d] + 1));
            i /= (size_t)(n_[d] + 1);
        }
        return n;
    }
    __host__ __device__ size_t c2i(const cell_index& c) const
    {
        size_t i = (size_t)c[0];
        for (int d = 1; d < Dim; ++d) i = i * (size_t)n_[d] + (size_t)c[d];
        return i;
    }

    template <typename Pos>
    __host__ __device__ bool inGrid(const Pos& p) const
    {
        for (int d = 0; d < Dim; ++d)
            if (!(p[d] >= lb_[d] && p[d] < ub_[d])) return false;
        return true;
    }

    __host__ __device__ size_t contentSize(var_loc loc) const
    {
        size_t n = 1;
        for (int d = 0; d < Dim; ++d) n *= (size_t)(loc == var_loc::nodeCentered ? n_[d] + 1 : n_[d]);
        return n;
    }

    __host__ __device__ std::array<double, Dim> nodePosition(const node_index& n) const
    {
        std::array<double, Dim> p{};
        for (int d = 0; d < Dim; ++d) p[d] = lb_[d] + n[d] * delta_[d];
        return p;
    }
    // Tecplot ASCII zone of a field with this grid (simple_grid.hpp:234-
    // 295): after the caller's "variables=" line, "zone i=..,j=..,F=point"
    // and one "x y value" line per node, or for cell-centred content
    // F=block with the node coordinates block by block, then the values;
    // SolutionTime when time is not NaN.  1D has no zone line.  vnum values
    // per element (content[i * vnum + v]): one column each for nodes, one
    // block each for cells.
    template <typename Scalar>
    void plot(std::ostream& fp, const Scalar* content, double time, var_loc loc, int vnum = 1) const
    {
        static const char ijk[3] = {'i', 'j', 'k'};
        if (Dim > 1) {
            fp << "\nzone ";
            for (int i = 0; i < Dim; ++i) fp << ijk[i] << "=" << (n_[Dim - 1 - i] + 1) << ",";
            if (loc == var_loc::cellCentered) {
                fp << "F=block, VARLOCATION=([" << Dim + 1;
                if (vnum > 1) fp << "-" << Dim + vnum;
                fp << "]=CELLCENTERED)" << std::endl;
            } else {
                fp << "F=point" << std::endl;
            }
        }
        if (!std::isnan(time)) fp << "AUXDATA TimeUnits=\"Seconds\" SolutionTime=" << time;
        const size_t nodes = contentSize(var_loc::nodeCentered);
        if (loc == var_loc::cellCentered) {
            for (int d = 0; d < Dim; ++d)
                for (size_t i = 0; i < nodes; ++i) fp << std::endl << lb_[d] + i2n(i)[d] * delta_[d];
            const size_t cells = contentSize(var_loc::cellCentered);
            for (int v = 0; v < vnum; ++v)
                for (size_t i = 0; i < cells; ++i) fp << std::endl << content[i * vnum + v];
        } else {
            for (size_t i = 0; i < nodes; ++i) {
                const auto p = nodePosition(i2n(i));
                fp << std::endl << p[0];
                for (int d = 1; d < Dim; ++d) fp << "\t" << p[d];
                for (int v = 0; v < vnum; ++v) fp << "\t" << content[i * vnum + v];
            }
        }
    }

    // centre of the cell of linear index i (last dimension fastest)
    __host__ __device__ std::array<double, Dim> cellCenter(size_t i) const
    {
        std::array<double, Dim> p{};
        for (int d = Dim - 1; d >= 0; --d) {
            const size_t c = i % (size_t)n_[d];
            i /= (size_t)n_[d];
            p[d] = lb_[d] + ((double)c + 0.5) * delta_[d];
        }
        return p;
    }

    // The C ABI's grid.  periodic[d] = 1 wraps particles into
    // [plo[d], plo[d] + plen[d]) (defaults: the reference periodic node
    // convention, period = cells + 1 nodes); 0 absorbs outside [lb, ub).
    psum_grid to_psum(const std::array<int, Dim>& periodic = {},
                      const double* plo = nullptr, const double* plen = nullptr) const
    {
        psum_grid g{};
        g.dim = Dim;
        for (int d = 0; d < Dim; ++d) {
            g.cells[d] = n_[d];
            g.lb[d] = lb_[d];
            g.ub[d] = ub_[d];
            g.delta[d] = delta_[d];
            g.delta_r[d] = delta_r_[d];
            g.periodic[d] = periodic[d];
            g.plo[d] = plo ? plo[d] : lb_[d] - 0.5 * delta_[d];
            g.plen[d] = plen ? plen[d] : (n_[d] + 1) * delta_[d];
        }
        return g;
    }

private:
    double lb_[Dim], ub_[Dim], delta_[Dim], delta_r_[Dim];
    int n_[Dim];
};

using grid1D = simple_grid<1>;
using grid2D = simple_grid<2>;
using grid3D = simple_grid<3>;

}  // namespace field
}  // namespace psum
