// Poisson_solver_1d: -d/dx (eps dphi/dx) = source on the nodes of a grid1D
// with Dirichlet nodes (fixed_boundary_1d_list) and Robin / Neumann end
// conditions (mixed_boundary_1d_list) (src/field_solver/Poisson_solver_1d.hpp:
// 21-299).  The rows are the reference's: the flux difference of the two
// half-cell gradients with eps at the half-cell points for interior nodes
// (neighbours taken cyclically), phi = value at fixed nodes, and
// v * phi +- g * dphi/dx averaged over the directions given at a mixed node,
// with second-order one-sided differences at the ends.  The "native"
// backend of the reference (a sparse LU on the host) is a banded LU with
// partial pivoting here; a 1D system is far too small for a GPU solve to
// pay for its launches.  Cyclic rows (a periodic interior end) make the
// band the whole matrix, so those systems are limited to 4096 nodes.
#pragma once

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "boundary_creator.hpp"
#include "grid.hpp"

namespace psum {
namespace field_solver {

class Poisson_solver_1d {
public:
    using Grid = field::grid1D;

    Poisson_solver_1d() = default;

    void init(const Grid& g, const mixed_boundary_1d_list& mbl, const fixed_boundary_1d_list& fbl,
              double epsilon = 8.854e-12)
    {
        init_("native", g, mbl, fbl, [epsilon](double) { return epsilon; });
    }
    void init(const Grid& g, const mixed_boundary_1d_list& mbl, const fixed_boundary_1d_list& fbl,
              const std::function<double(double)>& epsilon)
    {
        init_("native", g, mbl, fbl, epsilon);
    }
    void init(const Grid& g, const fixed_boundary_1d_list& fbl, double epsilon = 8.854e-12)
    {
        init_("native", g, {}, fbl, [epsilon](double) { return epsilon; });
    }
    void init(const Grid& g, const fixed_boundary_1d_list& fbl, const std::function<double(double)>& epsilon)
    {
        init_("native", g, {}, fbl, epsilon);
    }
    void init(const std::string& backend, const Grid& g, const mixed_boundary_1d_list& mbl,
              const fixed_boundary_1d_list& fbl, double epsilon = 8.854e-12)
    {
        init_(backend, g, mbl, fbl, [epsilon](double) { return epsilon; });
    }
    void init(const std::string& backend, const Grid& g, const fixed_boundary_1d_list& fbl,
              double epsilon = 8.854e-12)
    {
        init_(backend, g, {}, fbl, [epsilon](double) { return epsilon; });
    }

    // accepted for the reference's interface; the banded solve has no options
    void set_options(const std::string& options) { options_ = options; }

    // source becomes the right-hand side (mixed values added, fixed values
    // replaced), then phi = A^-1 source
    void solve(double* phi, double* source)
    {
        make_righthand_item(source);
        std::vector<double> x(source, source + n_);
        // row swaps interleaved with the elimination, as they were made
        for (int k = 0; k < n_; ++k) {
            if (piv_[k] != k) std::swap(x[k], x[piv_[k]]);
            for (int i = k + 1; i <= std::min(n_ - 1, k + kl_); ++i) x[i] -= at(i, k) * x[k];
        }
        for (int i = n_ - 1; i >= 0; --i) {
            double s = x[i];
            for (int j = i + 1; j <= std::min(n_ - 1, i + kl_ + ku_); ++j) s -= at(i, j) * x[j];
            x[i] = s / at(i, i);
        }
        std::copy(x.begin(), x.end(), phi);
    }
    void solve(std::vector<double>& phi, const std::vector<double>& source)
    {
        if ((int)source.size() != n_) throw std::invalid_argument("Poisson_solver_1d::solve: source size");
        std::vector<double> b = source;
        phi.resize(n_);
        solve(phi.data(), b.data());
    }

    void make_righthand_item(double* b) const
    {
        std::vector<double> mv;
        mbl_.evaluate(mv);
        for (size_t i = 0; i < addback_.size() && i < mv.size(); ++i) b[addback_[i]] += mv[i];
        std::vector<unsigned long long> fi;
        std::vector<double> fv;
        fbl_.evaluate(fi, fv);
        for (size_t i = 0; i < fi.size(); ++i) b[fi[i]] = fv[i];
    }

    // the assembled matrix as (row, column, value), row by row
    const std::vector<std::tuple<int, int, double>>& coeff_triplets() const { return coo_; }

private:
    void init_(const std::string& backend, const Grid& g, const mixed_boundary_1d_list& mbl_in,
               const fixed_boundary_1d_list& fbl, const std::function<double(double)>& eps)
    {
        if (backend != "native")
            throw std::invalid_argument("Poisson_solver_1d: only the native (host banded LU) backend solves 1D");
        grid_ = g;
        fbl_ = fbl;
        mbl_ = mbl_in;
        const int I = g.numCells<0>();
        n_ = I + 1;
        const std::vector<uint8_t> fixed = fbl_.mask((size_t)n_);
        std::vector<size_t> drop;
        for (int i = 0; i < n_; ++i)
            if (fixed[i] && mbl_.element_exists((size_t)i)) drop.push_back((size_t)i);
        mbl_.erase_at(drop);
        addback_.clear();
        for (const auto& nd : mbl_.get_boundary_nodes()) addback_.push_back(nd.element);

        const double dxr = g.del_r<0>(), dx = g.del<0>();
        auto cyc = [n = n_](int i) { return ((i % n) + n) % n; };
        coo_.clear();
        for (int i = 0; i < n_; ++i) {
            std::map<int, double> row;
            if (fixed[i]) {
                row[i] = 1.0;
            } else if (mbl_.element_exists((size_t)i)) {
                std::map<int, double> gx;
                if (i == I) {
                    gx[cyc(i - 2)] += dxr * +0.5;
                    gx[cyc(i - 1)] += dxr * -2;
                    gx[i] += dxr * +1.5;
                } else if (i == 0) {
                    gx[i] += dxr * -1.5;
                    gx[cyc(i + 1)] += dxr * +2;
                    gx[cyc(i + 2)] += dxr * -0.5;
                } else {
                    gx[cyc(i - 1)] += dxr * -0.5;
                    gx[cyc(i + 1)] += dxr * 0.5;
                }
                int dirs = 0;
                for (auto dir : {field::boundary_direction_1d::L, field::boundary_direction_1d::R}) {
                    if (!mbl_.element_exists((size_t)i, dir)) continue;
                    const auto [v, gc] = mbl_.evaluate_coeffs_at((size_t)i, dir);
                    const double sg = dir == field::boundary_direction_1d::R ? gc : -gc;
                    for (const auto& [c, a] : gx) row[c] += a * sg;
                    row[i] += 1.0 * v;
                    ++dirs;
                }
                for (auto& [c, a] : row) a *= 1.0 / dirs;
            } else {
                const double xi = g.nodePosition({i})[0];
                const double eR = eps((xi + g.nodePosition({i + 1})[0]) / 2);
                const double eL = eps((xi + g.nodePosition({i - 1})[0]) / 2);
                const double s = -1.0 / dx;
                // (grad_R eps_R - grad_L eps_L) * (-1 / dx)
                row[i] += (dxr * -1 * eR - dxr * eL) * s;
                row[cyc(i + 1)] += (dxr * eR) * s;
                row[cyc(i - 1)] += (-(dxr * -1 * eL)) * s;
            }
            for (const auto& [c, a] : row) coo_.emplace_back(i, c, a);
        }
        factor_();
    }

    // banded storage: row i keeps columns [i - kl, i + kl + ku] (room for
    // the fill-in of partial pivoting)
    double& at(int i, int j) { return band_[(size_t)i * width_ + (size_t)(j - i + kl_)]; }
    double at(int i, int j) const { return band_[(size_t)i * width_ + (size_t)(j - i + kl_)]; }

    void factor_()
    {
        kl_ = ku_ = 0;
        for (const auto& [i, j, a] : coo_) {
            kl_ = std::max(kl_, i - j);
            ku_ = std::max(ku_, j - i);
        }
        if ((kl_ > 2 || ku_ > 2) && n_ > 4096)
            throw std::invalid_argument("Poisson_solver_1d: cyclic rows above 4096 nodes are not supported");
        width_ = 2 * kl_ + ku_ + 1;
        band_.assign((size_t)n_ * width_, 0.0);
        for (const auto& [i, j, a] : coo_) at(i, j) += a;
        piv_.assign(n_, 0);
        for (int k = 0; k < n_; ++k) {
            int p = k;
            const int last = std::min(n_ - 1, k + kl_);
            for (int i = k + 1; i <= last; ++i)
                if (std::fabs(at(i, k)) > std::fabs(at(p, k))) p = i;
            piv_[k] = p;
            if (at(p, k) == 0.0) throw std::runtime_error("Poisson_solver_1d: singular matrix");
            const int jend = std::min(n_ - 1, k + kl_ + ku_);
            if (p != k)
                for (int j = k; j <= jend; ++j) std::swap(at(k, j), at(p, j));
            for (int i = k + 1; i <= last; ++i) {
                const double f = at(i, k) / at(k, k);
                at(i, k) = f;
                if (f != 0.0)
                    for (int j = k + 1; j <= jend; ++j) at(i, j) -= f * at(k, j);
            }
        }
    }

    Grid grid_{{0.0}, {1.0}, {1}};
    std::string options_;
    fixed_boundary_1d_list fbl_;
    mixed_boundary_1d_list mbl_;
    std::vector<size_t> addback_;
    std::vector<std::tuple<int, int, double>> coo_;
    int n_ = 0, kl_ = 0, ku_ = 0;
    size_t width_ = 1;
    std::vector<double> band_;
    std::vector<int> piv_;
};

}  // namespace field_solver
}  // namespace psum
