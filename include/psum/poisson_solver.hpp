// Poisson_solver_2d / _3d: the reference frontends (Poisson_solver_2d.hpp:22-411,
// Poisson_solver_3d.hpp:21-432) on a backend found by name through the solver
// registry (get_solver_funcs).  init assembles A = -div(eps grad) -- fixed
// (Dirichlet) rows, mixed (Robin/Neumann) sides eliminated through ghost
// values, Cartesian or (2D) cylindrical, constant or position-dependent
// permittivity -- with psum_poisson_assemble_ex and hands it to the backend;
// solve(phi, source) evaluates the boundary values, sets
// addback = M * mixed values and the Dirichlet replace entries, and calls the
// backend on device pointers (source is modified in place, phi is the warm
// start).  The grid extents are passed to the backend as "I .. J .. [K ..]"
// ahead of the caller's options.
#pragma once

#include <functional>
#include <string>
#include <type_traits>
#include <vector>

#include "../psum_field_solver.h"
#include "../psum_hip.h"
#include "boundary_creator.hpp"
#include "grid.hpp"

namespace psum {
namespace field_solver {

template <int Dim>
class Poisson_solver {
public:
    enum psolverDomainKind { Cylindrical, Cartesian };
    using Grid = field::simple_grid<Dim>;
    // permittivity as a function of position, eps(x, y) or eps(x, y, z)
    using EpsFunc = std::conditional_t<Dim == 2, std::function<double(double, double)>,
                                       std::function<double(double, double, double)>>;

    Poisson_solver() = default;
    Poisson_solver(const Poisson_solver&) = delete;
    Poisson_solver& operator=(const Poisson_solver&) = delete;
    ~Poisson_solver()
    {
        if (h_ && t_.free) t_.free(h_);
    }

    // backend options; applied at init, after the grid extents
    void set_options(const std::string& o) { options_ = o; }

    // backend: a registered solver name ("hip_multigrid_gpu", ...).  The
    // assembly couples neighbours cyclically over cells+1 nodes, as the
    // reference's does, so a side with neither fixed nor mixed nodes is periodic.
    void init(const std::string& backend, const Grid& g, psolverDomainKind kind,
              const fixed_boundary_list<Dim>& fbl, double eps = 8.854e-12)
    {
        init_impl(backend, g, kind, mixed_boundary_list<Dim>{}, fbl, eps, EpsFunc{});
    }
    void init(const std::string& backend, const Grid& g, psolverDomainKind kind,
              const fixed_boundary_list<Dim>& fbl, const EpsFunc& eps)
    {
        init_impl(backend, g, kind, mixed_boundary_list<Dim>{}, fbl, 0.0, eps);
    }
    void init(const std::string& backend, const Grid& g, psolverDomainKind kind,
              const mixed_boundary_list<Dim>& mbl, const fixed_boundary_list<Dim>& fbl,
              double eps = 8.854e-12)
    {
        init_impl(backend, g, kind, mbl, fbl, eps, EpsFunc{});
    }
    void init(const std::string& backend, const Grid& g, psolverDomainKind kind,
              const mixed_boundary_list<Dim>& mbl, const fixed_boundary_list<Dim>& fbl, const EpsFunc& eps)
    {
        init_impl(backend, g, kind, mbl, fbl, 0.0, eps);
    }

    // phi, source: device arrays of the node count
    void solve(double* phi, double* source)
    {
        if (!h_) throw std::logic_error("Poisson_solver: solve before init");
        std::vector<unsigned long long> idx;
        std::vector<double> val;
        fbl_.evaluate(idx, val);
        std::vector<double> mv;
        mbl_.evaluate(mv);
        // addback per row of M: 0 + sum_m M(r, m) * value(m), in column order
        std::vector<unsigned long long> arow;
        std::vector<double> aval;
        for (size_t q = 0; q < mr_.size();) {
            double s = 0.0;
            size_t e = q;
            for (; e < mr_.size() && mr_[e] == mr_[q]; ++e) s += mv_[e] * mv.at(mc_[e]);
            arow.push_back(mr_[q]);
            aval.push_back(s);
            q = e;
        }
        t_.set_source_addback(h_, arow.size(), arow.data(), aval.data());
        t_.set_source_replace(h_, idx.size(), idx.data(), val.data());
        t_.solve(h_, source, phi);
    }

    const Grid& getGrid() const { return grid_; }
    size_t size() const { return n_; }

private:
    static void check(psum_status s)
    {
        if (s != PSUM_OK) throw std::runtime_error(std::string("Poisson_solver: ") + psum_hip_last_error());
    }
    static double eps_trampoline(const double* p, void* user)
    {
        const EpsFunc& f = *static_cast<const EpsFunc*>(user);
        if constexpr (Dim == 2) return f(p[0], p[1]);
        else return f(p[0], p[1], p[2]);
    }

    void init_impl(const std::string& backend, const Grid& g, psolverDomainKind kind,
                   const mixed_boundary_list<Dim>& mbl, const fixed_boundary_list<Dim>& fbl, double eps,
                   const EpsFunc& eps_fn)
    {
        if (kind == Cylindrical && Dim != 2)
            throw std::invalid_argument("Poisson_solver: the cylindrical form is 2D (z, r)");
        t_ = get_solver_funcs(backend.c_str());
        if (!t_.init) throw std::invalid_argument("Poisson_solver: no backend named " + backend);
        if (h_) t_.free(h_);
        h_ = t_.init();
        if (!h_) throw std::runtime_error("Poisson_solver: backend init failed");
        grid_ = g;
        fbl_ = fbl;
        mbl_ = mbl;
        const psum_grid pg = g.to_psum();
        const size_t nn = g.contentSize(field::var_loc::nodeCentered);
        const std::vector<uint8_t> mask = fbl_.mask(nn);
        std::vector<psum_mixed_node> nodes;
        for (const auto& m : mbl_.get_boundary_nodes()) {
            psum_mixed_node q{};
            q.element = m.element;
            q.direction = (int32_t)m.direction;
            q.value_coef = m.coeffs.first;
            q.grad_coef = m.coeffs.second;
            nodes.push_back(q);
        }
        psum_poisson_spec spec{};
        spec.cylindrical = kind == Cylindrical ? 1 : 0;
        spec.eps = eps;
        if (eps_fn) {
            spec.eps_fn = &Poisson_solver::eps_trampoline;
            spec.eps_user = const_cast<EpsFunc*>(&eps_fn);
        }
        spec.fixed = mask.data();
        spec.mixed = nodes.empty() ? nullptr : nodes.data();
        spec.n_mixed = (int64_t)nodes.size();
        int64_t nnz = 0, mnnz = 0;
        check(psum_poisson_assemble_ex(&pg, &spec, 0, nullptr, nullptr, nullptr, &nnz, 0, nullptr, nullptr,
                                       nullptr, &mnnz));
        std::vector<unsigned long long> r(nnz), c(nnz);
        std::vector<double> v(nnz);
        mr_.assign(mnnz, 0);
        mc_.assign(mnnz, 0);
        mv_.assign(mnnz, 0.0);
        check(psum_poisson_assemble_ex(&pg, &spec, nnz, r.data(), c.data(), v.data(), &nnz, mnnz, mr_.data(),
                                       mc_.data(), mv_.data(), &mnnz));
        std::string dims = "I " + std::to_string(pg.cells[0] + 1) + " J " + std::to_string(pg.cells[1] + 1);
        if (Dim == 3) dims += " K " + std::to_string(pg.cells[2] + 1);
        if (!options_.empty()) dims += " " + options_;   // later keys win
        t_.set_options(h_, dims.c_str());
        t_.set_matrix(h_, nn, nn, (unsigned long long)nnz, r.data(), c.data(), v.data());
        n_ = nn;
    }

    func_table t_{};
    psum_field_solver_handle h_ = nullptr;
    std::string options_;
    Grid grid_;
    fixed_boundary_list<Dim> fbl_;
    mixed_boundary_list<Dim> mbl_;
    std::vector<unsigned long long> mr_, mc_;   // M = -B D^-1, row-major
    std::vector<double> mv_;
    size_t n_ = 0;
};

using Poisson_solver_2d = Poisson_solver<2>;
using Poisson_solver_3d = Poisson_solver<3>;

}  // namespace field_solver
}  // namespace psum
