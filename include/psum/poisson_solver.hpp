// Poisson_solver_2d / _3d: the reference frontends (Poisson_solver_2d.hpp:22-411,
// Poisson_solver_3d.hpp:21-432) for the Cartesian, constant-eps case, on a
// backend found by name through the solver registry (get_solver_funcs).
// init assembles A = -div(eps grad) with identity rows at the fixed nodes
// (psum_poisson_assemble) and hands it to the backend; solve(phi, source)
// sets the Dirichlet values as replace entries and calls the backend on
// device pointers (source is modified in place, phi is the warm start).
#pragma once

#include <string>
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

    Poisson_solver() = default;
    Poisson_solver(const Poisson_solver&) = delete;
    Poisson_solver& operator=(const Poisson_solver&) = delete;
    ~Poisson_solver()
    {
        if (h_ && t_.free) t_.free(h_);
    }

    void set_options(const std::string& o)
    {
        options_ = o;
        if (h_) t_.set_options(h_, options_.c_str());
    }

    // backend: a registered solver name ("hip_multigrid_gpu", ...); the
    // assembly couples neighbours cyclically over cells+1 nodes, as the
    // reference's does, so a grid without fixed edge nodes is periodic
    void init(const std::string& backend, const Grid& g, psolverDomainKind kind,
              const fixed_boundary_list<Dim>& fbl, double eps = 8.854e-12)
    {
        if (kind != Cartesian)
            throw std::invalid_argument("Poisson_solver: only the Cartesian form is part of this build");
        t_ = get_solver_funcs(backend.c_str());
        if (!t_.init) throw std::invalid_argument("Poisson_solver: no backend named " + backend);
        if (h_) t_.free(h_);
        h_ = t_.init();
        if (!h_) throw std::runtime_error("Poisson_solver: backend init failed");
        grid_ = g;
        fbl_ = fbl;
        const psum_grid pg = g.to_psum();
        const size_t nn = g.contentSize(field::var_loc::nodeCentered);
        const std::vector<uint8_t> mask = fbl_.mask(nn);
        int64_t nnz = 0;
        check(psum_poisson_assemble(&pg, eps, mask.data(), 0, nullptr, nullptr, nullptr, &nnz));
        std::vector<unsigned long long> r(nnz), c(nnz);
        std::vector<double> v(nnz);
        check(psum_poisson_assemble(&pg, eps, mask.data(), nnz, r.data(), c.data(), v.data(), &nnz));
        if (!options_.empty()) t_.set_options(h_, options_.c_str());
        t_.set_matrix(h_, nn, nn, (unsigned long long)nnz, r.data(), c.data(), v.data());
        n_ = nn;
    }

    // phi, source: device arrays of the node count
    void solve(double* phi, double* source)
    {
        if (!h_) throw std::logic_error("Poisson_solver: solve before init");
        std::vector<unsigned long long> idx;
        std::vector<double> val;
        fbl_.evaluate(idx, val);
        t_.set_source_addback(h_, 0, nullptr, nullptr);
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

    func_table t_{};
    psum_field_solver_handle h_ = nullptr;
    std::string options_;
    Grid grid_;
    fixed_boundary_list<Dim> fbl_;
    size_t n_ = 0;
};

using Poisson_solver_2d = Poisson_solver<2>;
using Poisson_solver_3d = Poisson_solver<3>;

}  // namespace field_solver
}  // namespace psum
