// Poisson_solver_1d (psum/poisson_solver_1d.hpp) on the reference's own
// cases (test/field_solver/test_psolver_1d.cpp): a linear Dirichlet
// solution, x^2 with a Neumann and with Robin ends, second-order convergence
// on sin(pi x), and the discrete residual of a system with cyclic rows.
// Host code only (no GPU call); prints one line per check and "ALL OK".
#include <cmath>
#include <cstdio>
#include <functional>
#include <vector>

#include "psum/psum.hpp"

using namespace psum::prelude;

static int fails = 0;
static void report(const char* what, bool ok, double v)
{
    std::printf("%-64s %s (%.3e)\n", what, ok ? "ok" : "FAIL", v);
    if (!ok) ++fails;
}

int main()
{
    // ---- Poisson_solver_1d: the reference's own cases (test_psolver_1d.cpp)
    {
        namespace bc = boundary_creator;
        using D1 = boundary_direction_1d;
        const double e0 = 8.854e-12;
        auto run = [&](Poisson_solver_1d& s, const grid1D& g, const std::function<double(double)>& rho_of,
                       const std::function<double(double)>& exact) {
            host_node_field1D<double> phi(g), rho(g);
            rho.for_each([&](size_t, double& v, const vec<1>& x) { v = rho_of(x[0]); });
            s.solve(phi.data(), rho.data());
            double e = 0;
            phi.for_each([&](size_t, double& v, const vec<1>& x) { e = std::max(e, std::fabs(v - exact(x[0]))); });
            return e;
        };
        grid1D g({0.0}, {1.0}, {100});
        Poisson_solver_1d s1;
        s1.init(g, fixed_boundary_1d_list{bc::Dirichlet_point(g, D1::L) = 0, bc::Dirichlet_point(g, D1::R) = 1});
        const double e1 = run(s1, g, [](double) { return 0.0; }, [](double x) { return x; });
        Poisson_solver_1d s2;
        s2.init(g, mixed_boundary_1d_list{bc::Neumann_point(g, D1::L) = 0},
                fixed_boundary_1d_list{bc::Dirichlet_point(g, D1::R) = 1});
        const double e2 = run(s2, g, [&](double) { return -e0 * 2.0; }, [](double x) { return x * x; });
        Poisson_solver_1d s3;
        s3.init(g, mixed_boundary_1d_list{bc::Robin_point(g, D1::L, {1.0, 1.0}) = 0.0,
                                          bc::Robin_point(g, D1::R, {1.0, 1.0}) = 3.0},
                fixed_boundary_1d_list{});
        const double e3 = run(s3, g, [&](double) { return -e0 * 2.0; }, [](double x) { return x * x; });
        // manufactured sin(pi x): second order
        std::vector<double> l2;
        for (int k = 1; k <= 6; k += 5) {
            grid1D gk({0.0}, {1.0}, {40 * k});
            Poisson_solver_1d sk;
            sk.init(gk, fixed_boundary_1d_list{bc::Dirichlet_point(gk, D1::L) = 0, bc::Dirichlet_point(gk, D1::R) = 0});
            host_node_field1D<double> phi(gk), rho(gk);
            // the reference test's own lambda shape: (i, value, position) with pos.x()
            rho.for_each([&](size_t, double& v, auto pos) {
                const double x = pos.x();
                v = -e0 * (-M_PI * M_PI * std::sin(M_PI * x));
            });
            sk.solve(phi.data(), rho.data());
            double s = 0;
            phi.for_each([&](size_t, double& v, const vec<1>& x) {
                s += (v - std::sin(M_PI * x[0])) * (v - std::sin(M_PI * x[0]));
            });
            l2.push_back(std::sqrt(s / phi.size()));
        }
        const double order = std::log(l2[0] / l2[1]) / std::log(6.0);
        // a periodic interior (cyclic rows) with one pinned node: phi = cos(2 pi x) shape
        grid1D gp({0.0}, {1.0}, {64});
        Poisson_solver_1d sp;
        sp.init(gp, fixed_boundary_1d_list{bc::Dirichlet_range(gp, 0.5, 0.5) = 0.0}, 1.0);
        host_node_field1D<double> pp(gp), rp(gp);
        const double per = 65.0 / 64.0;   // the cyclic row wraps the I + 1 nodes
        rp.for_each([&](size_t, double& v, const vec<1>& x) { v = std::sin(2 * M_PI * x[0] / per); });
        sp.solve(pp.data(), rp.data());
        // residual of the cyclic rows (the discrete equation holds)
        double res = 0;
        std::vector<double> rr(gp.contentSize(var_loc::nodeCentered));
        for (const auto& [i, j, a] : sp.coeff_triplets()) rr[i] += a * pp(j);
        rp.for_each([&](size_t i, double& v, const vec<1>& x) {
            if (std::fabs(x[0] - 0.5) > 1e-9) res = std::max(res, std::fabs(rr[i] - std::sin(2 * M_PI * x[0] / per)));
        });
        std::printf("  1D: linear %.2e, Neumann %.2e, Robin %.2e, order %.3f, cyclic residual %.2e\n", e1, e2, e3, order,
                    res);
        report("Poisson_solver_1d: Dirichlet / Neumann / Robin / order 2 / cyclic",
               e1 < 1e-10 && e2 < 1e-8 && e3 < 1e-8 && order > 1.9 && res < 1e-9, order);
    }
    std::printf(fails ? "FAILED %d\n" : "ALL OK\n", fails);
    return fails ? 1 : 0;
}
