// Host-side finite-difference Poisson assembly: the matrix the reference
// frontends build before handing it to a solver backend
// (Poisson_solver_2d.hpp:182-410, Poisson_solver_3d.hpp:178-431).
//
//   row(n) = -(1/V) * sum over faces f of (+/-) eps_f * S_f * (one-sided
//            difference across f)          in the face order E W N S (U D)
//
// with identity rows at fixed (Dirichlet) nodes, cyclic neighbours over the
// cells+1 nodes otherwise, Cartesian or cylindrical (z, r) areas/volumes,
// and the permittivity sampled at face midpoints.  A mixed (Robin/Neumann)
// node gets a ghost value beyond its side; the ghost's central-difference
// boundary equation  v*phi + g*dphi/dn = value  is eliminated, giving
// A' = A - B D^-1 C and the right-hand-side map M = -B D^-1
// (ghost_element_manager.hpp:56-117).
//
// Each row is built as a small index->coefficient list with the merge rule
// of the reference's local_equation (local_equation.hpp:29-61): scaling by
// an exact zero empties the list, and a coefficient that cancels below
// 1e-14 of the magnitudes it came from is removed.  The summation orders
// (faces in E W N S U D order; ghost contributions in ascending mixed-node
// order, first product assigned; A + MC with A's value first) are those of
// the reference's expression evaluation and its sparse products, so the
// entries are bit-identical to it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace psum_hip {
namespace {

constexpr double kPi = 3.14159265358979323846;

// index -> coefficient list kept sorted by index
class Terms {
public:
    Terms() = default;
    Terms(unsigned long long i0, double c0, unsigned long long i1, double c1)
    {
        put(i0, c0);
        if (find(i1) < 0) put(i1, c1);   // a repeated index keeps the first
    }
    int size() const { return n_; }
    unsigned long long index(int k) const { return idx_[k]; }
    double coef(int k) const { return val_[k]; }

    Terms scaled(double s) const
    {
        Terms t = *this;
        for (int k = 0; k < t.n_; ++k) t.val_[k] *= s;
        if (std::fabs(s) == 0) t.n_ = 0;
        return t;
    }
    // this += sign * o, with cancellation
    void merge(const Terms& o, bool subtract)
    {
        for (int k = 0; k < o.n_; ++k) {
            const double c = o.val_[k];
            const int f = find(o.idx_[k]);
            if (f < 0) {
                put(o.idx_[k], subtract ? -c : c);
                continue;
            }
            const double mag = std::fabs(c) + std::fabs(val_[f]);
            val_[f] = subtract ? val_[f] - c : val_[f] + c;
            if (std::fabs(val_[f]) < mag * 1e-14) drop(f);
        }
    }

private:
    static constexpr int kCap = 16;
    int find(unsigned long long i) const
    {
        for (int k = 0; k < n_; ++k)
            if (idx_[k] == i) return k;
        return -1;
    }
    void put(unsigned long long i, double c)
    {
        if (n_ == kCap) throw std::runtime_error("poisson assembly: stencil row overflow");
        int k = n_++;
        for (; k > 0 && idx_[k - 1] > i; --k) {
            idx_[k] = idx_[k - 1];
            val_[k] = val_[k - 1];
        }
        idx_[k] = i;
        val_[k] = c;
    }
    void drop(int k)
    {
        for (--n_; k < n_; ++k) {
            idx_[k] = idx_[k + 1];
            val_[k] = val_[k + 1];
        }
    }
    int n_ = 0;
    unsigned long long idx_[kCap];
    double val_[kCap];
};

struct Entry {
    unsigned long long r, c;
    double v;
};

bool row_major(const Entry& a, const Entry& b) { return a.r != b.r ? a.r < b.r : a.c < b.c; }

// side of a direction code: axis and whether it faces +axis
void side_of(int D, int dir, int& axis, bool& plus)
{
    if (D == 2) {
        axis = dir < 2 ? 1 : 0;   // N S along axis 1, E W along axis 0
    } else {
        axis = dir / 2;           // Xpos Xneg Ypos Yneg Zpos Zneg
    }
    plus = (dir % 2) == 0;
}

}  // namespace

void assemble_poisson(const psum_grid& g, const psum_poisson_spec& s, std::vector<Entry>& A,
                      std::vector<Entry>& M)
{
    const int D = g.dim;
    if (D < 2 || D > 3) throw std::invalid_argument("poisson assembly: 2D or 3D grid required");
    if (s.cylindrical && D != 2) throw std::invalid_argument("poisson assembly: the cylindrical form is 2D");
    const int n3[3] = {g.cells[0] + 1, g.cells[1] + 1, D > 2 ? g.cells[2] + 1 : 1};
    const int64_t nn = (int64_t)n3[0] * n3[1] * n3[2];
    const unsigned long long NN = (unsigned long long)nn;
    auto lin = [&](const int* c) { return (unsigned long long)(((int64_t)c[0] * n3[1] + c[1]) * n3[2] + c[2]); };
    auto fixed_at = [&](unsigned long long i) { return s.fixed && s.fixed[i] != 0; };
    auto node_pos = [&](const int* c, double* p) {
        for (int d = 0; d < D; ++d) p[d] = g.lb[d] + c[d] * g.delta[d];
    };

    // mixed nodes per element, in argument order; those on fixed nodes drop out
    const int64_t nm = s.mixed ? s.n_mixed : 0;
    std::vector<int64_t> first(nn, -1), next(nm > 0 ? nm : 1, -1), last(nn, -1);
    for (int64_t m = 0; m < nm; ++m) {
        const psum_mixed_node& q = s.mixed[m];
        if (q.element >= NN || q.direction < 0 || q.direction >= 2 * D)
            throw std::invalid_argument("poisson assembly: mixed node " + std::to_string(m) + " out of range");
        if (fixed_at(q.element)) continue;
        if (last[q.element] < 0) first[q.element] = m;
        else next[last[q.element]] = m;
        last[q.element] = m;
    }

    std::vector<Terms> ghost_eq(nm > 0 ? nm : 1);          // per mixed node
    std::vector<int64_t> ghost_of(nm > 0 ? nm : 1, -1);
    std::vector<Entry> inner, ghost_terms;                 // ghost_terms: (row, ghost, coef)
    inner.reserve((size_t)nn * (2 * D + 1));
    int64_t n_ghost = 0;
    const double dr[3] = {g.delta_r[0], g.delta_r[1], D > 2 ? g.delta_r[2] : 0.0};
    const int face_axis[6] = {0, 0, 1, 1, 2, 2};

    int c[3];
    for (c[0] = 0; c[0] < n3[0]; ++c[0])
        for (c[1] = 0; c[1] < n3[1]; ++c[1])
            for (c[2] = 0; c[2] < n3[2]; ++c[2]) {
                const unsigned long long row = lin(c);
                Terms eq;
                if (fixed_at(row)) {
                    eq = Terms(row, 1.0, row, 1.0);
                } else {
                    unsigned long long nb[3][2];   // [axis][minus, plus]
                    for (int a = 0; a < D; ++a)
                        for (int p = 0; p < 2; ++p) {
                            int o[3] = {c[0], c[1], c[2]};
                            o[a] = ((o[a] + (p ? 1 : -1)) % n3[a] + n3[a]) % n3[a];
                            nb[a][p] = lin(o);
                        }
                    for (int64_t m = first[row]; m >= 0; m = next[m]) {
                        int axis;
                        bool plus;
                        side_of(D, s.mixed[m].direction, axis, plus);
                        ghost_of[m] = n_ghost++;
                        nb[axis][plus] = NN + (unsigned long long)ghost_of[m];
                    }
                    for (int64_t m = first[row]; m >= 0; m = next[m]) {
                        int axis;
                        bool plus;
                        side_of(D, s.mixed[m].direction, axis, plus);
                        const Terms central(nb[axis][0], dr[axis] * -0.5, nb[axis][1], dr[axis] * 0.5);
                        Terms e = central.scaled(plus ? s.mixed[m].grad_coef : -s.mixed[m].grad_coef);
                        e.merge(Terms(row, 1.0, row, 1.0).scaled(s.mixed[m].value_coef), false);
                        ghost_eq[m] = e;
                    }
                    // face areas (minus, plus side) and the control volume
                    double area[3][2], vol;
                    if (s.cylindrical) {
                        double p0[3], pu[3], pl[3];
                        const int up[3] = {c[0], c[1] + 1, 0}, dn[3] = {c[0], c[1] - 1, 0};
                        node_pos(c, p0);
                        node_pos(up, pu);
                        node_pos(dn, pl);
                        const double ru = (p0[1] + pu[1]) / 2;
                        double rl = (p0[1] + pl[1]) / 2;
                        if (rl < 0) rl = 0;
                        area[1][1] = g.delta[0] * kPi * 2 * ru;
                        area[1][0] = g.delta[0] * kPi * 2 * rl;
                        area[0][0] = area[0][1] = kPi * (ru * ru - rl * rl);
                        vol = g.delta[0] * kPi * (ru * ru - rl * rl);
                    } else if (D == 2) {
                        area[1][0] = area[1][1] = g.delta[0];
                        area[0][0] = area[0][1] = g.delta[1];
                        vol = g.delta[0] * g.delta[1];
                    } else {
                        area[1][0] = area[1][1] = g.delta[0] * g.delta[2];
                        area[0][0] = area[0][1] = g.delta[1] * g.delta[2];
                        area[2][0] = area[2][1] = g.delta[0] * g.delta[1];
                        vol = g.delta[0] * g.delta[1] * g.delta[2];
                    }
                    for (int f = 0; f < 2 * D; ++f) {
                        const int a = face_axis[f];
                        const bool plus = (f % 2) == 0;
                        const Terms diff = plus ? Terms(row, dr[a] * -1, nb[a][1], dr[a])
                                                : Terms(nb[a][0], dr[a] * -1, row, dr[a]);
                        double eps = s.eps;
                        if (s.eps_fn) {
                            int o[3] = {c[0], c[1], c[2]};
                            o[a] += plus ? 1 : -1;
                            double pa[3], pb[3], mid[3];
                            node_pos(c, pa);
                            node_pos(o, pb);
                            for (int d = 0; d < D; ++d) mid[d] = (pa[d] + pb[d]) / 2;
                            eps = s.eps_fn(mid, s.eps_user);
                        }
                        const Terms flux = diff.scaled(area[a][plus ? 1 : 0]).scaled(eps);
                        if (f == 0) eq = flux;
                        else eq.merge(flux, !plus);
                    }
                    eq = eq.scaled(-1.0 / vol);
                }
                for (int k = 0; k < eq.size(); ++k) {
                    const unsigned long long col = eq.index(k);
                    if (col >= NN) ghost_terms.push_back({row, col - NN, eq.coef(k)});
                    else inner.push_back({row, col, eq.coef(k)});
                }
            }

    // D^-1: every ghost must appear in exactly one boundary equation
    std::vector<int64_t> owner(n_ghost > 0 ? n_ghost : 1, -1), seen(n_ghost > 0 ? n_ghost : 1, 0);
    std::vector<double> dinv(n_ghost > 0 ? n_ghost : 1, 0.0);
    for (int64_t m = 0; m < nm; ++m) {
        if (ghost_of[m] < 0) continue;
        const Terms& e = ghost_eq[m];
        for (int k = 0; k < e.size(); ++k)
            if (e.index(k) >= NN) {
                const int64_t q = (int64_t)(e.index(k) - NN);
                ++seen[q];
                owner[q] = m;
                dinv[q] = 1.0 / e.coef(k);
            }
    }
    for (int64_t q = 0; q < n_ghost; ++q)
        if (seen[q] != 1)
            throw std::invalid_argument(
                "poisson assembly: unsupported mixed boundary (a ghost value is not fixed by exactly one "
                "boundary equation; grad_coef must be non-zero)");

    // M = -B D^-1, row-major
    M.clear();
    for (const Entry& t : ghost_terms) M.push_back({t.r, (unsigned long long)owner[t.c], (-t.v) * dinv[t.c]});
    std::sort(M.begin(), M.end(), row_major);

    // A' = A + M C: per row of M, the C rows of its mixed nodes in ascending
    // order (inner is row-major by construction)
    std::vector<Entry> extra;   // entries of M C not in A
    size_t ia = 0;
    for (size_t q0 = 0; q0 < M.size();) {
        size_t q1 = q0;
        while (q1 < M.size() && M[q1].r == M[q0].r) ++q1;
        const unsigned long long row = M[q0].r;
        std::vector<std::pair<unsigned long long, double>> mc;
        for (size_t q = q0; q < q1; ++q) {
            const Terms& e = ghost_eq[M[q].c];
            for (int k = 0; k < e.size(); ++k) {
                if (e.index(k) >= NN) continue;
                const double x = M[q].v * e.coef(k);
                auto it = std::find_if(mc.begin(), mc.end(), [&](const auto& p) { return p.first == e.index(k); });
                if (it == mc.end()) mc.push_back({e.index(k), x});
                else it->second += x;
            }
        }
        while (ia < inner.size() && inner[ia].r < row) ++ia;
        for (const auto& p : mc) {
            size_t k = ia;
            while (k < inner.size() && inner[k].r == row && inner[k].c != p.first) ++k;
            if (k < inner.size() && inner[k].r == row) inner[k].v = inner[k].v + p.second;
            else extra.push_back({row, p.first, p.second});
        }
        q0 = q1;
    }
    A = std::move(inner);
    if (!extra.empty()) {
        A.insert(A.end(), extra.begin(), extra.end());
        std::sort(A.begin(), A.end(), row_major);
    }
}

// Cartesian constant-eps matrix of the stepper (no mixed nodes)
void assemble(const psum_grid& g, double eps, const std::vector<uint8_t>& fixed,
              std::vector<unsigned long long>& R, std::vector<unsigned long long>& C,
              std::vector<double>& V)
{
    psum_poisson_spec s{};
    s.eps = eps;
    s.fixed = fixed.empty() ? nullptr : fixed.data();
    std::vector<Entry> A, M;
    assemble_poisson(g, s, A, M);
    R.resize(A.size());
    C.resize(A.size());
    V.resize(A.size());
    for (size_t k = 0; k < A.size(); ++k) {
        R[k] = A[k].r;
        C[k] = A[k].c;
        V[k] = A[k].v;
    }
    std::vector<Entry>().swap(A);
}

}  // namespace psum_hip

extern "C" {

psum_status psum_poisson_assemble_ex(const psum_grid* g, const psum_poisson_spec* spec, int64_t cap,
                                     unsigned long long* rows, unsigned long long* cols, double* vals,
                                     int64_t* nnz, int64_t mcap, unsigned long long* mrows,
                                     unsigned long long* mcols, double* mvals, int64_t* mnnz)
{
    using namespace psum_hip;
    PSUM_ARG_CHECK(g && spec && grid_ok(g) && nnz && mnnz, "psum_poisson_assemble_ex: bad argument");
    PSUM_ARG_CHECK(spec->n_mixed >= 0 && (spec->n_mixed == 0 || spec->mixed),
                   "psum_poisson_assemble_ex: bad mixed node list");
    try {
        std::vector<Entry> A, M;
        assemble_poisson(*g, *spec, A, M);
        *nnz = (int64_t)A.size();
        *mnnz = (int64_t)M.size();
        if (rows && cols && vals) {
            if (cap < (int64_t)A.size()) throw std::invalid_argument("psum_poisson_assemble_ex: capacity too small");
            for (size_t k = 0; k < A.size(); ++k) {
                rows[k] = A[k].r;
                cols[k] = A[k].c;
                vals[k] = A[k].v;
            }
        }
        if (mrows && mcols && mvals) {
            if (mcap < (int64_t)M.size()) throw std::invalid_argument("psum_poisson_assemble_ex: M capacity too small");
            for (size_t k = 0; k < M.size(); ++k) {
                mrows[k] = M[k].r;
                mcols[k] = M[k].c;
                mvals[k] = M[k].v;
            }
        }
        return PSUM_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return PSUM_ERR_ARG;
    } catch (const std::exception& e) {
        set_error(e.what());
        return PSUM_ERR_UNSUPPORTED;
    }
}

psum_status psum_poisson_assemble(const psum_grid* g, double eps, const uint8_t* fixed, int64_t cap,
                                  unsigned long long* rows, unsigned long long* cols, double* vals,
                                  int64_t* nnz)
{
    psum_poisson_spec s{};
    s.eps = eps;
    s.fixed = fixed;
    int64_t mnnz = 0;
    return psum_poisson_assemble_ex(g, &s, cap, rows, cols, vals, nnz, 0, nullptr, nullptr, nullptr, &mnnz);
}

}  // extern "C"
