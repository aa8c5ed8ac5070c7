// Native step engine: the reference user loop (TSI/main.cpp:121-146,
// big_bang/stage_6.cpp:125-147) as one C++ object driving the HIP kernels
// on one stream, with all state resident in HBM.
//
// Per step s:
//   rho  = acc * 2^-S + rho_bg             (acc: exact int64 deposit of x_s)
//   all-periodic grids: rho -= mean(rho) in FP64, node 0 pinned to 0
//   phi  = HIP multigrid (FP64, or FP32 V-cycles in FP64 refinement), warm start
//   hist = field energy(phi), kinetic energy and momentum of v_s
//   acc  = 0; per species: gather grad(phi) at x_s, kick, drift, boundary,
//          deposit x_{s+1} into acc      (one fused kernel)
//   every sort_every steps: bin + stable sort + permute (+ drop absorbed)
// The deposit of x_{s+1} inside the push is the same integer sum as a
// separate deposit at the start of step s+1, so rho is bit-identical.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace psum_hip {
class HipMultigrid;
HipMultigrid* mg_create(bool mixed);
void mg_destroy(HipMultigrid* h);
void mg_set_options(HipMultigrid* h, const char* o);
void mg_set_matrix(HipMultigrid* h, unsigned long long n, unsigned long long nnz,
                   const unsigned long long* r, const unsigned long long* c, const double* v);
void mg_set_replace(HipMultigrid* h, unsigned long long n, const unsigned long long* i, const double* v);
void mg_set_stream(HipMultigrid* h, hipStream_t s);
void mg_solve(HipMultigrid* h, double* b, double* x, bool sync);
int mg_last_cycles(HipMultigrid* h);
double mg_last_rel(HipMultigrid* h);
void mg_solve_homogeneous(HipMultigrid* h, double* b, double* x);
int mg_predicted(HipMultigrid* h);
void mg_set_predicted(HipMultigrid* h, int p);
}  // namespace psum_hip

extern "C" {
psum_status psum_hip_adjoint_push(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                  void* const* lv, int64_t n, const void* phi, double qm, double dt,
                                  const double* scale_dev, int64_t* acc, void* stream);
int64_t psum_hip_adjoint_blocks(int64_t n);
psum_status psum_hip_adjoint_gather(const psum_grid* g, int dtype, void* const* x, void* const* lx,
                                    void* const* lv, int64_t n, const void* lrho, double w_dep,
                                    double dt, double* bmax, void* stream);
psum_status psum_hip_adjoint_scale(const double* bmax, int64_t nb, double factor, double* scale_dev,
                                   int accumulate, void* stream);
psum_status psum_hip_fixed_to_double_dev(const int64_t* acc, int64_t nn, const double* scale_dev,
                                         double* out, void* stream);
psum_status psum_hip_field_energy_grad(const psum_grid* g, const double* phi, double eps, double* out,
                                       void* stream);
}

using namespace psum_hip;

namespace {

constexpr int kT = 256;

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

__global__ __launch_bounds__(kT) void k_zero_i64(long long* __restrict__ a, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < n) a[i] = 0;
}

// partial sums of x over blocks (part[b])
__global__ __launch_bounds__(kT) void k_sum_part(const double* __restrict__ x, int64_t n,
                                                 double* __restrict__ part)
{
    __shared__ double ws[kT / 64];
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    double v = i < n ? x[i] : 0.0;
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int w = 0; w < kT / 64; ++w) s += ws[w];
        part[blockIdx.x] = s;
    }
}

__global__ __launch_bounds__(kT) void k_sum_final(const double* __restrict__ part, int64_t nb,
                                                  double* __restrict__ out)
{
    __shared__ double sh[kT];
    double s = 0.0;
    for (int64_t b = threadIdx.x; b < nb; b += kT) s += part[b];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = kT / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// warm start by linear extrapolation in time: phi <- 2 phi - phi_prev, phi_prev <- phi
__global__ __launch_bounds__(kT) void k_extrapolate(double* __restrict__ phi,
                                                    double* __restrict__ prev, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const double t = phi[i];
    phi[i] = 2.0 * t - prev[i];
    prev[i] = t;
}

// x -= sum / n  (FP64 neutralisation of a periodic charge density)
__global__ __launch_bounds__(kT) void k_sub_mean(double* __restrict__ x, int64_t n,
                                                 const double* __restrict__ sum)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    const double mean = sum[0] / (double)n;
    if (i < n) x[i] = x[i] - mean;
}

// lphi[0] -= sum(lphi): the consistent right-hand side of the adjoint of a
// pinned periodic solve
__global__ void k_sub_at0(double* __restrict__ x, const double* __restrict__ sum)
{
    if (threadIdx.x == 0) x[0] = x[0] - sum[0];
}

__global__ __launch_bounds__(kT) void k_zero_any(unsigned char* __restrict__ a, int64_t nbytes)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < nbytes) a[i] = 0;
}

// hist[5*s + k]: field energy, then species-summed kinetic energy and momentum
__global__ void k_hist_record(const double* __restrict__ fe, const double* __restrict__ red,
                              int nspecies, double* __restrict__ hist, int s)
{
    if (threadIdx.x != 0) return;
    hist[5 * s + 0] = fe[0];
    for (int k = 0; k < 4; ++k) {
        double t = 0.0;
        for (int q = 0; q < nspecies; ++q) t += red[4 * q + k];
        hist[5 * s + 1 + k] = t;
    }
}

__global__ void k_count_valid(const uint32_t* __restrict__ sorted_keys, int64_t n,
                              unsigned long long* __restrict__ out)
{
    // first index whose key is PSUM_KEY_INVALID (keys are sorted, invalid last)
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i >= n) return;
    const bool inv = sorted_keys[i] == PSUM_KEY_INVALID;
    const bool prev_ok = i == 0 || sorted_keys[i - 1] != PSUM_KEY_INVALID;
    if (inv && prev_ok) out[0] = (unsigned long long)i;
}

int fx_shift(double w)
{
    int e = 0;
    std::frexp(std::fabs(w), &e);
    return 40 - e;
}

struct Species {
    psum_species sp{};
    void* pos[3] = {nullptr, nullptr, nullptr};
    void* vel[3] = {nullptr, nullptr, nullptr};
    void* pos2[3] = {nullptr, nullptr, nullptr};
    void* vel2[3] = {nullptr, nullptr, nullptr};
    int64_t n = 0;
    double wdep = 0.0;
};

// Cartesian FD Poisson rows A = -div(eps grad) with identity rows at fixed
// nodes, in the value order of Poisson_solver_{2,3}d::_init_impl
// (Poisson_solver_2d.hpp:221-349, Poisson_solver_3d.hpp:228-350) for
// constant eps: interior row = (gE sE e - gW sW e + gN sN e - gS sS e
// [+ gU sU e - gD sD e]) * (-1/V), cyclic neighbours.
void assemble(const psum_grid& g, double eps, const std::vector<uint8_t>& fixed,
              std::vector<unsigned long long>& R, std::vector<unsigned long long>& C,
              std::vector<double>& V)
{
    const int D = g.dim;
    const int I = g.cells[0], J = D > 1 ? g.cells[1] : 0, K = D > 2 ? g.cells[2] : 0;
    const int64_t nJ = J + 1, nK = K + 1;
    auto cyc = [](int a, int n) { return ((a % n) + n) % n; };
    struct Term { unsigned long long i; double v; };
    for (int i = 0; i <= I; ++i)
        for (int j = 0; j <= J; ++j)
            for (int k = 0; k <= K; ++k) {
                const unsigned long long idx = (unsigned long long)(((int64_t)i * nJ + j) * nK + k);
                if (fixed[idx]) {
                    R.push_back(idx); C.push_back(idx); V.push_back(1.0);
                    continue;
                }
                std::vector<Term> t;
                auto acc = [&](unsigned long long q, double v, bool sub) {
                    for (auto& e : t)
                        if (e.i == q) {
                            const double old = std::fabs(v) + std::fabs(e.v);
                            e.v = sub ? e.v - v : e.v + v;
                            if (std::fabs(e.v) < old * 1e-14) {
                                e.v = 0.0;
                                e.i = ~0ull;   // erased (local_equation.hpp:35-38)
                            }
                            return;
                        }
                    t.push_back({q, sub ? -v : v});
                };
                struct Face { unsigned long long nb; double dr, s; bool plus; };
                std::vector<Face> faces;
                const double dxr = g.delta_r[0];
                if (D == 2) {
                    const double sNS = g.delta[0], sEW = g.delta[1];
                    const double dyr = g.delta_r[1];
                    faces = {{(unsigned long long)cyc(i + 1, I + 1) * nJ + j, dxr, sEW, true},
                             {(unsigned long long)cyc(i - 1, I + 1) * nJ + j, dxr, sEW, false},
                             {(unsigned long long)i * nJ + cyc(j + 1, J + 1), dyr, sNS, true},
                             {(unsigned long long)i * nJ + cyc(j - 1, J + 1), dyr, sNS, false}};
                } else {
                    const double sNS = g.delta[0] * g.delta[2], sEW = g.delta[1] * g.delta[2];
                    const double sUD = g.delta[0] * g.delta[1];
                    const double dyr = g.delta_r[1], dzr = g.delta_r[2];
                    faces = {{(unsigned long long)((cyc(i + 1, I + 1) * nJ + j) * nK + k), dxr, sEW, true},
                             {(unsigned long long)((cyc(i - 1, I + 1) * nJ + j) * nK + k), dxr, sEW, false},
                             {(unsigned long long)(((int64_t)i * nJ + cyc(j + 1, J + 1)) * nK + k), dyr, sNS, true},
                             {(unsigned long long)(((int64_t)i * nJ + cyc(j - 1, J + 1)) * nK + k), dyr, sNS, false},
                             {(unsigned long long)(((int64_t)i * nJ + j) * nK + cyc(k + 1, K + 1)), dzr, sUD, true},
                             {(unsigned long long)(((int64_t)i * nJ + j) * nK + cyc(k - 1, K + 1)), dzr, sUD, false}};
                }
                const double vol = D == 2 ? g.delta[0] * g.delta[1] : g.delta[0] * g.delta[1] * g.delta[2];
                for (size_t f = 0; f < faces.size(); ++f) {
                    const Face& F = faces[f];
                    // grad_plus = {idx: -dr, nb: dr}, grad_minus = {nb: -dr, idx: dr}
                    unsigned long long i0, i1;
                    double c0, c1;
                    if (F.plus) { i0 = idx; c0 = F.dr * -1; i1 = F.nb; c1 = F.dr; }
                    else { i0 = F.nb; c0 = F.dr * -1; i1 = idx; c1 = F.dr; }
                    const bool sub = (f % 2) == 1;    // E - W + N - S (+ U - D)
                    const double a0 = (c0 * F.s) * eps, a1 = (c1 * F.s) * eps;
                    if (f == 0) {
                        t.push_back({i0, a0});
                        if (i1 != i0) t.push_back({i1, a1});
                    } else {
                        acc(i0, a0, sub);
                        if (i1 != i0) acc(i1, a1, sub);
                    }
                }
                const double sc = -1.0 / vol;
                for (auto& e : t) {
                    if (e.i == ~0ull) continue;
                    R.push_back(idx); C.push_back(e.i); V.push_back(e.v * sc);
                }
            }
}

class Stepper {
public:
    explicit Stepper(const psum_step_config& c) : cfg_(c)
    {
        if (!grid_ok(&cfg_.grid) || cfg_.grid.dim < 2)
            throw std::invalid_argument("psum_stepper: grid must be 2D or 3D");
        if (cfg_.nspecies < 1 || cfg_.nspecies > 4) throw std::invalid_argument("psum_stepper: 1..4 species");
        if (cfg_.dtype != PSUM_F64 && cfg_.dtype != PSUM_F32) throw std::invalid_argument("psum_stepper: dtype");
        esz_ = cfg_.dtype == PSUM_F64 ? 8 : 4;
        nn_ = grid_nodes(cfg_.grid);
        all_periodic_ = true;
        for (int d = 0; d < cfg_.grid.dim; ++d) all_periodic_ &= cfg_.grid.periodic[d] != 0;
        S_TRY(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));

        // fixed nodes and the frontend's matrix
        std::vector<uint8_t> fixed(nn_, 0);
        if (cfg_.fixed) std::memcpy(fixed.data(), cfg_.fixed, nn_);
        if (all_periodic_) fixed[0] = 1;
        std::vector<unsigned long long> ri, ci, rep_idx;
        std::vector<double> vv, rep_val;
        assemble(cfg_.grid, cfg_.eps, fixed, ri, ci, vv);
        for (int64_t i = 0; i < nn_; ++i)
            if (fixed[i]) {
                rep_idx.push_back((unsigned long long)i);
                rep_val.push_back(cfg_.fixed_val && !(all_periodic_ && i == 0) ? cfg_.fixed_val[i] : 0.0);
            }
        mg_ = mg_create(cfg_.mixed_mg != 0);
        const int D = cfg_.grid.dim;
        std::string opt = "I " + std::to_string(cfg_.grid.cells[0] + 1) + " J " + std::to_string(cfg_.grid.cells[1] + 1);
        if (D == 3) opt += " K " + std::to_string(cfg_.grid.cells[2] + 1);
        opt += " depth " + std::to_string(cfg_.depth) + " pre_relax " + std::to_string(cfg_.pre) +
               " post_relax " + std::to_string(cfg_.post) + " max_cycles " + std::to_string(cfg_.max_cycles);
        if (!cfg_.mg_graph) opt += " graph 0";
        if (cfg_.tol > 0) {
            char buf[64];
            std::snprintf(buf, sizeof buf, " tol %.17g", cfg_.tol);
            opt += buf;
        }
        mg_set_options(mg_, opt.c_str());
        mg_set_matrix(mg_, (unsigned long long)nn_, ri.size(), ri.data(), ci.data(), vv.data());
        mg_set_replace(mg_, rep_idx.size(), rep_idx.data(), rep_val.data());
        mg_set_stream(mg_, st_);

        // species and the deposit scale
        double dV = 1.0;
        for (int d = 0; d < D; ++d) dV *= cfg_.grid.delta[d];
        double wmax = 0.0;
        for (int s = 0; s < cfg_.nspecies; ++s) {
            Species& sp = sp_[s];
            sp.sp = cfg_.species[s];
            if (sp.sp.capacity < sp.sp.n) sp.sp.capacity = sp.sp.n;
            sp.wdep = sp.sp.q * sp.sp.w / dV;
            wmax = std::max(wmax, std::fabs(sp.wdep));
            const size_t bytes = (size_t)std::max<int64_t>(1, sp.sp.capacity) * esz_;
            for (int d = 0; d < D; ++d) {
                S_TRY(hipMalloc(&sp.pos[d], bytes));
                S_TRY(hipMalloc(&sp.vel[d], bytes));
                if (cfg_.sort_every > 0) {
                    S_TRY(hipMalloc(&sp.pos2[d], bytes));
                    S_TRY(hipMalloc(&sp.vel2[d], bytes));
                }
            }
            max_cap_ = std::max(max_cap_, sp.sp.capacity);
        }
        scale_ = std::ldexp(1.0, fx_shift(wmax > 0 ? wmax : 1.0));

        S_TRY(hipMalloc(&acc_, nn_ * sizeof(int64_t)));
        S_TRY(hipMalloc(&b_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&phi_, nn_ * sizeof(double)));
        if (cfg_.extrapolate) S_TRY(hipMalloc(&phi_prev_, nn_ * sizeof(double)));
        S_TRY(hipMemsetAsync(phi_, 0, nn_ * sizeof(double), st_));
        if (cfg_.dtype == PSUM_F32) S_TRY(hipMalloc(&phi32_, nn_ * sizeof(float)));
        const int64_t nbp = std::max<int64_t>(1, psum_hip_push_blocks(max_cap_));
        S_TRY(hipMalloc(&diag_, (size_t)nbp * 4 * sizeof(double)));
        S_TRY(hipMalloc(&red_, 16 * sizeof(double)));
        S_TRY(hipMalloc(&redw_, psum_hip_reduce_workspace()));
        S_TRY(hipMalloc(&fe_, 4 * sizeof(double)));
        few_bytes_ = psum_hip_field_energy_workspace(&cfg_.grid);
        S_TRY(hipMalloc(&few_, few_bytes_));
        S_TRY(hipMalloc(&sum_part_, (size_t)nblk(nn_) * sizeof(double)));
        S_TRY(hipMalloc(&sum_, 8 * sizeof(double)));
        S_TRY(hipMalloc(&counters_, 8 * sizeof(int64_t)));
        S_TRY(hipMemsetAsync(counters_, 0, 8 * sizeof(int64_t), st_));
        if (cfg_.sort_every > 0) {
            S_TRY(hipMalloc(&keys_, (size_t)max_cap_ * 4 + 4));
            S_TRY(hipMalloc(&skeys_, (size_t)max_cap_ * 4 + 4));
            S_TRY(hipMalloc(&perm_, (size_t)max_cap_ * 4 + 4));
            sort_bytes_ = psum_hip_sort_workspace(max_cap_);
            S_TRY(hipMalloc(&sort_work_, sort_bytes_));
            S_TRY(hipMalloc(&nvalid_, sizeof(unsigned long long)));
        }
        S_TRY(hipStreamSynchronize(st_));
    }

    ~Stepper()
    {
        if (mg_) mg_destroy(mg_);
        for (auto& sp : sp_)
            for (int d = 0; d < 3; ++d)
                for (void* p : {sp.pos[d], sp.vel[d], sp.pos2[d], sp.vel2[d]})
                    if (p) (void)hipFree(p);
        for (void* p : {(void*)acc_, (void*)b_, (void*)phi_, (void*)phi_prev_, (void*)phi32_, (void*)diag_, (void*)red_,
                        (void*)fe_, few_, redw_, (void*)sum_part_, (void*)sum_, (void*)counters_,
                        (void*)keys_, (void*)skeys_, (void*)perm_, sort_work_, (void*)nvalid_, (void*)hist_})
            if (p) (void)hipFree(p);
        for (auto& e : ev_)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)lphi_, (void*)lrho_, (void*)adj_scale_, (void*)bmax_, lrho_t_, (void*)adj_acc_})
            if (p) (void)hipFree(p);
        for (void* p : lx_) if (p) (void)hipFree(p);
        for (void* p : lv_) if (p) (void)hipFree(p);
        for (auto& v : seg_x_) for (void* p : v) if (p) (void)hipFree(p);
        for (void* p : seg_phi_) if (p) (void)hipFree(p);
        if (st_) (void)hipStreamDestroy(st_);
    }

    void load(int s, const void* const* pos, const void* const* vel, int64_t n)
    {
        if (s < 0 || s >= cfg_.nspecies) throw std::invalid_argument("psum_stepper_load: bad species");
        Species& sp = sp_[s];
        if (n > sp.sp.capacity) throw std::invalid_argument("psum_stepper_load: n exceeds capacity");
        for (int d = 0; d < cfg_.grid.dim; ++d) {
            S_TRY(hipMemcpyAsync(sp.pos[d], pos[d], (size_t)n * esz_, hipMemcpyHostToDevice, st_));
            S_TRY(hipMemcpyAsync(sp.vel[d], vel[d], (size_t)n * esz_, hipMemcpyHostToDevice, st_));
        }
        sp.n = n;
        acc_valid_ = false;
        S_TRY(hipStreamSynchronize(st_));
    }

    void state(int s, void** pos, void** vel, int64_t* n)
    {
        if (s < 0 || s >= cfg_.nspecies) throw std::invalid_argument("psum_stepper_state: bad species");
        for (int d = 0; d < 3; ++d) {
            if (pos) pos[d] = sp_[s].pos[d];
            if (vel) vel[d] = sp_[s].vel[d];
        }
        if (n) *n = sp_[s].n;
    }

    void run(int steps, double* hist_host)
    {
        if (steps <= 0) return;
        if (hist_cap_ < steps) {
            if (hist_) (void)hipFree(hist_);
            S_TRY(hipMalloc(&hist_, (size_t)steps * 5 * sizeof(double)));
            hist_cap_ = steps;
        }
        const int D = cfg_.grid.dim;
        for (int s = 0; s < steps; ++s) {
            if (cfg_.sort_every > 0 && !no_sort_ && step_ % cfg_.sort_every == 0) sort_all();
            solve_field();
            CHK(psum_hip_field_energy(&cfg_.grid, phi_, cfg_.eps, fe_, few_, few_bytes_, st_));
            const void* ph = cfg_.dtype == PSUM_F32 ? (const void*)phi32_ : (const void*)phi_;
            hipLaunchKernelGGL(k_zero_i64, dim3(nblk(nn_)), dim3(kT), 0, st_, (long long*)acc_, nn_);
            for (int q = 0; q < cfg_.nspecies; ++q) {
                Species& sp = sp_[q];
                const double qm = sp.sp.q / sp.sp.m;
                if (timing_) S_TRY(hipEventRecord(ev_[2 * nev_], st_));
                CHK(psum_hip_gather_push(&cfg_.grid, cfg_.dtype, sp.pos, sp.vel, sp.n, ph, qm, cfg_.dt,
                                         sp.sp.m * sp.sp.w, sp.wdep, scale_, acc_, diag_, counters_, st_));
                if (timing_) {
                    S_TRY(hipEventRecord(ev_[2 * nev_ + 1], st_));
                    push_particles_ += sp.n;
                    if (++nev_ == kEv) flush_events();
                }
                CHK(psum_hip_reduce_diag(diag_, psum_hip_push_blocks(sp.n), red_ + 4 * q, redw_, st_));
                if (sp.n == 0) S_TRY(hipMemsetAsync(red_ + 4 * q, 0, 4 * sizeof(double), st_));
            }
            hipLaunchKernelGGL(k_hist_record, dim3(1), dim3(64), 0, st_, fe_, red_, cfg_.nspecies, hist_, s);
            S_TRY(hipGetLastError());
            ++step_;
            (void)D;
        }
        if (hist_host)
            S_TRY(hipMemcpyAsync(hist_host, hist_, (size_t)steps * 5 * sizeof(double), hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
        if (timing_) flush_events();
    }

    // HIP-event timing of every fused gather/push/deposit launch on the
    // engine's stream: total milliseconds, launch count, particles pushed.
    void set_timing(bool on)
    {
        if (on && !ev_[0])
            for (auto& e : ev_) S_TRY(hipEventCreate(&e));
        timing_ = on;
        push_ms_ = 0.0;
        push_launches_ = 0;
        push_particles_ = 0;
        nev_ = 0;
    }

    void push_timing(double* ms, int64_t* launches, int64_t* particles)
    {
        if (ms) *ms = push_ms_;
        if (launches) *launches = push_launches_;
        if (particles) *particles = push_particles_;
    }

    // rho from the exact deposit of the current positions, neutralised,
    // then phi by the multigrid (warm start); phi32 refreshed in FP32 mode
    void solve_field()
    {
        if (!acc_valid_) deposit_all();
        CHK(psum_hip_fixed_to_double(acc_, nn_, 1.0 / scale_, cfg_.rho_bg, b_, st_));
        if (all_periodic_) {
            hipLaunchKernelGGL(k_sum_part, dim3(nblk(nn_)), dim3(kT), 0, st_, b_, nn_, sum_part_);
            hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kT), 0, st_, sum_part_, (int64_t)nblk(nn_), sum_);
            hipLaunchKernelGGL(k_sub_mean, dim3(nblk(nn_)), dim3(kT), 0, st_, b_, nn_, sum_);
        }
        if (cfg_.extrapolate && have_prev_)
            hipLaunchKernelGGL(k_extrapolate, dim3(nblk(nn_)), dim3(kT), 0, st_, phi_, phi_prev_, nn_);
        mg_solve(mg_, b_, phi_, false);
        if (cfg_.extrapolate && !have_prev_) {
            S_TRY(hipMemcpyAsync(phi_prev_, phi_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
            have_prev_ = true;
        }
        cycles_ += mg_last_cycles(mg_);
        last_rel_ = mg_last_rel(mg_);
        if (cfg_.dtype == PSUM_F32) CHK(psum_hip_f64_to_f32(phi_, phi32_, nn_, st_));
    }

    // Loss L = field energy of phi solved from the current positions.
    double field_energy_now()
    {
        solve_field();
        CHK(psum_hip_field_energy(&cfg_.grid, phi_, cfg_.eps, fe_, few_, few_bytes_, st_));
        double e = 0.0;
        S_TRY(hipMemcpyAsync(&e, fe_, sizeof(double), hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
        return e;
    }

    double forward_loss(int steps)
    {
        no_sort_ = true;
        run(steps, nullptr);
        no_sort_ = false;
        return field_energy_now();
    }

    // ---------------------------------------------------------- adjoint
    struct Snapshot {
        std::vector<void*> x, v;   // per species * D
        double* phi = nullptr;
        double* phi_prev = nullptr;
        bool have_prev = false;
        int predicted = 1;
    };

    void snap_alloc(Snapshot& s)
    {
        const int D = cfg_.grid.dim;
        s.x.assign(cfg_.nspecies * D, nullptr);
        s.v.assign(cfg_.nspecies * D, nullptr);
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)std::max<int64_t>(1, sp_[q].n) * esz_;
                S_TRY(hipMalloc(&s.x[q * D + d], b));
                S_TRY(hipMalloc(&s.v[q * D + d], b));
            }
        S_TRY(hipMalloc(&s.phi, nn_ * sizeof(double)));
        if (cfg_.extrapolate) S_TRY(hipMalloc(&s.phi_prev, nn_ * sizeof(double)));
    }

    void snap_free(Snapshot& s)
    {
        for (void* p : s.x) if (p) (void)hipFree(p);
        for (void* p : s.v) if (p) (void)hipFree(p);
        if (s.phi) (void)hipFree(s.phi);
        if (s.phi_prev) (void)hipFree(s.phi_prev);
        s = Snapshot{};
    }

    void snap_save(Snapshot& s)
    {
        const int D = cfg_.grid.dim;
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)sp_[q].n * esz_;
                S_TRY(hipMemcpyAsync(s.x[q * D + d], sp_[q].pos[d], b, hipMemcpyDeviceToDevice, st_));
                S_TRY(hipMemcpyAsync(s.v[q * D + d], sp_[q].vel[d], b, hipMemcpyDeviceToDevice, st_));
            }
        S_TRY(hipMemcpyAsync(s.phi, phi_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        if (cfg_.extrapolate && have_prev_)
            S_TRY(hipMemcpyAsync(s.phi_prev, phi_prev_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        s.have_prev = have_prev_;
        s.predicted = mg_predicted(mg_);
    }

    void snap_restore(const Snapshot& s)
    {
        const int D = cfg_.grid.dim;
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)sp_[q].n * esz_;
                S_TRY(hipMemcpyAsync(sp_[q].pos[d], s.x[q * D + d], b, hipMemcpyDeviceToDevice, st_));
                S_TRY(hipMemcpyAsync(sp_[q].vel[d], s.v[q * D + d], b, hipMemcpyDeviceToDevice, st_));
            }
        S_TRY(hipMemcpyAsync(phi_, s.phi, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        if (cfg_.extrapolate && s.have_prev)
            S_TRY(hipMemcpyAsync(phi_prev_, s.phi_prev, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        have_prev_ = s.have_prev;
        mg_set_predicted(mg_, s.predicted);
        acc_valid_ = false;
    }

    // lrho = S^T lphi (lphi in lphi_, overwritten); result in lrho_ (double) and lrho_t_
    void adjoint_solve()
    {
        if (all_periodic_) {
            hipLaunchKernelGGL(k_sum_part, dim3(nblk(nn_)), dim3(kT), 0, st_, lphi_, nn_, sum_part_);
            hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kT), 0, st_, sum_part_, (int64_t)nblk(nn_), sum_);
            hipLaunchKernelGGL(k_sub_at0, dim3(1), dim3(64), 0, st_, lphi_, sum_);
        }
        mg_solve_homogeneous(mg_, lphi_, lrho_);
        if (all_periodic_) {
            S_TRY(hipMemcpyAsync(lphi_, lrho_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
            hipLaunchKernelGGL(k_sum_part, dim3(nblk(nn_)), dim3(kT), 0, st_, lphi_, nn_, sum_part_);
            hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kT), 0, st_, sum_part_, (int64_t)nblk(nn_), sum_);
            hipLaunchKernelGGL(k_sub_mean, dim3(nblk(nn_)), dim3(kT), 0, st_, lphi_, nn_, sum_);
        } else {
            S_TRY(hipMemcpyAsync(lphi_, lrho_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        }
        // lphi_ now holds lrho (zero-mean); lrho_ keeps the warm start
        if (cfg_.dtype == PSUM_F32) CHK(psum_hip_f64_to_f32(lphi_, (float*)lrho_t_, nn_, st_));
    }

    const void* lrho_view() const
    {
        return cfg_.dtype == PSUM_F32 ? (const void*)lrho_t_ : (const void*)lphi_;
    }

    void adjoint_gather_all(void* const* xs_of_species, double* bmax_scale_factor)
    {
        (void)bmax_scale_factor;
        const int D = cfg_.grid.dim;
        double maxfac = 0.0;
        for (int q = 0; q < cfg_.nspecies; ++q) {
            double dr = 0.0;
            for (int d = 0; d < D; ++d) dr = std::max(dr, cfg_.grid.delta_r[d]);
            maxfac = std::max(maxfac, std::fabs(sp_[q].sp.q / sp_[q].sp.m) * cfg_.dt * dr);
        }
        for (int q = 0; q < cfg_.nspecies; ++q) {
            void* const* x = xs_of_species + q * 3;
            CHK(psum_hip_adjoint_gather(&cfg_.grid, cfg_.dtype, x, &lx_[q * 3], &lv_[q * 3], sp_[q].n,
                                        lrho_view(), sp_[q].wdep, cfg_.dt, bmax_, st_));
            CHK(psum_hip_adjoint_scale(bmax_, psum_hip_adjoint_blocks(sp_[q].n), maxfac, adj_scale_,
                                       q > 0 ? 1 : 0, st_));
        }
    }

    void adjoint_push_all(void* const* xs_of_species, const void* phi_t)
    {
        hipLaunchKernelGGL(k_zero_i64, dim3(nblk(nn_)), dim3(kT), 0, st_, (long long*)adj_acc_, nn_);
        for (int q = 0; q < cfg_.nspecies; ++q) {
            void* const* x = xs_of_species + q * 3;
            CHK(psum_hip_adjoint_push(&cfg_.grid, cfg_.dtype, x, &lx_[q * 3], &lv_[q * 3], sp_[q].n,
                                      phi_t, sp_[q].sp.q / sp_[q].sp.m, cfg_.dt, adj_scale_,
                                      adj_acc_, st_));
        }
        CHK(psum_hip_fixed_to_double_dev(adj_acc_, nn_, adj_scale_, lphi_, st_));
    }

    // dL/dx0, dL/dv0 of L = field energy after `steps` steps from the current
    // state; forward checkpoints every `ckpt` steps, recompute per segment.
    double adjoint(int steps, int ckpt, void* const* gx, void* const* gv, double* t_fwd, double* t_bwd)
    {
        if (steps < 1 || ckpt < 1) throw std::invalid_argument("psum_stepper_adjoint: steps, ckpt >= 1");
        const int D = cfg_.grid.dim;
        no_sort_ = true;
        adj_buffers();
        hipEvent_t e0, e1, e2;
        S_TRY(hipEventCreate(&e0));
        S_TRY(hipEventCreate(&e1));
        S_TRY(hipEventCreate(&e2));
        S_TRY(hipEventRecord(e0, st_));
        const int nck = (steps + ckpt - 1) / ckpt;
        std::vector<Snapshot> cks(nck);
        for (auto& c : cks) snap_alloc(c);
        for (int j = 0; j < nck; ++j) {
            snap_save(cks[j]);
            run(std::min(ckpt, steps - j * ckpt), nullptr);
        }
        const double loss = field_energy_now();
        S_TRY(hipEventRecord(e1, st_));
        // final: lphi = dW/dphi(phi^N), lrho = S^T lphi, lx = sum w grad(lrho)(x^N), lv = 0
        CHK(psum_hip_field_energy_grad(&cfg_.grid, phi_, cfg_.eps, lphi_, st_));
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                hipLaunchKernelGGL(k_zero_any, dim3(nblk((int64_t)sp_[q].n * esz_)), dim3(kT), 0, st_, (unsigned char*)lx_[q * 3 + d], (int64_t)sp_[q].n * esz_);
                hipLaunchKernelGGL(k_zero_any, dim3(nblk((int64_t)sp_[q].n * esz_)), dim3(kT), 0, st_, (unsigned char*)lv_[q * 3 + d], (int64_t)sp_[q].n * esz_);
            }
        hipLaunchKernelGGL(k_zero_any, dim3(nblk(nn_ * 8)), dim3(kT), 0, st_, (unsigned char*)lrho_, nn_ * 8);
        adjoint_solve();
        {
            std::vector<void*> xn(cfg_.nspecies * 3, nullptr);
            for (int q = 0; q < cfg_.nspecies; ++q)
                for (int d = 0; d < D; ++d) xn[q * 3 + d] = sp_[q].pos[d];
            adjoint_gather_all(xn.data(), nullptr);
        }
        // reverse sweep, one segment at a time
        for (int j = nck - 1; j >= 0; --j) {
            const int s0 = j * ckpt, len = std::min(ckpt, steps - s0);
            snap_restore(cks[j]);
            for (int k = 0; k < len; ++k) {
                for (int q = 0; q < cfg_.nspecies; ++q)
                    for (int d = 0; d < D; ++d)
                        S_TRY(hipMemcpyAsync(seg_x_[k][q * 3 + d], sp_[q].pos[d], (size_t)sp_[q].n * esz_, hipMemcpyDeviceToDevice, st_));
                run(1, nullptr);   // solves phi^(s0+k) then pushes
                const void* ph = cfg_.dtype == PSUM_F32 ? (const void*)phi32_ : (const void*)phi_;
                S_TRY(hipMemcpyAsync(seg_phi_[k], ph, (size_t)nn_ * esz_, hipMemcpyDeviceToDevice, st_));
            }
            for (int k = len - 1; k >= 0; --k) {
                adjoint_push_all(seg_x_[k].data(), seg_phi_[k]);
                adjoint_solve();
                adjoint_gather_all(seg_x_[k].data(), nullptr);
            }
        }
        S_TRY(hipEventRecord(e2, st_));
        // outputs and restore
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                if (gx) S_TRY(hipMemcpyAsync(gx[q * 3 + d], lx_[q * 3 + d], (size_t)sp_[q].n * esz_, hipMemcpyDeviceToDevice, st_));
                if (gv) S_TRY(hipMemcpyAsync(gv[q * 3 + d], lv_[q * 3 + d], (size_t)sp_[q].n * esz_, hipMemcpyDeviceToDevice, st_));
            }
        snap_restore(cks[0]);
        S_TRY(hipStreamSynchronize(st_));
        float ms1 = 0, ms2 = 0;
        S_TRY(hipEventElapsedTime(&ms1, e0, e1));
        S_TRY(hipEventElapsedTime(&ms2, e1, e2));
        if (t_fwd) *t_fwd = ms1;
        if (t_bwd) *t_bwd = ms2;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        (void)hipEventDestroy(e2);
        for (auto& c : cks) snap_free(c);
        no_sort_ = false;
        return loss;
    }

    void adj_buffers()
    {
        if (lphi_) return;
        const int D = cfg_.grid.dim;
        S_TRY(hipMalloc(&lphi_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&lrho_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&lrho_t_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&adj_acc_, nn_ * sizeof(int64_t)));
        S_TRY(hipMalloc(&adj_scale_, 4 * sizeof(double)));
        int64_t nbmax = 1;
        for (int q = 0; q < cfg_.nspecies; ++q) nbmax = std::max(nbmax, psum_hip_adjoint_blocks(sp_[q].sp.capacity));
        S_TRY(hipMalloc(&bmax_, (size_t)nbmax * sizeof(double)));
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                S_TRY(hipMalloc(&lx_[q * 3 + d], (size_t)std::max<int64_t>(1, sp_[q].sp.capacity) * esz_));
                S_TRY(hipMalloc(&lv_[q * 3 + d], (size_t)std::max<int64_t>(1, sp_[q].sp.capacity) * esz_));
            }
    }

    void set_ckpt_capacity(int ckpt)
    {
        const int D = cfg_.grid.dim;
        for (auto& v : seg_x_) for (void* p : v) if (p) (void)hipFree(p);
        for (void* p : seg_phi_) if (p) (void)hipFree(p);
        seg_x_.assign(ckpt, std::vector<void*>(cfg_.nspecies * 3, nullptr));
        seg_phi_.assign(ckpt, nullptr);
        for (int k = 0; k < ckpt; ++k) {
            for (int q = 0; q < cfg_.nspecies; ++q)
                for (int d = 0; d < D; ++d)
                    S_TRY(hipMalloc(&seg_x_[k][q * 3 + d], (size_t)std::max<int64_t>(1, sp_[q].sp.capacity) * esz_));
            S_TRY(hipMalloc(&seg_phi_[k], (size_t)nn_ * esz_));
        }
    }

    void phi(double* out)
    {
        S_TRY(hipMemcpyAsync(out, phi_, nn_ * sizeof(double), hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
    }

    void stats(int64_t* cycles, double* rel, int64_t* counters)
    {
        if (cycles) *cycles = cycles_;
        if (rel) *rel = last_rel_;
        if (counters) {
            S_TRY(hipMemcpyAsync(counters, counters_, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, st_));
            S_TRY(hipStreamSynchronize(st_));
        }
    }

    void sync() { S_TRY(hipStreamSynchronize(st_)); }

private:
    bool no_sort_ = false;
    double *lphi_ = nullptr, *lrho_ = nullptr, *adj_scale_ = nullptr, *bmax_ = nullptr;
    void* lrho_t_ = nullptr;
    int64_t* adj_acc_ = nullptr;
    void* lx_[12] = {};
    void* lv_[12] = {};
    std::vector<std::vector<void*>> seg_x_;
    std::vector<void*> seg_phi_;
    static constexpr int kEv = 64;
    hipEvent_t ev_[2 * kEv] = {};
    int nev_ = 0;
    bool timing_ = false;
    double push_ms_ = 0.0;
    int64_t push_launches_ = 0, push_particles_ = 0;

    void flush_events()
    {
        if (nev_ == 0) return;
        S_TRY(hipEventSynchronize(ev_[2 * nev_ - 1]));
        for (int i = 0; i < nev_; ++i) {
            float ms = 0.f;
            S_TRY(hipEventElapsedTime(&ms, ev_[2 * i], ev_[2 * i + 1]));
            push_ms_ += ms;
            ++push_launches_;
        }
        nev_ = 0;
    }

    psum_step_config cfg_;
    int esz_ = 8;
    int64_t nn_ = 0;
    bool all_periodic_ = true;
    hipStream_t st_ = nullptr;
    HipMultigrid* mg_ = nullptr;
    Species sp_[4];
    int64_t max_cap_ = 0;
    double scale_ = 1.0;
    int64_t* acc_ = nullptr;
    double *b_ = nullptr, *phi_ = nullptr, *diag_ = nullptr, *red_ = nullptr, *fe_ = nullptr;
    double* phi_prev_ = nullptr;
    bool have_prev_ = false;
    void* redw_ = nullptr;
    float* phi32_ = nullptr;
    void* few_ = nullptr;
    size_t few_bytes_ = 0;
    double *sum_part_ = nullptr, *sum_ = nullptr, *hist_ = nullptr;
    int hist_cap_ = 0;
    int64_t* counters_ = nullptr;
    uint32_t *keys_ = nullptr, *skeys_ = nullptr, *perm_ = nullptr;
    void* sort_work_ = nullptr;
    size_t sort_bytes_ = 0;
    unsigned long long* nvalid_ = nullptr;
    bool acc_valid_ = false;
    int64_t step_ = 0;
    int64_t cycles_ = 0;
    double last_rel_ = NAN;

    static void S_TRY(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error(std::string("psum_stepper: ") + hipGetErrorString(e));
    }
    static void CHK(psum_status s)
    {
        if (s != PSUM_OK) throw std::runtime_error(std::string("psum_stepper: ") + psum_hip_last_error());
    }

    void deposit_all()
    {
        hipLaunchKernelGGL(k_zero_i64, dim3(nblk(nn_)), dim3(kT), 0, st_, (long long*)acc_, nn_);
        for (int q = 0; q < cfg_.nspecies; ++q) {
            Species& sp = sp_[q];
            CHK(psum_hip_deposit_cic(&cfg_.grid, cfg_.dtype, sp.pos, sp.n, sp.wdep, scale_, acc_, counters_, st_));
        }
        acc_valid_ = true;
    }

    void sort_all()
    {
        const int D = cfg_.grid.dim;
        const uint32_t nkeys = (uint32_t)nn_;
        for (int q = 0; q < cfg_.nspecies; ++q) {
            Species& sp = sp_[q];
            if (sp.n == 0) continue;
            CHK(psum_hip_bin(&cfg_.grid, cfg_.dtype, sp.pos, sp.n, keys_, st_));
            CHK(psum_hip_sort(keys_, sp.n, nkeys, skeys_, perm_, sort_work_, sort_bytes_, st_));
            void* src[6];
            void* dst[6];
            int nc = 0;
            for (int d = 0; d < D; ++d) { src[nc] = sp.pos[d]; dst[nc] = sp.pos2[d]; ++nc; }
            for (int d = 0; d < D; ++d) { src[nc] = sp.vel[d]; dst[nc] = sp.vel2[d]; ++nc; }
            CHK(psum_hip_permute(cfg_.dtype, nc, src, dst, perm_, sp.n, st_));
            for (int d = 0; d < D; ++d) {
                std::swap(sp.pos[d], sp.pos2[d]);
                std::swap(sp.vel[d], sp.vel2[d]);
            }
            // drop particles whose key is invalid (absorbed): they sort last
            unsigned long long nv = (unsigned long long)sp.n;
            S_TRY(hipMemcpyAsync(nvalid_, &nv, sizeof nv, hipMemcpyHostToDevice, st_));
            hipLaunchKernelGGL(k_count_valid, dim3(nblk(sp.n)), dim3(kT), 0, st_, skeys_, sp.n, nvalid_);
            S_TRY(hipMemcpyAsync(&nv, nvalid_, sizeof nv, hipMemcpyDeviceToHost, st_));
            S_TRY(hipStreamSynchronize(st_));
            sp.n = (int64_t)nv;
        }
    }
};

template <typename F>
psum_status guarded(F&& f)
{
    try {
        f();
        return PSUM_OK;
    } catch (const std::invalid_argument& e) {
        set_error(e.what());
        return PSUM_ERR_ARG;
    } catch (const std::exception& e) {
        set_error(e.what());
        return PSUM_ERR_HIP;
    }
}

}  // namespace

extern "C" {

void* psum_stepper_create(const psum_step_config* cfg)
{
    if (!cfg) {
        set_error("psum_stepper_create: null config");
        return nullptr;
    }
    Stepper* h = nullptr;
    if (guarded([&]() { h = new Stepper(*cfg); }) != PSUM_OK) return nullptr;
    return h;
}

psum_status psum_stepper_load(void* h, int s, const void* const* pos, const void* const* vel, int64_t n)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->load(s, pos, vel, n); });
}

psum_status psum_stepper_state(void* h, int s, void** pos, void** vel, int64_t* n)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->state(s, pos, vel, n); });
}

psum_status psum_stepper_run(void* h, int steps, double* hist)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->run(steps, hist); });
}

psum_status psum_stepper_phi(void* h, double* phi_host)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->phi(phi_host); });
}

psum_status psum_stepper_stats(void* h, int64_t* cycles, double* last_rel_res, int64_t* counters)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->stats(cycles, last_rel_res, counters); });
}

psum_status psum_stepper_sync(void* h)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->sync(); });
}

psum_status psum_stepper_set_timing(void* h, int on)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->set_timing(on != 0); });
}

psum_status psum_stepper_push_timing(void* h, double* ms, int64_t* launches, int64_t* particles)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->push_timing(ms, launches, particles); });
}

psum_status psum_stepper_forward_loss(void* h, int steps, double* loss)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { *loss = static_cast<Stepper*>(h)->forward_loss(steps); });
}

psum_status psum_stepper_adjoint(void* h, int steps, int ckpt, double* loss, void* const* grad_x,
                                 void* const* grad_v, double* t_fwd_ms, double* t_bwd_ms)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        s->set_ckpt_capacity(ckpt);
        const double l = s->adjoint(steps, ckpt, grad_x, grad_v, t_fwd_ms, t_bwd_ms);
        if (loss) *loss = l;
    });
}

void psum_stepper_destroy(void* h) { delete static_cast<Stepper*>(h); }

}  // extern "C"
