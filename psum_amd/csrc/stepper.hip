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
}  // namespace psum_hip

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

// x -= sum / n  (FP64 neutralisation of a periodic charge density)
__global__ __launch_bounds__(kT) void k_sub_mean(double* __restrict__ x, int64_t n,
                                                 const double* __restrict__ sum)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    const double mean = sum[0] / (double)n;
    if (i < n) x[i] = x[i] - mean;
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
        S_TRY(hipMemsetAsync(phi_, 0, nn_ * sizeof(double), st_));
        if (cfg_.dtype == PSUM_F32) S_TRY(hipMalloc(&phi32_, nn_ * sizeof(float)));
        const int64_t nbp = std::max<int64_t>(1, psum_hip_push_blocks(max_cap_));
        S_TRY(hipMalloc(&diag_, (size_t)nbp * 4 * sizeof(double)));
        S_TRY(hipMalloc(&red_, 16 * sizeof(double)));
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
        for (void* p : {(void*)acc_, (void*)b_, (void*)phi_, (void*)phi32_, (void*)diag_, (void*)red_,
                        (void*)fe_, few_, (void*)sum_part_, (void*)sum_, (void*)counters_,
                        (void*)keys_, (void*)skeys_, (void*)perm_, sort_work_, (void*)nvalid_, (void*)hist_})
            if (p) (void)hipFree(p);
        for (auto& e : ev_)
            if (e) (void)hipEventDestroy(e);
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
            if (!acc_valid_) deposit_all();
            if (cfg_.sort_every > 0 && step_ % cfg_.sort_every == 0) sort_all();
            CHK(psum_hip_fixed_to_double(acc_, nn_, 1.0 / scale_, cfg_.rho_bg, b_, st_));
            if (all_periodic_) {
                hipLaunchKernelGGL(k_sum_part, dim3(nblk(nn_)), dim3(kT), 0, st_, b_, nn_, sum_part_);
                hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(kT), 0, st_, sum_part_, (int64_t)nblk(nn_), sum_);
                hipLaunchKernelGGL(k_sub_mean, dim3(nblk(nn_)), dim3(kT), 0, st_, b_, nn_, sum_);
            }
            mg_solve(mg_, b_, phi_, false);
            cycles_ += mg_last_cycles(mg_);
            last_rel_ = mg_last_rel(mg_);
            CHK(psum_hip_field_energy(&cfg_.grid, phi_, cfg_.eps, fe_, few_, few_bytes_, st_));
            const void* ph = phi_;
            if (cfg_.dtype == PSUM_F32) {
                CHK(psum_hip_f64_to_f32(phi_, phi32_, nn_, st_));
                ph = phi32_;
            }
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
                CHK(psum_hip_reduce_diag(diag_, psum_hip_push_blocks(sp.n), red_ + 4 * q, st_));
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

void psum_stepper_destroy(void* h) { delete static_cast<Stepper*>(h); }

}  // extern "C"
