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
#include <memory>
#include <vector>

#include "common.hpp"
#include "mg_batch.hpp"
#include "psum/mas.hpp"

namespace psum_hip {
class HipMultigrid;
HipMultigrid* mg_create(bool mixed);
void mg_destroy(HipMultigrid* h);
void mg_set_options(HipMultigrid* h, const char* o);
void mg_set_tuning(HipMultigrid* h, const char* o);
void mg_set_smoothing(HipMultigrid* h, int pre, int post);
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
void mg_set_fixed_cycles(HipMultigrid* h, int cycles);
// Cartesian constant-eps FD Poisson matrix of the reference frontends with
// identity rows at fixed nodes (poisson.hip)
void assemble(const psum_grid& g, double eps, const std::vector<uint8_t>& fixed,
              std::vector<unsigned long long>& R, std::vector<unsigned long long>& C,
              std::vector<double>& V);
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

__global__ __launch_bounds__(kT) void k_zero_any(unsigned char* __restrict__ a, int64_t nbytes)
{
    const int64_t i = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (i < nbytes) a[i] = 0;
}

// hist[5*s + k], s = *count (then incremented; rows beyond cap are dropped):
// field energy, then species-summed kinetic energy and momentum.  The row
// index lives on the device so that a replayed graph appends.
__global__ void k_hist_record(const double* __restrict__ fe, const double* __restrict__ red,
                              int nspecies, double* __restrict__ hist, long long* __restrict__ count,
                              long long cap)
{
    if (threadIdx.x != 0) return;
    const long long s = count[0];
    count[0] = s + 1;
    if (s >= cap) return;
    hist[5 * s + 0] = fe[0];
    for (int k = 0; k < 4; ++k) {
        double t = 0.0;
        for (int q = 0; q < nspecies; ++q) t += red[4 * q + k];
        hist[5 * s + 1 + k] = t;
    }
}

__global__ void k_zero_f64(double* __restrict__ a, int n)
{
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0.0;
}

// First index whose key is PSUM_KEY_INVALID in sorted_keys[0, n) (n when
// none; keys are sorted, invalid last), found by one wave: the answer a lies
// in [lo, hi]; each round probes lo, lo + step, ... (step = ceil((hi - lo) /
// 64), one probe per lane, below hi) and keeps the gap before the first
// invalid probe, or the part past the last probe.
__global__ __launch_bounds__(64) void k_count_valid(const uint32_t* __restrict__ sorted_keys, int64_t n,
                                                    unsigned long long* __restrict__ out)
{
    const int lane = threadIdx.x;
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t step = (hi - lo + 63) / 64;
        const int64_t p = lo + (int64_t)lane * step;
        const bool probe = p < hi;
        const bool inv = probe && sorted_keys[p] == PSUM_KEY_INVALID;
        const unsigned long long m = __ballot(inv);
        if (m) {
            const int f = __ffsll((long long)m) - 1;
            const int64_t pf = lo + (int64_t)f * step;
            lo = f == 0 ? lo : pf - step + 1;
            hi = pf;
        } else {
            const int last = 63 - __clzll((long long)__ballot(probe));
            lo = lo + (int64_t)last * step + 1;
        }
    }
    if (lane == 0) out[0] = (unsigned long long)lo;
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
        S_TRY(hipMalloc(&hist_count_, sizeof(long long)));
        few_bytes_ = psum_hip_field_energy_workspace(&cfg_.grid);
        S_TRY(hipMalloc(&few_, few_bytes_));
        S_TRY(hipMalloc(&sum_part_, (size_t)nblk(nn_) * sizeof(double)));
        S_TRY(hipMalloc(&sum_, 8 * sizeof(double)));
        S_TRY(hipMalloc(&counters_, 8 * sizeof(int64_t)));
        S_TRY(hipMemsetAsync(counters_, 0, 8 * sizeof(int64_t), st_));
        if (cfg_.sort_every > 0) sort_buffers();
        S_TRY(hipStreamSynchronize(st_));
    }

    void sort_buffers()
    {
        for (int q = 0; q < cfg_.nspecies; ++q) {
            Species& sp = sp_[q];
            const size_t bytes = (size_t)std::max<int64_t>(1, sp.sp.capacity) * esz_;
            for (int d = 0; d < cfg_.grid.dim; ++d) {
                if (!sp.pos2[d]) S_TRY(hipMalloc(&sp.pos2[d], bytes));
                if (!sp.vel2[d]) S_TRY(hipMalloc(&sp.vel2[d], bytes));
            }
        }
        if (keys_) return;
        S_TRY(hipMalloc(&keys_, (size_t)max_cap_ * 4 + 4));
        S_TRY(hipMalloc(&skeys_, (size_t)max_cap_ * 4 + 4));
        S_TRY(hipMalloc(&perm_, (size_t)max_cap_ * 4 + 4));
        sort_bytes_ = psum_hip_sort_workspace(max_cap_);
        S_TRY(hipMalloc(&sort_work_, sort_bytes_));
        S_TRY(hipMalloc(&nvalid_, sizeof(unsigned long long)));
    }

    ~Stepper()
    {
        if (mg_) mg_destroy(mg_);
        for (auto& sp : sp_)
            for (int d = 0; d < 3; ++d)
                for (void* p : {sp.pos[d], sp.vel[d], sp.pos2[d], sp.vel2[d]})
                    if (p) (void)hipFree(p);
        for (void* p : {(void*)acc_, (void*)b_, (void*)phi_, (void*)phi_prev_, (void*)phi32_, (void*)diag_, (void*)red_,
                        (void*)fe_, (void*)hist_count_, few_, redw_, (void*)sum_part_, (void*)sum_, (void*)counters_,
                        (void*)keys_, (void*)skeys_, (void*)perm_, sort_work_, (void*)nvalid_, (void*)hist_})
            if (p) (void)hipFree(p);
        for (auto& e : ev_)
            if (e) (void)hipEventDestroy(e);
        for (void* p : {(void*)lphi_, (void*)lrho_, (void*)adj_scale_, (void*)bmax_, (void*)adj_acc_})
            if (p) (void)hipFree(p);
        for (void* p : lx_) if (p) (void)hipFree(p);
        for (void* p : lv_) if (p) (void)hipFree(p);
        for (auto& v : seg_x_) for (void* p : v) if (p) (void)hipFree(p);
        for (void* p : seg_phi_) if (p) (void)hipFree(p);
        for (auto& c : cks_) snap_free(c);
        for (auto& e : aev_) if (e) (void)hipEventDestroy(e);
        if (st_) (void)hipStreamDestroy(st_);
    }

    void load(int s, const void* const* pos, const void* const* vel, int64_t n)
    {
        if (s < 0 || s >= cfg_.nspecies) throw std::invalid_argument("psum_stepper_load: bad species");
        Species& sp = sp_[s];
        if (n > sp.sp.capacity) throw std::invalid_argument("psum_stepper_load: n exceeds capacity");
        for (int d = 0; d < cfg_.grid.dim; ++d) {
            S_TRY(hipMemcpyAsync(sp.pos[d], pos[d], (size_t)n * esz_, hipMemcpyDefault, st_));
            S_TRY(hipMemcpyAsync(sp.vel[d], vel[d], (size_t)n * esz_, hipMemcpyDefault, st_));
        }
        sp.n = n;
        acc_valid_ = false;
        step_ = 0;          // a new run: the sort cadence restarts
        S_TRY(hipMemsetAsync(counters_ + 3, 0, sizeof(int64_t), st_));   // the accumulator's overflow flag
        S_TRY(hipStreamSynchronize(st_));
    }

    // Throws when an add to the int64 charge accumulator wrapped since the
    // particles were loaded (counters [3]; the kernels count every wrap).
    void check_overflow(const int64_t* wraps_host)
    {
        if (*wraps_host > 0)
            throw std::overflow_error("psum_stepper: the int64 charge accumulator overflowed (" +
                                      std::to_string(*wraps_host) +
                                      " wrapped adds): a node collected more charge than 2^63 / 2^S holds");
    }
    void check_overflow()
    {
        int64_t w = 0;
        S_TRY(hipMemcpyAsync(&w, counters_ + 3, sizeof w, hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
        check_overflow(&w);
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

    void run(int steps, double* hist_host) { run_steps(steps, hist_host, true, true); }

    // diagnostics = false skips the field energy and the reductions of the
    // history (recompute and recording passes of the adjoint)
    void run_steps(int steps, double* hist_host, bool sync, bool diagnostics)
    {
        if (steps <= 0) return;
        reserve_hist(steps);
        hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(kT), 0, st_, (long long*)hist_count_, (int64_t)1);
        for (int s = 0; s < steps; ++s) {
            if (!no_sort_) sort_due_species();
            step_body(diagnostics);
        }
        if (hist_host)
            S_TRY(hipMemcpyAsync(hist_host, hist_, (size_t)steps * 5 * sizeof(double), hipMemcpyDeviceToHost, st_));
        if (!sync) return;
        int64_t wraps = 0;
        S_TRY(hipMemcpyAsync(&wraps, counters_ + 3, sizeof wraps, hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
        if (timing_) flush_events();
        check_overflow(&wraps);
    }

    void reserve_hist(int64_t rows)
    {
        if (hist_cap_ >= rows) return;
        if (hist_) (void)hipFree(hist_);
        S_TRY(hipMalloc(&hist_, (size_t)rows * 5 * sizeof(double)));
        hist_cap_ = rows;
    }

    // One step on the engine's stream with nothing the host waits for when
    // the solver runs a fixed cycle count: graph-capturable.  History rows
    // go to the device-side row counter.
    void step_body(bool diagnostics)
    {
        solve_field();
        step_after_solve(diagnostics);
    }

    // the rest of a step once phi is solved: field energy, gather + push +
    // deposit of every species, diagnostics
    void step_after_solve(bool diagnostics)
    {
        {
            if (diagnostics) CHK(psum_hip_field_energy(&cfg_.grid, phi_, cfg_.eps, fe_, few_, few_bytes_, st_));
            const void* ph = cfg_.dtype == PSUM_F32 ? (const void*)phi32_ : (const void*)phi_;
            hipLaunchKernelGGL(k_zero_i64, dim3(nblk(nn_)), dim3(kT), 0, st_, (long long*)acc_, nn_);
            for (int q = 0; q < cfg_.nspecies; ++q) {
                Species& sp = sp_[q];
                const double qm = sp.sp.q / sp.sp.m;
                if (timing_) S_TRY(hipEventRecord(ev_[2 * nev_], st_));
                if (order_ > 1)
                    CHK(psum_hip_gather_push_bspline(&cfg_.grid, cfg_.dtype, order_, sp.pos, sp.vel, sp.n, ph, qm,
                                                     cfg_.dt, sp.sp.m * sp.sp.w, diag_, counters_, st_));
                else
                    CHK(psum_hip_gather_push(&cfg_.grid, cfg_.dtype, sp.pos, sp.vel, sp.n, ph, qm, cfg_.dt,
                                             sp.sp.m * sp.sp.w, sp.wdep, scale_, acc_, diag_, counters_, st_));
                if (timing_) {
                    S_TRY(hipEventRecord(ev_[2 * nev_ + 1], st_));
                    push_particles_ += sp.n;
                    ev_sp_[nev_] = q;
                    if (++nev_ == kEv) flush_events();
                }
                if (diagnostics) {
                    if (sp.n > 0) CHK(psum_hip_reduce_diag(diag_, psum_hip_push_blocks(sp.n), red_ + 4 * q, redw_, st_));
                    else hipLaunchKernelGGL(k_zero_f64, dim3(1), dim3(64), 0, st_, red_ + 4 * q, 4);
                }
            }
            // higher orders deposit the new positions after the push (the CIC
            // push deposits them itself)
            if (order_ > 1)
                for (int q = 0; q < cfg_.nspecies; ++q)
                    CHK(psum_hip_deposit_bspline(&cfg_.grid, cfg_.dtype, order_, sp_[q].pos, sp_[q].n, sp_[q].wdep,
                                                 scale_, acc_, counters_, st_));
            if (diagnostics)
                hipLaunchKernelGGL(k_hist_record, dim3(1), dim3(64), 0, st_, fe_, red_, cfg_.nspecies, hist_,
                                   (long long*)hist_count_, (long long)hist_cap_);
            S_TRY(hipGetLastError());
            ++step_;
        }
    }

    // HIP-event timing of every fused gather/push/deposit launch on the
    // engine's stream: total milliseconds, launch count, particles pushed.
    void set_timing(bool on)
    {
        if (on && !ev_[0])
            for (auto& e : ev_) S_TRY(hipEventCreate(&e));
        timing_ = on;
        push_ms_ = 0.0;
        for (double& t : push_ms_sp_) t = 0.0;
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
        const double* bs = solve_field_begin();
        mg_solve_core(mg_, b_, bs, phi_);
        solve_field_end();
    }

    // solve_field in three parts, so that a batch can run the solver cores of
    // several instances as batched launches: begin (rho from the deposit,
    // neutralisation, extrapolation, the solver's replace / pin; returns the
    // right-hand side the cycles see) and end (gauge, bookkeeping, FP32 copy)
    const double* solve_field_begin()
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
        return mg_solve_begin(mg_, b_);
    }

    void solve_field_end()
    {
        mg_solve_end(mg_, b_, phi_);
        if (cfg_.extrapolate && !have_prev_) {
            S_TRY(hipMemcpyAsync(phi_prev_, phi_, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
            have_prev_ = true;
        }
        cycles_ += mg_last_cycles(mg_);
        max_cyc_seen_ = std::max(max_cyc_seen_, mg_last_cycles(mg_));
        last_rel_ = mg_last_rel(mg_);
        if (cfg_.dtype == PSUM_F32) CHK(psum_hip_f64_to_f32(phi_, phi32_, nn_, st_));
    }

    // batch hooks: whether the solve's core can be recorded, and its ops
    bool solve_recordable() const { return mg_recordable(mg_); }
    void record_solve_core(std::vector<MgRecOp>& ops)
    {
        mg_record_core(mg_, b_, mg_core_rhs(mg_, b_), phi_, ops);
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
        if (steps > 0) run_steps(steps, nullptr, true, false);
        no_sort_ = false;
        return field_energy_now();
    }

    // ---------------------------------------------------------- adjoint
    // Reverse mode over the step, for L = field energy after N steps.
    // record(): forward from the current state with a checkpoint (x, v, phi,
    // extrapolation history, MG cycle predictor) every `ckpt` steps.
    // reverse(): seeds lphi = dL/dphi^N and sweeps the segments backwards;
    // each segment is recomputed from its checkpoint (bit-identical: the
    // deposit is an exact integer sum and nothing else reorders), keeping
    // x^n and phi^n, then the reverse kernels run step by step:
    //   adjoint push (lx, lv, lphi) -> lrho = S^T lphi -> adjoint gather (lx).
    // Sorting is suspended between record() and reverse().
    struct Snapshot {
        std::vector<void*> x, v;   // index s*3 + d
        double* phi = nullptr;
        double* phi_prev = nullptr;
        bool have_prev = false;
        int predicted = 1;
        int64_t n[4] = {0, 0, 0, 0};
    };

    void snap_alloc(Snapshot& s)
    {
        const int D = cfg_.grid.dim;
        s.x.assign(cfg_.nspecies * 3, nullptr);
        s.v.assign(cfg_.nspecies * 3, nullptr);
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)std::max<int64_t>(1, sp_[q].n) * esz_;
                S_TRY(hipMalloc(&s.x[q * 3 + d], b));
                S_TRY(hipMalloc(&s.v[q * 3 + d], b));
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
        for (int q = 0; q < cfg_.nspecies; ++q) {
            s.n[q] = sp_[q].n;
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)sp_[q].n * esz_;
                S_TRY(hipMemcpyAsync(s.x[q * 3 + d], sp_[q].pos[d], b, hipMemcpyDeviceToDevice, st_));
                S_TRY(hipMemcpyAsync(s.v[q * 3 + d], sp_[q].vel[d], b, hipMemcpyDeviceToDevice, st_));
            }
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
        for (int q = 0; q < cfg_.nspecies; ++q) {
            sp_[q].n = s.n[q];
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)sp_[q].n * esz_;
                S_TRY(hipMemcpyAsync(sp_[q].pos[d], s.x[q * 3 + d], b, hipMemcpyDeviceToDevice, st_));
                S_TRY(hipMemcpyAsync(sp_[q].vel[d], s.v[q * 3 + d], b, hipMemcpyDeviceToDevice, st_));
            }
        }
        S_TRY(hipMemcpyAsync(phi_, s.phi, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        if (cfg_.extrapolate && s.have_prev)
            S_TRY(hipMemcpyAsync(phi_prev_, s.phi_prev, nn_ * sizeof(double), hipMemcpyDeviceToDevice, st_));
        have_prev_ = s.have_prev;
        mg_set_predicted(mg_, s.predicted);
        acc_valid_ = false;
    }

    // ---- checkpoint files (MAS).  Blocks follow the reference's save()
    // naming for a particle group and a node field (particle_group.hpp:272,
    // device_vector.hpp:294-309, tag_sl.hpp:62-98, device_field.hpp:327-331):
    //   species<q>.data.device, species<q>.data.content.{position,velocity}
    //   (n x D x 1, the particle dtype), phi.location, phi.grid.{lower,upper,
    //   cells}, phi.content.{device,content}; plus psum.* scalars holding the
    // step counter, the solver's cycle predictors and the extrapolation
    // state, so that a restored stepper continues bit-identically.
    template <typename T>
    void save_species(psum::serialization::mas_file& f, int q, const std::string& dev)
    {
        const int D = cfg_.grid.dim;
        const int64_t n = sp_[q].n;
        std::vector<T> comp((size_t)std::max<int64_t>(n, 1)), inter((size_t)std::max<int64_t>(n * D, 1));
        const std::string name = "species" + std::to_string(q);
        f.writeData(name + ".data.device", dev);
        for (int which = 0; which < 2; ++which) {
            for (int d = 0; d < D; ++d) {
                S_TRY(hipMemcpy(comp.data(), which ? sp_[q].vel[d] : sp_[q].pos[d], (size_t)n * sizeof(T),
                                hipMemcpyDeviceToHost));
                for (int64_t i = 0; i < n; ++i) inter[(size_t)(i * D + d)] = comp[(size_t)i];
            }
            f.writeData<T>(name + (which ? ".data.content.velocity" : ".data.content.position"), inter.data(),
                           {(size_t)n, (size_t)D, 1});
        }
    }

    template <typename T>
    void restore_species(psum::serialization::mas_file& f, int q)
    {
        const int D = cfg_.grid.dim;
        const std::string name = "species" + std::to_string(q);
        std::vector<size_t> sp_shape, sv_shape;
        const std::vector<T> pos = f.read<T>(name + ".data.content.position", &sp_shape);
        const std::vector<T> vel = f.read<T>(name + ".data.content.velocity", &sv_shape);
        if (sp_shape.size() != 3 || sp_shape[1] != (size_t)D || sp_shape[2] != 1 || sv_shape != sp_shape)
            throw std::invalid_argument("checkpoint: " + name + " has the wrong shape");
        const int64_t n = (int64_t)sp_shape[0];
        if (n > sp_[q].sp.capacity) throw std::invalid_argument("checkpoint: " + name + " exceeds the capacity");
        std::vector<T> comp((size_t)std::max<int64_t>(n, 1));
        for (int which = 0; which < 2; ++which)
            for (int d = 0; d < D; ++d) {
                const std::vector<T>& src = which ? vel : pos;
                for (int64_t i = 0; i < n; ++i) comp[(size_t)i] = src[(size_t)(i * D + d)];
                S_TRY(hipMemcpy(which ? sp_[q].vel[d] : sp_[q].pos[d], comp.data(), (size_t)n * sizeof(T),
                                hipMemcpyHostToDevice));
            }
        sp_[q].n = n;
    }

    void save_checkpoint(const char* path)
    {
        using psum::serialization::mas_file;
        S_TRY(hipStreamSynchronize(st_));
        int dev = 0;
        S_TRY(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        S_TRY(hipGetDeviceProperties(&prop, dev));
        const std::string devname = std::string("[#0]") + prop.name;
        mas_file f(path, mas_file::replaceMode);
        const int D = cfg_.grid.dim;
        f.writeScalar<int64_t>("psum.step", step_);
        f.writeScalar<int32_t>("psum.dtype", cfg_.dtype);
        f.writeScalar<int32_t>("psum.nspecies", cfg_.nspecies);
        f.writeScalar<int32_t>("psum.mg_predicted", mg_predicted(mg_));
        f.writeScalar<int32_t>("psum.adjoint_predicted", adj_predicted_);
        f.writeScalar<int32_t>("psum.have_prev", have_prev_ ? 1 : 0);
        for (int q = 0; q < cfg_.nspecies; ++q) {
            if (cfg_.dtype == PSUM_F64) save_species<double>(f, q, devname);
            else save_species<float>(f, q, devname);
        }
        std::vector<double> lower(D), upper(D);
        std::vector<int32_t> cells(D);
        for (int d = 0; d < D; ++d) {
            lower[d] = cfg_.grid.lb[d];
            upper[d] = cfg_.grid.ub[d];
            cells[d] = cfg_.grid.cells[d];
        }
        f.writeScalar<int32_t>("phi.location", 0);   // node centred
        f.writeData<double>("phi.grid.lower", lower.data(), {(size_t)D, 1});
        f.writeData<double>("phi.grid.upper", upper.data(), {(size_t)D, 1});
        f.writeData<int32_t>("phi.grid.cells", cells.data(), {(size_t)D, 1});
        std::vector<double> h((size_t)nn_);
        S_TRY(hipMemcpy(h.data(), phi_, (size_t)nn_ * sizeof(double), hipMemcpyDeviceToHost));
        f.writeData("phi.content.device", devname);
        f.writeData<double>("phi.content.content", h.data(), {(size_t)nn_, 1});
        if (cfg_.extrapolate && have_prev_) {
            S_TRY(hipMemcpy(h.data(), phi_prev_, (size_t)nn_ * sizeof(double), hipMemcpyDeviceToHost));
            f.writeData<double>("psum.phi_prev", h.data(), {(size_t)nn_, 1});
        }
    }

    void restore_checkpoint(const char* path)
    {
        using psum::serialization::mas_file;
        mas_file f(path, mas_file::autoMode);
        auto scalar_i32 = [&](const char* k) { return f.read<int32_t>(k).at(0); };
        if (scalar_i32("psum.dtype") != cfg_.dtype) throw std::invalid_argument("checkpoint: dtype differs");
        if (scalar_i32("psum.nspecies") != cfg_.nspecies)
            throw std::invalid_argument("checkpoint: species count differs");
        const std::vector<int32_t> cells = f.read<int32_t>("phi.grid.cells");
        if ((int)cells.size() != cfg_.grid.dim) throw std::invalid_argument("checkpoint: grid dimension differs");
        for (int d = 0; d < cfg_.grid.dim; ++d)
            if (cells[d] != cfg_.grid.cells[d]) throw std::invalid_argument("checkpoint: grid differs");
        const std::vector<double> phi = f.read<double>("phi.content.content");
        if ((int64_t)phi.size() != nn_) throw std::invalid_argument("checkpoint: phi size differs");
        S_TRY(hipStreamSynchronize(st_));
        for (int q = 0; q < cfg_.nspecies; ++q) {
            if (cfg_.dtype == PSUM_F64) restore_species<double>(f, q);
            else restore_species<float>(f, q);
        }
        S_TRY(hipMemcpy(phi_, phi.data(), (size_t)nn_ * sizeof(double), hipMemcpyHostToDevice));
        have_prev_ = scalar_i32("psum.have_prev") != 0;
        if (cfg_.extrapolate && have_prev_) {
            const std::vector<double> prev = f.read<double>("psum.phi_prev");
            if ((int64_t)prev.size() != nn_) throw std::invalid_argument("checkpoint: phi_prev size differs");
            S_TRY(hipMemcpy(phi_prev_, prev.data(), (size_t)nn_ * sizeof(double), hipMemcpyHostToDevice));
        } else {
            have_prev_ = false;
        }
        mg_set_predicted(mg_, scalar_i32("psum.mg_predicted"));
        adj_predicted_ = scalar_i32("psum.adjoint_predicted");
        step_ = f.read<int64_t>("psum.step").at(0);
        acc_valid_ = false;
    }

    // lrho = S^T lphi (lphi_ is overwritten).  S^T is the homogeneous solve
    // of the same pinned / Dirichlet system (rows of pinned and fixed nodes
    // take 0; no addback), warm started from the previous lrho.  Its result
    // may carry a constant on periodic grids, which the gradient gather
    // annihilates (the CIC gradient weights sum to zero).
    void adjoint_solve()
    {
        // the adjoint solves keep their own cycle predictor, so that their
        // sequence does not depend on how the recompute is segmented
        const int fwd_pred = mg_predicted(mg_);
        mg_set_predicted(mg_, adj_predicted_);
        mg_solve_homogeneous(mg_, lphi_, lrho_);
        adj_predicted_ = mg_predicted(mg_);
        mg_set_predicted(mg_, fwd_pred);
        adj_cycles_ += mg_last_cycles(mg_);
    }

    // The reverse sweep keeps lx, lv, phi and lrho in FP64 for either particle
    // precision: FP32 positions use the PSUM_F32_ADJ64 kernels.
    int adj_dtype() const { return cfg_.dtype == PSUM_F32 ? PSUM_F32_ADJ64 : PSUM_F64; }
    static constexpr size_t kAdjBytes = sizeof(double);

    // lx += w_dep grad(lrho)(x) for every species, and the fixed-point scale
    // of the next adjoint push from the bound sum_e |lv_e + dt lx_e|
    void adjoint_gather_all(void* const* xs)
    {
        const int D = cfg_.grid.dim;
        double dr = 0.0;
        for (int d = 0; d < D; ++d) dr = std::max(dr, cfg_.grid.delta_r[d]);
        double maxfac = 0.0;
        for (int q = 0; q < cfg_.nspecies; ++q)
            maxfac = std::max(maxfac, std::fabs(sp_[q].sp.q / sp_[q].sp.m) * cfg_.dt * dr);
        for (int q = 0; q < cfg_.nspecies; ++q) {
            CHK(psum_hip_adjoint_gather(&cfg_.grid, adj_dtype(), xs + q * 3, &lx_[q * 3], &lv_[q * 3],
                                        sp_[q].n, lrho_, sp_[q].wdep, cfg_.dt, bmax_, st_));
            CHK(psum_hip_adjoint_scale(bmax_, psum_hip_adjoint_blocks(sp_[q].n), maxfac, adj_scale_,
                                       q > 0 ? 1 : 0, st_));
        }
    }

    void adjoint_push_all(void* const* xs, const void* phi_t)
    {
        hipLaunchKernelGGL(k_zero_i64, dim3(nblk(nn_)), dim3(kT), 0, st_, (long long*)adj_acc_, nn_);
        for (int q = 0; q < cfg_.nspecies; ++q)
            CHK(psum_hip_adjoint_push(&cfg_.grid, adj_dtype(), xs + q * 3, &lx_[q * 3], &lv_[q * 3],
                                      sp_[q].n, phi_t, sp_[q].sp.q / sp_[q].sp.m, cfg_.dt,
                                      adj_scale_, adj_acc_, st_));
        CHK(psum_hip_fixed_to_double_dev(adj_acc_, nn_, adj_scale_, lphi_, st_));
    }

    void zero_dev(void* p, int64_t bytes)
    {
        if (bytes > 0)
            hipLaunchKernelGGL(k_zero_any, dim3(nblk(bytes)), dim3(kT), 0, st_, (unsigned char*)p, bytes);
    }

    double record(int steps, int ckpt, double* t_ms)
    {
        require_cic("psum_stepper_record");
        if (steps < 0 || ckpt < 1) throw std::invalid_argument("psum_stepper_record: steps >= 0, ckpt >= 1");
        release_record();
        adj_buffers();
        set_ckpt_capacity(std::min(ckpt, std::max(steps, 1)));
        no_sort_ = true;
        rec_steps_ = steps;
        rec_ckpt_ = ckpt;
        S_TRY(hipEventRecord(aev_[0], st_));
        const int nck = std::max(1, (steps + ckpt - 1) / ckpt);
        cks_.assign(nck, Snapshot{});
        for (auto& c : cks_) snap_alloc(c);
        for (int j = 0; j < nck; ++j) {
            snap_save(cks_[j]);
            const int len = std::min(ckpt, steps - j * ckpt);
            if (len > 0) run_steps(len, nullptr, false, false);
        }
        const double loss = field_energy_now();
        S_TRY(hipEventRecord(aev_[1], st_));
        S_TRY(hipEventSynchronize(aev_[1]));
        float ms = 0;
        S_TRY(hipEventElapsedTime(&ms, aev_[0], aev_[1]));
        if (t_ms) *t_ms = ms;
        recorded_ = true;
        return loss;
    }

    // seed: dL_total/dL (the upstream gradient); gx/gv[s*3+d]: device outputs
    void reverse(double seed, void* const* gx, void* const* gv, double* t_ms)
    {
        require_cic("psum_stepper_reverse");
        if (!recorded_) throw std::invalid_argument("psum_stepper_reverse: no recorded forward");
        const int D = cfg_.grid.dim;
        const int steps = rec_steps_, ckpt = rec_ckpt_;
        S_TRY(hipEventRecord(aev_[0], st_));
        adj_cycles_ = 0;
        adj_predicted_ = 1;
        // final: lphi = seed dW/dphi(phi^N) (phi_ holds phi^N after record)
        CHK(psum_hip_field_energy_grad(&cfg_.grid, phi_, cfg_.eps * seed, lphi_, st_));
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                zero_dev(lx_[q * 3 + d], sp_[q].n * (int64_t)kAdjBytes);
                zero_dev(lv_[q * 3 + d], sp_[q].n * (int64_t)kAdjBytes);
            }
        zero_dev(lrho_, nn_ * (int64_t)sizeof(double));
        adjoint_solve();
        {
            std::vector<void*> xn(cfg_.nspecies * 3, nullptr);
            for (int q = 0; q < cfg_.nspecies; ++q)
                for (int d = 0; d < D; ++d) xn[q * 3 + d] = sp_[q].pos[d];
            adjoint_gather_all(xn.data());
        }
        const int nck = (int)cks_.size();
        for (int j = nck - 1; j >= 0; --j) {
            const int s0 = j * ckpt, len = std::min(ckpt, steps - s0);
            if (len <= 0) continue;
            snap_restore(cks_[j]);
            for (int k = 0; k < len; ++k) {
                for (int q = 0; q < cfg_.nspecies; ++q)
                    for (int d = 0; d < D; ++d)
                        S_TRY(hipMemcpyAsync(seg_x_[k][q * 3 + d], sp_[q].pos[d], (size_t)sp_[q].n * esz_,
                                             hipMemcpyDeviceToDevice, st_));
                run_steps(1, nullptr, false, false);   // phi^(s0+k), then the push
                S_TRY(hipMemcpyAsync(seg_phi_[k], phi_, (size_t)nn_ * kAdjBytes, hipMemcpyDeviceToDevice, st_));
            }
            for (int k = len - 1; k >= 0; --k) {
                adjoint_push_all(seg_x_[k].data(), seg_phi_[k]);
                adjoint_solve();
                adjoint_gather_all(seg_x_[k].data());
            }
        }
        // gradients in the stepper's dtype (FP32 steppers: one rounding of
        // the FP64 adjoint state)
        auto out = [&](void* dst, const void* src, int64_t n) {
            if (cfg_.dtype == PSUM_F32)
                CHK(psum_hip_f64_to_f32(static_cast<const double*>(src), static_cast<float*>(dst), n, st_));
            else
                S_TRY(hipMemcpyAsync(dst, src, (size_t)n * kAdjBytes, hipMemcpyDeviceToDevice, st_));
        };
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                if (gx && gx[q * 3 + d]) out(gx[q * 3 + d], lx_[q * 3 + d], sp_[q].n);
                if (gv && gv[q * 3 + d]) out(gv[q * 3 + d], lv_[q * 3 + d], sp_[q].n);
            }
        S_TRY(hipEventRecord(aev_[1], st_));
        snap_restore(cks_[0]);
        S_TRY(hipStreamSynchronize(st_));
        float ms = 0;
        S_TRY(hipEventElapsedTime(&ms, aev_[0], aev_[1]));
        if (t_ms) *t_ms = ms;
        release_record();
    }

    void release_record()
    {
        for (auto& c : cks_) snap_free(c);
        cks_.clear();
        recorded_ = false;
        no_sort_ = false;
    }

    void adj_buffers()
    {
        if (lphi_) return;
        const int D = cfg_.grid.dim;
        S_TRY(hipMalloc(&lphi_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&lrho_, nn_ * sizeof(double)));
        S_TRY(hipMalloc(&adj_acc_, nn_ * sizeof(int64_t)));
        S_TRY(hipMalloc(&adj_scale_, 4 * sizeof(double)));
        int64_t nbmax = 1;
        for (int q = 0; q < cfg_.nspecies; ++q)
            nbmax = std::max(nbmax, psum_hip_adjoint_blocks(sp_[q].sp.capacity));
        S_TRY(hipMalloc(&bmax_, (size_t)nbmax * sizeof(double)));
        for (int q = 0; q < cfg_.nspecies; ++q)
            for (int d = 0; d < D; ++d) {
                const size_t b = (size_t)std::max<int64_t>(1, sp_[q].sp.capacity) * kAdjBytes;
                S_TRY(hipMalloc(&lx_[q * 3 + d], b));
                S_TRY(hipMalloc(&lv_[q * 3 + d], b));
            }
        if (!aev_[0]) {
            S_TRY(hipEventCreate(&aev_[0]));
            S_TRY(hipEventCreate(&aev_[1]));
        }
    }

    void set_ckpt_capacity(int ckpt)
    {
        if ((int)seg_x_.size() >= ckpt) return;
        const int D = cfg_.grid.dim;
        for (int k = (int)seg_x_.size(); k < ckpt; ++k) {
            seg_x_.emplace_back(cfg_.nspecies * 3, nullptr);
            seg_phi_.push_back(nullptr);
            for (int q = 0; q < cfg_.nspecies; ++q)
                for (int d = 0; d < D; ++d)
                    S_TRY(hipMalloc(&seg_x_[k][q * 3 + d], (size_t)std::max<int64_t>(1, sp_[q].sp.capacity) * esz_));
            S_TRY(hipMalloc(&seg_phi_[k], (size_t)nn_ * kAdjBytes));
        }
    }

    int64_t adjoint_cycles() const { return adj_cycles_; }

    // ---------------------------------------------------------- batch hooks
    void set_sort_shift(int b) { sort_shift_ = b; }
    int sort_shift() const { return sort_shift_; }
    void set_sort_every(int k) { cfg_.sort_every = k; }
    // species q sorts every k steps (k > 0), or on the stepper's interval
    // (k = 0)
    void set_species_sort_every(int q, int k)
    {
        if (q < 0 || q >= cfg_.nspecies || k < 0)
            throw std::invalid_argument("psum_stepper_set_species_sort: species in range, interval >= 0");
        sp_every_[q] = k;
    }
    int species_sort_every(int q) const { return sp_every_[q] > 0 ? sp_every_[q] : cfg_.sort_every; }
    bool sort_due(int q) const
    {
        const int k = species_sort_every(q);
        return k > 0 && step_ % k == 0 && sp_[q].n > 0;
    }
    bool any_sort_due() const
    {
        for (int q = 0; q < cfg_.nspecies; ++q)
            if (sort_due(q)) return true;
        return false;
    }
    // the species whose interval divides the step count sort before the step
    void sort_due_species()
    {
        for (int q = 0; q < cfg_.nspecies; ++q)
            if (sort_due(q)) sort_species(q);
    }
    void invalidate_deposit() { acc_valid_ = false; }
    void set_interp_order(int o)
    {
        if (o < 1 || o > 4) throw std::invalid_argument("psum_stepper_set_interp_order: order must be 1..4");
        order_ = o;
        acc_valid_ = false;
    }
    void mg_tuning(const char* o)
    {
        mg_set_tuning(mg_, o);
        ++tuning_epoch_;
    }
    void set_smoothing(int pre, int post)
    {
        mg_set_smoothing(mg_, pre, post);
        ++tuning_epoch_;
    }
    // For each candidate (pre[c], post[c]): a sort (so that every candidate
    // sees the same particle order), one untimed step that checks the
    // residual after every cycle (the timed solves then start from the cycle
    // count this candidate needs), and `steps` steps timed with HIP events
    // around the whole step (no sorts, tolerance-controlled solves).  Keeps
    // the fastest.  The state advances.
    void tune_smoothing(const int* pre, const int* post, int n, int steps, int* best_pre, int* best_post,
                        double* ms_out)
    {
        if (n < 1 || steps < 1 || !pre || !post) throw std::invalid_argument("tune_smoothing: n >= 1, steps >= 1");
        if (!aev_[0]) {
            S_TRY(hipEventCreate(&aev_[0]));
            S_TRY(hipEventCreate(&aev_[1]));
        }
        double best = 1e300;
        int bc = 0;
        for (int c = 0; c < n; ++c) {
            if (cfg_.sort_every > 0 && !no_sort_) sort_all();
            set_smoothing(pre[c], post[c]);
            mg_set_predicted(mg_, 1);
            step_body(false);
            S_TRY(hipEventRecord(aev_[0], st_));
            for (int s = 0; s < steps; ++s) step_body(false);
            S_TRY(hipEventRecord(aev_[1], st_));
            S_TRY(hipEventSynchronize(aev_[1]));
            float ms = 0;
            S_TRY(hipEventElapsedTime(&ms, aev_[0], aev_[1]));
            const double per = (double)ms / steps;
            if (ms_out) ms_out[c] = per;
            if (per < best) {
                best = per;
                bc = c;
            }
        }
        set_smoothing(pre[bc], post[bc]);
        if (best_pre) *best_pre = pre[bc];
        if (best_post) *best_post = post[bc];
    }
    // changes with every mg_tuning or set_smoothing call (a batch re-captures
    // its graphs)
    int64_t tuning_epoch() const { return tuning_epoch_; }
    void require_cic(const char* what) const
    {
        if (order_ != 1) throw std::invalid_argument(std::string(what) + ": the adjoint covers CIC (order 1) only");
    }
    int sort_every() const { return cfg_.sort_every; }
    int nspecies() const { return cfg_.nspecies; }
    void set_stable_buffers(bool on) { stable_buffers_ = on; }
    void set_fixed_cycles(int k)
    {
        mg_set_fixed_cycles(mg_, k);
        fixed_cycles_ = k;
    }
    int fixed_cycles() const { return fixed_cycles_; }
    int max_cycles_seen() const { return max_cyc_seen_; }
    void reset_cycles_seen() { max_cyc_seen_ = 0; }
    hipStream_t stream() const { return st_; }
    int64_t steps_done() const { return step_; }
    void count_step() { ++step_; }
    void uncount_step() { --step_; }
    // total push milliseconds since set_timing(true), after a stream sync;
    // ms_sp: the same per species
    void flush_timing(double* ms, double* ms_sp = nullptr)
    {
        flush_events();
        if (ms) *ms = push_ms_;
        if (ms_sp)
            for (int q = 0; q < cfg_.nspecies; ++q) ms_sp[q] = push_ms_sp_[q];
    }
    bool acc_valid() const { return acc_valid_; }
    void prepare_capture()
    {
        if (!acc_valid_) deposit_all();
        if (timing_) set_timing(false);
    }
    void hist_reset(int64_t rows)
    {
        reserve_hist(rows);
        hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(kT), 0, st_, (long long*)hist_count_, (int64_t)1);
    }
    int64_t hist_read(double* out, int64_t max_rows)
    {
        long long rows = 0;
        S_TRY(hipMemcpyAsync(&rows, hist_count_, sizeof rows, hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
        const int64_t r = std::min<int64_t>(std::min<int64_t>(rows, hist_cap_), max_rows);
        if (out && r > 0) {
            S_TRY(hipMemcpyAsync(out, hist_, (size_t)r * 5 * sizeof(double), hipMemcpyDeviceToHost, st_));
            S_TRY(hipStreamSynchronize(st_));
        }
        return r;
    }
    int64_t particles() const
    {
        int64_t n = 0;
        for (int q = 0; q < cfg_.nspecies; ++q) n += sp_[q].n;
        return n;
    }
    // Autotuner: `warmup` tolerance-controlled steps calibrate the fixed
    // V-cycle count (the most any solve needed, plus one); then for each sort
    // bin shift 0..2 a timed sort of each species and `probe` timed steps give
    // per species t_sort and the push time t_push(j) j steps after its sort,
    // and its interval K in {5, 10, 20, 50, 100, 200} minimising
    // (t_sort + sum_{j<K} t_push(j)) / K (t_push beyond the probe window
    // extrapolated linearly from its second half).  The shift with the
    // smallest summed cost is kept with its per-species intervals; the
    // returned sort_every is the shortest of them.  The state has advanced.
    void tune(int warmup, int probe, psum_batch_tuning* out)
    {
        if (warmup < 1 || probe < 2) throw std::invalid_argument("tune: warmup >= 1, probe >= 2");
        static const int kCand[] = {5, 10, 20, 50, 100, 200};
        const int ns = cfg_.nspecies;
        // the replayed step warm-starts every solve from the previous
        // potential: the first warm-up step (a solve from a cold start) does
        // not enter the calibration
        if (warmup > 1) run_steps(1, nullptr, true, false);
        reset_cycles_seen();
        run_steps(warmup > 1 ? warmup - 1 : 1, nullptr, true, false);
        const int cycles = max_cycles_seen() + 1;
        int best_shift = 0, best_k[4] = {0, 0, 0, 0};
        double best_cost = 1e300, best_sort = 0.0, best_push = 0.0;
        for (int shift = 0; shift <= 2; ++shift) {
            set_sort_shift(shift);
            double t_sort[4] = {0, 0, 0, 0};
            for (int q = 0; q < ns; ++q) t_sort[q] = timed_sort(q);
            std::vector<double> tp((size_t)probe * ns);
            for (int j = 0; j < probe; ++j) {
                set_timing(true);
                step_body(false);
                S_TRY(hipStreamSynchronize(st_));
                double ms = 0.0, ms_sp[4] = {0, 0, 0, 0};
                flush_timing(&ms, ms_sp);
                for (int q = 0; q < ns; ++q) tp[(size_t)j * ns + q] = ms_sp[q];
            }
            set_timing(false);
            double cost = 0.0, sort_sum = 0.0, push0 = 0.0;
            int ks[4] = {0, 0, 0, 0};
            for (int q = 0; q < ns; ++q) {
                const int h0 = probe / 2;
                double sx = 0, sy = 0, sxx = 0, sxy = 0;
                for (int j = h0; j < probe; ++j) {
                    const double y = tp[(size_t)j * ns + q];
                    sx += j; sy += y; sxx += (double)j * j; sxy += j * y;
                }
                const double m = probe - h0;
                const double den = m * sxx - sx * sx;
                const double slope = den > 0 ? std::max(0.0, (m * sxy - sx * sy) / den) : 0.0;
                const double icpt = (sy - slope * sx) / m;
                double cq = 1e300;
                for (int K : kCand) {
                    double sum = t_sort[q];
                    for (int j = 0; j < K; ++j) sum += j < probe ? tp[(size_t)j * ns + q] : icpt + slope * j;
                    if (sum / K < cq) {
                        cq = sum / K;
                        ks[q] = K;
                    }
                }
                cost += cq;
                sort_sum += t_sort[q];
                push0 += tp[q];
            }
            if (cost < best_cost) {
                best_cost = cost;
                best_shift = shift;
                for (int q = 0; q < ns; ++q) best_k[q] = ks[q];
                best_sort = sort_sum;
                best_push = push0;
            }
        }
        set_sort_shift(best_shift);
        int kmin = 0;
        for (int q = 0; q < ns; ++q) {
            set_species_sort_every(q, best_k[q]);
            if (best_k[q] > 0 && (kmin == 0 || best_k[q] < kmin)) kmin = best_k[q];
        }
        if (out) {
            out->fixed_cycles = cycles;
            out->sort_every = kmin;
            out->sort_shift = best_shift;
            out->pad_ = 0;
            out->sort_ms = best_sort;
            out->push_ms = best_push;
            out->push_sort_cost_ms = best_cost;
        }
    }

    // HIP-event time of one sort pass of species q (all species: q < 0), ms
    float timed_sort(int q = -1)
    {
        if (!aev_[0]) {
            S_TRY(hipEventCreate(&aev_[0]));
            S_TRY(hipEventCreate(&aev_[1]));
        }
        S_TRY(hipEventRecord(aev_[0], st_));
        if (q < 0) sort_all();
        else sort_species(q);
        S_TRY(hipEventRecord(aev_[1], st_));
        S_TRY(hipEventSynchronize(aev_[1]));
        float ms = 0;
        S_TRY(hipEventElapsedTime(&ms, aev_[0], aev_[1]));
        return ms;
    }

    void phi(double* out)
    {
        S_TRY(hipMemcpyAsync(out, phi_, nn_ * sizeof(double), hipMemcpyDeviceToHost, st_));
        S_TRY(hipStreamSynchronize(st_));
    }

    // rho of the current positions (the deposit the next step solves with,
    // before neutralisation and fixed-node values): acc * 2^-S + rho_bg.
    // Also the deposit scale 2^S.
    void rho(double* out, double* scale)
    {
        if (!acc_valid_) deposit_all();
        if (out) {
            CHK(psum_hip_fixed_to_double(acc_, nn_, 1.0 / scale_, cfg_.rho_bg, b_, st_));
            S_TRY(hipMemcpyAsync(out, b_, nn_ * sizeof(double), hipMemcpyDeviceToHost, st_));
        }
        if (scale) *scale = scale_;
        S_TRY(hipStreamSynchronize(st_));
    }

    void stats(int64_t* cycles, double* rel, int64_t* counters)
    {
        if (cycles) *cycles = cycles_;
        if (rel) *rel = last_rel_;
        if (counters) {
            S_TRY(hipMemcpyAsync(counters, counters_, 4 * sizeof(int64_t), hipMemcpyDeviceToHost, st_));   // [3]: int64 wraps
            S_TRY(hipStreamSynchronize(st_));
        }
    }

    void sync() { S_TRY(hipStreamSynchronize(st_)); }

private:
    bool no_sort_ = false;
    double *lphi_ = nullptr, *lrho_ = nullptr, *adj_scale_ = nullptr, *bmax_ = nullptr;
    int64_t* adj_acc_ = nullptr;
    void* lx_[12] = {};
    void* lv_[12] = {};
    std::vector<std::vector<void*>> seg_x_;
    std::vector<void*> seg_phi_;
    std::vector<Snapshot> cks_;
    bool recorded_ = false;
    int rec_steps_ = 0, rec_ckpt_ = 1;
    int64_t adj_cycles_ = 0;
    int adj_predicted_ = 1;
    int sort_shift_ = 0;
    bool stable_buffers_ = false;
    int fixed_cycles_ = 0;
    int max_cyc_seen_ = 0;
    hipEvent_t aev_[2] = {nullptr, nullptr};
    static constexpr int kEv = 64;
    hipEvent_t ev_[2 * kEv] = {};
    int ev_sp_[kEv] = {};          // species of each event pair
    double push_ms_sp_[4] = {};    // push milliseconds per species since set_timing(true)
    int sp_every_[4] = {};         // per-species sort interval; 0: cfg_.sort_every
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
            push_ms_sp_[ev_sp_[i]] += ms;
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
    int64_t tuning_epoch_ = 0;
    Species sp_[4];
    int64_t max_cap_ = 0;
    double scale_ = 1.0;
    int64_t* acc_ = nullptr;
    double *b_ = nullptr, *phi_ = nullptr, *diag_ = nullptr, *red_ = nullptr, *fe_ = nullptr;
    void* hist_count_ = nullptr;      // device row counter of hist_
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
    int order_ = 1;   // interpolation order: 1 CIC, 2..4 B-splines
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
            if (order_ > 1)
                CHK(psum_hip_deposit_bspline(&cfg_.grid, cfg_.dtype, order_, sp.pos, sp.n, sp.wdep, scale_, acc_,
                                             counters_, st_));
            else
                CHK(psum_hip_deposit_cic(&cfg_.grid, cfg_.dtype, sp.pos, sp.n, sp.wdep, scale_, acc_, counters_, st_));
        }
        acc_valid_ = true;
    }

public:
    void sort_all()
    {
        for (int q = 0; q < cfg_.nspecies; ++q) sort_species(q);
    }

    void sort_species(int q)
    {
        const int D = cfg_.grid.dim;
        sort_buffers();
        // keys: the 2^shift-cell bin refined by the velocity signs, so that
        // co-moving particles stay adjacent between sorts (long equal-cell
        // runs for the deposit even when beams interpenetrate)
        const uint32_t nbins = sort_shift_ ? bin_count(cfg_.grid, sort_shift_) : (uint32_t)nn_;
        const bool vel_keys = ((uint64_t)nbins << D) < (uint64_t)PSUM_KEY_INVALID;
        const uint32_t nkeys = vel_keys ? nbins << D : nbins;
        {
            Species& sp = sp_[q];
            if (sp.n == 0) return;
            CHK(bin_keys(&cfg_.grid, cfg_.dtype, sp.pos, sp.n, keys_, sort_shift_, st_, vel_keys ? sp.vel : nullptr));
            CHK(psum_hip_sort(keys_, sp.n, nkeys, skeys_, perm_, sort_work_, sort_bytes_, st_));
            void* src[6];
            void* dst[6];
            int nc = 0;
            for (int d = 0; d < D; ++d) { src[nc] = sp.pos[d]; dst[nc] = sp.pos2[d]; ++nc; }
            for (int d = 0; d < D; ++d) { src[nc] = sp.vel[d]; dst[nc] = sp.vel2[d]; ++nc; }
            CHK(psum_hip_permute(cfg_.dtype, nc, src, dst, perm_, sp.n, st_));
            if (stable_buffers_) {
                // keep the arrays a captured graph points at
                for (int d = 0; d < D; ++d) {
                    S_TRY(hipMemcpyAsync(sp.pos[d], sp.pos2[d], (size_t)sp.n * esz_, hipMemcpyDeviceToDevice, st_));
                    S_TRY(hipMemcpyAsync(sp.vel[d], sp.vel2[d], (size_t)sp.n * esz_, hipMemcpyDeviceToDevice, st_));
                }
            } else {
                for (int d = 0; d < D; ++d) {
                    std::swap(sp.pos[d], sp.pos2[d]);
                    std::swap(sp.vel[d], sp.vel2[d]);
                }
            }
            // drop particles whose key is invalid (absorbed): they sort last
            unsigned long long nv = (unsigned long long)sp.n;
            S_TRY(hipMemcpyAsync(nvalid_, &nv, sizeof nv, hipMemcpyHostToDevice, st_));
            hipLaunchKernelGGL(k_count_valid, dim3(1), dim3(64), 0, st_, skeys_, sp.n, nvalid_);
            S_TRY(hipMemcpyAsync(&nv, nvalid_, sizeof nv, hipMemcpyDeviceToHost, st_));
            S_TRY(hipStreamSynchronize(st_));
            sp.n = (int64_t)nv;
        }
    }
};

// ------------------------------------------------------------------ batch
// Independent instances stepped together.  One step of every instance is
// captured into a single hipGraph (fork/join across the instances' streams,
// so that small kernels of different instances run concurrently) and
// replayed every step.  Inside the graph the solvers run a fixed V-cycle
// count; sorting runs between replays at per-instance intervals.  tune()
// is the autotuner: it calibrates each instance's cycle count (the most a
// tolerance-controlled warm-up needed, plus one) and picks its sort bin size
// and sort interval from measured sort and push times.
class Batch {
public:
    Batch(const psum_step_config* cfgs, int n)
    {
        if (n < 1) throw std::invalid_argument("psum_batch_create: need at least one instance");
        for (int i = 0; i < n; ++i) {
            psum_step_config c = cfgs[i];
            c.mg_graph = 0;       // the batch graph captures the solver's kernels itself
            c.extrapolate = 0;
            inst_.emplace_back(new Stepper(c));
            inst_.back()->set_stable_buffers(true);
        }
        S_TRY(hipStreamCreate(&origin_));
        S_TRY(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        join_.resize(n);
        for (auto& e : join_) S_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        join2_.resize(n);
        for (auto& e : join2_) S_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        S_TRY(hipEventCreateWithFlags(&fork2_, hipEventDisableTiming));
        plan_ = mg_batch_plan_create();
        S_TRY(hipEventCreate(&t0_));
        S_TRY(hipEventCreate(&t1_));
    }

    ~Batch()
    {
        release_lanes();
        for (auto& e : join_) if (e) (void)hipEventDestroy(e);
        for (auto& e : join2_) if (e) (void)hipEventDestroy(e);
        for (hipEvent_t e : {fork_, fork2_, t0_, t1_}) if (e) (void)hipEventDestroy(e);
        if (plan_) mg_batch_plan_destroy(plan_);
        if (origin_) (void)hipStreamDestroy(origin_);
    }

    int size() const { return (int)inst_.size(); }
    // batched solver cores on (default) or off: off, each instance's solver
    // launches its own kernels inside the captured step
    void set_solver_batching(bool on)
    {
        batch_mg_ = on;
        drop();
    }
    // graph launches per step of the batched solver cores (0 when the step
    // does not batch them), and the ops they replace
    void solver_launches(int64_t* launches, int64_t* ops)
    {
        if (!captured_) capture();
        const bool b = batched_solve();
        if (launches) *launches = b ? mg_batch_plan_launches(plan_) : 0;
        if (ops) *ops = b ? mg_batch_plan_ops(plan_) : 0;
    }
    void set_lanes(int lanes)
    {
        if (lanes < 1 || lanes > 64) throw std::invalid_argument("psum_batch_set_lanes: lanes must be 1..64");
        lanes_ = lanes;
        release_lanes();
    }
    Stepper& at(int i)
    {
        if (i < 0 || i >= size()) throw std::invalid_argument("psum_batch: instance out of range");
        return *inst_[i];
    }

    void load(int i, int s, const void* const* pos, const void* const* vel, int64_t n)
    {
        at(i).load(s, pos, vel, n);
        drop();
    }

    void configure(int i, int cycles, int sort_every, int sort_shift)
    {
        Stepper& st = at(i);
        st.set_fixed_cycles(cycles);
        st.set_sort_every(sort_every);
        st.set_sort_shift(sort_shift);
        drop();
    }

    // Per instance: `warmup` tolerance-controlled steps (cycle calibration),
    // then for each bin shift a timed sort and `probe` timed steps.  The sort
    // interval K minimises (t_sort + sum_{j<K} t_push(j)) / K over
    // K in {5, 10, 20, 50, 100, 200}, with t_push(j) beyond the probe window
    // extrapolated linearly from its second half.
    void tune(int warmup, int probe, psum_batch_tuning* out)
    {
        for (int i = 0; i < size(); ++i) {
            psum_batch_tuning t{};
            at(i).tune(warmup, probe, &t);
            configure(i, t.fixed_cycles, t.sort_every, t.sort_shift);
            if (out) out[i] = t;
        }
    }

    // replay the captured step `steps` times; sorts due before a step run
    // eagerly on the instance's stream first
    void run(int steps, double* t_ms)
    {
        for (auto& p : inst_)
            if (p->fixed_cycles() < 1)
                throw std::invalid_argument("psum_batch_run: every instance needs a fixed cycle count "
                                            "(psum_batch_configure or psum_batch_tune)");
        int64_t epoch = 0;
        for (auto& p : inst_) epoch += p->tuning_epoch();
        if (captured_ && epoch != captured_epoch_) drop();
        if (!captured_) {
            capture();
            captured_epoch_ = epoch;
        }
        // instances reloaded since the capture need their deposit redone
        for (int i = 0; i < size(); ++i)
            if (!inst_[i]->acc_valid()) {
                inst_[i]->prepare_capture();
                S_TRY(hipEventRecord(join_[i], inst_[i]->stream()));
                S_TRY(hipStreamWaitEvent(origin_, join_[i], 0));
            }
        S_TRY(hipEventRecord(t0_, origin_));
        S_TRY(hipEventRecord(fork_, origin_));
        for (auto st : lane_st_) S_TRY(hipStreamWaitEvent(st, fork_, 0));
        for (int s = 0; s < steps; ++s) {
            for (size_t l = 0; l < lane_st_.size(); ++l) {
                // eager sorts due before this step run on the instances'
                // streams after the lane's previous replay
                bool forked = false;
                for (int i : lane_inst_[l]) {
                    Stepper& p = *inst_[i];
                    if (p.any_sort_due()) {
                        if (!forked) {
                            S_TRY(hipEventRecord(lane_fork_[l], lane_st_[l]));
                            forked = true;
                        }
                        S_TRY(hipStreamWaitEvent(p.stream(), lane_fork_[l], 0));
                        p.sort_due_species();
                        S_TRY(hipEventRecord(join_[i], p.stream()));
                        S_TRY(hipStreamWaitEvent(lane_st_[l], join_[i], 0));
                    }
                }
                S_TRY(hipGraphLaunch(lane_exec_[l], lane_st_[l]));
            }
            for (auto& p : inst_) p->count_step();
        }
        for (size_t l = 0; l < lane_st_.size(); ++l) {
            S_TRY(hipEventRecord(lane_join_[l], lane_st_[l]));
            S_TRY(hipStreamWaitEvent(origin_, lane_join_[l], 0));
        }
        S_TRY(hipEventRecord(t1_, origin_));
        S_TRY(hipEventSynchronize(t1_));
        float ms = 0;
        S_TRY(hipEventElapsedTime(&ms, t0_, t1_));
        if (t_ms) *t_ms = ms;
        for (auto& p : inst_) p->check_overflow();
    }

    void hist_reset(int64_t rows)
    {
        for (auto& p : inst_) p->hist_reset(rows);
        for (auto& p : inst_) S_TRY(hipStreamSynchronize(p->stream()));
        drop();   // the history capacity is baked into the graph
    }

private:
    static void S_TRY(hipError_t e)
    {
        if (e != hipSuccess) throw std::runtime_error(std::string("psum_batch: ") + hipGetErrorString(e));
    }

    std::vector<std::unique_ptr<Stepper>> inst_;
    hipStream_t origin_ = nullptr;
    hipEvent_t fork_ = nullptr, t0_ = nullptr, t1_ = nullptr;
    std::vector<hipEvent_t> join_;
    // lanes: the instances are split into `lanes_` groups balanced by
    // particle count; each group's step is one captured graph replayed on
    // the group's own stream, so groups run concurrently on the hardware
    // queues (branches inside one graph run in order)
    int lanes_ = 1;
    std::vector<hipStream_t> lane_st_;
    std::vector<hipEvent_t> lane_fork_, lane_join_;
    std::vector<std::vector<int>> lane_inst_;
    std::vector<hipGraphExec_t> lane_exec_;
    bool captured_ = false;
    int64_t captured_epoch_ = 0;     // sum of the instances' tuning epochs at capture
    // batched solver cores (one lane): the recorded ops of each instance and
    // their merged launches; batch_mg_ = 0 steps each solver on its own
    bool batch_mg_ = true;
    std::vector<std::vector<MgRecOp>> rec_ops_;
    MgBatchPlan* plan_ = nullptr;
    hipEvent_t fork2_ = nullptr;
    std::vector<hipEvent_t> join2_;

    void drop()
    {
        for (auto& e : lane_exec_)
            if (e) (void)hipGraphExecDestroy(e);
        lane_exec_.clear();
        captured_ = false;
    }

    void release_lanes()
    {
        drop();
        for (auto& st : lane_st_) if (st) (void)hipStreamDestroy(st);
        for (auto& e : lane_fork_) if (e) (void)hipEventDestroy(e);
        for (auto& e : lane_join_) if (e) (void)hipEventDestroy(e);
        lane_st_.clear();
        lane_fork_.clear();
        lane_join_.clear();
        lane_inst_.clear();
    }

    // greedy balance by particle count, largest first
    void plan_lanes()
    {
        release_lanes();
        const int L = std::max(1, std::min(lanes_, size()));
        std::vector<int> order(size());
        for (int i = 0; i < size(); ++i) order[i] = i;
        std::sort(order.begin(), order.end(),
                  [&](int a, int b) { return inst_[a]->particles() > inst_[b]->particles(); });
        std::vector<int64_t> load(L, 0);
        lane_inst_.assign(L, {});
        for (int i : order) {
            const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            lane_inst_[l].push_back(i);
            load[l] += std::max<int64_t>(1, inst_[i]->particles());
        }
        for (int l = 0; l < L; ++l) {
            hipStream_t st = nullptr;
            S_TRY(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
            lane_st_.push_back(st);
            hipEvent_t f = nullptr, j = nullptr;
            S_TRY(hipEventCreateWithFlags(&f, hipEventDisableTiming));
            S_TRY(hipEventCreateWithFlags(&j, hipEventDisableTiming));
            lane_fork_.push_back(f);
            lane_join_.push_back(j);
        }
    }

    // whether one step runs the instances' solver cores as batched launches:
    // one lane and every instance's solve recordable
    bool batched_solve() const
    {
        if (lanes_ != 1 || !batch_mg_) return false;
        for (auto& p : inst_)
            if (!p->solve_recordable()) return false;
        return true;
    }

    // One step of every instance, captured on the single lane's stream: each
    // instance's part before its solve on its own stream (joined back), the
    // batched solver cores of all instances on the lane stream, then each
    // instance's part after its solve.
    void capture_batched()
    {
        rec_ops_.assign(inst_.size(), {});
        std::vector<const std::vector<MgRecOp>*> per;
        for (size_t i = 0; i < inst_.size(); ++i) {
            inst_[i]->record_solve_core(rec_ops_[i]);
            per.push_back(&rec_ops_[i]);
        }
        mg_batch_plan_build(plan_, per);
        lane_exec_.assign(1, nullptr);
        hipStream_t ls = lane_st_[0];
        S_TRY(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
        // an exception below ends the capture and drops the partial graph
        CaptureGuard guard(ls);
        S_TRY(hipEventRecord(lane_fork_[0], ls));
        for (size_t i = 0; i < inst_.size(); ++i) {
            Stepper& st = *inst_[i];
            S_TRY(hipStreamWaitEvent(st.stream(), lane_fork_[0], 0));
            (void)st.solve_field_begin();
            S_TRY(hipEventRecord(join_[i], st.stream()));
            S_TRY(hipStreamWaitEvent(ls, join_[i], 0));
        }
        mg_batch_plan_launch(plan_, ls);
        S_TRY(hipEventRecord(fork2_, ls));
        for (size_t i = 0; i < inst_.size(); ++i) {
            Stepper& st = *inst_[i];
            S_TRY(hipStreamWaitEvent(st.stream(), fork2_, 0));
            st.solve_field_end();
            st.step_after_solve(true);
            S_TRY(hipEventRecord(join2_[i], st.stream()));
            S_TRY(hipStreamWaitEvent(ls, join2_[i], 0));
        }
        hipGraph_t graph = nullptr;
        guard.release();
        S_TRY(hipStreamEndCapture(ls, &graph));
        const hipError_t e = hipGraphInstantiate(&lane_exec_[0], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        S_TRY(e);
    }

    void capture()
    {
        if (lane_st_.empty()) plan_lanes();
        for (auto& p : inst_) p->prepare_capture();
        for (auto& p : inst_) S_TRY(hipStreamSynchronize(p->stream()));
        if (batched_solve()) {
            capture_batched();
            captured_ = true;
            for (auto& p : inst_) p->uncount_step();
            return;
        }
        lane_exec_.assign(lane_st_.size(), nullptr);
        for (size_t l = 0; l < lane_st_.size(); ++l) {
            hipStream_t ls = lane_st_[l];
            S_TRY(hipStreamBeginCapture(ls, hipStreamCaptureModeThreadLocal));
            // an exception below ends the capture and drops the partial graph
            CaptureGuard guard(ls);
            S_TRY(hipEventRecord(lane_fork_[l], ls));
            for (int i : lane_inst_[l]) {
                Stepper& st = *inst_[i];
                S_TRY(hipStreamWaitEvent(st.stream(), lane_fork_[l], 0));
                st.step_body(true);
                S_TRY(hipEventRecord(join_[i], st.stream()));
                S_TRY(hipStreamWaitEvent(ls, join_[i], 0));
            }
            hipGraph_t graph = nullptr;
            guard.release();
            S_TRY(hipStreamEndCapture(ls, &graph));
            const hipError_t e = hipGraphInstantiate(&lane_exec_[l], graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            S_TRY(e);
        }
        captured_ = true;
        // step_body counted one step per instance while capturing
        for (auto& p : inst_) p->uncount_step();
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

psum_status psum_stepper_rho(void* h, double* rho_host, double* scale)
{
    if (!h) { set_error("psum_stepper_rho: null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->rho(rho_host, scale); });
}

psum_status psum_stepper_step_count(void* h, int64_t* steps)
{
    if (!h || !steps) { set_error("psum_stepper_step_count: null argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { *steps = static_cast<Stepper*>(h)->steps_done(); });
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

psum_status psum_stepper_record(void* h, int steps, int ckpt, double* loss, double* t_ms)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        const double l = static_cast<Stepper*>(h)->record(steps, ckpt, t_ms);
        if (loss) *loss = l;
    });
}

psum_status psum_stepper_reverse(void* h, double seed, void* const* grad_x, void* const* grad_v,
                                 double* t_ms, int64_t* solve_cycles)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        s->reverse(seed, grad_x, grad_v, t_ms);
        if (solve_cycles) *solve_cycles = s->adjoint_cycles();
    });
}

psum_status psum_stepper_adjoint(void* h, int steps, int ckpt, double* loss, void* const* grad_x,
                                 void* const* grad_v, double* t_fwd_ms, double* t_bwd_ms)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        const double l = s->record(steps, ckpt, t_fwd_ms);
        s->reverse(1.0, grad_x, grad_v, t_bwd_ms);
        if (loss) *loss = l;
    });
}

psum_status psum_stepper_sort(void* h)
{
    if (!h) { set_error("psum_stepper_sort: null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        s->sort_all();
        s->invalidate_deposit();
    });
}

psum_status psum_stepper_set_interp_order(void* h, int order)
{
    if (!h) { set_error("psum_stepper_set_interp_order: null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->set_interp_order(order); });
}

psum_status psum_stepper_set_smoothing(void* h, int pre, int post)
{
    if (!h) { set_error("psum_stepper_set_smoothing: null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->set_smoothing(pre, post); });
}

psum_status psum_stepper_tune_smoothing(void* h, const int* pre, const int* post, int n, int steps, int* best_pre,
                                        int* best_post, double* ms)
{
    if (!h) { set_error("psum_stepper_tune_smoothing: null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->tune_smoothing(pre, post, n, steps, best_pre, best_post, ms); });
}

psum_status psum_stepper_mg_tuning(void* h, const char* options)
{
    if (!h || !options) { set_error("psum_stepper_mg_tuning: null argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->mg_tuning(options); });
}

psum_status psum_stepper_save(void* h, const char* path)
{
    if (!h || !path) { set_error("psum_stepper_save: null argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->save_checkpoint(path); });
}

psum_status psum_stepper_restore(void* h, const char* path)
{
    if (!h || !path) { set_error("psum_stepper_restore: null argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->restore_checkpoint(path); });
}

void psum_stepper_destroy(void* h) { delete static_cast<Stepper*>(h); }

psum_status psum_stepper_tune(void* h, int warmup, int probe, psum_batch_tuning* out)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->tune(warmup, probe, out); });
}

psum_status psum_stepper_set_schedule(void* h, int fixed_cycles, int sort_every, int sort_shift)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    if (sort_shift < 0 || sort_shift > 8 || sort_every < 0) {
        set_error("psum_stepper_set_schedule: sort_every >= 0, sort_shift in 0..8");
        return PSUM_ERR_ARG;
    }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        if (fixed_cycles > 0) s->set_fixed_cycles(fixed_cycles);
        s->set_sort_every(sort_every);
        s->set_sort_shift(sort_shift);
    });
}

psum_status psum_stepper_set_species_sort(void* h, int species, int every)
{
    if (!h) { set_error("null stepper"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Stepper*>(h)->set_species_sort_every(species, every); });
}

psum_status psum_stepper_species_sort(void* h, int species, int* every)
{
    if (!h || !every) { set_error("psum_stepper_species_sort: null argument"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        Stepper* s = static_cast<Stepper*>(h);
        if (species < 0 || species >= s->nspecies())
            throw std::invalid_argument("psum_stepper_species_sort: species out of range");
        *every = s->species_sort_every(species);
    });
}

void* psum_batch_create(const psum_step_config* cfgs, int n)
{
    if (!cfgs || n < 1) {
        set_error("psum_batch_create: need n >= 1 configs");
        return nullptr;
    }
    Batch* b = nullptr;
    if (guarded([&]() { b = new Batch(cfgs, n); }) != PSUM_OK) return nullptr;
    return b;
}

psum_status psum_batch_load(void* h, int inst, int s, const void* const* pos, const void* const* vel,
                            int64_t n)
{
    if (!h) { set_error("null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->load(inst, s, pos, vel, n); });
}

psum_status psum_batch_configure(void* h, int inst, int fixed_cycles, int sort_every, int sort_shift)
{
    if (!h) { set_error("null batch"); return PSUM_ERR_ARG; }
    if (sort_shift < 0 || sort_shift > 8 || sort_every < 0) {
        set_error("psum_batch_configure: sort_every >= 0, sort_shift in 0..8");
        return PSUM_ERR_ARG;
    }
    return guarded([&]() { static_cast<Batch*>(h)->configure(inst, fixed_cycles, sort_every, sort_shift); });
}

psum_status psum_batch_tune(void* h, int warmup, int probe, psum_batch_tuning* out)
{
    if (!h) { set_error("null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->tune(warmup, probe, out); });
}

psum_status psum_batch_reset_history(void* h, int64_t rows)
{
    if (!h || rows < 0) { set_error("psum_batch_reset_history: bad argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->hist_reset(rows); });
}

psum_status psum_batch_set_lanes(void* h, int lanes)
{
    if (!h) { set_error("null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->set_lanes(lanes); });
}

psum_status psum_batch_set_solver_batching(void* h, int on)
{
    if (!h) { set_error("psum_batch_set_solver_batching: null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->set_solver_batching(on != 0); });
}

psum_status psum_batch_solver_launches(void* h, int64_t* launches, int64_t* ops)
{
    if (!h) { set_error("psum_batch_solver_launches: null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->solver_launches(launches, ops); });
}

psum_status psum_batch_run(void* h, int steps, double* t_ms)
{
    if (!h || steps < 0) { set_error("psum_batch_run: bad argument"); return PSUM_ERR_ARG; }
    return guarded([&]() { static_cast<Batch*>(h)->run(steps, t_ms); });
}

psum_status psum_batch_history(void* h, int inst, double* out, int64_t max_rows, int64_t* rows)
{
    if (!h) { set_error("null batch"); return PSUM_ERR_ARG; }
    return guarded([&]() {
        const int64_t r = static_cast<Batch*>(h)->at(inst).hist_read(out, max_rows);
        if (rows) *rows = r;
    });
}

void* psum_batch_instance(void* h, int inst)
{
    if (!h) return nullptr;
    Batch* b = static_cast<Batch*>(h);
    if (inst < 0 || inst >= b->size()) {
        set_error("psum_batch_instance: out of range");
        return nullptr;
    }
    return &b->at(inst);
}

void psum_batch_destroy(void* h) { delete static_cast<Batch*>(h); }

}  // extern "C"
