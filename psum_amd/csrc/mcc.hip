// Monte Carlo collisions between the push and the next deposit: a sequence
// of collision processes (one incident species against a background target
// each) applied to every valid particle, with per-particle LCG random
// streams.  Follows particle_collision/foundation.hpp:17-107 (random_choose,
// collision_processor::execute, sequence_executor), mcc_model.hpp:184-218
// (execute_mcc_model), random/rander.hpp:40-139 and rand_functions.hpp:15-28,
// and the collision actions of application/benchmark/CCP/custom_mccm.hpp:
// 29-104 as channel kinds.
//
// Work is elementwise over particles and HBM-bound: one thread per particle,
// 256-thread blocks.  A particle that does not collide only has its seed
// rewritten; velocities are stored only when an action changed them.
// Ionization products are placed deterministically in (particle, process)
// order: a counting pass stores each particle's rank among its block's
// products (uint16) and the block totals, a scan of the totals gives block
// offsets, and the apply pass writes products at offset + rank.  Without an
// ionization channel the apply pass runs alone.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "common.hpp"

namespace psum_hip {
namespace {

constexpr int kMccThreads = 256;
constexpr double kPi = 3.1415926535897932;

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline hipStream_t S(void* s) { return static_cast<hipStream_t>(s); }

// the particle's random stream (rander_core<1664525, 1013904223> and
// rander::operator(), random/rander.hpp:40-74)
struct Lcg {
    uint32_t x;
    __device__ __forceinline__ double operator()()
    {
        x = x * 1664525u + 1013904223u;
        return (double)x / 4294967295.0;
    }
};

// RandV_spherical (rand_functions.hpp:15-19)
__device__ __forceinline__ void rand_spherical(Lcg& R, double* o)
{
    const double r1 = 2 * R() - 1;
    const double r2 = 2 * kPi * R();
    const double s = sqrt(1 - r1 * r1);
    o[0] = r1;
    o[1] = s * sin(r2);
    o[2] = s * cos(r2);
}

// RandV_Maxwell (rand_functions.hpp:21-28): per component the sin draw,
// then the log draw, x then y then z
__device__ __forceinline__ void rand_maxwell(Lcg& R, double temperature, double mass, double kb, double* o)
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double a = sin(2 * kPi * R());
        o[c] = a * sqrt(-log(R()));
    }
    const double s = sqrt(2.0 * kb * temperature / mass);
#pragma unroll
    for (int c = 0; c < 3; ++c) o[c] = o[c] * s;
}

__device__ __forceinline__ double sqn(const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

// x = ln(v2 * arg_scale) is shared by consecutive channels of one scale
// (the same value the reference computes inside each cross-section call)
__device__ double cross_section(const psum_mcc_channel& c, double v2, double& scale, double& x)
{
    if (c.nseg == 0) return c.coef[0][4];
    if (c.arg_scale != scale) {
        scale = c.arg_scale;
        x = log(v2 * scale);
    }
    int k = 0;
    while (k < c.nseg - 1 && !(x < c.seg_hi[k])) ++k;
    const double* a = c.coef[k];
    const double y = a[0] * (((x * x) * x) * x) + a[1] * ((x * x) * x) + a[2] * (x * x) + a[3] * x + a[4];
    return c.exp_form ? exp(y) : y;
}

struct Product {
    double ve[3], vi[3];
    uint32_t se, si;
};

// one collision action on v, drawing from its own copy of the stream;
// returns whether an ionization made a product pair
template <bool PRODUCTS>
__device__ bool collide(const psum_mcc_process& P, const psum_mcc_channel& c, Lcg R, double* v, Product* pr)
{
    double r[3];
    switch (c.kind) {
    case PSUM_MCC_ELASTIC: {
        rand_spherical(R, r);
        const double s = sqrt(sqn(v));
        for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        return false;
    }
    case PSUM_MCC_EXCITATION: {
        rand_spherical(R, r);
        const double v2n = sqn(v) - 2 * P.unit_charge * c.threshold_ev / P.incident_mass;
        if (v2n > 0) {
            const double s = sqrt(v2n);
            for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        }
        return false;
    }
    case PSUM_MCC_IONIZATION: {
        const double energy = sqn(v) * P.incident_mass * 0.5;
        const double e2 = (energy - c.threshold_ev * P.unit_charge) / 2;
        if (e2 < 0) return false;
        rand_spherical(R, r);
        const double s = sqrt(2 * e2 / P.incident_mass);
        for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        if constexpr (PRODUCTS) {
            rand_spherical(R, r);
            for (int k = 0; k < 3; ++k) pr->ve[k] = r[k] * s;
            pr->se = (uint32_t)(R() * 2147483647.0);
            rand_maxwell(R, P.target_temperature, P.target_mass, P.boltzmann, pr->vi);
            pr->si = (uint32_t)(R() * 2147483647.0);
        }
        return true;
    }
    case PSUM_MCC_ION_ISOTROPIC: {
        double va[3], mean[3], d[3];
        rand_maxwell(R, P.target_temperature, P.target_mass, P.boltzmann, va);
        for (int k = 0; k < 3; ++k) {
            mean[k] = (v[k] + va[k]) * 0.5;
            d[k] = v[k] - va[k];
        }
        const double rel = sqrt(sqn(d));
        rand_spherical(R, r);
        for (int k = 0; k < 3; ++k) v[k] = mean[k] + rel * 0.5 * r[k];
        return false;
    }
    case PSUM_MCC_ION_BACKWARD:
        rand_maxwell(R, P.target_temperature, P.target_mass, P.boltzmann, v);
        return false;
    default:
        return false;
    }
}

// collision_processor::execute for one process: returns the channel that
// collided or -1; the seed keeps the state after the process's own draws
template <bool PRODUCTS>
__device__ int run_process(const psum_mcc_process& P, double dens, double dt, double* v, uint32_t& seed,
                           bool& made, Product* pr)
{
    Lcg R{seed};
    const double v2 = sqn(v);
    const double vm = sqrt(v2);
    const double r = R();
    double area[PSUM_MCC_MAX_CHAN];
    double sum = 0, scale = 0, x = 0;      // arg_scale 0 is never a valid scale
#pragma unroll
    for (int k = 0; k < PSUM_MCC_MAX_CHAN; ++k)
        if (k < P.nchan) {
            area[k] = cross_section(P.chan[k], v2, scale, x);
            sum += area[k];
        }
    int idx = -1;
    if (sum > 0) {
        const double thr = r * sum;
        double acc = 0;
#pragma unroll
        for (int k = 0; k < PSUM_MCC_MAX_CHAN; ++k)
            if (k < P.nchan && idx < 0) {
                acc += area[k];
                if (thr <= acc) idx = k;
            }
    }
    int hit = -1;
    made = false;
    if (idx != -1 && R() < sum * dt * dens * vm) {
        hit = idx;
        made = collide<PRODUCTS>(P, P.chan[idx], R, v, pr);
    }
    seed = R.x;
    return hit;
}

// generate_interp with CIC weights (field_interp.hpp:78-91); 0 outside the
// single-wrap range
template <int D>
__device__ double density_at(const DevGrid& g, const double* __restrict__ f, const double* x)
{
    int cw[3];
    double s[3];
    if (!base_cell<D>(g, x, cw, s)) return 0.0;
    double W[3][2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        W[d][0] = 1.0 - s[d];
        W[d][1] = s[d];
    }
    double res = 0;
#pragma unroll
    for (int i = 0; i < (1 << D); ++i) {
        const double c = cic_coef<double, D>(W, i);
        if (fabs(c) > 1e-10) res += f[support_index<D>(g, cw, i)] * c;
    }
    return res;
}

struct MccArgs {
    DevGrid g;
    const psum_mcc_process* procs;
    int nproc;
    const double* dens[PSUM_MCC_MAX_PROC];
    void* pos[3];
    void* vel[3];
    uint32_t* seed;
    int64_t n;
    double dt;
    void* epos[3];
    void* evel[3];
    uint32_t* eseed;
    int64_t en;
    void* ipos[3];
    void* ivel[3];
    uint32_t* iseed;
    int64_t in;
    uint16_t* rank;          // per particle: products of lower threads in its block
    uint32_t* boff;          // per block: product count (count pass), offset (apply pass)
    int32_t* outcome;
};

// MODE 0: apply, no products possible; 1: count products; 2: apply and
// write products at boff[block] + rank[i]
template <typename T, int D, int MODE>
__global__ __launch_bounds__(kMccThreads) void k_mcc(MccArgs a)
{
    const int64_t i = (int64_t)blockIdx.x * kMccThreads + threadIdx.x;
    int made_n = 0;
    if (i < a.n) {
        double x[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) x[d] = (double)static_cast<const T*>(a.pos[d])[i];
        if (!isfinite(x[0])) {
            if (MODE != 1 && a.outcome)
                for (int k = 0; k < a.nproc; ++k) a.outcome[i * a.nproc + k] = -2;
        } else {
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = (double)static_cast<const T*>(a.vel[c])[i];
            uint32_t seed = a.seed[i];
            bool changed = false;
            const int64_t base = MODE == 2 ? (int64_t)a.boff[blockIdx.x] + a.rank[i] : 0;
            for (int k = 0; k < a.nproc; ++k) {
                const psum_mcc_process& P = a.procs[k];
                const double dens = a.dens[k] ? density_at<D>(a.g, a.dens[k], x) : P.density;
                bool made = false;
                Product pr;
                const int hit = run_process<MODE == 2>(P, dens, a.dt, v, seed, made, &pr);
                changed |= hit >= 0;
                if (MODE != 1 && a.outcome) a.outcome[i * a.nproc + k] = hit;
                if (MODE == 2 && made) {
                    const int64_t je = a.en + base + made_n, ji = a.in + base + made_n;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const T xd = static_cast<const T*>(a.pos[d])[i];
                        static_cast<T*>(a.epos[d])[je] = xd;
                        static_cast<T*>(a.ipos[d])[ji] = xd;
                    }
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        static_cast<T*>(a.evel[c])[je] = (T)pr.ve[c];
                        static_cast<T*>(a.ivel[c])[ji] = (T)pr.vi[c];
                    }
                    a.eseed[je] = pr.se;
                    a.iseed[ji] = pr.si;
                }
                made_n += made ? 1 : 0;
            }
            if (MODE != 1) {
                if (changed)
#pragma unroll
                    for (int c = 0; c < 3; ++c) static_cast<T*>(a.vel[c])[i] = (T)v[c];
                a.seed[i] = seed;
            }
        }
    }
    if (MODE == 1) {
        // block-exclusive rank of each particle's products and the block total
        __shared__ int wsum[kMccThreads / 64];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        int incl = made_n;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int t = __shfl_up(incl, off, 64);
            if (lane >= off) incl += t;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kMccThreads / 64; ++w) {
            before += w < wv ? wsum[w] : 0;
            total += wsum[w];
        }
        if (i < a.n) a.rank[i] = (uint16_t)(before + incl - made_n);
        if (threadIdx.x == 0) a.boff[blockIdx.x] = (uint32_t)total;
    }
}

bool has_ionization(const psum_mcc_process* procs, int nproc)
{
    for (int k = 0; k < nproc; ++k)
        for (int c = 0; c < procs[k].nchan; ++c)
            if (procs[k].chan[c].kind == PSUM_MCC_IONIZATION) return true;
    return false;
}

template <typename T, int D>
psum_status launch_mcc(MccArgs a, bool products, void* scan_work, int64_t* produced,
                       const psum_mcc_species* e, const psum_mcc_species* ion, hipStream_t st)
{
    const unsigned nb = (unsigned)((a.n + kMccThreads - 1) / kMccThreads);
    if (!products) {
        hipLaunchKernelGGL((k_mcc<T, D, 0>), dim3(nb), dim3(kMccThreads), 0, st, a);
        PSUM_HIP_TRY(hipGetLastError());
        if (produced) *produced = 0;
        return PSUM_OK;
    }
    hipLaunchKernelGGL((k_mcc<T, D, 1>), dim3(nb), dim3(kMccThreads), 0, st, a);
    PSUM_HIP_TRY(hipGetLastError());
    // nb+1 entries: entry nb becomes the total
    psum_status s = exclusive_scan_u32(a.boff, a.boff, (int64_t)nb + 1, scan_work, st);
    if (s != PSUM_OK) return s;
    uint32_t total = 0;
    PSUM_HIP_TRY(hipMemcpyAsync(&total, a.boff + nb, 4, hipMemcpyDeviceToHost, st));
    PSUM_HIP_TRY(hipStreamSynchronize(st));
    if (produced) *produced = total;
    if (e->n + (int64_t)total > e->cap || ion->n + (int64_t)total > ion->cap) {
        set_error("psum_hip_mcc: products exceed the capacity of the electron or ion arrays "
                  "(*produced = the number needed; nothing was modified)");
        return PSUM_ERR_ARG;
    }
    hipLaunchKernelGGL((k_mcc<T, D, 2>), dim3(nb), dim3(kMccThreads), 0, st, a);
    PSUM_HIP_TRY(hipGetLastError());
    return PSUM_OK;
}

}  // namespace
}  // namespace psum_hip

using namespace psum_hip;

extern "C" {

size_t psum_hip_mcc_workspace(int64_t n, int nproc)
{
    if (n < 0) n = 0;
    const int64_t nb = (n + kMccThreads - 1) / kMccThreads;
    return align256(sizeof(psum_mcc_process) * (size_t)(nproc > 0 ? nproc : 1)) +
           align256((size_t)n * sizeof(uint16_t)) + align256((size_t)(nb + 1) * sizeof(uint32_t)) +
           scan_workspace(nb + 1);
}

psum_status psum_hip_mcc(const psum_grid* g, int dtype, const psum_mcc_process* procs, int nproc,
                         const double* const* density, const psum_mcc_species* incident,
                         const psum_mcc_species* electrons, const psum_mcc_species* ions, double dt,
                         int64_t* produced, int32_t* outcome, void* work, size_t work_bytes, void* stream)
{
    if (produced) *produced = 0;
    PSUM_ARG_CHECK(grid_ok(g), "psum_hip_mcc: invalid grid");
    PSUM_ARG_CHECK(procs && nproc >= 1 && nproc <= PSUM_MCC_MAX_PROC, "psum_hip_mcc: nproc must be 1..4");
    PSUM_ARG_CHECK(incident && incident->n >= 0, "psum_hip_mcc: null incident species");
    PSUM_ARG_CHECK(dtype == PSUM_F64 || dtype == PSUM_F32, "psum_hip_mcc: bad dtype");
    for (int k = 0; k < nproc; ++k) {
        const psum_mcc_process& P = procs[k];
        PSUM_ARG_CHECK(P.nchan >= 1 && P.nchan <= PSUM_MCC_MAX_CHAN, "psum_hip_mcc: nchan must be 1..8");
        for (int c = 0; c < P.nchan; ++c) {
            const psum_mcc_channel& ch = P.chan[c];
            PSUM_ARG_CHECK(ch.kind >= PSUM_MCC_ELASTIC && ch.kind <= PSUM_MCC_ION_BACKWARD,
                           "psum_hip_mcc: unknown channel kind");
            PSUM_ARG_CHECK(ch.nseg >= 0 && ch.nseg <= PSUM_MCC_MAX_SEG, "psum_hip_mcc: nseg must be 0..12");
            PSUM_ARG_CHECK(ch.nseg == 0 || ch.arg_scale > 0, "psum_hip_mcc: arg_scale must be positive");
        }
    }
    const int64_t n = incident->n;
    PSUM_ARG_CHECK(work_bytes >= psum_hip_mcc_workspace(n, nproc), "psum_hip_mcc: workspace too small");
    if (n == 0) return PSUM_OK;
    // products are counted in uint32; a launch covers < 2^32 work-items
    PSUM_ARG_CHECK(n * nproc <= (int64_t)0xFFFFFFFF && n <= (int64_t)0xFFFFFF00,
                   "psum_hip_mcc: too many particles for one call");
    const int D = g->dim;
    for (int d = 0; d < D; ++d) PSUM_ARG_CHECK(incident->pos[d], "psum_hip_mcc: null position array");
    for (int c = 0; c < 3; ++c) PSUM_ARG_CHECK(incident->vel[c], "psum_hip_mcc: null velocity array");
    PSUM_ARG_CHECK(incident->seed && work, "psum_hip_mcc: null seed array or workspace");
    const bool products = has_ionization(procs, nproc);
    if (products) {
        PSUM_ARG_CHECK(electrons && ions, "psum_hip_mcc: an ionization channel needs electron and ion arrays");
        for (const psum_mcc_species* s : {electrons, ions}) {
            PSUM_ARG_CHECK(s->n >= 0 && s->cap >= s->n, "psum_hip_mcc: product species count above capacity");
            for (int d = 0; d < D; ++d) PSUM_ARG_CHECK(s->pos[d], "psum_hip_mcc: null product position array");
            for (int c = 0; c < 3; ++c) PSUM_ARG_CHECK(s->vel[c], "psum_hip_mcc: null product velocity array");
            PSUM_ARG_CHECK(s->seed, "psum_hip_mcc: null product seed array");
        }
        PSUM_ARG_CHECK(ions != incident && ions->seed != incident->seed && ions->seed != electrons->seed,
                       "psum_hip_mcc: ion products need arrays of their own");
        if (electrons == incident || electrons->seed == incident->seed)
            PSUM_ARG_CHECK(electrons->n >= n, "psum_hip_mcc: products appended to the incident species "
                                              "must land after its particles");
    }
    hipStream_t st = S(stream);
    const int64_t nb = (n + kMccThreads - 1) / kMccThreads;
    char* w = static_cast<char*>(work);
    psum_mcc_process* dprocs = reinterpret_cast<psum_mcc_process*>(w);
    w += align256(sizeof(psum_mcc_process) * (size_t)nproc);
    uint16_t* rank = reinterpret_cast<uint16_t*>(w);
    w += align256((size_t)n * sizeof(uint16_t));
    uint32_t* boff = reinterpret_cast<uint32_t*>(w);
    w += align256((size_t)(nb + 1) * sizeof(uint32_t));
    void* scan_work = w;
    PSUM_HIP_TRY(hipMemcpyAsync(dprocs, procs, sizeof(psum_mcc_process) * (size_t)nproc, hipMemcpyHostToDevice, st));

    MccArgs a{};
    a.g = to_dev(*g);
    a.procs = dprocs;
    a.nproc = nproc;
    for (int k = 0; k < nproc; ++k) a.dens[k] = density ? density[k] : nullptr;
    for (int d = 0; d < 3; ++d) {
        a.pos[d] = incident->pos[d];
        a.vel[d] = incident->vel[d];
    }
    a.seed = incident->seed;
    a.n = n;
    a.dt = dt;
    if (products) {
        for (int d = 0; d < 3; ++d) {
            a.epos[d] = electrons->pos[d];
            a.evel[d] = electrons->vel[d];
            a.ipos[d] = ions->pos[d];
            a.ivel[d] = ions->vel[d];
        }
        a.eseed = electrons->seed;
        a.en = electrons->n;
        a.iseed = ions->seed;
        a.in = ions->n;
    }
    a.rank = rank;
    a.boff = boff;
    a.outcome = outcome;
    const bool f64 = dtype == PSUM_F64;
#define MCC_LAUNCH(T, DD) launch_mcc<T, DD>(a, products, scan_work, produced, electrons, ions, st)
    if (D == 1) return f64 ? MCC_LAUNCH(double, 1) : MCC_LAUNCH(float, 1);
    if (D == 2) return f64 ? MCC_LAUNCH(double, 2) : MCC_LAUNCH(float, 2);
    return f64 ? MCC_LAUNCH(double, 3) : MCC_LAUNCH(float, 3);
#undef MCC_LAUNCH
}

}  // extern "C"
