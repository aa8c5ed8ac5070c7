/*
 * Monte Carlo collisions of one incident species against background
 * targets: the collision_processor of the reference applied per particle,
 * with the collision actions of its He CCP model.  TEST INFRASTRUCTURE ONLY
 * (see psum_oracle.h).
 *
 *   per-particle random stream  random/rander.hpp:40-74,116-139 (uint32
 *                               LCG x = x*1664525 + 1013904223, R() =
 *                               x / UINT32_MAX, the particle's seed is the
 *                               LCG state)
 *   process                     particle_collision/foundation.hpp:67-98:
 *                               density at x, v2, |v|, channel choice
 *                               (:17-45) with R(), collision test
 *                               R() < ((sum*dt)*dens)*|v| (only when a
 *                               channel was chosen); the seed keeps the
 *                               state after these draws -- the wrapper
 *                               of the process writes it back last, over
 *                               the state the collision action advanced
 *   sequence of processes       foundation.hpp:100-107 (in order, each
 *                               from the seed the previous one left)
 *   density                     field_interp.hpp:78-91 (CIC, |coef| > 1e-10)
 *                               or a constant
 *   cross sections              CCP/HeCCS.h form: piecewise polynomials of
 *                               degree <= 4 in x = ln(v2 * arg_scale),
 *                               optionally exponentiated; x^k as repeated
 *                               products (the reference calls pow)
 *   actions                     CCP/custom_mccm.hpp:29-104: isotropic
 *                               elastic, excitation (|v|^2 loses
 *                               2*e*thr/m), ionization (energy split in
 *                               half, a new electron and a Maxwellian
 *                               ion), ion isotropic scattering in the
 *                               centre-of-mass frame of a Maxwellian
 *                               atom, ion backward (charge exchange)
 *   random vectors              random/rand_functions.hpp:15-28
 *                               (spherical; Maxwell with the draws taken
 *                               left to right: sin draw, then log draw,
 *                               x then y then z)
 *   invalid particles           particle_group.hpp:195-220 (skipped, no
 *                               draws)
 *
 * Products are written in (particle, process) order; the reference appends
 * them through an atomic push_back (mcc_model.hpp:184-187), whose order is
 * not defined.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

/* planted mistakes for the pin tests: 1 = the seed keeps the state the
 * collision action advanced, 2 = ionization gives each electron the whole
 * remaining energy, 3 = Maxwell velocities miss the sqrt(-log) factor */
static int g_mcc_plant = 0;
void orc_mcc_plant(int which) { g_mcc_plant = which; }

static const double kPi = 3.1415926535897932;

double orc_lcg_next(uint32_t* x)
{
    *x = *x * 1664525u + 1013904223u;
    return (double)*x / 4294967295.0;
}

void orc_rand_spherical(uint32_t* x, double* o)
{
    const double r1 = 2 * orc_lcg_next(x) - 1;
    const double r2 = 2 * kPi * orc_lcg_next(x);
    o[0] = r1;
    o[1] = sqrt(1 - r1 * r1) * sin(r2);
    o[2] = sqrt(1 - r1 * r1) * cos(r2);
}

void orc_rand_maxwell(uint32_t* x, double temperature, double mass, double boltzmann, double* o)
{
    for (int c = 0; c < 3; ++c) {
        const double a = sin(2 * kPi * orc_lcg_next(x));
        const double l = orc_lcg_next(x);
        o[c] = g_mcc_plant == 3 ? a : a * sqrt(-log(l));
    }
    const double s = sqrt(2.0 * boltzmann * temperature / mass);
    for (int c = 0; c < 3; ++c) o[c] = o[c] * s;
}

double orc_mcc_sigma(const orc_mcc_channel* c, double v2)
{
    if (c->nseg == 0) return c->coef[0][4];
    const double x = log(v2 * c->arg_scale);
    int k = 0;
    while (k < c->nseg - 1 && !(x < c->seg_hi[k])) ++k;
    const double* a = c->coef[k];
    const double y = a[0] * (((x * x) * x) * x) + a[1] * ((x * x) * x) + a[2] * (x * x) + a[3] * x + a[4];
    return c->exp_form ? exp(y) : y;
}

static double sqn(const double* v) { return v[0] * v[0] + v[1] * v[1] + v[2] * v[2]; }

/* one collision action on v with its own copy of the stream; returns 1
 * when an ionization made a product pair (pe/pi velocities, se/si seeds) */
static int collide(const orc_mcc_process* P, const orc_mcc_channel* c, uint32_t* x, double* v,
                   double* pe, uint32_t* se, double* pi, uint32_t* si)
{
    double r[3];
    switch (c->kind) {
    case ORC_MCC_ELASTIC: {
        orc_rand_spherical(x, r);
        const double s = sqrt(sqn(v));
        for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        return 0;
    }
    case ORC_MCC_EXCITATION: {
        orc_rand_spherical(x, r);
        const double v2n = sqn(v) - 2 * P->unit_charge * c->threshold_ev / P->incident_mass;
        if (v2n > 0) {
            const double s = sqrt(v2n);
            for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        }
        return 0;
    }
    case ORC_MCC_IONIZATION: {
        const double energy = sqn(v) * P->incident_mass * 0.5;
        const double rest = energy - c->threshold_ev * P->unit_charge;
        const double e2 = g_mcc_plant == 2 ? rest : rest / 2;
        if (e2 < 0) return 0;
        orc_rand_spherical(x, r);
        const double s = sqrt(2 * e2 / P->incident_mass);
        for (int k = 0; k < 3; ++k) v[k] = r[k] * s;
        orc_rand_spherical(x, r);
        for (int k = 0; k < 3; ++k) pe[k] = r[k] * s;
        *se = (uint32_t)(orc_lcg_next(x) * 2147483647.0);
        orc_rand_maxwell(x, P->target_temperature, P->target_mass, P->boltzmann, pi);
        *si = (uint32_t)(orc_lcg_next(x) * 2147483647.0);
        return 1;
    }
    case ORC_MCC_ION_ISOTROPIC: {
        double va[3], mean[3], d[3];
        orc_rand_maxwell(x, P->target_temperature, P->target_mass, P->boltzmann, va);
        for (int k = 0; k < 3; ++k) {
            mean[k] = (v[k] + va[k]) * 0.5;
            d[k] = v[k] - va[k];
        }
        const double rel = sqrt(sqn(d));
        orc_rand_spherical(x, r);
        for (int k = 0; k < 3; ++k) v[k] = mean[k] + rel * 0.5 * r[k];
        return 0;
    }
    case ORC_MCC_ION_BACKWARD:
        orc_rand_maxwell(x, P->target_temperature, P->target_mass, P->boltzmann, v);
        return 0;
    }
    return 0;
}

/* one process on one particle; returns the channel that collided or -1 */
static int run_process(const orc_mcc_process* P, double dens, double dt, double* v, uint32_t* seed,
                       int* made, double* pe, uint32_t* se, double* pi, uint32_t* si)
{
    uint32_t x = *seed;
    const double v2 = sqn(v);
    const double vm = sqrt(v2);
    const double r = orc_lcg_next(&x);
    double area[ORC_MCC_MAX_CHAN];
    double sum = 0;
    for (int k = 0; k < P->nchan; ++k) {
        area[k] = orc_mcc_sigma(&P->chan[k], v2);
        sum += area[k];
    }
    int idx = -1;
    if (sum > 0) {
        const double thr = r * sum;
        double acc = 0;
        for (int k = 0; k < P->nchan; ++k) {
            acc += area[k];
            if (thr <= acc) { idx = k; break; }
        }
    }
    int hit = -1;
    *made = 0;
    if (idx != -1 && orc_lcg_next(&x) < sum * dt * dens * vm) {
        hit = idx;
        uint32_t y = x;
        *made = collide(P, &P->chan[idx], &y, v, pe, se, pi, si);
        if (g_mcc_plant == 1) x = y;
    }
    *seed = x;
    return hit;
}

static double density_at(const orc_grid* g, const orc_mcc_process* P, const double* field, const double* x)
{
    if (!field) return P->density;
    int64_t idx[8];
    double coef[8];
    if (!orc_cic_weights(g, x, idx, coef)) return 0.0;
    double res = 0;
    for (int i = 0; i < (1 << g->dim); ++i)
        if (fabs(coef[i]) > 1e-10) res += field[idx[i]] * coef[i];
    return res;
}

#define ORC_T double
#define ORC_SUF f64
#include "mcc_impl.h"
#undef ORC_T
#undef ORC_SUF
#define ORC_T float
#define ORC_SUF f32
#include "mcc_impl.h"
#undef ORC_T
#undef ORC_SUF
