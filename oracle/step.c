/*
 * CPU restatement of one electrostatic PIC step as the reference's user
 * loops compose it (TSI/main.cpp:121-146, big_bang/stage_6.cpp:125-147):
 * deposit -> Poisson solve through the MG backend -> gather + push.
 * TEST INFRASTRUCTURE ONLY: the history parity tests and the CPU baseline
 * timing of bench.py use it; the product never calls it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

/* Fixed-point exponent for the deposit: term = llrint(w*coef * 2^S) with
 * S = 40 - e, 2^(e-1) <= |w| < 2^e, so one particle contributes at most 2^40
 * and a node may collect up to 2^22 full particles before int64 overflows. */
/* interpolation order of the loop: 1 = CIC, 2..4 = B-splines (TSI's
 * interp_order, TSI/main.cpp:29,54,64) */
static int g_order = 1;
void orc_pic_set_order(int order) { g_order = order; }

/* planted mistakes for the whole-step pin tests (0 = none): 1 = the deposit
 * takes the charge with the wrong sign, 2 = the background charge density
 * enters with the wrong sign, 3 = the push gathers the potential of the
 * previous step, 4 = the periodic neutralisation (mean subtraction) is
 * dropped. */
static int g_step_plant = 0;
void orc_pic_plant(int which) { g_step_plant = which; }

/* Test probe: when set, every step records the charge-weighted Fourier sums
 * sum_s q_s w_s sum_p (cos(k.x_p), sin(k.x_p)) of the positions the step's
 * solve sees, for nk wave vectors kvec[j*D..j*D+D-1], into
 * out[(step*nk + j)*2 + {0,1}]. */
static int g_probe_nk = 0;
static const double* g_probe_k = NULL;
static double* g_probe_out = NULL;
void orc_pic_set_probe(int nk, const double* kvec, double* out)
{
    g_probe_nk = (kvec && out) ? nk : 0;
    g_probe_k = kvec;
    g_probe_out = out;
}

static void probe_record(int D, int nspecies, double*** pos, const int64_t* n, const double* q,
                         const double* w, int st)
{
    for (int j = 0; j < g_probe_nk; ++j) {
        const double* k = g_probe_k + (size_t)j * D;
        double c = 0.0, s = 0.0;
        for (int sp = 0; sp < nspecies; ++sp) {
            double cs = 0.0, ss = 0.0;
            for (int64_t p = 0; p < n[sp]; ++p) {
                double ph = 0.0;
                for (int d = 0; d < D; ++d) ph += k[d] * pos[sp][d][p];
                if (!isfinite(ph)) continue;
                cs += cos(ph);
                ss += sin(ph);
            }
            c += q[sp] * w[sp] * cs;
            s += q[sp] * w[sp] * ss;
        }
        g_probe_out[((size_t)st * g_probe_nk + j) * 2 + 0] = c;
        g_probe_out[((size_t)st * g_probe_nk + j) * 2 + 1] = s;
    }
}

static int fx_shift(double w)
{
    int e = 0;
    frexp(fabs(w), &e);
    return 40 - e;
}

int64_t orc_pic_run_multi(const orc_grid* g, int nspecies, double*** pos, double*** vel,
                          const int64_t* n, const double* q, const double* m, const double* w,
                          double rho_bg, double eps, double dt, int steps, int depth, int pre,
                          int post, double tol, int max_cycles, int dep_mode,
                          const uint8_t* fixed, const double* fixed_val, double* hist,
                          double* phi_out)
{
    const int D = g->dim;
    const int64_t nn = orc_num_nodes(g);
    int all_periodic = 1;
    for (int d = 0; d < D; ++d) all_periodic &= g->periodic[d] ? 1 : 0;

    uint8_t* fx = (uint8_t*)calloc((size_t)nn, 1);
    if (fixed) memcpy(fx, fixed, (size_t)nn);
    if (all_periodic) fx[0] = 1;                 /* pinned node, TSI/main.cpp:105-113 */
    const int64_t nnz = orc_poisson_coo(g, eps, fx, NULL, NULL, NULL);
    uint64_t* r = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nnz);
    uint64_t* c = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nnz);
    double* v = (double*)malloc(sizeof(double) * (size_t)nnz);
    orc_poisson_coo(g, eps, fx, r, c, v);
    const int I = g->cells[0] + 1, J = D > 1 ? g->cells[1] + 1 : 1, K = D > 2 ? g->cells[2] + 1 : 1;
    void* mg = orc_mg_create(nn, nnz, r, c, v, I, J, K, depth, pre, post, 1);
    free(r); free(c); free(v);
    if (!mg) { free(fx); return -1; }

    double dV = 1.0;
    for (int d = 0; d < D; ++d) dV *= g->delta[d];
    double wdep[8], wmax = 0.0;
    for (int s = 0; s < nspecies; ++s) {
        wdep[s] = (g_step_plant == 1 ? -q[s] : q[s]) * w[s] / dV;
        if (fabs(wdep[s]) > wmax) wmax = fabs(wdep[s]);
    }
    const double scale = ldexp(1.0, fx_shift(wmax > 0 ? wmax : 1.0));

    int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    double* b = (double*)malloc(sizeof(double) * (size_t)nn);
    double* phi = (double*)calloc((size_t)nn, sizeof(double));
    double* phi_prev = g_step_plant == 3 ? (double*)calloc((size_t)nn, sizeof(double)) : NULL;
    const double bg = g_step_plant == 2 ? -rho_bg : rho_bg;
    int64_t total_cycles = 0;
    for (int st = 0; st < steps; ++st) {
        if (g_probe_nk) probe_record(D, nspecies, pos, n, q, w, st);
        if (dep_mode == 0) {
            memset(acc, 0, sizeof(int64_t) * (size_t)nn);
            for (int s = 0; s < nspecies; ++s)
                if (g_order > 1) orc_deposit_bs_f64(g, g_order, (const double* const*)pos[s], n[s], wdep[s], scale, acc);
                else orc_deposit_fx_f64(g, (const double* const*)pos[s], n[s], wdep[s], scale, acc);
            for (int64_t i = 0; i < nn; ++i) b[i] = (double)acc[i] * (1.0 / scale) + bg;
        } else {
            memset(b, 0, sizeof(double) * (size_t)nn);
            for (int s = 0; s < nspecies; ++s)
                orc_deposit_ref_f64(g, (const double* const*)pos[s], n[s], wdep[s], b);
            for (int64_t i = 0; i < nn; ++i) b[i] = b[i] + bg;
        }
        if (all_periodic && g_step_plant != 4) {
            double sum = 0.0;
            for (int64_t i = 0; i < nn; ++i) sum += b[i];
            const double mean = sum / (double)nn;
            for (int64_t i = 0; i < nn; ++i) b[i] = b[i] - mean;
        }
        for (int64_t i = 0; i < nn; ++i)                               /* set_source_replace */
            if (fx[i]) b[i] = (fixed_val && !(all_periodic && i == 0)) ? fixed_val[i] : 0.0;
        double bn = 0.0;
        for (int64_t i = 0; i < nn; ++i) bn += b[i] * b[i];
        bn = sqrt(bn);
        for (int cyc = 0; cyc < max_cycles; ++cyc) {
            orc_mg_vcycles(mg, b, phi, 1);
            ++total_cycles;
            if (orc_mg_residual_norm(mg, b, phi) <= tol * bn) break;
        }
        const double fe = orc_field_energy(g, phi, eps);
        double tot[4] = {0, 0, 0, 0};
        const double* phi_push = phi_prev ? phi_prev : phi;
        for (int s = 0; s < nspecies; ++s) {
            double diag[4];
            if (g_order > 1) orc_gather_push_bs_f64(g, g_order, pos[s], vel[s], n[s], phi_push, q[s] / m[s], dt, m[s] * w[s], diag);
            else orc_gather_push_f64(g, pos[s], vel[s], n[s], phi_push, q[s] / m[s], dt, m[s] * w[s], diag);
            for (int k = 0; k < 4; ++k) tot[k] += diag[k];
        }
        if (phi_prev) memcpy(phi_prev, phi, sizeof(double) * (size_t)nn);
        if (hist) {
            hist[5 * st + 0] = fe;
            for (int k = 0; k < 4; ++k) hist[5 * st + 1 + k] = tot[k];
        }
    }
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)nn);
    orc_mg_free(mg);
    free(acc); free(b); free(phi); free(fx); free(phi_prev);
    return total_cycles;
}

int64_t orc_pic_run_f64(const orc_grid* g, double** pos, double** vel, int64_t n,
                        double q, double m, double w, double rho_bg, double eps, double dt,
                        int steps, int depth, int pre, int post, double tol, int max_cycles,
                        int dep_mode, const uint8_t* fixed, double* hist, double* phi_out)
{
    double** P[1] = {pos};
    double** V[1] = {vel};
    return orc_pic_run_multi(g, 1, P, V, &n, &q, &m, &w, rho_bg, eps, dt, steps, depth, pre, post,
                             tol, max_cycles, dep_mode, fixed, NULL, hist, phi_out);
}
