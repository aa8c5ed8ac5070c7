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
static int fx_shift(double w)
{
    int e = 0;
    frexp(fabs(w), &e);
    return 40 - e;
}

int64_t orc_pic_run_f64(const orc_grid* g, double** pos, double** vel, int64_t n,
                        double q, double m, double w, double rho_bg, double eps, double dt,
                        int steps, int depth, int pre, int post, double tol, int max_cycles,
                        int dep_mode, const uint8_t* fixed, double* hist, double* phi_out)
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
    const double wdep = q * w / dV;
    const double scale = ldexp(1.0, fx_shift(wdep));
    const double qm = q / m;

    int64_t* acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    double* b = (double*)malloc(sizeof(double) * (size_t)nn);
    double* phi = (double*)calloc((size_t)nn, sizeof(double));
    int64_t total_cycles = 0;
    for (int s = 0; s < steps; ++s) {
        if (dep_mode == 0) {
            memset(acc, 0, sizeof(int64_t) * (size_t)nn);
            orc_deposit_fx_f64(g, (const double* const*)pos, n, wdep, scale, acc);
            for (int64_t i = 0; i < nn; ++i) b[i] = (double)acc[i] * (1.0 / scale) + rho_bg;
        } else {
            memset(b, 0, sizeof(double) * (size_t)nn);
            orc_deposit_ref_f64(g, (const double* const*)pos, n, wdep, b);
            for (int64_t i = 0; i < nn; ++i) b[i] = b[i] + rho_bg;
        }
        if (all_periodic) {
            double sum = 0.0;
            for (int64_t i = 0; i < nn; ++i) sum += b[i];
            const double mean = sum / (double)nn;
            for (int64_t i = 0; i < nn; ++i) b[i] = b[i] - mean;
        }
        for (int64_t i = 0; i < nn; ++i) if (fx[i]) b[i] = 0.0;      /* set_source_replace */
        double bn = 0.0;
        for (int64_t i = 0; i < nn; ++i) bn += b[i] * b[i];
        bn = sqrt(bn);
        for (int cyc = 0; cyc < max_cycles; ++cyc) {
            orc_mg_vcycles(mg, b, phi, 1);
            ++total_cycles;
            if (orc_mg_residual_norm(mg, b, phi) <= tol * bn) break;
        }
        double diag[4];
        const double fe = orc_field_energy(g, phi, eps);
        orc_gather_push_f64(g, pos, vel, n, phi, qm, dt, m * w, diag);
        if (hist) {
            hist[5 * s + 0] = fe;
            hist[5 * s + 1] = diag[0];
            hist[5 * s + 2] = diag[1];
            hist[5 * s + 3] = diag[2];
            hist[5 * s + 4] = diag[3];
        }
    }
    if (phi_out) memcpy(phi_out, phi, sizeof(double) * (size_t)nn);
    orc_mg_free(mg);
    free(acc); free(b); free(phi); free(fx);
    return total_cycles;
}
