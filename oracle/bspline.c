/*
 * Higher-order (B-spline) particle/field coupling of the reference:
 * quadratic, cubic and quartic B-spline interpolation, its gradient, and
 * the deposit (add_back) with those weights.  TEST INFRASTRUCTURE ONLY (see
 * psum_oracle.h).
 *
 *   1D weights          simple_interpolation.hpp:105-184
 *   base node, support  simple_grid.hpp:316-385 (gen_interp / gen_interp_diff:
 *                       odd support -> nearest node nN = round, shift -S/2;
 *                       even support -> corner node cN = floor, shift -(S/2-1);
 *                       one periodic wrap per dimension)
 *   tensor product      simple_interpolation.hpp:36-90 (factors taken from
 *                       dimension D-1 down to 0; derivative factors scaled by
 *                       delta_r)
 *   interp / add_back   field_interp.hpp:63-113 (|coef| > 1e-10)
 *   order selection     higher_order_interp.hpp:79-248, TSI/main.cpp:29,54,64
 *
 * The weights are evaluated in FP64 for both particle precisions, as in the
 * reference.  (1.5-a)^4, (2.5-a)^4 and their cubes, which the reference takes
 * from std::pow, are evaluated as repeated products here and in the HIP
 * kernels alike, so that both sides round identically.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

/* planted mistakes for the pin tests: 1 = odd orders use the corner node
 * (floor) instead of the nearest node, 2 = gradient factors miss delta_r,
 * 3 = the quartic's middle piece drops its (1.5-a)^4/24 term */
static int g_bs_plant = 0;
void orc_bspline_plant(int which) { g_bs_plant = which; }

static double p4(double r) { return ((r * r) * r) * r; }
static double p3(double r) { return (r * r) * r; }

static double b3(double t)
{
    const double a = fabs(t);
    if (a <= 1.0) return 2.0 / 3.0 - a * a + a * a * a / 2.0;
    if (a <= 2.0) { const double r = 2.0 - a; return r * r * r / 6.0; }
    return 0.0;
}
static double d3(double t)
{
    const double a = fabs(t);
    if (a <= 1.0) return -2.0 * t + 1.5 * t * a;
    if (a <= 2.0) { const double r = 2.0 - a; return -r * r / 2.0 * (t < 0 ? -1.0 : 1.0); }
    return 0.0;
}
static double b4(double t)
{
    const double a = fabs(t);
    if (a <= 0.5) return 115.0 / 192.0 - a * a * (5.0 / 8.0) + a * a * a * a / 4.0;
    if (a <= 1.5) {
        const double v = a - 0.5;
        const double tail = g_bs_plant == 3 ? 0.0 : p4(1.5 - a) / 24.0;
        return 0.5 - (2.0 * v / 3.0 - v * v * v / 3.0 + v * v * v * v / 8.0) - tail;
    }
    if (a <= 2.5) return p4(2.5 - a) / 24.0;
    return 0.0;
}
static double d4(double t)
{
    const double a = fabs(t);
    const double sign = t < 0 ? -1.0 : 1.0;
    if (a <= 0.5) return sign * (-a * 1.25 + a * a * a);
    if (a <= 1.5) {
        const double v = a - 0.5;
        return sign * (-(2.0 / 3.0 - v * v + v * v * v / 2.0) + p3(1.5 - a) / 6.0);
    }
    if (a <= 2.5) return sign * (-p3(2.5 - a) / 6.0);
    return 0.0;
}

/* 1D weights and derivatives at normalised offset s, support order+1 */
void orc_bspline_1d(int order, double s, double* c, double* dc)
{
    if (order == 1) {
        c[0] = 1 - s; c[1] = s;
        if (dc) { dc[0] = -1; dc[1] = 1; }
    } else if (order == 2) {
        c[0] = (0.5 - s) * (0.5 - s) * 0.5;
        c[1] = 0.75 - s * s;
        c[2] = (0.5 + s) * (0.5 + s) * 0.5;
        if (dc) { dc[0] = s - 0.5; dc[1] = -2.0 * s; dc[2] = s + 0.5; }
    } else if (order == 3) {
        c[0] = b3(s + 1.0); c[1] = b3(s); c[2] = b3(s - 1.0); c[3] = b3(s - 2.0);
        if (dc) { dc[0] = d3(s + 1.0); dc[1] = d3(s); dc[2] = d3(s - 1.0); dc[3] = d3(s - 2.0); }
    } else {
        c[0] = b4(s + 2.0); c[1] = b4(s + 1.0); c[2] = b4(s); c[3] = b4(s - 1.0); c[4] = b4(s - 2.0);
        if (dc) {
            dc[0] = d4(s + 2.0); dc[1] = d4(s + 1.0); dc[2] = d4(s); dc[3] = d4(s - 1.0); dc[4] = d4(s - 2.0);
        }
    }
}

/* support of x: idx[S^D], coef[S^D], gcoef[d * 125 + i] (may be NULL).
 * Returns S^D, or 0 when the support leaves the single-wrap range or x is
 * not finite. */
int orc_bspline_weights(const orc_grid* g, int order, const double* x, int64_t* idx, double* coef,
                        double* gcoef)
{
    const int D = g->dim;
    const int S = order + 1;
    const int nearest = (S % 2) == 1 && g_bs_plant != 1;
    const int shift = (S % 2) == 1 ? -(S / 2) : -(S / 2 - 1);
    int base[3];
    double c1[3][5], d1[3][5];
    for (int d = 0; d < D; ++d) {
        if (!isfinite(x[d])) return 0;
        const double t = (x[d] - g->lb[d]) * g->delta_r[d];
        const double f = nearest ? round(t) : floor(t);
        const int w = g->cells[d] + 1;
        if (!(f + shift >= (double)(-w) && f + shift + (S - 1) <= (double)(2 * w - 1))) return 0;
        base[d] = (int)f;
        const double bp = g->lb[d] + (double)base[d] * g->delta[d];
        const double s = (x[d] - bp) * g->delta_r[d];
        orc_bspline_1d(order, s, c1[d], d1[d]);
        if (g_bs_plant != 2)
            for (int k = 0; k < S; ++k) d1[d][k] *= g->delta_r[d];
    }
    int total = 1;
    for (int d = 0; d < D; ++d) total *= S;
    for (int i = 0; i < total; ++i) {
        int node[3] = {base[0], base[1], base[2]};
        int t = i;
        for (int j = 0; j < D; ++j) {
            const int dd = D - 1 - j;
            node[dd] += t % S + shift;
            const int w = g->cells[dd] + 1;
            if (node[dd] < 0) node[dd] += w;
            else if (node[dd] >= w) node[dd] -= w;
            t /= S;
        }
        int64_t lin = node[0];
        for (int d = 1; d < D; ++d) lin = lin * (g->cells[d] + 1) + node[d];
        idx[i] = lin;
        double cf = 1;
        t = i;
        for (int j = 0; j < D; ++j) { cf *= c1[D - 1 - j][t % S]; t /= S; }
        coef[i] = cf;
        if (gcoef)
            for (int dg = 0; dg < D; ++dg) {
                double gf = 1;
                t = i;
                for (int j = 0; j < D; ++j) {
                    const int dd = D - 1 - j;
                    gf *= dd == dg ? d1[dd][t % S] : c1[dd][t % S];
                    t /= S;
                }
                gcoef[dg * 125 + i] = gf;
            }
    }
    return total;
}

#define ORC_T double
#define ORC_SUF f64
#include "bspline_impl.h"
#undef ORC_T
#undef ORC_SUF
#define ORC_T float
#define ORC_SUF f32
#include "bspline_impl.h"
#undef ORC_T
#undef ORC_SUF

int64_t orc_deposit_bs_ref(const orc_grid* g, int order, const double* const* pos, int64_t n, double w,
                           double* rho)
{
    const int D = g->dim;
    int64_t skipped = 0;
    for (int64_t p = 0; p < n; ++p) {
        double x[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) x[d] = pos[d][p];
        if (!isfinite(x[0])) continue;
        int64_t idx[125];
        double coef[125];
        const int S = orc_bspline_weights(g, order, x, idx, coef, NULL);
        if (!S) { ++skipped; continue; }
        for (int i = 0; i < S; ++i)
            if (fabs(coef[i]) > 1e-10) rho[idx[i]] += w * coef[i];
    }
    return skipped;
}
