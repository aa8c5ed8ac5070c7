/*
 * CPU restatement of the reference's particle-side hot path: cell index,
 * CIC weights, deposit, gather + push, compaction, and the stable sort the
 * build defines.  TEST INFRASTRUCTURE ONLY (see psum_oracle.h).
 *
 * Build flags: -ffp-contract=off, no fast-math; all FP operations below
 * are single IEEE operations in the order written.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

/* Base node and fractional offset per dimension, in FP64.
 *   cN_d   = int(floor((x_d - lb_d) * delta_r_d))     simple_grid.hpp:89-95
 *   base_d = lb_d + cN_d * delta_d                    simple_grid.hpp:201-206
 *   s_d    = (x_d - base_d) * delta_r_d               simple_grid.hpp:327-328
 * The support node base+o (o in {0,1}) is wrapped once over cells+1 nodes
 * (simple_grid.hpp:338-343); the reference result is defined only when that
 * single wrap lands inside the node range, i.e. cN_d in [-w, 2w-2] with
 * w = cells+1.  Outside that range (and for non-finite x) return 0. */
int orc_base(const orc_grid* g, double x0, double x1, double x2, int* c, double* s)
{
    const double x[3] = {x0, x1, x2};
    for (int d = 0; d < g->dim; ++d) {
        const double t = (x[d] - g->lb[d]) * g->delta_r[d];
        const double f = floor(t);
        const int w = g->cells[d] + 1;
        if (!(f >= (double)(-w) && f <= (double)(2 * w - 2))) return 0;
        const int ci = (int)f;
        const double base = g->lb[d] + (double)ci * g->delta[d];
        s[d] = (x[d] - base) * g->delta_r[d];
        c[d] = ci;
    }
    return 1;
}

int64_t orc_num_nodes(const orc_grid* g)
{
    int64_t nn = 1;
    for (int d = 0; d < g->dim; ++d) nn *= (int64_t)(g->cells[d] + 1);
    return nn;
}

/* n2i of the wrapped base node (simple_grid.hpp:107-112). */
uint32_t orc_key(const orc_grid* g, double x0, double x1, double x2)
{
    if (!isfinite(x0)) return ORC_KEY_INVALID;
    int c[3];
    double s[3];
    if (!orc_base(g, x0, x1, x2, c, s)) return ORC_KEY_INVALID;
    uint64_t lin = 0;
    for (int d = 0; d < g->dim; ++d) {
        const int w = g->cells[d] + 1;
        int nd = c[d];
        if (nd < 0) nd += w;
        else if (nd >= w) nd -= w;
        lin = lin * (uint64_t)w + (uint64_t)nd;
    }
    return (uint32_t)lin;
}

#define ORC_T double
#define ORC_SUF f64
#include "particles_impl.h"
#undef ORC_T
#undef ORC_SUF

#define ORC_T float
#define ORC_SUF f32
#include "particles_impl.h"
#undef ORC_T
#undef ORC_SUF

void orc_stable_sort(const uint32_t* keys, int64_t n, uint32_t nkeys, uint32_t* perm)
{
    /* counting sort: bucket nkeys holds ORC_KEY_INVALID */
    int64_t* cnt = (int64_t*)calloc((size_t)nkeys + 2, sizeof(int64_t));
    for (int64_t p = 0; p < n; ++p) {
        const uint32_t k = keys[p] == ORC_KEY_INVALID ? nkeys : keys[p];
        cnt[k + 1]++;
    }
    for (uint64_t k = 0; k <= nkeys; ++k) cnt[k + 1] += cnt[k];
    for (int64_t p = 0; p < n; ++p) {
        const uint32_t k = keys[p] == ORC_KEY_INVALID ? nkeys : keys[p];
        perm[cnt[k]++] = (uint32_t)p;
    }
    free(cnt);
}

int orc_cic_weights(const orc_grid* g, const double* x, int64_t* idx, double* coef)
{
    return orc_support_f64(g, x, idx, coef, NULL);
}

int orc_cic_grad_weights(const orc_grid* g, const double* x, int64_t* idx, double* gcoef)
{
    double coef[8];
    return orc_support_f64(g, x, idx, coef, gcoef);
}

int64_t orc_deposit_ref_f64(const orc_grid* g, const double* const* pos, int64_t n,
                            double w, double* rho)
{
    const int D = g->dim;
    const int S = 1 << D;
    int64_t skipped = 0;
    for (int64_t p = 0; p < n; ++p) {
        double x[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) x[d] = pos[d][p];
        if (!isfinite(x[0])) continue;
        int64_t idx[8];
        double coef[8];
        if (!orc_support_f64(g, x, idx, coef, NULL)) { ++skipped; continue; }
        for (int i = 0; i < S; ++i)
            if (fabs(coef[i]) > 1e-10) rho[idx[i]] += w * coef[i];   /* atomic_add.hpp:39 */
    }
    return skipped;
}

double orc_field_energy(const orc_grid* g, const double* phi, double eps)
{
    const int D = g->dim;
    int64_t nn[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) nn[d] = g->cells[d] + 1;
    double vol = 1.0;
    for (int d = 0; d < D; ++d) vol *= g->delta[d];
    double e = 0.0;
    for (int64_t i = 0; i < nn[0]; ++i)
        for (int64_t j = 0; j < nn[1]; ++j)
            for (int64_t k = 0; k < nn[2]; ++k) {
                const int64_t id[3] = {i, j, k};
                const int64_t lin = (i * nn[1] + j) * nn[2] + k;
                double acc = 0.0;
                for (int d = 0; d < D; ++d) {
                    int64_t nb[3] = {i, j, k};
                    nb[d] = id[d] + 1;
                    if (nb[d] == nn[d]) {
                        if (!g->periodic[d]) continue;
                        nb[d] = 0;
                    }
                    const int64_t ln = (nb[0] * nn[1] + nb[1]) * nn[2] + nb[2];
                    const double gd = (phi[ln] - phi[lin]) * g->delta_r[d];
                    acc += gd * gd;
                }
                e += acc;
            }
    return 0.5 * eps * e * vol;
}
