/*
 * Precision-generic body of the oracle's B-spline deposit and gather+push
 * (test infrastructure only).  Included twice from bspline.c with
 *   ORC_T = double | float, ORC_SUF = f64 | f32.
 * Weights are FP64 in both (orc_bspline_weights).  The gather accumulates
 * phi * coef in FP64 (the reference's Value type) and rounds the gradient
 * to ORC_T once; the kick and drift are ORC_T operations.
 */
#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_FN(name) ORC_CAT(name, ORC_SUF)

/* acc[node] += llrint((w * coef) * scale) for |coef| > 1e-10 */
int64_t ORC_FN(orc_deposit_bs)(const orc_grid* g, int order, const ORC_T* const* pos, int64_t n, double w,
                               double scale, int64_t* acc)
{
    const int D = g->dim;
    const int64_t nn = orc_num_nodes(g);
    int64_t skipped = 0;
#pragma omp parallel reduction(+ : skipped)
    {
        int64_t* priv = (int64_t*)calloc((size_t)nn, sizeof(int64_t));
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            double x[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) x[d] = (double)pos[d][p];
            if (!isfinite(x[0])) continue;
            int64_t idx[125];
            double coef[125];
            const int S = orc_bspline_weights(g, order, x, idx, coef, NULL);
            if (!S) { ++skipped; continue; }
            for (int i = 0; i < S; ++i)
                if (fabs(coef[i]) > 1e-10) priv[idx[i]] += (int64_t)llrint((w * coef[i]) * scale);
        }
#pragma omp critical
        for (int64_t i = 0; i < nn; ++i) acc[i] += priv[i];
        free(priv);
    }
    return skipped;
}

int64_t ORC_FN(orc_gather_push_bs)(const orc_grid* g, int order, ORC_T** pos, ORC_T** vel, int64_t n,
                                   const ORC_T* phi, ORC_T qm, ORC_T dt, double mass, double* diag)
{
    const int D = g->dim;
    int64_t absorbed = 0;
    double ke = 0, px = 0, py = 0, pz = 0;
#pragma omp parallel for schedule(static) reduction(+ : absorbed, ke, px, py, pz)
    for (int64_t p = 0; p < n; ++p) {
        ORC_T x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) { x[d] = pos[d][p]; v[d] = vel[d][p]; }
        if (!isfinite((double)x[0])) continue;
        double v2 = 0;
        for (int d = 0; d < D; ++d) v2 += (double)v[d] * (double)v[d];
        ke += 0.5 * mass * v2;
        px += mass * (double)v[0];
        if (D > 1) py += mass * (double)v[1];
        if (D > 2) pz += mass * (double)v[2];

        double xd[3] = {(double)x[0], (double)x[1], (double)x[2]};
        int64_t idx[125];
        double coef[125], gc[375];
        ORC_T grad[3] = {0, 0, 0};
        const int S = orc_bspline_weights(g, order, xd, idx, coef, gc);
        if (S)
            for (int d = 0; d < D; ++d) {
                double r = 0;
                for (int i = 0; i < S; ++i)
                    if (fabs(gc[d * 125 + i]) > 1e-10) r += (double)phi[idx[i]] * gc[d * 125 + i];
                grad[d] = (ORC_T)r;
            }
        for (int d = 0; d < D; ++d) {
            const ORC_T e = -grad[d];
            v[d] = v[d] + (e * qm) * dt;
            x[d] = x[d] + v[d] * dt;
        }
        int dead = 0;
        for (int d = 0; d < D; ++d)
            if (g->periodic[d]) {
                const ORC_T lo = (ORC_T)g->plo[d];
                const ORC_T hi = (ORC_T)(g->plo[d] + g->plen[d]);
                const ORC_T L = (ORC_T)g->plen[d];
                for (int k = 0; k < 2; ++k) {
                    if (x[d] < lo) x[d] = x[d] + L;
                    else if (x[d] >= hi) x[d] = x[d] - L;
                }
            }
        for (int d = 0; d < D; ++d)
            if (!g->periodic[d]) {
                const double xx = (double)x[d];
                if (!(xx >= g->lb[d] && xx < g->ub[d])) dead = 1;
            }
        for (int d = 0; d < D; ++d) { pos[d][p] = x[d]; vel[d][p] = v[d]; }
        if (dead) { pos[0][p] = (ORC_T)NAN; ++absorbed; }
    }
    if (diag) { diag[0] = ke; diag[1] = px; diag[2] = py; diag[3] = pz; }
    return absorbed;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT2
