/*
 * Precision-generic body of the oracle's particle functions (test
 * infrastructure only).  Included twice from particles.c with
 *   ORC_T   = double | float   (particle state and weight arithmetic)
 *   ORC_SUF = f64    | f32
 * Index and fractional-offset arithmetic is always FP64, as in the
 * reference (simple_grid.hpp:89-95 computes cN in double).
 */

#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_FN(name) ORC_CAT(name, ORC_SUF)

/* CIC support of x in precision ORC_T.  coef[i] and, when gcoef != NULL,
 * gcoef[d*8+i] follow interp_tensor / interp_tensor_diff
 * (simple_interpolation.hpp:36-54,61-90): support point i takes digit j of
 * i (base 2, least significant first) as the offset of dimension D-1-j and
 * the coefficient is the product, starting from 1, of the per-dimension
 * factors for dimensions D-1, D-2, ..., 0. */
static int ORC_FN(orc_support)(const orc_grid* g, const ORC_T* x, int64_t* idx,
                               ORC_T* coef, ORC_T* gcoef)
{
    const int D = g->dim;
    int c[3];
    double s[3];
    if (!orc_base(g, x[0], D > 1 ? x[1] : 0, D > 2 ? x[2] : 0, c, s)) return 0;
    ORC_T W[3][2], G[3][2];
    for (int d = 0; d < D; ++d) {
        ORC_T sd = (ORC_T)s[d];
        W[d][0] = (ORC_T)1 - sd;        /* linear_interp_1D: {1 - x, x} */
        W[d][1] = sd;
        ORC_T dr = (ORC_T)g->delta_r[d]; /* diff_linear_interp_1D * c_dx */
        G[d][0] = (ORC_T)-1 * dr;
        G[d][1] = (ORC_T)1 * dr;
    }
    const int S = 1 << D;
    for (int i = 0; i < S; ++i) {
        int node[3] = {c[0], c[1], c[2]};
        int t = i;
        for (int j = 0; j < D; ++j) {
            const int dd = D - 1 - j;
            node[dd] += t % 2;
            const int w = g->cells[dd] + 1;
            if (node[dd] < 0) node[dd] += w;
            else if (node[dd] >= w) node[dd] -= w;
            t /= 2;
        }
        int64_t lin = node[0];
        for (int d = 1; d < D; ++d) lin = lin * (g->cells[d] + 1) + node[d];
        idx[i] = lin;
        ORC_T cf = (ORC_T)1;
        t = i;
        for (int j = 0; j < D; ++j) {
            cf = cf * W[D - 1 - j][t % 2];
            t /= 2;
        }
        coef[i] = cf;
        if (gcoef) {
            for (int d = 0; d < D; ++d) {
                ORC_T gf = (ORC_T)1;
                t = i;
                for (int j = 0; j < D; ++j) {
                    const int dd = D - 1 - j;
                    gf = gf * (dd == d ? G[dd][t % 2] : W[dd][t % 2]);
                    t /= 2;
                }
                gcoef[d * 8 + i] = gf;
            }
        }
    }
    return 1;
}

void ORC_FN(orc_cell_keys)(const orc_grid* g, const ORC_T* const* pos, int64_t n, uint32_t* keys)
{
    const int D = g->dim;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        ORC_T x[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) x[d] = pos[d][p];
        keys[p] = orc_key(g, (double)x[0], (double)x[1], (double)x[2]);
    }
}

int64_t ORC_FN(orc_deposit_fx)(const orc_grid* g, const ORC_T* const* pos, int64_t n,
                               double w, double scale, int64_t* acc)
{
    /* Integer accumulation is exact and order independent, so per-thread
     * private grids summed afterwards give the same bits as a sequential
     * loop for any thread count. */
    const int D = g->dim;
    const int S = 1 << D;
    const int64_t nn = orc_num_nodes(g);
    int64_t skipped = 0;
#pragma omp parallel reduction(+ : skipped)
    {
        int64_t* priv = (int64_t*)calloc((size_t)nn, sizeof(int64_t));
#pragma omp for schedule(static)
        for (int64_t p = 0; p < n; ++p) {
            ORC_T x[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) x[d] = pos[d][p];
            if (!isfinite((double)x[0])) continue;          /* validator: x NaN = invalid */
            int64_t idx[8];
            ORC_T coef[8];
            if (!ORC_FN(orc_support)(g, x, idx, coef, NULL)) { ++skipped; continue; }
            for (int i = 0; i < S; ++i) {
                const double cf = (double)coef[i];
                if (fabs(cf) > 1e-10) {                        /* field_interp.hpp:70 */
                    const double v = w * cf;                   /* fetch_add(w * c) */
                    priv[idx[i]] += (int64_t)llrint(v * scale);
                }
            }
        }
#pragma omp critical
        for (int64_t i = 0; i < nn; ++i) acc[i] += priv[i];
        free(priv);
    }
    return skipped;
}

int64_t ORC_FN(orc_gather_push)(const orc_grid* g, ORC_T** pos, ORC_T** vel, int64_t n,
                                const ORC_T* phi, ORC_T qm, ORC_T dt, double mass, double* diag)
{
    const int D = g->dim;
    const int S = 1 << D;
    int64_t absorbed = 0;
    double ke = 0, px = 0, py = 0, pz = 0;
#pragma omp parallel for schedule(static) reduction(+ : absorbed, ke, px, py, pz)
    for (int64_t p = 0; p < n; ++p) {
        ORC_T x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) { x[d] = pos[d][p]; v[d] = vel[d][p]; }
        if (!isfinite((double)x[0])) continue;
        /* diagnostics of the velocity entering the step */
        double v2 = 0;
        for (int d = 0; d < D; ++d) v2 += (double)v[d] * (double)v[d];
        ke += 0.5 * mass * v2;
        px += mass * (double)v[0];
        if (D > 1) py += mass * (double)v[1];
        if (D > 2) pz += mass * (double)v[2];

        int64_t idx[8];
        ORC_T coef[8], gc[24];
        ORC_T grad[3] = {0, 0, 0};
        if (ORC_FN(orc_support)(g, x, idx, coef, gc)) {
            for (int d = 0; d < D; ++d) {                 /* field_interp.hpp:96-111 */
                ORC_T r = 0;
                for (int i = 0; i < S; ++i)
                    if (fabs((double)gc[d * 8 + i]) > 1e-10) r = r + phi[idx[i]] * gc[d * 8 + i];
                grad[d] = r;
            }
        }
        for (int d = 0; d < D; ++d) {
            const ORC_T e = -grad[d];
            v[d] = v[d] + (e * qm) * dt;                  /* TSI/main.cpp:67 */
            x[d] = x[d] + v[d] * dt;                      /* TSI/main.cpp:68 */
        }
        int dead = 0;
        for (int d = 0; d < D; ++d) {
            if (g->periodic[d]) {
                const ORC_T lo = (ORC_T)g->plo[d];
                const ORC_T hi = (ORC_T)(g->plo[d] + g->plen[d]);
                const ORC_T L = (ORC_T)g->plen[d];
                for (int k = 0; k < 2; ++k) {             /* TSI/main.cpp:71-72, bounded */
                    if (x[d] < lo) x[d] = x[d] + L;
                    else if (x[d] >= hi) x[d] = x[d] - L;
                }
            }
        }
        for (int d = 0; d < D; ++d) {
            if (!g->periodic[d]) {                        /* simple_grid.hpp:168-174 */
                const double xd = (double)x[d];
                if (!(xd >= g->lb[d] && xd < g->ub[d])) dead = 1;
            }
        }
        for (int d = 0; d < D; ++d) { pos[d][p] = x[d]; vel[d][p] = v[d]; }
        if (dead) { pos[0][p] = (ORC_T)NAN; ++absorbed; }  /* validator.hpp:20-22 */
    }
    if (diag) { diag[0] = ke; diag[1] = px; diag[2] = py; diag[3] = pz; }
    return absorbed;
}

int64_t ORC_FN(orc_compact)(int dim, ORC_T* const* pos, ORC_T* const* vel, int64_t n,
                            ORC_T** out_pos, ORC_T** out_vel)
{
    int64_t m = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (!isfinite((double)pos[0][p])) continue;
        for (int d = 0; d < dim; ++d) { out_pos[d][m] = pos[d][p]; out_vel[d][m] = vel[d][p]; }
        ++m;
    }
    return m;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT2
