/*
 * CPU restatement of the reference's Poisson assembly and its
 * geometric-transfer Galerkin multigrid, with plain CSR matrices in place of
 * Eigen.  TEST INFRASTRUCTURE ONLY (see psum_oracle.h).
 *
 *   orc_poisson_coo      Poisson_solver_2d.hpp:182-410 / Poisson_solver_3d.hpp:178-431
 *                        (Cartesian, constant eps, fixed nodes, no mixed boundary),
 *                        local_equation.hpp:11-80 merge rules
 *   transfers            multigrid_utilities.hpp:28-104
 *   setup                multigrid_core.hpp:49-150
 *   V-cycle              multigrid_core.hpp:163-272
 *   coarse solve         sycl_inner_solver.hpp:52-99 (dense inverse, row dot products)
 *   mixed solve          sycl_multigrid32f_gpu.cpp:106-160
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

#include "leq.h"

int64_t orc_poisson_coo(const orc_grid* g, double eps, const uint8_t* fixed,
                        uint64_t* rows, uint64_t* cols, double* vals)
{
    const int D = g->dim;
    const int I = g->cells[0], J = D > 1 ? g->cells[1] : 0, K = D > 2 ? g->cells[2] : 0;
    const int64_t nJ = J + 1, nK = K + 1;
    const int64_t n_nodes = (int64_t)(I + 1) * nJ * nK;
    trip* t = (trip*)malloc(sizeof(trip) * (size_t)n_nodes * 7);
    int64_t nt = 0;
    const double dx_r = g->delta_r[0], dy_r = D > 1 ? g->delta_r[1] : 0, dz_r = D > 2 ? g->delta_r[2] : 0;
#define CYC(a, n) ((((a) % (n)) + (n)) % (n))
    for (int i = 0; i <= I; ++i)
        for (int j = 0; j <= J; ++j)
            for (int k = 0; k <= K; ++k) {
                const uint64_t idx = (uint64_t)(((int64_t)i * nJ + j) * nK + k);
                leq fin; fin.n = 0;
                if (fixed[idx]) {
                    leq_insert(&fin, idx, 1.0);
                } else if (D == 2) {
                    const uint64_t iN = (uint64_t)i * nJ + CYC(j + 1, J + 1);
                    const uint64_t iS = (uint64_t)i * nJ + CYC(j - 1, J + 1);
                    const uint64_t iE = (uint64_t)CYC(i + 1, I + 1) * nJ + j;
                    const uint64_t iW = (uint64_t)CYC(i - 1, I + 1) * nJ + j;
                    leq gN = leq_pair(idx, dy_r * -1, iN, dy_r);
                    leq gS = leq_pair(iS, dy_r * -1, idx, dy_r);
                    leq gE = leq_pair(idx, dx_r * -1, iE, dx_r);
                    leq gW = leq_pair(iW, dx_r * -1, idx, dx_r);
                    const double sNS = g->delta[0], sEW = g->delta[1];
                    const double vol = g->delta[0] * g->delta[1];
                    fin = leq_scale(leq_scale(gE, sEW), eps);
                    leq tW = leq_scale(leq_scale(gW, sEW), eps); leq_accum(&fin, &tW, 1);
                    leq tN = leq_scale(leq_scale(gN, sNS), eps); leq_accum(&fin, &tN, 0);
                    leq tS = leq_scale(leq_scale(gS, sNS), eps); leq_accum(&fin, &tS, 1);
                    fin = leq_scale(fin, -1.0 / vol);
                } else {
                    const uint64_t iN = ((uint64_t)i * nJ + CYC(j + 1, J + 1)) * nK + k;
                    const uint64_t iS = ((uint64_t)i * nJ + CYC(j - 1, J + 1)) * nK + k;
                    const uint64_t iE = ((uint64_t)CYC(i + 1, I + 1) * nJ + j) * nK + k;
                    const uint64_t iW = ((uint64_t)CYC(i - 1, I + 1) * nJ + j) * nK + k;
                    const uint64_t iU = ((uint64_t)i * nJ + j) * nK + CYC(k + 1, K + 1);
                    const uint64_t iD = ((uint64_t)i * nJ + j) * nK + CYC(k - 1, K + 1);
                    leq gN = leq_pair(idx, dy_r * -1, iN, dy_r);
                    leq gS = leq_pair(iS, dy_r * -1, idx, dy_r);
                    leq gE = leq_pair(idx, dx_r * -1, iE, dx_r);
                    leq gW = leq_pair(iW, dx_r * -1, idx, dx_r);
                    leq gU = leq_pair(idx, dz_r * -1, iU, dz_r);
                    leq gD = leq_pair(iD, dz_r * -1, idx, dz_r);
                    const double sNS = g->delta[0] * g->delta[2];
                    const double sEW = g->delta[1] * g->delta[2];
                    const double sUD = g->delta[0] * g->delta[1];
                    const double vol = g->delta[0] * g->delta[1] * g->delta[2];
                    fin = leq_scale(leq_scale(gE, sEW), eps);
                    leq tW = leq_scale(leq_scale(gW, sEW), eps); leq_accum(&fin, &tW, 1);
                    leq tN = leq_scale(leq_scale(gN, sNS), eps); leq_accum(&fin, &tN, 0);
                    leq tS = leq_scale(leq_scale(gS, sNS), eps); leq_accum(&fin, &tS, 1);
                    leq tU = leq_scale(leq_scale(gU, sUD), eps); leq_accum(&fin, &tU, 0);
                    leq tD = leq_scale(leq_scale(gD, sUD), eps); leq_accum(&fin, &tD, 1);
                    fin = leq_scale(fin, -1.0 / vol);
                }
                for (int q = 0; q < fin.n; ++q) {
                    t[nt].r = idx; t[nt].c = fin.idx[q]; t[nt].v = fin.val[q]; ++nt;
                }
            }
#undef CYC
    if (rows) {
        qsort(t, (size_t)nt, sizeof(trip), trip_colmajor);
        for (int64_t q = 0; q < nt; ++q) { rows[q] = t[q].r; cols[q] = t[q].c; vals[q] = t[q].v; }
    }
    free(t);
    return nt;
}

/* ------------------------------------------------------------------ */
/* CSR matrices */
typedef struct { int64_t n, m; int64_t* rp; int64_t* ci; double* v; } csr;

static void csr_free(csr* a) { free(a->rp); free(a->ci); free(a->v); memset(a, 0, sizeof(*a)); }

static int cmp_i64(const void* a, const void* b)
{
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return x < y ? -1 : x > y;
}

static int cmp_pair(const void* a, const void* b)
{
    const int64_t* x = (const int64_t*)a; const int64_t* y = (const int64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return x[1] < y[1] ? -1 : x[1] > y[1];
}

/* triplets -> CSR, duplicates summed (Eigen setFromTriplets) */
static csr csr_from_trips(int64_t n, int64_t m, int64_t nnz, const uint64_t* r, const uint64_t* c, const double* v)
{
    csr a; a.n = n; a.m = m;
    a.rp = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    for (int64_t q = 0; q < nnz; ++q) a.rp[r[q] + 1]++;
    for (int64_t i = 0; i < n; ++i) a.rp[i + 1] += a.rp[i];
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n + 1));
    memcpy(pos, a.rp, sizeof(int64_t) * (size_t)(n + 1));
    int64_t* ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    double* vv = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    for (int64_t q = 0; q < nnz; ++q) { const int64_t p = pos[r[q]]++; ci[p] = (int64_t)c[q]; vv[p] = v[q]; }
    free(pos);
    /* sort each row by (column, triplet position); duplicates are summed in
     * triplet order */
    a.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    a.v = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    int64_t out = 0;
    int64_t* order = NULL; int64_t cap = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t b = a.rp[i], e = a.rp[i + 1];
        a.rp[i] = out;
        if (e - b > cap) { cap = e - b; order = (int64_t*)realloc(order, sizeof(int64_t) * (size_t)cap * 2); }
        for (int64_t q = b; q < e; ++q) { order[2 * (q - b)] = ci[q]; order[2 * (q - b) + 1] = q; }
        qsort(order, (size_t)(e - b), 2 * sizeof(int64_t), cmp_pair);
        for (int64_t q = 0; q < e - b;) {
            int64_t q2 = q;
            double s = vv[order[2 * q + 1]];
            while (q2 + 1 < e - b && order[2 * (q2 + 1)] == order[2 * q]) { ++q2; s += vv[order[2 * q2 + 1]]; }
            a.ci[out] = order[2 * q]; a.v[out] = s; ++out;
            q = q2 + 1;
        }
    }
    a.rp[n] = out;
    free(order); free(ci); free(vv);
    return a;
}

static void csr_spmv(const csr* a, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < a->n; ++i) {
        double s = 0.0;
        for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) s += a->v[q] * x[a->ci[q]];
        y[i] = s;
    }
}

static csr csr_transpose(const csr* a)
{
    csr t; t.n = a->m; t.m = a->n;
    const int64_t nnz = a->rp[a->n];
    t.rp = (int64_t*)calloc((size_t)t.n + 1, sizeof(int64_t));
    t.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    t.v = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    for (int64_t q = 0; q < nnz; ++q) t.rp[a->ci[q] + 1]++;
    for (int64_t i = 0; i < t.n; ++i) t.rp[i + 1] += t.rp[i];
    int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * (size_t)(t.n + 1));
    memcpy(pos, t.rp, sizeof(int64_t) * (size_t)(t.n + 1));
    for (int64_t i = 0; i < a->n; ++i)
        for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) {
            const int64_t p = pos[a->ci[q]]++;
            t.ci[p] = i; t.v[p] = a->v[q];
        }
    free(pos);
    return t;
}

/* C = A*B (Gustavson, columns sorted, exact zeros dropped = Eigen pruned()) */
static csr csr_mul(const csr* a, const csr* b)
{
    csr c; c.n = a->n; c.m = b->m;
    c.rp = (int64_t*)calloc((size_t)c.n + 1, sizeof(int64_t));
    int64_t cap = a->rp[a->n] * 4 + 16, nnz = 0;
    c.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)cap);
    c.v = (double*)malloc(sizeof(double) * (size_t)cap);
    double* acc = (double*)calloc((size_t)b->m, sizeof(double));
    int64_t* mark = (int64_t*)malloc(sizeof(int64_t) * (size_t)b->m);
    for (int64_t j = 0; j < b->m; ++j) mark[j] = -1;
    int64_t* cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)b->m);
    for (int64_t i = 0; i < a->n; ++i) {
        int64_t nc = 0;
        for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) {
            const int64_t k = a->ci[q];
            const double av = a->v[q];
            for (int64_t r = b->rp[k]; r < b->rp[k + 1]; ++r) {
                const int64_t j = b->ci[r];
                if (mark[j] != i) { mark[j] = i; acc[j] = 0.0; cols[nc++] = j; }
                acc[j] += av * b->v[r];
            }
        }
        qsort(cols, (size_t)nc, sizeof(int64_t), cmp_i64);
        if (nnz + nc > cap) {
            while (nnz + nc > cap) cap *= 2;
            c.ci = (int64_t*)realloc(c.ci, sizeof(int64_t) * (size_t)cap);
            c.v = (double*)realloc(c.v, sizeof(double) * (size_t)cap);
        }
        for (int64_t q = 0; q < nc; ++q) {
            if (acc[cols[q]] == 0.0) continue;
            c.ci[nnz] = cols[q]; c.v[nnz] = acc[cols[q]]; ++nnz;
        }
        c.rp[i + 1] = nnz;
    }
    free(acc); free(mark); free(cols);
    return c;
}

static csr csr_copy(const csr* a)
{
    csr c; c.n = a->n; c.m = a->m;
    const int64_t nnz = a->rp[a->n];
    c.rp = (int64_t*)malloc(sizeof(int64_t) * (size_t)(a->n + 1));
    c.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz ? nnz : 1));
    c.v = (double*)malloc(sizeof(double) * (size_t)(nnz ? nnz : 1));
    memcpy(c.rp, a->rp, sizeof(int64_t) * (size_t)(a->n + 1));
    memcpy(c.ci, a->ci, sizeof(int64_t) * (size_t)nnz);
    memcpy(c.v, a->v, sizeof(double) * (size_t)nnz);
    return c;
}

static double csr_coeff(const csr* a, int64_t i, int64_t j)
{
    for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) if (a->ci[q] == j) return a->v[q];
    return 0.0;
}

/* ------------------------------------------------------------------ */
/* structured transfers (multigrid_utilities.hpp:28-104) */
typedef struct { int64_t n; double* x; } axis;

static axis axis_coarsen(axis f)
{
    axis c; c.n = 0; c.x = (double*)malloc(sizeof(double) * (size_t)(f.n / 2 + 2));
    for (int64_t i = 0; i < f.n; i += 2) c.x[c.n++] = f.x[i];
    if (c.x[c.n - 1] != f.x[f.n - 1]) c.x[c.n++] = f.x[f.n - 1];
    return c;
}

static void axis_volumes(axis a, double* w)
{
    for (int64_t i = 0; i < a.n; ++i) w[i] = 1.0;
    if (a.n == 1) return;
    for (int64_t i = 0; i < a.n; ++i)
        w[i] = ((i ? a.x[i] - a.x[i - 1] : 0.0) + (i + 1 < a.n ? a.x[i + 1] - a.x[i] : 0.0)) * 0.5;
}

/* fine point x on the coarse axis: 1 or 2 (index, weight) pairs */
static int axis_interp(double x, axis c, int64_t* ix, double* w)
{
    int64_t hi = 0;
    while (hi < c.n && c.x[hi] < x) ++hi;      /* lower_bound */
    if (c.x[hi] == x) { ix[0] = hi; w[0] = 1.0; return 1; }
    const double t = (x - c.x[hi - 1]) / (c.x[hi] - c.x[hi - 1]);
    ix[0] = hi - 1; w[0] = 1.0 - t; ix[1] = hi; w[1] = t;
    return 2;
}

/* ------------------------------------------------------------------ */
typedef struct {
    int L;                      /* number of transfers; levels 0..L */
    int pre, post;
    csr A0;                     /* unscaled fine matrix as given */
    int singular;               /* pinned-periodic matrix solved in consistent singular form */
    int64_t pin;                /* the pinned node */
    csr As;                     /* unscaled singular fine matrix (singular mode) */
    csr* ah;                    /* scaled level matrices, L+1 */
    csr* relax;                 /* L */
    csr* R; csr* P;             /* L */
    double** dinv;              /* L+1 */
    double* cinv;               /* dense coarse inverse, row-major */
    int64_t nc;
    /* FP32 copies for the mixed solve */
    float **ah_v32, **relax_v32, **R_v32, **P_v32, **dinv32, *cinv32;
} orc_mg;

static void dense_inverse(const csr* a, double* inv)
{
    const int64_t n = a->n;
    double* m = (double*)calloc((size_t)(n * n), sizeof(double));
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) m[i * n + a->ci[q]] = a->v[q];
        for (int64_t j = 0; j < n; ++j) inv[i * n + j] = (i == j) ? 1.0 : 0.0;
    }
    /* Gauss-Jordan with partial pivoting */
    for (int64_t c = 0; c < n; ++c) {
        int64_t p = c;
        for (int64_t r = c + 1; r < n; ++r) if (fabs(m[r * n + c]) > fabs(m[p * n + c])) p = r;
        if (p != c)
            for (int64_t j = 0; j < n; ++j) {
                double t = m[c * n + j]; m[c * n + j] = m[p * n + j]; m[p * n + j] = t;
                t = inv[c * n + j]; inv[c * n + j] = inv[p * n + j]; inv[p * n + j] = t;
            }
        const double d = m[c * n + c];
        for (int64_t j = 0; j < n; ++j) { m[c * n + j] /= d; inv[c * n + j] /= d; }
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < n; ++r) {
            if (r == c) continue;
            const double f = m[r * n + c];
            if (f == 0.0) continue;
            for (int64_t j = 0; j < n; ++j) { m[r * n + j] -= f * m[c * n + j]; inv[r * n + j] -= f * inv[c * n + j]; }
        }
    }
    free(m);
}

static float* to_f32(const double* x, int64_t n)
{
    float* y = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1));
    for (int64_t i = 0; i < n; ++i) y[i] = (float)x[i];
    return y;
}

/* Axis roles in periodic-aware mode: an axis is wrapped when the matrix
 * couples its first and last node.  A wrapped axis with an even node count
 * >= 4 coarsens periodically (even nodes kept, odd nodes interpolated
 * between their neighbours with wrap-around, uniform volumes); a wrapped
 * axis that cannot be halved is kept unchanged (identity transfer).  Other
 * axes follow the reference coarsening. */
static void detect_wrap(const csr* a, const int* sz, int* wrapped)
{
    wrapped[0] = wrapped[1] = wrapped[2] = 0;
    for (int64_t r = 0; r < a->n; ++r)
        for (int64_t q = a->rp[r]; q < a->rp[r + 1]; ++q) {
            int64_t x = r, y = a->ci[q];
            int rc[3], cc[3];
            rc[2] = (int)(x % sz[2]); x /= sz[2]; rc[1] = (int)(x % sz[1]); rc[0] = (int)(x / sz[1]);
            cc[2] = (int)(y % sz[2]); y /= sz[2]; cc[1] = (int)(y % sz[1]); cc[0] = (int)(y / sz[1]);
            for (int d = 0; d < 3; ++d)
                if (sz[d] >= 3 && abs(rc[d] - cc[d]) == sz[d] - 1) wrapped[d] = 1;
        }
}

/* A pinned-periodic matrix: every axis of size > 1 wrapped, exactly one
 * identity row (the pinned node p), and every other row summing to zero.
 * Its pinned solution equals the solution of the consistent singular system
 * with row p restored (by symmetry from column p) and b'_p = b_p - sum(b),
 * shifted so that x_p = b_p.  Returns p or -1. */
static int64_t detect_pinned_periodic(const csr* a, const int* sz, const int* wrapped)
{
    for (int d = 0; d < 3; ++d) if (sz[d] > 1 && !wrapped[d]) return -1;
    int64_t pin = -1;
    for (int64_t r = 0; r < a->n; ++r) {
        const int64_t len = a->rp[r + 1] - a->rp[r];
        if (len == 1 && a->ci[a->rp[r]] == r && a->v[a->rp[r]] == 1.0) {
            if (pin >= 0) return -1;
            pin = r;
            continue;
        }
        double s = 0.0, m = 0.0;
        for (int64_t q = a->rp[r]; q < a->rp[r + 1]; ++q) { s += a->v[q]; m += fabs(a->v[q]); }
        if (fabs(s) > 1e-12 * m) return -1;
    }
    return pin;
}

static csr restore_pinned_row(const csr* a, int64_t p)
{
    const int64_t n = a->n, nnz = a->rp[n];
    uint64_t* tr = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nnz + 16));
    uint64_t* tc = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nnz + 16));
    double* tv = (double*)malloc(sizeof(double) * (size_t)(nnz + 16));
    int64_t nt = 0;
    double off = 0.0;
    for (int64_t r = 0; r < n; ++r) {
        if (r == p) continue;
        for (int64_t q = a->rp[r]; q < a->rp[r + 1]; ++q) {
            tr[nt] = (uint64_t)r; tc[nt] = (uint64_t)a->ci[q]; tv[nt] = a->v[q]; ++nt;
            if (a->ci[q] == p) {       /* A symmetric: A[p][r] = A[r][p] */
                tr[nt] = (uint64_t)p; tc[nt] = (uint64_t)r; tv[nt] = a->v[q]; ++nt;
                off += a->v[q];
            }
        }
    }
    tr[nt] = (uint64_t)p; tc[nt] = (uint64_t)p; tv[nt] = -off; ++nt;
    csr s = csr_from_trips(n, n, nt, tr, tc, tv);
    free(tr); free(tc); free(tv);
    return s;
}

void* orc_mg_create(int64_t n, int64_t nnz, const uint64_t* rows, const uint64_t* cols,
                    const double* vals, int I, int J, int K, int depth, int pre, int post,
                    int mode)
{
    if ((int64_t)I * J * K != n || depth < 1) return NULL;
    orc_mg* h = (orc_mg*)calloc(1, sizeof(orc_mg));
    h->pre = pre; h->post = post;
    h->A0 = csr_from_trips(n, n, nnz, rows, cols, vals);
    h->pin = -1;
    const int sz[3] = {I, J, K};
    int wrapped[3] = {0, 0, 0};
    if (mode) {
        detect_wrap(&h->A0, sz, wrapped);
        h->pin = detect_pinned_periodic(&h->A0, sz, wrapped);
        if (h->pin >= 0) {
            h->singular = 1;
            h->As = restore_pinned_row(&h->A0, h->pin);
        }
    }

    /* transfers */
    csr Rs[32], Ps[32];
    int L = 0;
    axis ax[3];
    int per[3];
    for (int d = 0; d < 3; ++d) {
        ax[d].n = sz[d]; ax[d].x = (double*)malloc(sizeof(double) * (size_t)sz[d]);
        for (int i = 0; i < sz[d]; ++i) ax[d].x[i] = i;
        per[d] = mode && wrapped[d];
    }
    for (int level = 0; level < depth && level < 32; ++level) {
        axis co[3];
        double* vol[3];
        int64_t* wi[3]; double* ww[3]; int* wn[3];
        for (int d = 0; d < 3; ++d) {
            const int64_t nf = ax[d].n;
            vol[d] = (double*)malloc(sizeof(double) * (size_t)nf);
            wi[d] = (int64_t*)malloc(sizeof(int64_t) * 2 * (size_t)nf);
            ww[d] = (double*)malloc(sizeof(double) * 2 * (size_t)nf);
            wn[d] = (int*)malloc(sizeof(int) * (size_t)nf);
            if (per[d]) {
                const int halve = (nf % 2 == 0) && nf >= 4;
                const int64_t nc = halve ? nf / 2 : nf;
                co[d].n = nc;
                co[d].x = (double*)malloc(sizeof(double) * (size_t)nc);
                for (int64_t i = 0; i < nc; ++i) co[d].x[i] = halve ? ax[d].x[2 * i] : ax[d].x[i];
                for (int64_t i = 0; i < nf; ++i) {
                    vol[d][i] = 1.0;
                    if (!halve || i % 2 == 0) { wn[d][i] = 1; wi[d][2 * i] = halve ? i / 2 : i; ww[d][2 * i] = 1.0; }
                    else {
                        wn[d][i] = 2;
                        wi[d][2 * i] = (i - 1) / 2; ww[d][2 * i] = 0.5;
                        wi[d][2 * i + 1] = ((i + 1) / 2) % nc; ww[d][2 * i + 1] = 0.5;
                    }
                }
            } else {
                co[d] = axis_coarsen(ax[d]);
                axis_volumes(ax[d], vol[d]);
                for (int64_t i = 0; i < nf; ++i) wn[d][i] = axis_interp(ax[d].x[i], co[d], wi[d] + 2 * i, ww[d] + 2 * i);
            }
        }
        const int64_t ni = ax[0].n, nj = ax[1].n, nk = ax[2].n;
        const int64_t ci = co[0].n, cj = co[1].n, ck = co[2].n;
        int stop = (ni == ci && nj == cj && nk == ck);
        if (!stop) {
            const int64_t nf = ni * nj * nk, ncs = ci * cj * ck;
            uint64_t* tr = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nf * 8);
            uint64_t* tc = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nf * 8);
            double* tv = (double*)malloc(sizeof(double) * (size_t)nf * 8);
            double* fvol = (double*)malloc(sizeof(double) * (size_t)nf);
            int64_t nt = 0;
            for (int64_t i = 0; i < ni; ++i) for (int64_t j = 0; j < nj; ++j) for (int64_t k = 0; k < nk; ++k) {
                const int64_t row = (i * nj + j) * nk + k;
                fvol[row] = vol[0][i] * vol[1][j] * vol[2][k];
                for (int a = 0; a < wn[0][i]; ++a)
                    for (int b = 0; b < wn[1][j]; ++b)
                        for (int c = 0; c < wn[2][k]; ++c) {
                            tr[nt] = (uint64_t)row;
                            tc[nt] = (uint64_t)((wi[0][2 * i + a] * cj + wi[1][2 * j + b]) * ck + wi[2][2 * k + c]);
                            tv[nt] = ww[0][2 * i + a] * ww[1][2 * j + b] * ww[2][2 * k + c];
                            ++nt;
                        }
            }
            csr P = csr_from_trips(nf, ncs, nt, tr, tc, tv);
            csr R = csr_transpose(&P);
            double* cvol = (double*)calloc((size_t)ncs, sizeof(double));
            csr_spmv(&R, fvol, cvol);                                 /* P^T * fine_volume */
            for (int64_t c = 0; c < R.n; ++c)
                for (int64_t q = R.rp[c]; q < R.rp[c + 1]; ++q) R.v[q] *= fvol[R.ci[q]] / cvol[c];
            Ps[L] = P; Rs[L] = R; ++L;
            free(tr); free(tc); free(tv); free(fvol); free(cvol);
        }
        for (int d = 0; d < 3; ++d) { free(vol[d]); free(wi[d]); free(ww[d]); free(wn[d]); free(ax[d].x); ax[d] = co[d]; }
        if (stop) break;
    }
    for (int d = 0; d < 3; ++d) free(ax[d].x);
    if (L == 0) { csr_free(&h->A0); free(h); return NULL; }
    h->L = L;
    h->R = (csr*)malloc(sizeof(csr) * (size_t)L);
    h->P = (csr*)malloc(sizeof(csr) * (size_t)L);
    for (int l = 0; l < L; ++l) { h->R[l] = Rs[l]; h->P[l] = Ps[l]; }

    /* multigrid_core::setup */
    h->ah = (csr*)calloc((size_t)L + 1, sizeof(csr));
    h->dinv = (double**)calloc((size_t)L + 1, sizeof(double*));
    for (int l = 0; l <= L; ++l) {
        csr a;
        if (l == 0) {
            if (h->singular) {
                a = csr_copy(&h->As);
            } else {
                a = csr_from_trips(n, n, nnz, rows, cols, vals);
            }
        } else {
            csr ra = csr_mul(&h->R[l - 1], &h->ah[l - 1]);
            a = csr_mul(&ra, &h->P[l - 1]);
            csr_free(&ra);
        }
        h->dinv[l] = (double*)malloc(sizeof(double) * (size_t)a.n);
        for (int64_t i = 0; i < a.n; ++i) h->dinv[l][i] = 1 / csr_coeff(&a, i, i);
        for (int64_t i = 0; i < a.n; ++i)
            for (int64_t q = a.rp[i]; q < a.rp[i + 1]; ++q) a.v[q] *= h->dinv[l][i];
        h->ah[l] = a;
    }
    h->nc = h->ah[L].n;
    h->cinv = (double*)malloc(sizeof(double) * (size_t)(h->nc * h->nc));
    if (h->singular) {
        /* singular coarsest operator: pin coarse node 0 (identity row, b_0
         * ignored), i.e. invert with row 0 replaced and drop column 0 */
        csr c = h->ah[L];
        csr pinned; pinned.n = c.n; pinned.m = c.m;
        pinned.rp = (int64_t*)calloc((size_t)c.n + 1, sizeof(int64_t));
        pinned.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(c.rp[c.n] + 1));
        pinned.v = (double*)malloc(sizeof(double) * (size_t)(c.rp[c.n] + 1));
        int64_t o = 0;
        for (int64_t i = 0; i < c.n; ++i) {
            if (i == 0) { pinned.ci[o] = 0; pinned.v[o] = 1.0; ++o; }
            else for (int64_t q = c.rp[i]; q < c.rp[i + 1]; ++q) { pinned.ci[o] = c.ci[q]; pinned.v[o] = c.v[q]; ++o; }
            pinned.rp[i + 1] = o;
        }
        dense_inverse(&pinned, h->cinv);
        csr_free(&pinned);
        for (int64_t i = 0; i < h->nc; ++i) h->cinv[i * h->nc] = 0.0;
    } else {
        dense_inverse(&h->ah[L], h->cinv);
    }
    h->relax = (csr*)calloc((size_t)L, sizeof(csr));
    for (int l = 0; l < L; ++l) {
        /* relax = I - A*2/3 on the union pattern (diagonal always present) */
        const csr* a = &h->ah[l];
        csr r; r.n = a->n; r.m = a->m;
        r.rp = (int64_t*)calloc((size_t)a->n + 1, sizeof(int64_t));
        r.ci = (int64_t*)malloc(sizeof(int64_t) * (size_t)(a->rp[a->n] + a->n));
        r.v = (double*)malloc(sizeof(double) * (size_t)(a->rp[a->n] + a->n));
        int64_t o = 0;
        for (int64_t i = 0; i < a->n; ++i) {
            int diag_done = 0;
            for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) {
                const int64_t j = a->ci[q];
                if (!diag_done && j > i) { r.ci[o] = i; r.v[o] = 1.0; ++o; diag_done = 1; }
                const double s = a->v[q] * 2.0 / 3;
                if (j == i) { r.ci[o] = j; r.v[o] = 1.0 - s; diag_done = 1; }
                else { r.ci[o] = j; r.v[o] = 0.0 - s; }
                ++o;
            }
            if (!diag_done) { r.ci[o] = i; r.v[o] = 1.0; ++o; }
            r.rp[i + 1] = o;
        }
        h->relax[l] = r;
    }
    /* FP32 copies */
    h->ah_v32 = (float**)calloc((size_t)L + 1, sizeof(float*));
    h->dinv32 = (float**)calloc((size_t)L + 1, sizeof(float*));
    h->relax_v32 = (float**)calloc((size_t)L, sizeof(float*));
    h->R_v32 = (float**)calloc((size_t)L, sizeof(float*));
    h->P_v32 = (float**)calloc((size_t)L, sizeof(float*));
    for (int l = 0; l <= L; ++l) {
        h->ah_v32[l] = to_f32(h->ah[l].v, h->ah[l].rp[h->ah[l].n]);
        h->dinv32[l] = to_f32(h->dinv[l], h->ah[l].n);
    }
    for (int l = 0; l < L; ++l) {
        h->relax_v32[l] = to_f32(h->relax[l].v, h->relax[l].rp[h->relax[l].n]);
        h->R_v32[l] = to_f32(h->R[l].v, h->R[l].rp[h->R[l].n]);
        h->P_v32[l] = to_f32(h->P[l].v, h->P[l].rp[h->P[l].n]);
    }
    h->cinv32 = to_f32(h->cinv, h->nc * h->nc);
    return h;
}

int orc_mg_levels(void* hv, int64_t* sizes)
{
    orc_mg* h = (orc_mg*)hv;
    if (sizes) for (int l = 0; l <= h->L; ++l) sizes[l] = h->ah[l].n;
    return h->L + 1;
}

#define MG_T double
#define MG_SUF f64
#include "mg_impl.h"
#undef MG_T
#undef MG_SUF
#define MG_T float
#define MG_SUF f32
#include "mg_impl.h"
#undef MG_T
#undef MG_SUF

/* singular mode: b'_p = b_p - sum(b) (sequential sum) */
static double* singular_rhs(const orc_mg* h, const double* b)
{
    const int64_t n = h->A0.n;
    double* bp = (double*)malloc(sizeof(double) * (size_t)n);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) { bp[i] = b[i]; s += b[i]; }
    bp[h->pin] = b[h->pin] - s;
    return bp;
}

/* gauge: x -= x_p - b_p, so the pinned row holds exactly */
static void gauge_fix(const orc_mg* h, const double* b, double* x)
{
    const double shift = x[h->pin] - b[h->pin];
    for (int64_t i = 0; i < h->A0.n; ++i) x[i] = x[i] - shift;
}

void orc_mg_vcycles(void* hv, const double* b, double* x, int cycles)
{
    orc_mg* h = (orc_mg*)hv;
    if (!h->singular) {
        vcycle_f64(h, b, x, cycles);
        return;
    }
    double* bp = singular_rhs(h, b);
    vcycle_f64(h, bp, x, cycles);
    free(bp);
    gauge_fix(h, b, x);
}

int orc_mg_is_singular(void* hv, int64_t* pin)
{
    orc_mg* h = (orc_mg*)hv;
    if (pin) *pin = h->pin;
    return h->singular;
}

double orc_mg_residual_norm(void* hv, const double* b, const double* x)
{
    orc_mg* h = (orc_mg*)hv;
    const int64_t n = h->A0.n;
    double* ax = (double*)malloc(sizeof(double) * (size_t)n);
    csr_spmv(&h->A0, x, ax);
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) { const double r = b[i] - ax[i]; s += r * r; }
    free(ax);
    return sqrt(s);
}

void orc_mg_solve_mixed(void* hv, const double* b, double* x, int max_cycles, int corr_interval)
{
    orc_mg* h = (orc_mg*)hv;
    const int64_t n = h->A0.n;
    double* ax = (double*)malloc(sizeof(double) * (size_t)n);
    float* db = (float*)malloc(sizeof(float) * (size_t)n);
    float* dx = (float*)malloc(sizeof(float) * (size_t)n);
    const csr* A = h->singular ? &h->As : &h->A0;
    double* bp = h->singular ? singular_rhs(h, b) : NULL;
    const double* bb = h->singular ? bp : b;
    for (int cycle = 0; cycle < max_cycles;) {
        const int batch = corr_interval < max_cycles - cycle ? corr_interval : max_cycles - cycle;
        csr_spmv(A, x, ax);
        for (int64_t i = 0; i < n; ++i) db[i] = (float)(bb[i] - ax[i]);
        memset(dx, 0, sizeof(float) * (size_t)n);
        vcycle_f32(h, db, dx, batch);
        for (int64_t i = 0; i < n; ++i) x[i] += (double)dx[i];
        cycle += batch;
    }
    if (h->singular) { gauge_fix(h, b, x); free(bp); }
    free(ax); free(db); free(dx);
}

void orc_mg_free(void* hv)
{
    orc_mg* h = (orc_mg*)hv;
    if (!h) return;
    for (int l = 0; l <= h->L; ++l) { csr_free(&h->ah[l]); free(h->dinv[l]); free(h->ah_v32[l]); free(h->dinv32[l]); }
    for (int l = 0; l < h->L; ++l) {
        csr_free(&h->relax[l]); csr_free(&h->R[l]); csr_free(&h->P[l]);
        free(h->relax_v32[l]); free(h->R_v32[l]); free(h->P_v32[l]);
    }
    csr_free(&h->A0);
    if (h->singular) csr_free(&h->As);
    free(h->ah); free(h->dinv); free(h->relax); free(h->R); free(h->P);
    free(h->ah_v32); free(h->dinv32); free(h->relax_v32); free(h->R_v32); free(h->P_v32);
    free(h->cinv); free(h->cinv32);
    free(h);
}
