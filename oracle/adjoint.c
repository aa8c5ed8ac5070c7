/*
 * Reverse-mode (adjoint) and forward-mode (tangent-linear) derivatives of
 * the oracle's PIC step, in FP64.  TEST INFRASTRUCTURE ONLY (see
 * psum_oracle.h): the checker for the HIP adjoint kernels and stepper.
 *
 * The reference has no derivative code; the forward map differentiated
 * here is orc_pic_run_multi's step (TSI/main.cpp:59-75 and :121-146):
 *   rho^n = sum_p w_dep c_k(x_p^n) + rho_bg      (fixed-point deposit)
 *   phi^n = S(rho^n)                              (neutralise, pin / fix, solve)
 *   v^{n+1} = v^n - qm dt grad_interp(phi^n)(x^n);  x^{n+1} = x^n + dt v^{n+1}
 * and the loss L = W(phi^N), W = orc_field_energy.  CIC weights are
 * piecewise multilinear, so the derivatives are those of the smooth piece
 * the inputs lie in (cell crossings are measure-zero events; the tests
 * check finite differences only where no particle changes cell).
 *
 * Linear pieces:
 *   d rho_k / d x_p          = w_dep grad c_k(x_p)
 *   d v' / d x   (e, f)      = -qm dt H_ef,  H_ef = sum_k phi_k d_e d_f c_k
 *                              (H_ee = 0: c_k is linear in each coordinate)
 *   d v'_e / d phi_k         = -qm dt d_e c_k(x)
 *   S (linear part)          = solve(A', E (I - J) drho) on periodic grids
 *                              (J = mean, E zeroes the pinned node 0),
 *                              solve(A', drho with fixed nodes zeroed) otherwise
 *   S^T                      : periodic:  u = solve(A', mu), mu = lphi with
 *                              mu_0 -= sum(lphi), lrho = u - mean(u); row 0 of
 *                              A' is the identity with right-hand side 0, so
 *                              mu_0 is never read and the solve takes lphi
 *                              with entry 0 zeroed;
 *                              walls: lrho = solve(A', lphi, fixed nodes zeroed)
 *                              (A' is A with identity rows at pinned / fixed
 *                              nodes; its interior block is symmetric)
 *   dW/dphi_i                = eps dV sum_{edges (i,j)} delta_r_d^2 (phi_i - phi_j)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>
#include "psum_oracle.h"

int orc_base(const orc_grid* g, double x0, double x1, double x2, int* c, double* s);
uint32_t orc_key(const orc_grid* g, double x0, double x1, double x2);

/* planted mistakes for the tests that show the checks are sharp (0 = none) */
static int g_plant = 0;
void orc_adjoint_plant(int which) { g_plant = which; }

typedef struct {
    const orc_pic_problem* P;
    int D;
    int64_t nn;
    int all_periodic;
    uint8_t* fx;
    void* mg;
    double wdep[ORC_MAX_SPECIES];
    double scale;
    int64_t* acc;
} ctx_t;

static int fx_shift(double w)
{
    int e = 0;
    frexp(fabs(w), &e);
    return 40 - e;
}

static int ctx_init(ctx_t* c, const orc_pic_problem* P)
{
    memset(c, 0, sizeof(*c));
    c->P = P;
    const orc_grid* g = P->g;
    c->D = g->dim;
    c->nn = orc_num_nodes(g);
    c->all_periodic = 1;
    for (int d = 0; d < c->D; ++d) c->all_periodic &= g->periodic[d] ? 1 : 0;
    c->fx = (uint8_t*)calloc((size_t)c->nn, 1);
    if (P->fixed) memcpy(c->fx, P->fixed, (size_t)c->nn);
    if (c->all_periodic) c->fx[0] = 1;
    const int64_t nnz = orc_poisson_coo(g, P->eps, c->fx, NULL, NULL, NULL);
    uint64_t* r = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nnz);
    uint64_t* cc = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)nnz);
    double* v = (double*)malloc(sizeof(double) * (size_t)nnz);
    orc_poisson_coo(g, P->eps, c->fx, r, cc, v);
    const int I = g->cells[0] + 1, J = c->D > 1 ? g->cells[1] + 1 : 1;
    const int K = c->D > 2 ? g->cells[2] + 1 : 1;
    c->mg = orc_mg_create(c->nn, nnz, r, cc, v, I, J, K, P->depth, P->pre, P->post, 1);
    free(r); free(cc); free(v);
    if (!c->mg) return 0;
    double dV = 1.0;
    for (int d = 0; d < c->D; ++d) dV *= g->delta[d];
    double wmax = 0.0;
    for (int s = 0; s < P->nspecies; ++s) {
        c->wdep[s] = P->q[s] * P->w[s] / dV;
        if (fabs(c->wdep[s]) > wmax) wmax = fabs(c->wdep[s]);
    }
    c->scale = ldexp(1.0, fx_shift(wmax > 0 ? wmax : 1.0));
    c->acc = (int64_t*)malloc(sizeof(int64_t) * (size_t)c->nn);
    return 1;
}

static void ctx_free(ctx_t* c)
{
    if (c->mg) orc_mg_free(c->mg);
    free(c->fx);
    free(c->acc);
}

/* V-cycles on b (already with replaced entries) from the warm start x until
 * ||b - A x|| <= tol ||b|| or max_cycles */
static void mg_solve(ctx_t* c, const double* b, double* x)
{
    double bn = 0.0;
    for (int64_t i = 0; i < c->nn; ++i) bn += b[i] * b[i];
    bn = sqrt(bn);
    if (bn == 0.0) {
        memset(x, 0, sizeof(double) * (size_t)c->nn);
        return;
    }
    for (int cyc = 0; cyc < c->P->max_cycles; ++cyc) {
        orc_mg_vcycles(c->mg, b, x, 1);
        if (orc_mg_residual_norm(c->mg, b, x) <= c->P->tol * bn) break;
    }
}

static double sum_of(const double* a, int64_t n)
{
    double s = 0.0;
    for (int64_t i = 0; i < n; ++i) s += a[i];
    return s;
}

/* forward field: phi (warm start in, solution out) from the positions */
static void solve_field(ctx_t* c, double*** pos, double* b, double* phi)
{
    const orc_pic_problem* P = c->P;
    memset(c->acc, 0, sizeof(int64_t) * (size_t)c->nn);
    for (int s = 0; s < P->nspecies; ++s)
        orc_deposit_fx_f64(P->g, (const double* const*)pos[s], P->n[s], c->wdep[s], c->scale, c->acc);
    for (int64_t i = 0; i < c->nn; ++i) b[i] = (double)c->acc[i] * (1.0 / c->scale) + P->rho_bg;
    if (c->all_periodic) {
        const double mean = sum_of(b, c->nn) / (double)c->nn;
        for (int64_t i = 0; i < c->nn; ++i) b[i] = b[i] - mean;
    }
    for (int64_t i = 0; i < c->nn; ++i)
        if (c->fx[i])
            b[i] = (P->fixed_val && !(c->all_periodic && i == 0)) ? P->fixed_val[i] : 0.0;
    mg_solve(c, b, phi);
}

/* linear part of S: dphi = S_lin drho (drho overwritten) */
static void solve_linear(ctx_t* c, double* drho, double* dphi)
{
    if (c->all_periodic) {
        const double mean = sum_of(drho, c->nn) / (double)c->nn;
        for (int64_t i = 0; i < c->nn; ++i) drho[i] = drho[i] - mean;
    }
    for (int64_t i = 0; i < c->nn; ++i)
        if (c->fx[i]) drho[i] = 0.0;
    memset(dphi, 0, sizeof(double) * (size_t)c->nn);
    mg_solve(c, drho, dphi);
}

/* S_lin^T: lrho = S_lin^T lphi (lphi overwritten) */
static void solve_adjoint(ctx_t* c, double* lphi, double* lrho)
{
    for (int64_t i = 0; i < c->nn; ++i)
        if (c->fx[i]) lphi[i] = 0.0;
    memset(lrho, 0, sizeof(double) * (size_t)c->nn);
    mg_solve(c, lphi, lrho);
    if (c->all_periodic) {
        const double mean = sum_of(lrho, c->nn) / (double)c->nn;
        for (int64_t i = 0; i < c->nn; ++i) lrho[i] = lrho[i] - mean;
    }
}

/* CIC factors: support indices, per-dimension weights W and derivatives G */
static int support(const orc_grid* g, const double* x, int64_t* idx, double W[3][2], double G[3][2])
{
    double coef[8];
    if (!orc_cic_weights(g, x, idx, coef)) return 0;
    int c[3];
    double s[3];
    orc_base(g, x[0], g->dim > 1 ? x[1] : 0, g->dim > 2 ? x[2] : 0, c, s);
    for (int d = 0; d < g->dim; ++d) {
        W[d][0] = 1.0 - s[d];
        W[d][1] = s[d];
        G[d][0] = -g->delta_r[d];
        G[d][1] = g->delta_r[d];
    }
    return 1;
}

/* offset bit of dimension d in support point i: dimension D-1-j takes bit j */
static int bit_of(int D, int i, int d) { return (i >> (D - 1 - d)) & 1; }

/* d_e c_i */
static double grad_c(int D, double W[3][2], double G[3][2], int i, int e)
{
    double r = 1.0;
    for (int d = 0; d < D; ++d) r *= (d == e ? G[d][bit_of(D, i, d)] : W[d][bit_of(D, i, d)]);
    return r;
}

/* d_e d_f c_i, e != f */
static double hess_c(int D, double W[3][2], double G[3][2], int i, int e, int f)
{
    double r = 1.0;
    for (int d = 0; d < D; ++d)
        r *= ((d == e || d == f) ? G[d][bit_of(D, i, d)] : W[d][bit_of(D, i, d)]);
    return r;
}

static void energy_grad(const orc_grid* g, const double* phi, double eps, double* out)
{
    const int D = g->dim;
    int64_t w[3] = {1, 1, 1};
    for (int d = 0; d < D; ++d) w[d] = g->cells[d] + 1;
    double vol = 1.0;
    for (int d = 0; d < D; ++d) vol *= g->delta[d];
    const int64_t nn = orc_num_nodes(g);
    for (int64_t i = 0; i < nn; ++i) out[i] = 0.0;
    /* W = 1/2 eps vol sum_{forward edges} (dr (phi_j - phi_i))^2 */
    for (int64_t a = 0; a < w[0]; ++a)
        for (int64_t b = 0; b < w[1]; ++b)
            for (int64_t k = 0; k < w[2]; ++k) {
                const int64_t id[3] = {a, b, k};
                const int64_t lin = (a * w[1] + b) * w[2] + k;
                for (int d = 0; d < D; ++d) {
                    int64_t nb[3] = {a, b, k};
                    nb[d] = id[d] + 1;
                    if (nb[d] == w[d]) {
                        if (!g->periodic[d]) continue;
                        nb[d] = 0;
                    }
                    const int64_t ln = (nb[0] * w[1] + nb[1]) * w[2] + nb[2];
                    const double t = eps * vol * g->delta_r[d] * g->delta_r[d] * (phi[ln] - phi[lin]);
                    out[ln] += t;
                    out[lin] -= t;
                }
            }
}

static double*** alloc_state(const orc_pic_problem* P, double*** src)
{
    double*** s = (double***)calloc((size_t)P->nspecies, sizeof(double**));
    for (int q = 0; q < P->nspecies; ++q) {
        s[q] = (double**)calloc(3, sizeof(double*));
        for (int d = 0; d < P->g->dim; ++d) {
            s[q][d] = (double*)malloc(sizeof(double) * (size_t)(P->n[q] > 0 ? P->n[q] : 1));
            if (src) memcpy(s[q][d], src[q][d], sizeof(double) * (size_t)P->n[q]);
            else memset(s[q][d], 0, sizeof(double) * (size_t)P->n[q]);
        }
    }
    return s;
}

static void copy_state(const orc_pic_problem* P, double*** dst, double*** src)
{
    for (int q = 0; q < P->nspecies; ++q)
        for (int d = 0; d < P->g->dim; ++d)
            memcpy(dst[q][d], src[q][d], sizeof(double) * (size_t)P->n[q]);
}

static void free_state(const orc_pic_problem* P, double*** s)
{
    if (!s) return;
    for (int q = 0; q < P->nspecies; ++q) {
        for (int d = 0; d < 3; ++d) free(s[q][d]);
        free(s[q]);
    }
    free(s);
}

static void push_all(ctx_t* c, double*** x, double*** v, const double* phi)
{
    const orc_pic_problem* P = c->P;
    for (int s = 0; s < P->nspecies; ++s)
        orc_gather_push_f64(P->g, x[s], v[s], P->n[s], phi, P->q[s] / P->m[s], P->dt,
                            P->m[s] * P->w[s], NULL);
}

static uint64_t fnv(uint64_t h, uint64_t v)
{
    for (int k = 0; k < 8; ++k) {
        h ^= (v >> (8 * k)) & 0xff;
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t hash_keys(const orc_pic_problem* P, double*** x, uint64_t h)
{
    for (int s = 0; s < P->nspecies; ++s)
        for (int64_t p = 0; p < P->n[s]; ++p) {
            const orc_grid* g = P->g;
            double xd[3] = {0, 0, 0};
            for (int d = 0; d < g->dim; ++d) xd[d] = x[s][d][p];
            const uint32_t k = orc_key(g, xd[0], xd[1], xd[2]);
            h = fnv(h, k);
        }
    return h;
}

double orc_pic_loss(const orc_pic_problem* P, double*** pos, double*** vel, uint64_t* key_hash)
{
    ctx_t c;
    if (!ctx_init(&c, P)) return NAN;
    double*** x = alloc_state(P, pos);
    double*** v = alloc_state(P, vel);
    double* b = (double*)malloc(sizeof(double) * (size_t)c.nn);
    double* phi = (double*)calloc((size_t)c.nn, sizeof(double));
    uint64_t h = 1469598103934665603ull;
    for (int st = 0; st < P->steps; ++st) {
        h = hash_keys(P, x, h);
        solve_field(&c, x, b, phi);
        push_all(&c, x, v, phi);
    }
    h = hash_keys(P, x, h);
    solve_field(&c, x, b, phi);
    const double L = orc_field_energy(P->g, phi, P->eps);
    if (key_hash) *key_hash = h;
    free(b); free(phi);
    free_state(P, x); free_state(P, v);
    ctx_free(&c);
    return L;
}

double orc_pic_adjoint(const orc_pic_problem* P, double*** pos, double*** vel, double*** gx,
                       double*** gv)
{
    ctx_t c;
    if (!ctx_init(&c, P)) return NAN;
    const int D = c.D, N = P->steps;
    const int64_t nn = c.nn;
    /* forward, keeping x^n (n = 0..N) and phi^n (n = 0..N) */
    double**** X = (double****)calloc((size_t)N + 1, sizeof(double***));
    double** PHI = (double**)calloc((size_t)N + 1, sizeof(double*));
    double*** v = alloc_state(P, vel);
    double* b = (double*)malloc(sizeof(double) * (size_t)nn);
    double* phi = (double*)calloc((size_t)nn, sizeof(double));
    X[0] = alloc_state(P, pos);
    for (int st = 0; st <= N; ++st) {
        solve_field(&c, X[st], b, phi);
        PHI[st] = (double*)malloc(sizeof(double) * (size_t)nn);
        memcpy(PHI[st], phi, sizeof(double) * (size_t)nn);
        if (st == N) break;
        X[st + 1] = alloc_state(P, X[st]);
        push_all(&c, X[st + 1], v, phi);
    }
    const double L = orc_field_energy(P->g, PHI[N], P->eps);

    double*** lx = alloc_state(P, NULL);
    double*** lv = alloc_state(P, NULL);
    double* lphi = (double*)malloc(sizeof(double) * (size_t)nn);
    double* lrho = (double*)malloc(sizeof(double) * (size_t)nn);
    energy_grad(P->g, PHI[N], P->eps, lphi);
    for (int st = N; st >= 0; --st) {
        /* lrho = S^T lphi; lx += w_dep grad(lrho)(x^st) */
        solve_adjoint(&c, lphi, lrho);
        for (int s = 0; s < P->nspecies; ++s)
            for (int64_t p = 0; p < P->n[s]; ++p) {
                double xp[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) xp[d] = X[st][s][d][p];
                if (!isfinite(xp[0])) continue;
                int64_t idx[8];
                double W[3][2], G[3][2];
                if (!support(P->g, xp, idx, W, G)) continue;
                for (int f = 0; f < D; ++f) {
                    double t = 0.0;
                    for (int i = 0; i < (1 << D); ++i) t += lrho[idx[i]] * grad_c(D, W, G, i, f);
                    lx[s][f][p] += c.wdep[s] * t;
                }
            }
        if (st == 0) break;
        /* reverse of the push of step st-1 (from x^{st-1} with phi^{st-1}) */
        const int n0 = st - 1;
        for (int64_t i = 0; i < nn; ++i) lphi[i] = 0.0;
        for (int s = 0; s < P->nspecies; ++s) {
            const double qmdt = P->q[s] / P->m[s] * P->dt;
            for (int64_t p = 0; p < P->n[s]; ++p) {
                double xp[3] = {0, 0, 0}, lvt[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) xp[d] = X[n0][s][d][p];
                if (!isfinite(xp[0])) continue;
                for (int d = 0; d < D; ++d) lvt[d] = lv[s][d][p] + P->dt * lx[s][d][p];
                int64_t idx[8];
                double W[3][2], G[3][2];
                if (support(P->g, xp, idx, W, G)) {
                    for (int f = 0; f < D; ++f) {
                        double a = 0.0;
                        for (int e = 0; e < D; ++e) {
                            if (e == f) continue;
                            double h = 0.0;
                            for (int i = 0; i < (1 << D); ++i)
                                h += PHI[g_plant == 3 ? n0 + 1 : n0][idx[i]] * hess_c(D, W, G, i, e, f);
                            a += lvt[e] * h;
                        }
                        if (g_plant != 1) lx[s][f][p] -= qmdt * a;
                    }
                    for (int i = 0; i < (1 << D); ++i) {
                        double t = 0.0;
                        for (int e = 0; e < D; ++e) t += lvt[e] * grad_c(D, W, G, i, e);
                        lphi[idx[i]] += (g_plant == 2 ? qmdt : -qmdt) * t;
                    }
                }
                for (int d = 0; d < D; ++d) lv[s][d][p] = lvt[d];
            }
        }
    }
    if (gx) for (int s = 0; s < P->nspecies; ++s) for (int d = 0; d < D; ++d)
        memcpy(gx[s][d], lx[s][d], sizeof(double) * (size_t)P->n[s]);
    if (gv) for (int s = 0; s < P->nspecies; ++s) for (int d = 0; d < D; ++d)
        memcpy(gv[s][d], lv[s][d], sizeof(double) * (size_t)P->n[s]);
    for (int st = 0; st <= N; ++st) { free_state(P, X[st]); free(PHI[st]); }
    free(X); free(PHI);
    free_state(P, v); free_state(P, lx); free_state(P, lv);
    free(b); free(phi); free(lphi); free(lrho);
    ctx_free(&c);
    return L;
}

/* drho = sum_p w_dep grad c_k(x_p) . dx_p over every species; particles are
 * split over OpenMP threads with private grids summed in thread order */
static void tangent_deposit(ctx_t* c, double*** x, double*** dx, double* drho)
{
    const orc_pic_problem* P = c->P;
    const int D = c->D;
    const int64_t nn = c->nn;
    for (int64_t i = 0; i < nn; ++i) drho[i] = 0.0;
#pragma omp parallel
    {
        double* priv = (double*)calloc((size_t)nn, sizeof(double));
        for (int s = 0; s < P->nspecies; ++s) {
#pragma omp for schedule(static)
            for (int64_t p = 0; p < P->n[s]; ++p) {
                double xp[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) xp[d] = x[s][d][p];
                if (!isfinite(xp[0])) continue;
                int64_t idx[8];
                double W[3][2], G[3][2];
                if (!support(P->g, xp, idx, W, G)) continue;
                for (int i = 0; i < (1 << D); ++i) {
                    double t = 0.0;
                    for (int e = 0; e < D; ++e) t += grad_c(D, W, G, i, e) * dx[s][e][p];
                    priv[idx[i]] += c->wdep[s] * t;
                }
            }
        }
#pragma omp for ordered schedule(static, 1)
        for (int t = 0; t < omp_get_num_threads(); ++t) {
#pragma omp ordered
            for (int64_t i = 0; i < nn; ++i) drho[i] += priv[i];
        }
        free(priv);
    }
}

double orc_pic_tangent_multi(const orc_pic_problem* P, double*** pos, double*** vel, int ndir,
                             double**** dpos, double**** dvel, double* dL, double* loss)
{
    ctx_t c;
    if (ndir < 1 || !ctx_init(&c, P)) return NAN;
    const int D = c.D, N = P->steps;
    const int64_t nn = c.nn;
    double*** x = alloc_state(P, pos);
    double*** v = alloc_state(P, vel);
    double**** dx = (double****)malloc(sizeof(double***) * (size_t)ndir);
    double**** dv = (double****)malloc(sizeof(double***) * (size_t)ndir);
    for (int j = 0; j < ndir; ++j) {
        dx[j] = alloc_state(P, dpos[j]);
        dv[j] = alloc_state(P, dvel[j]);
        dL[j] = 0.0;
    }
    double* b = (double*)malloc(sizeof(double) * (size_t)nn);
    double* phi = (double*)calloc((size_t)nn, sizeof(double));
    double* drho = (double*)malloc(sizeof(double) * (size_t)nn);
    double* dphi = (double*)malloc(sizeof(double) * (size_t)nn * (size_t)ndir);
    double* egrad = (double*)malloc(sizeof(double) * (size_t)nn);
    for (int st = 0; st <= N; ++st) {
        solve_field(&c, x, b, phi);
        for (int j = 0; j < ndir; ++j) {
            tangent_deposit(&c, x, dx[j], drho);
            solve_linear(&c, drho, dphi + (size_t)j * nn);
        }
        if (st == N) {
            if (loss) *loss = orc_field_energy(P->g, phi, P->eps);
            energy_grad(P->g, phi, P->eps, egrad);
            for (int j = 0; j < ndir; ++j) {
                const double* dp = dphi + (size_t)j * nn;
                double acc = 0.0;
                for (int64_t i = 0; i < nn; ++i) acc += egrad[i] * dp[i];
                dL[j] = acc;
            }
            break;
        }
        /* dv' = dv - qm dt (H dx + sum_k dphi_k grad c_k); dx' = dx + dt dv' */
        for (int s = 0; s < P->nspecies; ++s) {
            const double qmdt = P->q[s] / P->m[s] * P->dt;
#pragma omp parallel for schedule(static)
            for (int64_t p = 0; p < P->n[s]; ++p) {
                double xp[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) xp[d] = x[s][d][p];
                if (!isfinite(xp[0])) continue;
                int64_t idx[8];
                double W[3][2], G[3][2];
                if (!support(P->g, xp, idx, W, G)) continue;
                double hes[3][3];
                for (int e = 0; e < D; ++e)
                    for (int f = 0; f < D; ++f) {
                        hes[e][f] = 0.0;
                        if (f == e) continue;
                        for (int i = 0; i < (1 << D); ++i) hes[e][f] += phi[idx[i]] * hess_c(D, W, G, i, e, f);
                    }
                for (int j = 0; j < ndir; ++j) {
                    const double* dp = dphi + (size_t)j * nn;
                    double ddv[3] = {0, 0, 0};
                    for (int e = 0; e < D; ++e) {
                        double t = 0.0;
                        for (int i = 0; i < (1 << D); ++i) t += dp[idx[i]] * grad_c(D, W, G, i, e);
                        for (int f = 0; f < D; ++f) {
                            if (f == e) continue;
                            t += hes[e][f] * dx[j][s][f][p];
                        }
                        ddv[e] = -qmdt * t;
                    }
                    for (int e = 0; e < D; ++e) {
                        dv[j][s][e][p] += ddv[e];
                        dx[j][s][e][p] += P->dt * dv[j][s][e][p];
                    }
                }
            }
        }
        push_all(&c, x, v, phi);
        for (int s = 0; s < P->nspecies; ++s)       /* absorbed particles drop out */
            for (int64_t p = 0; p < P->n[s]; ++p)
                if (!isfinite(x[s][0][p]))
                    for (int j = 0; j < ndir; ++j)
                        for (int d = 0; d < D; ++d) { dx[j][s][d][p] = 0.0; dv[j][s][d][p] = 0.0; }
    }
    free_state(P, x); free_state(P, v);
    for (int j = 0; j < ndir; ++j) { free_state(P, dx[j]); free_state(P, dv[j]); }
    free(dx); free(dv);
    free(b); free(phi); free(drho); free(dphi); free(egrad);
    ctx_free(&c);
    return dL[0];
}

double orc_pic_tangent(const orc_pic_problem* P, double*** pos, double*** vel, double*** dpos,
                       double*** dvel, double* loss)
{
    double dL = 0.0;
    double*** dp[1] = {dpos};
    double*** dv[1] = {dvel};
    orc_pic_tangent_multi(P, pos, vel, 1, dp, dv, &dL, loss);
    return dL;
}
