/*
 * poisson.c -- general finite-difference Poisson assembly of the reference
 * frontends: fixed (Dirichlet) rows, mixed (Robin/Neumann) boundary nodes
 * through ghost elements that are eliminated as A' = A - B D^-1 C with the
 * right-hand-side map M = -B D^-1, the cylindrical (z, r) form, and a
 * position-dependent permittivity.  TEST INFRASTRUCTURE ONLY (see
 * psum_oracle.h).
 *
 *   orc_poisson_assemble_ex   Poisson_solver_2d.hpp:182-349 (node loop, ghost
 *                             creation, ghost equations, cylindrical surfaces
 *                             and volume), Poisson_solver_3d.hpp:211-350,
 *                             ghost_element_manager.hpp:29-117 (contribution
 *                             extraction, elimination), local_equation.hpp
 *   orc_poisson_rhs           Poisson_solver_2d.hpp:71-103 (make_righthand_item:
 *                             addback = M * values, then fixed-node replace)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"
#include "leq.h"

/* the double nearest pi (M_PI of the C library) */
#define ORC_PI 3.14159265358979323846

/* planted mistakes for the pin tests: 1 = the minus-side ghost equations
 * keep +grad_coef, 2 = the axis face radius is not clamped at 0, 3 = M has
 * the wrong sign */
static int g_plant = 0;
void orc_poisson_plant(int which) { g_plant = which; }

static void node_pos(const orc_grid* g, const int* n, double* p)
{
    for (int d = 0; d < g->dim; ++d) p[d] = g->lb[d] + n[d] * g->delta[d];
}

/* permittivity at the midpoint of nodes n and n + e_axis*step */
static double eps_mid(const orc_grid* g, const int* n, int axis, int step, double eps_c,
                      orc_eps_fn f, void* user)
{
    if (!f) return eps_c;
    int m[3] = {n[0], n[1], n[2]};
    m[axis] += step;
    double a[3], b[3], p[3];
    node_pos(g, n, a);
    node_pos(g, m, b);
    for (int d = 0; d < g->dim; ++d) p[d] = (a[d] + b[d]) / 2;
    return f(p, user);
}

int orc_poisson_assemble_ex(const orc_grid* g, int cylindrical, double eps_c, orc_eps_fn epsf,
                            void* user, const uint8_t* fixed, const orc_mixed_node* mb, int64_t nmb,
                            int64_t cap, uint64_t* rows, uint64_t* cols, double* vals, int64_t* nnz,
                            int64_t mcap, uint64_t* mrows, uint64_t* mcols, double* mvals,
                            int64_t* mnnz)
{
    const int D = g->dim;
    if (D < 2 || D > 3 || (cylindrical && D != 2)) return -1;
    const int I = g->cells[0], J = g->cells[1], K = D > 2 ? g->cells[2] : 0;
    const int64_t nJ = J + 1, nK = K + 1;
    const int64_t nn = (int64_t)(I + 1) * nJ * nK;
    const int ndir = 2 * D;
    int rc = 0;

    /* mixed nodes at fixed elements are dropped: fixed covers mixed
     * (Poisson_solver_2d.hpp:193-202) */
    int64_t* head = (int64_t*)malloc(sizeof(int64_t) * (size_t)nn);
    int64_t* nxt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nmb ? nmb : 1));
    int64_t* ghost = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nmb ? nmb : 1));
    leq* geq = (leq*)calloc((size_t)(nmb ? nmb : 1), sizeof(leq));
    for (int64_t q = 0; q < nn; ++q) head[q] = -1;
    for (int64_t m = nmb - 1; m >= 0; --m) {
        ghost[m] = -1;
        if (mb[m].element >= (uint64_t)nn || mb[m].direction < 0 || mb[m].direction >= ndir) { rc = -1; goto done; }
        if (fixed && fixed[mb[m].element]) { nxt[m] = -2; continue; }
        nxt[m] = head[mb[m].element];
        head[mb[m].element] = m;
    }

    trip* at = (trip*)malloc(sizeof(trip) * (size_t)nn * 7);
    trip* bt = (trip*)malloc(sizeof(trip) * (size_t)(nmb ? nmb : 1));
    int64_t nat = 0, nbt = 0, ng = 0;
    const double dr[3] = {g->delta_r[0], g->delta_r[1], D > 2 ? g->delta_r[2] : 0.0};
#define CYC(a, n) ((((a) % (n)) + (n)) % (n))
    for (int i = 0; i <= I; ++i)
        for (int j = 0; j <= J; ++j)
            for (int k = 0; k <= K; ++k) {
                const int n3[3] = {i, j, k};
                const uint64_t idx = (uint64_t)(((int64_t)i * nJ + j) * nK + k);
                /* neighbours [axis][0 = minus, 1 = plus], cyclic */
                uint64_t nb[3][2];
                for (int a = 0; a < D; ++a)
                    for (int s = 0; s < 2; ++s) {
                        int m3[3] = {i, j, k};
                        const int len = a == 0 ? I + 1 : (a == 1 ? J + 1 : K + 1);
                        m3[a] = CYC(m3[a] + (s ? 1 : -1), len);
                        nb[a][s] = (uint64_t)(((int64_t)m3[0] * nJ + m3[1]) * nK + (D > 2 ? m3[2] : 0));
                    }
                leq fin; fin.n = 0;
                const leq self = leq_pair(idx, 1.0, idx, 1.0);
                if (fixed && fixed[idx]) {
                    fin = self;
                } else {
                    /* ghosts replace the neighbour on the boundary side
                     * (2D N,S,E,W = axis 1 +,-, axis 0 +,-; 3D Xpos..Zneg) */
                    for (int64_t m = head[idx]; m >= 0; m = nxt[m]) {
                        int axis, plus;
                        const int dir = mb[m].direction;
                        if (D == 2) { axis = dir < 2 ? 1 : 0; plus = (dir % 2) == 0; }
                        else { axis = dir / 2; plus = (dir % 2) == 0; }
                        ghost[m] = ng++;
                        nb[axis][plus] = (uint64_t)nn + (uint64_t)ghost[m];
                    }
                    leq grad_c[3];
                    for (int a = 0; a < D; ++a) grad_c[a] = leq_pair(nb[a][0], dr[a] * -0.5, nb[a][1], dr[a] * 0.5);
                    for (int64_t m = head[idx]; m >= 0; m = nxt[m]) {
                        int axis, plus;
                        const int dir = mb[m].direction;
                        if (D == 2) { axis = dir < 2 ? 1 : 0; plus = (dir % 2) == 0; }
                        else { axis = dir / 2; plus = (dir % 2) == 0; }
                        const double gc = (plus || g_plant == 1) ? mb[m].grad_coef : -mb[m].grad_coef;
                        geq[m] = leq_scale(grad_c[axis], gc);
                        leq lv = leq_scale(self, mb[m].value_coef);
                        leq_accum(&geq[m], &lv, 0);
                    }
                    /* one-sided gradients and face fluxes: E - W + N - S (+ U - D) */
                    double s_m[3], s_p[3], vol;
                    if (cylindrical) {
                        double a[3], b[3], c[3];
                        int up[3] = {i, j + 1, 0}, lo[3] = {i, j - 1, 0};
                        node_pos(g, n3, a);
                        node_pos(g, up, b);
                        node_pos(g, lo, c);
                        const double r_up = (a[1] + b[1]) / 2;
                        double r_lo = (a[1] + c[1]) / 2;
                        if (r_lo < 0 && g_plant != 2) r_lo = 0;
                        s_p[1] = g->delta[0] * ORC_PI * 2 * r_up;
                        s_m[1] = g->delta[0] * ORC_PI * 2 * r_lo;
                        s_p[0] = s_m[0] = ORC_PI * (r_up * r_up - r_lo * r_lo);
                        vol = g->delta[0] * ORC_PI * (r_up * r_up - r_lo * r_lo);
                    } else if (D == 2) {
                        s_p[1] = s_m[1] = g->delta[0];
                        s_p[0] = s_m[0] = g->delta[1];
                        vol = g->delta[0] * g->delta[1];
                    } else {
                        s_p[1] = s_m[1] = g->delta[0] * g->delta[2];
                        s_p[0] = s_m[0] = g->delta[1] * g->delta[2];
                        s_p[2] = s_m[2] = g->delta[0] * g->delta[1];
                        vol = g->delta[0] * g->delta[1] * g->delta[2];
                    }
                    /* face order of the reference sum: E, W, N, S, U, D */
                    const int order_axis[6] = {0, 0, 1, 1, 2, 2};
                    for (int f = 0; f < 2 * D; ++f) {
                        const int a = order_axis[f], plus = (f % 2) == 0;
                        const leq gr = plus ? leq_pair(idx, dr[a] * -1, nb[a][1], dr[a])
                                            : leq_pair(nb[a][0], dr[a] * -1, idx, dr[a]);
                        const double e = eps_mid(g, n3, a, plus ? 1 : -1, eps_c, epsf, user);
                        const leq t = leq_scale(leq_scale(gr, plus ? s_p[a] : s_m[a]), e);
                        if (f == 0) fin = t;
                        else leq_accum(&fin, &t, plus ? 0 : 1);
                    }
                    fin = leq_scale(fin, -1.0 / vol);
                }
                for (int q = 0; q < fin.n; ++q) {
                    if (fin.idx[q] >= (uint64_t)nn) {
                        bt[nbt].r = idx; bt[nbt].c = fin.idx[q] - (uint64_t)nn; bt[nbt].v = fin.val[q]; ++nbt;
                    } else {
                        at[nat].r = idx; at[nat].c = fin.idx[q]; at[nat].v = fin.val[q]; ++nat;
                    }
                }
            }
#undef CYC

    /* elimination: D^-1 must be permutation-like (one ghost term per ghost) */
    {
        int64_t* dcount = (int64_t*)calloc((size_t)(ng ? ng : 1), sizeof(int64_t));
        int64_t* deq = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ng ? ng : 1));
        double* dinv = (double*)malloc(sizeof(double) * (size_t)(ng ? ng : 1));
        for (int64_t m = 0; m < nmb; ++m) {
            if (ghost[m] < 0) continue;
            for (int q = 0; q < geq[m].n; ++q)
                if (geq[m].idx[q] >= (uint64_t)nn) {
                    const int64_t gq = (int64_t)(geq[m].idx[q] - (uint64_t)nn);
                    dcount[gq]++;
                    deq[gq] = m;
                    dinv[gq] = 1.0 / geq[m].val[q];
                }
        }
        for (int64_t q = 0; q < ng; ++q)
            if (dcount[q] != 1) rc = -1;
        if (rc == 0) {
            /* M(r, m) = (-B(r, g)) * D^-1(g, m) */
            trip* mt = (trip*)malloc(sizeof(trip) * (size_t)(nbt ? nbt : 1));
            for (int64_t q = 0; q < nbt; ++q) {
                mt[q].r = bt[q].r;
                mt[q].c = (uint64_t)deq[bt[q].c];
                mt[q].v = (g_plant == 3 ? bt[q].v : -bt[q].v) * dinv[bt[q].c];
            }
            /* row-major M */
            for (int64_t a = 1; a < nbt; ++a) {
                trip x = mt[a];
                int64_t b = a - 1;
                while (b >= 0 && (mt[b].r > x.r || (mt[b].r == x.r && mt[b].c > x.c))) { mt[b + 1] = mt[b]; --b; }
                mt[b + 1] = x;
            }
            /* A' = A + M C; (M C)(r, c) sums over m ascending, first term
             * assigned (sparse product order) */
            trip* out = (trip*)malloc(sizeof(trip) * (size_t)(nat + nbt * 8 + 1));
            int64_t no = 0;
            memcpy(out, at, sizeof(trip) * (size_t)nat);
            no = nat;
            qsort(out, (size_t)no, sizeof(trip), trip_colmajor);
            int64_t r0 = 0;
            while (r0 < nbt) {
                int64_t r1 = r0;
                while (r1 < nbt && mt[r1].r == mt[r0].r) ++r1;
                /* row mt[r0].r: columns of M C from the ghost equations */
                leq mc; mc.n = 0;
                for (int64_t q = r0; q < r1; ++q) {
                    const leq* e = &geq[mt[q].c];
                    for (int t = 0; t < e->n; ++t) {
                        if (e->idx[t] >= (uint64_t)nn) continue;
                        const double x = mt[q].v * e->val[t];
                        const int f = leq_find(&mc, e->idx[t]);
                        if (f < 0) leq_insert(&mc, e->idx[t], x);
                        else mc.val[f] += x;
                    }
                }
                for (int t = 0; t < mc.n; ++t) {
                    /* add into an existing A entry, else a new one */
                    int64_t lo = 0, hi = no;
                    const trip key = {mt[r0].r, mc.idx[t], 0.0};
                    while (lo < hi) {
                        const int64_t mid = (lo + hi) / 2;
                        if (trip_colmajor(&out[mid], &key) < 0) lo = mid + 1; else hi = mid;
                    }
                    if (lo < no && out[lo].r == key.r && out[lo].c == key.c) {
                        out[lo].v = out[lo].v + mc.val[t];
                    } else {
                        memmove(out + lo + 1, out + lo, sizeof(trip) * (size_t)(no - lo));
                        out[lo] = key;
                        out[lo].v = mc.val[t];
                        ++no;
                    }
                }
                r0 = r1;
            }
            *nnz = no;
            *mnnz = nbt;
            if (rows) {
                if (cap < no) rc = -2;
                else for (int64_t q = 0; q < no; ++q) { rows[q] = out[q].r; cols[q] = out[q].c; vals[q] = out[q].v; }
            }
            if (mrows && rc == 0) {
                if (mcap < nbt) rc = -2;
                else for (int64_t q = 0; q < nbt; ++q) { mrows[q] = mt[q].r; mcols[q] = mt[q].c; mvals[q] = mt[q].v; }
            }
            free(out);
            free(mt);
        }
        free(dcount);
        free(deq);
        free(dinv);
    }
    free(at);
    free(bt);
done:
    free(head);
    free(nxt);
    free(ghost);
    free(geq);
    return rc;
}

void orc_poisson_rhs(int64_t n, const double* source, int64_t mnnz, const uint64_t* mrows,
                     const uint64_t* mcols, const double* mvals, const double* mixed_values,
                     int64_t nfix, const uint64_t* fix_idx, const double* fix_val, double* b)
{
    for (int64_t q = 0; q < n; ++q) b[q] = source[q];
    /* per row of M: 0.0 + sum_m M(r,m) * value(m) in column order */
    int64_t q = 0;
    while (q < mnnz) {
        int64_t e = q;
        double acc = 0.0;
        while (e < mnnz && mrows[e] == mrows[q]) { acc += mvals[e] * mixed_values[mcols[e]]; ++e; }
        b[mrows[q]] += acc;
        q = e;
    }
    for (int64_t f = 0; f < nfix; ++f) b[fix_idx[f]] = fix_val[f];
}
