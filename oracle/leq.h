/*
 * leq.h -- local_equation restatement shared by the oracle's Poisson
 * assemblers.  TEST INFRASTRUCTURE ONLY (see psum_oracle.h).
 */
#ifndef PSUM_ORACLE_LEQ_H
#define PSUM_ORACLE_LEQ_H

#include <math.h>
#include <stdint.h>

/* ------------------------------------------------------------------ */
/* local_equation: ordered (index -> coeff) map with the reference's
 * cancellation rule (local_equation.hpp:29-61). */
typedef struct { int n; uint64_t idx[16]; double val[16]; } leq;

static inline int leq_find(const leq* e, uint64_t i)
{
    for (int k = 0; k < e->n; ++k) if (e->idx[k] == i) return k;
    return -1;
}
static inline void leq_insert(leq* e, uint64_t i, double v)
{
    int k = e->n;
    while (k > 0 && e->idx[k - 1] > i) { e->idx[k] = e->idx[k - 1]; e->val[k] = e->val[k - 1]; --k; }
    e->idx[k] = i; e->val[k] = v; e->n++;
}
static inline void leq_erase(leq* e, int k)
{
    for (; k + 1 < e->n; ++k) { e->idx[k] = e->idx[k + 1]; e->val[k] = e->val[k + 1]; }
    e->n--;
}
/* initializer-list construction keeps the first of duplicate keys */
static inline leq leq_pair(uint64_t i0, double v0, uint64_t i1, double v1)
{
    leq e; e.n = 0;
    leq_insert(&e, i0, v0);
    if (leq_find(&e, i1) < 0) leq_insert(&e, i1, v1);
    return e;
}
static inline leq leq_scale(leq e, double s)
{
    for (int k = 0; k < e.n; ++k) e.val[k] *= s;
    if (fabs(s) == 0) e.n = 0;
    return e;
}
static inline void leq_accum(leq* e, const leq* o, int sub)
{
    for (int k = 0; k < o->n; ++k) {
        const int f = leq_find(e, o->idx[k]);
        if (f < 0) { leq_insert(e, o->idx[k], sub ? -o->val[k] : o->val[k]); continue; }
        const double old_scale = fabs(o->val[k]) + fabs(e->val[f]);
        if (sub) e->val[f] -= o->val[k]; else e->val[f] += o->val[k];
        if (fabs(e->val[f]) < old_scale * 1e-14) leq_erase(e, f);
    }
}

typedef struct { uint64_t r, c; double v; } trip;
static inline int trip_colmajor(const void* a, const void* b)
{
    const trip* x = (const trip*)a; const trip* y = (const trip*)b;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    if (x->r != y->r) return x->r < y->r ? -1 : 1;
    return 0;
}

#endif
