/*
 * Precision-generic V-cycle of the oracle's multigrid (test infrastructure
 * only).  Included from mg.c with MG_T = double | float.  Restates
 * multigrid_core::v_cycle_multi / solve_sequence / smooth
 * (multigrid_core.hpp:163-272) and matrix_traits::multiply_add
 * (sparse_linear_trait.hpp:31-35: out = (M x)*ans + b*(b_coeff*ans)).
 */
#define MG_CAT2(a, b) a##_##b
#define MG_CAT(a, b) MG_CAT2(a, b)
#define MG_FN(name) MG_CAT(name, MG_SUF)

/* values of matrix `m` (double CSR) in precision MG_T */
static const MG_T* MG_FN(vals)(const csr* m, const float* v32)
{
    if (sizeof(MG_T) == sizeof(double)) return (const MG_T*)(const void*)m->v;
    return (const MG_T*)(const void*)v32;
}

static void MG_FN(spmv)(const csr* a, const MG_T* av, const MG_T* x, MG_T* y)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < a->n; ++i) {
        MG_T s = 0;
        for (int64_t q = a->rp[i]; q < a->rp[i + 1]; ++q) s = s + av[q] * x[a->ci[q]];
        y[i] = s;
    }
}

static void MG_FN(vcycle)(orc_mg* h, const MG_T* in, MG_T* out, int cycles)
{
    const int L = h->L;
    MG_T** b = (MG_T**)malloc(sizeof(MG_T*) * (size_t)(L + 1));
    MG_T** x = (MG_T**)malloc(sizeof(MG_T*) * (size_t)(L + 1));
    MG_T** buf = (MG_T**)malloc(sizeof(MG_T*) * (size_t)(L + 1));
    MG_T** tmp = (MG_T**)malloc(sizeof(MG_T*) * (size_t)(L + 1));
    const MG_T* dinv[40];
    for (int l = 0; l <= L; ++l) {
        const int64_t n = h->ah[l].n;
        b[l] = (MG_T*)calloc((size_t)n, sizeof(MG_T));
        x[l] = (MG_T*)calloc((size_t)n, sizeof(MG_T));
        buf[l] = (MG_T*)calloc((size_t)n, sizeof(MG_T));
        tmp[l] = (MG_T*)calloc((size_t)n, sizeof(MG_T));
        dinv[l] = sizeof(MG_T) == sizeof(double) ? (const MG_T*)(const void*)h->dinv[l]
                                                 : (const MG_T*)(const void*)h->dinv32[l];
    }
    const MG_T* cinv = sizeof(MG_T) == sizeof(double) ? (const MG_T*)(const void*)h->cinv
                                                      : (const MG_T*)(const void*)h->cinv32;
    const MG_T two3 = (MG_T)(2.0 / 3);
    const int64_t n0 = h->ah[0].n;
    for (int64_t i = 0; i < n0; ++i) b[0][i] = in[i] * dinv[0][i];   /* hadamard_to */
    for (int64_t i = 0; i < n0; ++i) x[0][i] = out[i];               /* warm start */

    for (int cyc = 0; cyc < cycles; ++cyc) {
        for (int l = 0; l < L; ++l) {
            const int64_t n = h->ah[l].n;
            for (int s = 0; s < h->pre; ++s) {                          /* smooth */
                MG_FN(spmv)(&h->relax[l], MG_FN(vals)(&h->relax[l], h->relax_v32[l]), x[l], tmp[l]);
                for (int64_t i = 0; i < n; ++i) buf[l][i] = tmp[l][i] * (MG_T)1 + b[l][i] * (two3 * (MG_T)1);
                MG_T* t = x[l]; x[l] = buf[l]; buf[l] = t;
            }
            /* residual: buf = (A x)*(-1) + b*((-1)*(-1)) */
            MG_FN(spmv)(&h->ah[l], MG_FN(vals)(&h->ah[l], h->ah_v32[l]), x[l], tmp[l]);
            for (int64_t i = 0; i < n; ++i) buf[l][i] = tmp[l][i] * (MG_T)-1 + b[l][i] * (MG_T)1;
            MG_FN(spmv)(&h->R[l], MG_FN(vals)(&h->R[l], h->R_v32[l]), buf[l], b[l + 1]);
            const int64_t nc = h->ah[l + 1].n;
            for (int64_t i = 0; i < nc; ++i) b[l + 1][i] = b[l + 1][i] * dinv[l + 1][i];
            for (int64_t i = 0; i < nc; ++i) x[l + 1][i] = 0;
        }
        /* coarsest: dense inverse times b */
        {
            const int64_t nc = h->nc;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < nc; ++i) {
                MG_T s = 0;
                for (int64_t j = 0; j < nc; ++j) s = s + cinv[i * nc + j] * b[L][j];
                x[L][i] = s;
            }
        }
        for (int l = L - 1; l >= 0; --l) {
            const int64_t n = h->ah[l].n;
            /* prolong-add: x_l = (P x_{l+1})*1 + x_l*(1*1) */
            MG_FN(spmv)(&h->P[l], MG_FN(vals)(&h->P[l], h->P_v32[l]), x[l + 1], tmp[l]);
            for (int64_t i = 0; i < n; ++i) x[l][i] = tmp[l][i] * (MG_T)1 + x[l][i] * (MG_T)1;
            for (int s = 0; s < h->post; ++s) {
                MG_FN(spmv)(&h->relax[l], MG_FN(vals)(&h->relax[l], h->relax_v32[l]), x[l], tmp[l]);
                for (int64_t i = 0; i < n; ++i) buf[l][i] = tmp[l][i] * (MG_T)1 + b[l][i] * (two3 * (MG_T)1);
                MG_T* t = x[l]; x[l] = buf[l]; buf[l] = t;
            }
        }
    }
    for (int64_t i = 0; i < n0; ++i) out[i] = x[0][i];
    for (int l = 0; l <= L; ++l) { free(b[l]); free(x[l]); free(buf[l]); free(tmp[l]); }
    free(b); free(x); free(buf); free(tmp);
}

#undef MG_FN
#undef MG_CAT
#undef MG_CAT2
