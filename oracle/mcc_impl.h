/*
 * Precision-generic body of the oracle's MCC step (test infrastructure
 * only), included twice from mcc.c with ORC_T = double | float.  The
 * arithmetic is FP64 for both; results are rounded to ORC_T on store.
 */
#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_FN(name) ORC_CAT(name, ORC_SUF)

int64_t ORC_FN(orc_mcc)(const orc_grid* g, const orc_mcc_process* procs, int nproc,
                        const double* const* density, ORC_T** pos, ORC_T** vel, uint32_t* seed, int64_t n,
                        double dt, ORC_T** e_pos, ORC_T** e_vel, uint32_t* e_seed, ORC_T** i_pos,
                        ORC_T** i_vel, uint32_t* i_seed, int64_t cap, int32_t* outcome)
{
    const int D = g->dim;
    int64_t made_total = 0;
    for (int64_t p = 0; p < n; ++p) {
        double x[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) x[d] = (double)pos[d][p];
        if (!isfinite(x[0])) {
            if (outcome)
                for (int k = 0; k < nproc; ++k) outcome[p * nproc + k] = -2;
            continue;
        }
        double v[3];
        for (int c = 0; c < 3; ++c) v[c] = (double)vel[c][p];
        uint32_t s = seed[p];
        for (int k = 0; k < nproc; ++k) {
            const double dens = density_at(g, &procs[k], density ? density[k] : NULL, x);
            int made = 0;
            double pe[3], pi[3];
            uint32_t se = 0, si = 0;
            const int hit = run_process(&procs[k], dens, dt, v, &s, &made, pe, &se, pi, &si);
            if (outcome) outcome[p * nproc + k] = hit;
            if (made) {
                if (made_total < cap) {
                    for (int d = 0; d < D; ++d) {
                        e_pos[d][made_total] = pos[d][p];
                        i_pos[d][made_total] = pos[d][p];
                    }
                    for (int c = 0; c < 3; ++c) {
                        e_vel[c][made_total] = (ORC_T)pe[c];
                        i_vel[c][made_total] = (ORC_T)pi[c];
                    }
                    e_seed[made_total] = se;
                    i_seed[made_total] = si;
                }
                ++made_total;
            }
        }
        for (int c = 0; c < 3; ++c) vel[c][p] = (ORC_T)v[c];
        seed[p] = s;
    }
    return made_total;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT2
