/*
 * Precision-generic body of the oracle's mesh boundary (test
 * infrastructure only), included twice from mesh.c with ORC_T = double |
 * float.  Geometry and actions are FP64; results are rounded to ORC_T.
 */
#define ORC_CAT2(a, b) a##_##b
#define ORC_CAT(a, b) ORC_CAT2(a, b)
#define ORC_FN(name) ORC_CAT(name, ORC_SUF)

/* boundary_router::deal for every valid particle: segment old -> pos,
 * material action of the hit triangle's id (ORC_MAT_*).  Returns the
 * number of hits; hit_id[p] (may be NULL) = id hit or -1. */
int64_t ORC_FN(orc_mesh_boundary)(const orc_mesh* m, int fast, const ORC_T* const* old, ORC_T** pos, ORC_T** vel,
                                  uint32_t* seed, int64_t n, double temperature, double mass, int32_t* hit_id)
{
    int64_t hits = 0;
    for (int64_t p = 0; p < n; ++p) {
        if (hit_id) hit_id[p] = -1;
        if (!isfinite((double)pos[0][p])) continue;
        double a[3], b[3], v[3], k, nrm[3];
        int32_t id;
        int finite = 1;
        for (int c = 0; c < 3; ++c) {
            a[c] = (double)old[c][p];
            b[c] = (double)pos[c][p];
            v[c] = (double)vel[c][p];
            finite = finite && isfinite(a[c]) && isfinite(b[c]);
        }
        /* a segment with a non-finite end is not traced */
        if (!finite || !orc_mesh_detect(m, a, b, fast, &k, nrm, &id)) continue;
        ++hits;
        if (hit_id) hit_id[p] = id;
        double hit[3];
        for (int c = 0; c < 3; ++c) hit[c] = a[c] + k * (b[c] - a[c]);
        orc_material_action(id, b, v, &seed[p], hit, nrm, temperature, mass);
        for (int c = 0; c < 3; ++c) {
            pos[c][p] = (ORC_T)b[c];
            vel[c][p] = (ORC_T)v[c];
        }
    }
    return hits;
}

#undef ORC_FN
#undef ORC_CAT
#undef ORC_CAT2
