/*
 * Triangle-mesh particle boundary: the reference's triangle_mesh_trigger
 * (uniform cell partition of the mesh's box, segment traversal, Moller-
 * Trumbore segment/triangle test) and boundary_router with the four
 * materials of basic_material_set.  TEST INFRASTRUCTURE ONLY (see
 * psum_oracle.h).
 *
 *   segment/triangle test     particle_boundary/geometry.hpp:208-265
 *   cell traversal            geometry.hpp:24-206 (trace_to_idxs_fast,
 *                             trace_to_idxs_begin / _next)
 *   box/triangle overlap      geometry.hpp:267-350
 *   partition build           triangle_mesh_trigger.hpp:171-275 (normals
 *                             recomputed, zero-area triangles dropped,
 *                             cells of the clamped bounding box kept when
 *                             the cell box overlaps the triangle), the
 *                             automatic box :358-392
 *   detect                    triangle_mesh_trigger.hpp:54-155 (first
 *                             traversed cell with a hit; smallest lambda
 *                             in that cell, first on ties); with a
 *                             distance field, the skip of :66-92
 *   distance field            geometry.hpp:343-460 (compute_udf)
 *   router + materials        boundary_router.hpp:32-58,
 *                             basic_material_set.hpp:17-107 with
 *                             RandV_Cosine / RandV_halfMaxw
 *                             (random/rand_functions.hpp:31-53; the two
 *                             log draws of RandV_halfMaxw taken left to
 *                             right)
 *
 * Where the reference reads outside its partition (a traversal cell index
 * that rounding puts outside the grid) this restatement treats the cell as
 * empty, and cell coordinates are range-checked in double before the int
 * conversion.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "psum_oracle.h"

#define SMALL 1e-20

int orc_seg_tri(const double* t, const double* p1, const double* p2, double* lambda)
{
    const double e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
    const double e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
    const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
    const double hx = dy * e2z - dz * e2y;
    const double hy = dz * e2x - dx * e2z;
    const double hz = dx * e2y - dy * e2x;
    const double a = e1x * hx + e1y * hy + e1z * hz;
    *lambda = 0.0;
    if (fabs(a) < SMALL) return 0;
    const double f = 1.0 / a;
    const double ox = p1[0] - t[0], oy = p1[1] - t[1], oz = p1[2] - t[2];
    const double u = f * (ox * hx + oy * hy + oz * hz);
    if (u < 0.0 || u > 1.0) return 0;
    const double qx = oy * e1z - oz * e1y;
    const double qy = oz * e1x - ox * e1z;
    const double qz = ox * e1y - oy * e1x;
    const double v = f * (dx * qx + dy * qy + dz * qz);
    if (v < 0.0 || u + v > 1.0) return 0;
    const double l = f * (e2x * qx + e2y * qy + e2z * qz);
    *lambda = l;
    return l > 0.0 && l <= 1.0;
}

static int cell_of(const orc_trace* s)
{
    return (s->kx * s->ih + s->bx) * s->J * s->K + (s->ky * s->jh + s->by) * s->K + (s->kz * s->kh + s->bz);
}

int orc_trace_begin(int I, int J, int K, const double* a, const double* b, orc_trace* s)
{
    s->I = I; s->J = J; s->K = K;
    s->p[0] = a[0]; s->p[1] = a[1]; s->p[2] = a[2];
    s->d[0] = b[0] - a[0]; s->d[1] = b[1] - a[1]; s->d[2] = b[2] - a[2];
    s->t = 0;
    s->finished = 0;
    s->kx = 1; s->bx = 0; s->ky = 1; s->by = 0; s->kz = 1; s->bz = 0;
    if (s->d[0] < 0) { s->kx = -1; s->bx = I - 1; s->p[0] = I - s->p[0]; s->d[0] *= -1; }
    if (s->d[1] < 0) { s->ky = -1; s->by = J - 1; s->p[1] = J - s->p[1]; s->d[1] *= -1; }
    if (s->d[2] < 0) { s->kz = -1; s->bz = K - 1; s->p[2] = K - s->p[2]; s->d[2] *= -1; }
    for (int c = 0; c < 3; ++c)
        if (s->d[c] == 0) s->d[c] = SMALL;
    const double t0 = -s->p[0] / s->d[0];
    const double t1 = -s->p[1] / s->d[1];
    const double t2 = -s->p[2] / s->d[2];
    double mt = t0;
    if (t1 > mt) mt = t1;
    if (t2 > mt) mt = t2;
    if (mt >= 1) { s->finished = 1; return -1; }
    s->t = mt < 0 ? 0 : mt;
    for (int c = 0; c < 3; ++c) s->p[c] += s->t * s->d[c];
    const double fi = floor(s->p[0]), fj = floor(s->p[1]), fk = floor(s->p[2]);
    if (!(fi < I) || !(fj < J) || !(fk < K)) { s->finished = 1; return -1; }
    s->ih = (int)fi; s->jh = (int)fj; s->kh = (int)fk;
    if (mt > 0) {
        if (s->t == t0) s->ih = 0;
        if (s->t == t1) s->jh = 0;
        if (s->t == t2) s->kh = 0;
    }
    s->last = cell_of(s);
    return s->last;
}

int orc_trace_next(orc_trace* s)
{
    for (;;) {
        if (s->finished) return -1;
        const double t0 = (double)(s->ih + 1) - s->p[0];
        const double t1 = (double)(s->jh + 1) - s->p[1];
        const double t2 = (double)(s->kh + 1) - s->p[2];
        const double v0 = t0 / s->d[0], v1 = t1 / s->d[1], v2 = t2 / s->d[2];
        double step;
        int axis;
        if (v0 < v1) {
            if (v0 < v2) { step = v0; axis = 0; } else { step = v2; axis = 2; }
        } else {
            if (v1 < v2) { step = v1; axis = 1; } else { step = v2; axis = 2; }
        }
        s->t += step;
        for (int c = 0; c < 3; ++c) s->p[c] += step * s->d[c];
        if (axis == 0) ++s->ih; else if (axis == 1) ++s->jh; else ++s->kh;
        if (!(s->t < 1)) { s->finished = 1; return -1; }
        if (s->ih >= s->I || s->jh >= s->J || s->kh >= s->K) { s->finished = 1; return -1; }
        const int idx = cell_of(s);
        if (idx != s->last) { s->last = idx; return idx; }
    }
}

static int seg_box(const double* box, const double* a, const double* b)
{
    const double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
    const double p1[3] = {(a[0] - box[0]) / dx, (a[1] - box[1]) / dy, (a[2] - box[2]) / dz};
    const double p2[3] = {(b[0] - box[0]) / dx, (b[1] - box[1]) / dy, (b[2] - box[2]) / dz};
    orc_trace s;
    return orc_trace_begin(1, 1, 1, p1, p2, &s) >= 0;
}

int orc_box_tri_overlap(const double* t, const double* box)
{
    for (int v = 0; v < 3; ++v) {
        const double* p = t + 3 * v;
        if (p[0] > box[0] && p[0] < box[3] && p[1] > box[1] && p[1] < box[4] && p[2] > box[2] && p[2] < box[5])
            return 1;
    }
    const double xs[2] = {box[0], box[3]}, ys[2] = {box[1], box[4]}, zs[2] = {box[2], box[5]};
    double lam;
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[0], ys[i / 2], zs[i % 2]}, b[3] = {xs[1], ys[i / 2], zs[i % 2]};
        if (orc_seg_tri(t, a, b, &lam)) return 1;
    }
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[i / 2], ys[0], zs[i % 2]}, b[3] = {xs[i / 2], ys[1], zs[i % 2]};
        if (orc_seg_tri(t, a, b, &lam)) return 1;
    }
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[i / 2], ys[i % 2], zs[0]}, b[3] = {xs[i / 2], ys[i % 2], zs[1]};
        if (orc_seg_tri(t, a, b, &lam)) return 1;
    }
    if (seg_box(box, t, t + 3)) return 1;
    if (seg_box(box, t, t + 6)) return 1;
    if (seg_box(box, t + 6, t + 3)) return 1;
    return 0;
}

void orc_mesh_auto_box(const double* tri, int64_t ntri, double* box)
{
    double lo[3] = {tri[0], tri[1], tri[2]}, hi[3] = {tri[0], tri[1], tri[2]};
    for (int64_t i = 0; i < ntri; ++i)
        for (int v = 0; v < 3; ++v)
            for (int c = 0; c < 3; ++c) {
                const double x = tri[9 * i + 3 * v + c];
                if (x < lo[c]) lo[c] = x;
                if (x > hi[c]) hi[c] = x;
            }
    const double fl[3] = {0.1003, 0.1027, 0.1047}, fh[3] = {0.1057, 0.1059, 0.1069};
    for (int c = 0; c < 3; ++c) {
        box[c] = lo[c] - fl[c] * (hi[c] - lo[c]);
        box[3 + c] = hi[c] + fh[c] * (hi[c] - lo[c]);
    }
}

int64_t orc_mesh_build(const double* tri, const int32_t* ids, int64_t ntri, int I, int J, int K, const double* box,
                       double* tri12, int32_t* kept, uint32_t* cell_start, uint32_t* cell_tri, int64_t cap)
{
    const int64_t ncell = (int64_t)I * J * K;
    const double span[3] = {box[3] - box[0], box[4] - box[1], box[5] - box[2]};
    const double rd[3] = {(double)I / span[0], (double)J / span[1], (double)K / span[2]};
    const int dims[3] = {I, J, K};
    /* per-cell lists in triangle order: count, then fill */
    uint32_t* cnt = (uint32_t*)calloc((size_t)ncell + 1, sizeof(uint32_t));
    int64_t total = 0;
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            cell_start[0] = 0;
            for (int64_t c = 0; c < ncell; ++c) cell_start[c + 1] = cell_start[c] + cnt[c];
            memset(cnt, 0, (size_t)ncell * sizeof(uint32_t));
        }
        for (int64_t i = 0; i < ntri; ++i) {
            const double* t = tri + 9 * i;
            const double e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
            const double e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
            const double nx = e1y * e2z - e1z * e2y;
            const double ny = e1z * e2x - e1x * e2z;
            const double nz = e1x * e2y - e1y * e2x;
            const double nrm = sqrt(nx * nx + ny * ny + nz * nz);
            if (pass == 0) kept[i] = !(nrm < SMALL);
            if (nrm < SMALL) continue;
            if (pass == 0) {
                memcpy(tri12 + 12 * i, t, 9 * sizeof(double));
                tri12[12 * i + 9] = nx / nrm;
                tri12[12 * i + 10] = ny / nrm;
                tri12[12 * i + 11] = nz / nrm;
            }
            int lo[3], hi[3];
            for (int c = 0; c < 3; ++c) {
                const double mn = fmin(t[c], fmin(t[3 + c], t[6 + c]));
                const double mx = fmax(t[c], fmax(t[3 + c], t[6 + c]));
                const double a = floor((mn - box[c]) * rd[c]), b = ceil((mx - box[c]) * rd[c]);
                lo[c] = a < 0 ? 0 : (a >= dims[c] ? dims[c] - 1 : (int)a);
                hi[c] = b < 0 ? 0 : (b >= dims[c] ? dims[c] - 1 : (int)b);
            }
            const double dx = span[0] / (double)I, dy = span[1] / (double)J, dz = span[2] / (double)K;
            for (int a = lo[0]; a <= hi[0]; ++a)
                for (int b = lo[1]; b <= hi[1]; ++b)
                    for (int c = lo[2]; c <= hi[2]; ++c) {
                        const int64_t cell = (int64_t)a * J * K + (int64_t)b * K + c;
                        const double bx0 = box[0] + a * dx, by0 = box[1] + b * dy, bz0 = box[2] + c * dz;
                        const double cb[6] = {bx0, by0, bz0, bx0 + dx, by0 + dy, bz0 + dz};
                        if (!orc_box_tri_overlap(t, cb)) continue;
                        if (pass == 0) {
                            ++cnt[cell];
                            ++total;
                        } else {
                            const int64_t at = (int64_t)cell_start[cell] + cnt[cell]++;
                            if (at < cap) cell_tri[at] = (uint32_t)i;
                        }
                    }
        }
    }
    free(cnt);
    (void)ids;
    return total;
}

static int detect_cell(const orc_mesh* m, int64_t cell, const double* p1, const double* p2, double* k, double* nrm,
                       int32_t* id)
{
    if (cell < 0 || cell >= (int64_t)m->I * m->J * m->K) return 0;
    const uint32_t b = m->cell_start[cell], e = m->cell_start[cell + 1];
    double best = 1.0;
    int64_t bi = -1;
    for (uint32_t j = b; j < e; ++j) {
        double lam;
        if (orc_seg_tri(m->tri12 + 12 * (int64_t)m->cell_tri[j], p1, p2, &lam))
            if (bi < 0 || lam < best) { best = lam; bi = m->cell_tri[j]; }
    }
    if (bi < 0) return 0;
    *k = best;
    nrm[0] = m->tri12[12 * bi + 9];
    nrm[1] = m->tri12[12 * bi + 10];
    nrm[2] = m->tri12[12 * bi + 11];
    *id = m->ids[bi];
    return 1;
}

/* linear cell of grid coordinates g from the floors; -1 when a floor is not
 * a representable int */
static int64_t fast_cell(const orc_mesh* m, const double* g)
{
    double f[3];
    for (int c = 0; c < 3; ++c) {
        f[c] = floor(g[c]);
        if (!(f[c] >= -2147483648.0 && f[c] <= 2147483647.0)) return -1;
    }
    return ((int64_t)f[0] * m->J + (int64_t)f[1]) * m->K + (int64_t)f[2];
}

int orc_mesh_detect(const orc_mesh* m, const double* p1, const double* p2, int fast, double* k, double* nrm,
                    int32_t* id)
{
    double g1[3] = {(p1[0] - m->box[0]) * m->rd[0], (p1[1] - m->box[1]) * m->rd[1], (p1[2] - m->box[2]) * m->rd[2]};
    const double g2[3] = {(p2[0] - m->box[0]) * m->rd[0], (p2[1] - m->box[1]) * m->rd[1], (p2[2] - m->box[2]) * m->rd[2]};
    if (fast) {
        /* trace_to_idxs_fast: both cell indices come from the floors whether
         * or not the point is inside (:25-37) */
        const int in1 = g1[0] >= 0 && g1[0] < m->I && g1[1] >= 0 && g1[1] < m->J && g1[2] >= 0 && g1[2] < m->K;
        const int in2 = g2[0] >= 0 && g2[0] < m->I && g2[1] >= 0 && g2[1] < m->J && g2[2] >= 0 && g2[2] < m->K;
        const int64_t c1 = fast_cell(m, g1), c2 = fast_cell(m, g2);
        if (in1 && detect_cell(m, c1, p1, p2, k, nrm, id)) return 1;
        if (in2 && c1 != c2) return detect_cell(m, c2, p1, p2, k, nrm, id);
        return 0;
    }
    if (m->udf) {
        /* distance-field skip (:66-92): from a start cell whose field
         * exceeds one cell, a step shorter than the field cannot reach a
         * triangle, and a longer one starts its traversal that far along */
        double f[3];
        int in = 1;
        const int dims[3] = {m->I, m->J, m->K};
        for (int c = 0; c < 3; ++c) {
            f[c] = floor(g1[c]);
            in = in && f[c] >= 0 && f[c] < dims[c];
        }
        if (in) {
            const float u = m->udf[((int64_t)f[0] * m->J + (int64_t)f[1]) * m->K + (int64_t)f[2]];
            if (u > 1.0f) {
                const double d[3] = {g2[0] - g1[0], g2[1] - g1[1], g2[2] - g1[2]};
                const double sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const float u2 = u * u;
                if (sq < (double)u2) return 0;
                const double len = sqrt(sq);
                for (int c = 0; c < 3; ++c) g1[c] = g1[c] + (double)u * d[c] / len;
            }
        }
    }
    orc_trace s;
    const int first = orc_trace_begin(m->I, m->J, m->K, g1, g2, &s);
    if (first < 0) return 0;
    for (int idx = first; !s.finished; idx = orc_trace_next(&s))
        if (detect_cell(m, idx, p1, p2, k, nrm, id)) return 1;
    return 0;
}

void orc_compute_udf(int I, int J, int K, const uint8_t* occupied, float* udf, int udf_k, int max_loop)
{
    const int64_t n = (int64_t)I * J * K;
    int any = 0;
    for (int64_t c = 0; c < n && !any; ++c) any = occupied[c] != 0;
    if (!any) {
        const int lo = I < J ? (I < K ? I : K) : (J < K ? J : K);
        for (int64_t c = 0; c < n; ++c) udf[c] = (float)(lo - 1);
        return;
    }
    const float cap = (float)(udf_k - 1);
    for (int i = 0; i < I; ++i)
        for (int j = 0; j < J; ++j)
            for (int k = 0; k < K; ++k) {
                const int64_t c = ((int64_t)i * J + j) * K + k;
                if (occupied[c]) { udf[c] = 0.0f; continue; }
                float best = 1e30f;
                for (int di = -udf_k; di <= udf_k; ++di)
                    for (int dj = -udf_k; dj <= udf_k; ++dj)
                        for (int dk = -udf_k; dk <= udf_k; ++dk) {
                            const int a = i + di, b = j + dj, e = k + dk;
                            if (a < 0 || a >= I || b < 0 || b >= J || e < 0 || e >= K) continue;
                            if (!occupied[((int64_t)a * J + b) * K + e]) continue;
                            /* cells apart by d share no point closer than d - 1 per axis */
                            const int li = abs(di) - 1 > 0 ? abs(di) - 1 : 0;
                            const int lj = abs(dj) - 1 > 0 ? abs(dj) - 1 : 0;
                            const int lk = abs(dk) - 1 > 0 ? abs(dk) - 1 : 0;
                            const float d = sqrtf((float)(li * li + lj * lj + lk * lk));
                            if (d < best) best = d;
                        }
                float v = best < 1e29f ? best : cap;
                udf[c] = v < cap ? v : cap;
            }
    float* nu = (float*)malloc((size_t)n * sizeof(float));
    for (int loop = 0;; ++loop) {
        int updated = 0;
        for (int i = 0; i < I; ++i)
            for (int j = 0; j < J; ++j)
                for (int k = 0; k < K; ++k) {
                    const int64_t c = ((int64_t)i * J + j) * K + k;
                    const float cur = udf[c];
                    nu[c] = cur;
                    if (!(cur > 0.0f)) continue;
                    int raise = 1, nb = 0;
                    for (int di = -1; di <= 1; ++di)
                        for (int dj = -1; dj <= 1; ++dj)
                            for (int dk = -1; dk <= 1; ++dk) {
                                if (!di && !dj && !dk) continue;
                                const int a = i + di, b = j + dj, e = k + dk;
                                if (a < 0 || a >= I || b < 0 || b >= J || e < 0 || e >= K) continue;
                                ++nb;
                                if (udf[((int64_t)a * J + b) * K + e] < cur) raise = 0;
                            }
                    if (raise && nb > 0) {
                        nu[c] = cur + 1.0f;
                        updated = 1;
                    }
                }
        memcpy(udf, nu, (size_t)n * sizeof(float));
        if (!updated || loop + 1 > max_loop) break;
    }
    free(nu);
}

/* planted mistakes: 1 = specular reflection keeps the normal velocity
 * component, 2 = traversal visits only the start cell */
static int g_mesh_plant = 0;
void orc_mesh_plant(int which) { g_mesh_plant = which; }

static double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
static double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* the unit vector across `axis` used by RandV_Cosine / RandV_halfMaxw */
static void bi_axis(uint32_t* x, const double* axis, double* bi)
{
    double r[3];
    orc_rand_spherical(x, r);
    bi[0] = axis[1] * r[2] - axis[2] * r[1];
    bi[1] = axis[2] * r[0] - axis[0] * r[2];
    bi[2] = axis[0] * r[1] - axis[1] * r[0];
    const double inv = 1 / norm3(bi);
    for (int c = 0; c < 3; ++c) bi[c] = bi[c] * inv;
}

void orc_material_action(int mat, double* pos, double* vel, uint32_t* seed, const double* hit, const double* n,
                         double temperature, double mass)
{
    double over[3];
    for (int c = 0; c < 3; ++c) over[c] = pos[c] - hit[c];
    switch (mat) {
    case ORC_MAT_DIFFUSE: {
        double bi[3], dir[3];
        bi_axis(seed, n, bi);
        const double r1 = sqrt(orc_lcg_next(seed));
        const double r2 = sqrt(1 - r1 * r1);
        for (int c = 0; c < 3; ++c) dir[c] = n[c] * r1 + bi[c] * r2;
        const double on = norm3(over), vn = norm3(vel);
        for (int c = 0; c < 3; ++c) {
            pos[c] = hit[c] + dir[c] * on * 0.001;
            vel[c] = dir[c] * vn;
        }
        break;
    }
    case ORC_MAT_SPECULAR: {
        const double od = 2.0 * dot3(over, n);
        const double vd = g_mesh_plant == 1 ? 0.0 : 2.0 * dot3(vel, n);
        for (int c = 0; c < 3; ++c) {
            pos[c] = hit[c] + (over[c] - od * n[c]) * 0.001;
            vel[c] = vel[c] - vd * n[c];
        }
        break;
    }
    case ORC_MAT_MAXWELL: {
        const double ot = norm3(over) / norm3(vel);
        double u[3], bi[3], nv[3];
        const double nn = norm3(n);
        for (int c = 0; c < 3; ++c) u[c] = n[c] / nn;
        bi_axis(seed, u, bi);
        const double a = sqrt(-log(orc_lcg_next(seed)));
        const double b = sqrt(-log(orc_lcg_next(seed)));
        const double s = sqrt(2.0 * 1.3805e-23 * temperature / mass);
        for (int c = 0; c < 3; ++c) nv[c] = (bi[c] * a + u[c] * b) * s;
        for (int c = 0; c < 3; ++c) {
            pos[c] = hit[c] + nv[c] * ot;
            vel[c] = nv[c];
        }
        break;
    }
    case ORC_MAT_ABSORB:
        pos[0] = NAN;
        break;
    }
}

#define ORC_T double
#define ORC_SUF f64
#include "mesh_impl.h"
#undef ORC_T
#undef ORC_SUF
#define ORC_T float
#define ORC_SUF f32
#include "mesh_impl.h"
#undef ORC_T
#undef ORC_SUF
