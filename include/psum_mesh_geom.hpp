// Geometry and device-side actions of the triangle-mesh particle boundary,
// shared by the host-side partition build, the psum_hip_mesh_* kernels and
// the header-only C++ API (include/psum/particle_boundary.hpp): the
// segment/triangle test, the cell traversal of a segment over the uniform
// partition, the box/triangle overlap (particle_boundary/geometry.hpp:
// 24-350), detect over a partition (triangle_mesh_trigger.hpp:54-155) and the
// four materials (basic_material_set.hpp:31-105).  FP64 throughout; compile
// without contraction (-ffp-contract=off) for results identical to the
// oracle's.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "psum_hip.h"

namespace psum_hip {
namespace mesh {

constexpr double kSmall = 1e-20;   // finite_small (geometry.hpp:16)

// Moller-Trumbore segment/triangle test (line_seg_tri_intersect_test,
// geometry.hpp:208-265): hit when the segment parameter lambda is in (0, 1]
__host__ __device__ inline bool seg_tri(const double* t, const double* p1, const double* p2, double& lambda)
{
    const double e1x = t[3] - t[0], e1y = t[4] - t[1], e1z = t[5] - t[2];
    const double e2x = t[6] - t[0], e2y = t[7] - t[1], e2z = t[8] - t[2];
    const double dx = p2[0] - p1[0], dy = p2[1] - p1[1], dz = p2[2] - p1[2];
    const double hx = dy * e2z - dz * e2y;
    const double hy = dz * e2x - dx * e2z;
    const double hz = dx * e2y - dy * e2x;
    const double a = e1x * hx + e1y * hy + e1z * hz;
    lambda = 0.0;
    if (fabs(a) < kSmall) return false;
    const double f = 1.0 / a;
    const double ox = p1[0] - t[0], oy = p1[1] - t[1], oz = p1[2] - t[2];
    const double u = f * (ox * hx + oy * hy + oz * hz);
    if (u < 0.0 || u > 1.0) return false;
    const double qx = oy * e1z - oz * e1y;
    const double qy = oz * e1x - ox * e1z;
    const double qz = ox * e1y - oy * e1x;
    const double v = f * (dx * qx + dy * qy + dz * qz);
    if (v < 0.0 || u + v > 1.0) return false;
    lambda = f * (e2x * qx + e2y * qy + e2z * qz);
    return lambda > 0.0 && lambda <= 1.0;
}

// Cell traversal of a segment in grid coordinates (trace_to_idxs_begin /
// _next, geometry.hpp:40-206): each axis is mirrored so the direction is
// non-negative, the start is moved to the entry face, and every step
// crosses the nearest cell face.  Cells come out in order along the
// segment; linear index (i*J + j)*K + k.  Cell coordinates are
// range-checked in double before the int conversion.
struct Trace {
    int I, J, K;
    double p[3], d[3];
    int ih, jh, kh;
    double t;
    int kx, ky, kz, bx, by, bz;
    int last;
    bool finished;

    __host__ __device__ int cell() const
    {
        return (kx * ih + bx) * J * K + (ky * jh + by) * K + (kz * kh + bz);
    }

    // first cell, or -1 (finished) when the segment misses the grid
    __host__ __device__ int begin(int I_, int J_, int K_, const double* a, const double* b)
    {
        I = I_; J = J_; K = K_;
        for (int c = 0; c < 3; ++c) {
            p[c] = a[c];
            d[c] = b[c] - a[c];
        }
        t = 0;
        finished = false;
        kx = ky = kz = 1;
        bx = by = bz = 0;
        if (d[0] < 0) { kx = -1; bx = I - 1; p[0] = I - p[0]; d[0] *= -1; }
        if (d[1] < 0) { ky = -1; by = J - 1; p[1] = J - p[1]; d[1] *= -1; }
        if (d[2] < 0) { kz = -1; bz = K - 1; p[2] = K - p[2]; d[2] *= -1; }
        for (int c = 0; c < 3; ++c)
            if (d[c] == 0) d[c] = kSmall;
        const double t0 = -p[0] / d[0];
        const double t1 = -p[1] / d[1];
        const double t2 = -p[2] / d[2];
        double mt = t0;
        if (t1 > mt) mt = t1;
        if (t2 > mt) mt = t2;
        if (mt >= 1) { finished = true; return -1; }
        t = mt < 0 ? 0 : mt;
        for (int c = 0; c < 3; ++c) p[c] += t * d[c];
        const double fi = floor(p[0]), fj = floor(p[1]), fk = floor(p[2]);
        if (!(fi < I) || !(fj < J) || !(fk < K)) { finished = true; return -1; }
        ih = (int)fi;
        jh = (int)fj;
        kh = (int)fk;
        if (mt > 0) {
            if (t == t0) ih = 0;
            if (t == t1) jh = 0;
            if (t == t2) kh = 0;
        }
        last = cell();
        return last;
    }

    // next cell, or -1 (finished) at the segment's end or the grid's far side
    __host__ __device__ int next()
    {
        for (;;) {
            if (finished) return -1;
            const double v0 = ((double)(ih + 1) - p[0]) / d[0];
            const double v1 = ((double)(jh + 1) - p[1]) / d[1];
            const double v2 = ((double)(kh + 1) - p[2]) / d[2];
            double step;
            int axis;
            if (v0 < v1) {
                if (v0 < v2) { step = v0; axis = 0; } else { step = v2; axis = 2; }
            } else {
                if (v1 < v2) { step = v1; axis = 1; } else { step = v2; axis = 2; }
            }
            t += step;
            for (int c = 0; c < 3; ++c) p[c] += step * d[c];
            if (axis == 0) ++ih;
            else if (axis == 1) ++jh;
            else ++kh;
            if (!(t < 1)) { finished = true; return -1; }
            if (ih >= I || jh >= J || kh >= K) { finished = true; return -1; }
            const int idx = cell();
            if (idx != last) {
                last = idx;
                return idx;
            }
        }
    }
};

// line_seg_box_overlap_test (geometry.hpp:267-290)
__host__ __device__ inline bool seg_box(const double* box, const double* a, const double* b)
{
    const double dx = box[3] - box[0], dy = box[4] - box[1], dz = box[5] - box[2];
    const double p1[3] = {(a[0] - box[0]) / dx, (a[1] - box[1]) / dy, (a[2] - box[2]) / dz};
    const double p2[3] = {(b[0] - box[0]) / dx, (b[1] - box[1]) / dy, (b[2] - box[2]) / dz};
    Trace s;
    return s.begin(1, 1, 1, p1, p2) >= 0;
}

// box_tri_overlap_test (geometry.hpp:292-350): a vertex strictly inside,
// a box edge through the triangle, or a triangle edge through the box
__host__ __device__ inline bool box_tri_overlap(const double* t, const double* box)
{
    for (int v = 0; v < 3; ++v) {
        const double* q = t + 3 * v;
        if (q[0] > box[0] && q[0] < box[3] && q[1] > box[1] && q[1] < box[4] && q[2] > box[2] && q[2] < box[5])
            return true;
    }
    const double xs[2] = {box[0], box[3]}, ys[2] = {box[1], box[4]}, zs[2] = {box[2], box[5]};
    double lam;
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[0], ys[i / 2], zs[i % 2]}, b[3] = {xs[1], ys[i / 2], zs[i % 2]};
        if (seg_tri(t, a, b, lam)) return true;
    }
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[i / 2], ys[0], zs[i % 2]}, b[3] = {xs[i / 2], ys[1], zs[i % 2]};
        if (seg_tri(t, a, b, lam)) return true;
    }
    for (int i = 0; i < 4; ++i) {
        const double a[3] = {xs[i / 2], ys[i % 2], zs[0]}, b[3] = {xs[i / 2], ys[i % 2], zs[1]};
        if (seg_tri(t, a, b, lam)) return true;
    }
    return seg_box(box, t, t + 3) || seg_box(box, t, t + 6) || seg_box(box, t + 6, t + 3);
}

// The partition as kernels see it (psum_mesh_get_view): cells of an
// I x J x K grid over box, CSR lists of triangle indices, triangles as 12
// doubles (vertices, unit normal) and an id each.
struct View {
    int I, J, K, fast;
    double box[6];
    double rd[3];            // cells per unit length
    const uint32_t* cell_start;
    const uint32_t* cell_tri;
    const double* tri12;
    const int32_t* ids;
    const float* udf;        // distance field of the full traversal, or null
};

inline View view_of(const psum_mesh_view& v)
{
    View m;
    m.I = v.I;
    m.J = v.J;
    m.K = v.K;
    m.fast = v.fast;
    for (int c = 0; c < 6; ++c) m.box[c] = v.box[c];
    for (int c = 0; c < 3; ++c) m.rd[c] = v.rd[c];
    m.cell_start = v.cell_start;
    m.cell_tri = v.cell_tri;
    m.tri12 = v.tri12;
    m.ids = v.ids;
    m.udf = v.udf;
    return m;
}

// smallest segment parameter among the triangles listed in one cell (first
// on ties); an index outside the partition is an empty cell
__device__ inline bool detect_cell(const View& m, int cell, const double* p1, const double* p2, double& k,
                                   double* nrm, int32_t& id)
{
    if (cell < 0 || cell >= m.I * m.J * m.K) return false;
    const uint32_t b = m.cell_start[cell], e = m.cell_start[cell + 1];
    double best = 1.0;
    int64_t bi = -1;
    for (uint32_t j = b; j < e; ++j) {
        const uint32_t ti = m.cell_tri[j];
        double t9[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) t9[c] = m.tri12[12 * (int64_t)ti + c];
        double lam;
        if (seg_tri(t9, p1, p2, lam))
            if (bi < 0 || lam < best) {
                best = lam;
                bi = ti;
            }
    }
    if (bi < 0) return false;
    k = best;
    for (int c = 0; c < 3; ++c) nrm[c] = m.tri12[12 * bi + 9 + c];
    id = m.ids[bi];
    return true;
}

// linear cell of grid coordinates from the floors (trace_to_idxs_fast);
// -1 when a floor is not a representable int
__device__ inline int64_t fast_cell(const View& m, const double* g)
{
    double f[3];
    for (int c = 0; c < 3; ++c) {
        f[c] = floor(g[c]);
        if (!(f[c] >= -2147483648.0 && f[c] <= 2147483647.0)) return -1;
    }
    return ((int64_t)f[0] * m.J + (int64_t)f[1]) * m.K + (int64_t)f[2];
}

// triangle_mesh_trigger_acc::detect (:54-101): the first cell along the
// segment with a hit (restrict mode) or the end cells only (fast mode)
__device__ inline bool detect(const View& m, const double* p1, const double* p2, double& k, double* nrm,
                              int32_t& id)
{
    double g1[3], g2[3];
    for (int c = 0; c < 3; ++c) {
        g1[c] = (p1[c] - m.box[c]) * m.rd[c];
        g2[c] = (p2[c] - m.box[c]) * m.rd[c];
    }
    if (m.fast) {
        const bool in1 = g1[0] >= 0 && g1[0] < m.I && g1[1] >= 0 && g1[1] < m.J && g1[2] >= 0 && g1[2] < m.K;
        const bool in2 = g2[0] >= 0 && g2[0] < m.I && g2[1] >= 0 && g2[1] < m.J && g2[2] >= 0 && g2[2] < m.K;
        const int64_t c1 = fast_cell(m, g1), c2 = fast_cell(m, g2);
        if (in1 && detect_cell(m, (int)c1, p1, p2, k, nrm, id)) return true;
        if (in2 && c1 != c2) return detect_cell(m, (int)c2, p1, p2, k, nrm, id);
        return false;
    }
    if (m.udf) {
        // distance-field skip (:66-92): from a start cell whose field
        // exceeds one cell, a step shorter than the field cannot reach a
        // triangle, and a longer one starts its traversal that far along
        double f[3];
        const int dims[3] = {m.I, m.J, m.K};
        bool in = true;
        for (int c = 0; c < 3; ++c) {
            f[c] = floor(g1[c]);
            in = in && f[c] >= 0 && f[c] < dims[c];
        }
        if (in) {
            const float u = m.udf[((int64_t)f[0] * m.J + (int64_t)f[1]) * m.K + (int64_t)f[2]];
            if (u > 1.0f) {
                const double d[3] = {g2[0] - g1[0], g2[1] - g1[1], g2[2] - g1[2]};
                const double sq = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const float u2 = u * u;
                if (sq < (double)u2) return false;
                const double len = sqrt(sq);
                for (int c = 0; c < 3; ++c) g1[c] = g1[c] + (double)u * d[c] / len;
            }
        }
    }
    Trace s;
    const int first = s.begin(m.I, m.J, m.K, g1, g2);
    if (first < 0) return false;
    for (int idx = first; !s.finished; idx = s.next())
        if (detect_cell(m, idx, p1, p2, k, nrm, id)) return true;
    return false;
}

// the particle's random stream (random/rander.hpp:40-74)
struct Lcg {
    uint32_t x;
    __device__ __forceinline__ double operator()()
    {
        x = x * 1664525u + 1013904223u;
        return (double)x / 4294967295.0;
    }
};

__device__ __forceinline__ double norm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
__device__ __forceinline__ double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// unit vector across `axis` (the biAxis of RandV_Cosine / RandV_halfMaxw,
// rand_functions.hpp:31-53): axis x (isotropic unit vector), normalised
__device__ inline void bi_axis(Lcg& R, const double* axis, double* bi)
{
    const double r1 = 2 * R() - 1;
    const double r2 = 2 * 3.1415926535897932 * R();
    const double s = sqrt(1 - r1 * r1);
    const double r[3] = {r1, s * sin(r2), s * cos(r2)};
    bi[0] = axis[1] * r[2] - axis[2] * r[1];
    bi[1] = axis[2] * r[0] - axis[0] * r[2];
    bi[2] = axis[0] * r[1] - axis[1] * r[0];
    const double inv = 1 / norm3(bi);
    for (int c = 0; c < 3; ++c) bi[c] = bi[c] * inv;
}

// basic_material_set::acc_type::action (basic_material_set.hpp:31-105) on
// the particle's position, velocity and seed at the hit point with the
// triangle's unit normal
__device__ inline void material_action(int mat, double* pos, double* vel, uint32_t& seed, const double* hit,
                                       const double* n, double temperature, double mass)
{
    double over[3];
    for (int c = 0; c < 3; ++c) over[c] = pos[c] - hit[c];
    switch (mat) {
    case PSUM_MAT_DIFFUSE: {
        Lcg R{seed};
        double bi[3], dir[3];
        bi_axis(R, n, bi);
        const double r1 = sqrt(R());
        const double r2 = sqrt(1 - r1 * r1);
        for (int c = 0; c < 3; ++c) dir[c] = n[c] * r1 + bi[c] * r2;
        const double on = norm3(over), vn = norm3(vel);
        for (int c = 0; c < 3; ++c) {
            pos[c] = hit[c] + dir[c] * on * 0.001;
            vel[c] = dir[c] * vn;
        }
        seed = R.x;
        break;
    }
    case PSUM_MAT_SPECULAR: {
        const double od = 2.0 * dot3(over, n);
        const double vd = 2.0 * dot3(vel, n);
        for (int c = 0; c < 3; ++c) {
            pos[c] = hit[c] + (over[c] - od * n[c]) * 0.001;
            vel[c] = vel[c] - vd * n[c];
        }
        break;
    }
    case PSUM_MAT_MAXWELL_REFLECT: {
        Lcg R{seed};
        const double ot = norm3(over) / norm3(vel);
        double u[3], bi[3];
        const double nn = norm3(n);
        for (int c = 0; c < 3; ++c) u[c] = n[c] / nn;
        bi_axis(R, u, bi);
        const double a = sqrt(-log(R()));
        const double b = sqrt(-log(R()));
        const double s = sqrt(2.0 * 1.3805e-23 * temperature / mass);
        for (int c = 0; c < 3; ++c) {
            const double nv = (bi[c] * a + u[c] * b) * s;
            pos[c] = hit[c] + nv * ot;
            vel[c] = nv;
        }
        seed = R.x;
        break;
    }
    case PSUM_MAT_ABSORB:
        pos[0] = __builtin_nan("");
        break;
    default:
        break;
    }
}

}  // namespace mesh
}  // namespace psum_hip
