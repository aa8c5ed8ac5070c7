// Geometry of the triangle-mesh particle boundary, shared by the host-side
// partition build and the device kernels (same arithmetic on both, FP64,
// no contraction): the segment/triangle test, the cell traversal of a
// segment over the uniform partition, and the box/triangle overlap.
// Follows particle_boundary/geometry.hpp:24-350.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

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

}  // namespace mesh
}  // namespace psum_hip
