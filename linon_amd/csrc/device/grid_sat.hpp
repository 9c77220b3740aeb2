// Geometry of the uniform grid and the triangle/cell overlap test shared by the host build
// (grid.cpp) and the device build (grid_build.hip).  Both evaluate the test in double
// precision with FMA contraction off, so they decide every (face, cell) pair identically and
// produce the same cell lists.
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace linon {

// Cell (x, y, z) spans [lo + x * cs, lo + (x + 1) * cs) per axis; a face is listed in every cell
// whose box, inflated by `margin`, it overlaps.
struct GridGeom {
    float lo[3];
    float hi[3];
    int dim[3];
    double cs;      // cell size
    double margin;  // conservative inflation of every cell box
};

__host__ __device__ inline double gs_min(double a, double b) { return a < b ? a : b; }
__host__ __device__ inline double gs_max(double a, double b) { return a > b ? a : b; }
__host__ __device__ inline double gs_abs(double a) { return a < 0.0 ? -a : a; }

// Separating-axis test of a triangle against an axis-aligned box (centre c, half sizes half):
// the 9 edge-cross-axis directions, the 3 box normals and the triangle normal
// (Akenine-Moeller, "Fast 3D triangle-box overlap testing", 2001).
__host__ __device__ inline bool tri_box_overlap(const double c[3], const double half[3], const double v0[3],
                                                const double v1[3], const double v2[3]) {
#pragma clang fp contract(off)
    double p[3][3], e[3][3];
    for (int k = 0; k < 3; ++k) {
        p[0][k] = v0[k] - c[k];
        p[1][k] = v1[k] - c[k];
        p[2][k] = v2[k] - c[k];
    }
    for (int k = 0; k < 3; ++k) {
        e[0][k] = p[1][k] - p[0][k];
        e[1][k] = p[2][k] - p[1][k];
        e[2][k] = p[0][k] - p[2][k];
    }
    for (int j = 0; j < 3; ++j) {
        const double* f = e[j];
        const double axes[3][3] = {{0.0, -f[2], f[1]}, {f[2], 0.0, -f[0]}, {-f[1], f[0], 0.0}};
        for (int i = 0; i < 3; ++i) {
            const double* a = axes[i];
            const double q0 = a[0] * p[0][0] + a[1] * p[0][1] + a[2] * p[0][2];
            const double q1 = a[0] * p[1][0] + a[1] * p[1][1] + a[2] * p[1][2];
            const double q2 = a[0] * p[2][0] + a[1] * p[2][1] + a[2] * p[2][2];
            const double r = half[0] * gs_abs(a[0]) + half[1] * gs_abs(a[1]) + half[2] * gs_abs(a[2]);
            const double mn = gs_min(q0, gs_min(q1, q2)), mx = gs_max(q0, gs_max(q1, q2));
            if (mn > r || mx < -r) return false;
        }
    }
    for (int k = 0; k < 3; ++k) {
        const double mn = gs_min(p[0][k], gs_min(p[1][k], p[2][k]));
        const double mx = gs_max(p[0][k], gs_max(p[1][k], p[2][k]));
        if (mn > half[k] || mx < -half[k]) return false;
    }
    const double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                         e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    const double d = n[0] * p[0][0] + n[1] * p[0][1] + n[2] * p[0][2];
    const double r = half[0] * gs_abs(n[0]) + half[1] * gs_abs(n[1]) + half[2] * gs_abs(n[2]);
    return gs_abs(d) <= r;
}

// Inclusive cell range [c0, c1] per axis of a face's margin-inflated bounding box.
__host__ __device__ inline void face_cell_range(const GridGeom& g, const double v[3][3], int c0[3], int c1[3]) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) {
        const double tmin = gs_min(v[0][k], gs_min(v[1][k], v[2][k])) - g.margin;
        const double tmax = gs_max(v[0][k], gs_max(v[1][k], v[2][k])) + g.margin;
        const int a = (int)floor((tmin - (double)g.lo[k]) / g.cs);
        const int b = (int)floor((tmax - (double)g.lo[k]) / g.cs);
        c0[k] = a < 0 ? 0 : a;
        c1[k] = b > g.dim[k] - 1 ? g.dim[k] - 1 : b;
    }
}

// Whether the face overlaps cell (x, y, z) (box inflated by the margin).
__host__ __device__ inline bool face_in_cell(const GridGeom& g, const double v[3][3], int x, int y, int z) {
#pragma clang fp contract(off)
    const double h = 0.5 * g.cs + g.margin;
    const double half[3] = {h, h, h};
    const double c[3] = {(double)g.lo[0] + (x + 0.5) * g.cs, (double)g.lo[1] + (y + 0.5) * g.cs,
                         (double)g.lo[2] + (z + 0.5) * g.cs};
    return tri_box_overlap(c, half, v[0], v[1], v[2]);
}

}  // namespace linon
