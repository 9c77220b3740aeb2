// Host-side build of the uniform grid the ray-cast kernel queries: per-cell face lists
// (conservative separating-axis overlap, every cell box inflated by a margin) and a
// Chebyshev distance field over cell occupancy, from which the kernel derives a lower bound
// on the distance to the mesh and skips segments that cannot reach any face.
#include "grid.hpp"

#include <algorithm>
#include <cmath>
#include <thread>

namespace linon {

void pack_faces(const float* xyz, const uint32_t* faces, uint32_t nf, std::vector<float>& out) {
    out.assign((size_t)nf * 12, 0.0f);
    for (uint32_t i = 0; i < nf; ++i) {
        const float* a = xyz + 3 * (size_t)faces[3 * i];
        const float* b = xyz + 3 * (size_t)faces[3 * i + 1];
        const float* c = xyz + 3 * (size_t)faces[3 * i + 2];
        float* o = out.data() + 12 * (size_t)i;
        for (int k = 0; k < 3; ++k) {
            o[k] = a[k];
            o[4 + k] = b[k] - a[k];  // f32 subtraction, as v_in[1] - v_in[0] in the shader
            o[8 + k] = c[k] - a[k];
        }
    }
}

static inline void sub3(const double* a, const double* b, double* o) {
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}

// Separating-axis test of a triangle against an axis-aligned box (centre c, half sizes half):
// the 9 edge-cross-axis directions, the 3 box normals and the triangle normal.
bool tri_box_overlap(const double c[3], const double half[3], const double v0[3], const double v1[3],
                     const double v2[3]) {
    double p[3][3];
    sub3(v0, c, p[0]);
    sub3(v1, c, p[1]);
    sub3(v2, c, p[2]);
    double e[3][3];
    sub3(p[1], p[0], e[0]);
    sub3(p[2], p[1], e[1]);
    sub3(p[0], p[2], e[2]);
    for (int j = 0; j < 3; ++j) {
        const double* f = e[j];
        const double axes[3][3] = {{0.0, -f[2], f[1]}, {f[2], 0.0, -f[0]}, {-f[1], f[0], 0.0}};
        for (int i = 0; i < 3; ++i) {
            const double* a = axes[i];
            double q0 = a[0] * p[0][0] + a[1] * p[0][1] + a[2] * p[0][2];
            double q1 = a[0] * p[1][0] + a[1] * p[1][1] + a[2] * p[1][2];
            double q2 = a[0] * p[2][0] + a[1] * p[2][1] + a[2] * p[2][2];
            double r = half[0] * std::fabs(a[0]) + half[1] * std::fabs(a[1]) + half[2] * std::fabs(a[2]);
            double mn = std::min(q0, std::min(q1, q2)), mx = std::max(q0, std::max(q1, q2));
            if (mn > r || mx < -r) return false;
        }
    }
    for (int k = 0; k < 3; ++k) {
        double mn = std::min(p[0][k], std::min(p[1][k], p[2][k]));
        double mx = std::max(p[0][k], std::max(p[1][k], p[2][k]));
        if (mn > half[k] || mx < -half[k]) return false;
    }
    double n[3] = {e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
                   e[0][0] * e[1][1] - e[0][1] * e[1][0]};
    double d = n[0] * p[0][0] + n[1] * p[0][1] + n[2] * p[0][2];
    double r = half[0] * std::fabs(n[0]) + half[1] * std::fabs(n[1]) + half[2] * std::fabs(n[2]);
    return std::fabs(d) <= r;
}

template <class Fn>
static void parallel_for(size_t n, Fn fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 4096 || nt == 1) {
        fn(0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    size_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        size_t b = t * chunk, e = std::min(n, b + chunk);
        if (b >= e) break;
        th.emplace_back(fn, b, e, t);
    }
    for (auto& x : th) x.join();
}

// min over j of max(|i - j|, g[j]) along one line of n samples with stride s, searched
// outwards from i and stopped once the offset alone reaches the best value found.
static void minmax_line(const uint8_t* in, uint8_t* out, int n, size_t s) {
    for (int i = 0; i < n; ++i) {
        int best = in[(size_t)i * s];
        for (int dlt = 1; dlt < best; ++dlt) {
            int g = 255;
            if (i - dlt >= 0) g = std::min<int>(g, in[(size_t)(i - dlt) * s]);
            if (i + dlt < n) g = std::min<int>(g, in[(size_t)(i + dlt) * s]);
            best = std::min(best, std::max(dlt, g));
        }
        out[(size_t)i * s] = (uint8_t)best;
    }
}

bool build_grid(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                GridBuild& g, std::string& err) {
    g = GridBuild();
    if (resolution < 4 || resolution > 1024) {
        err = "grid resolution must be in [4, 1024]";
        return false;
    }
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (uint32_t i = 0; i < 3 * (size_t)nf; ++i) {
        if (faces[i] >= nv) {
            err = "face index out of range";
            return false;
        }
        const float* v = xyz + 3 * (size_t)faces[i];
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(v[k])) {
                err = "non-finite vertex coordinate";
                return false;
            }
            lo[k] = std::min(lo[k], (double)v[k]);
            hi[k] = std::max(hi[k], (double)v[k]);
        }
    }
    if (nf == 0) {
        for (int k = 0; k < 3; ++k) {
            lo[k] = -1.0;
            hi[k] = 1.0;
        }
    }
    double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    ext = std::max(ext, 1e-6);
    const double margin = 1e-4 * ext + 1e-6;
    const double pad = 2.0 * margin + 1e-3 * ext;
    const double cs = (ext + 2.0 * pad) / resolution;
    for (int k = 0; k < 3; ++k) {
        double l = lo[k] - pad;
        int n = std::max(1, (int)std::ceil((hi[k] + pad - l) / cs));
        g.dim[k] = n;
        g.lo[k] = (float)l;
        g.hi[k] = (float)(l + n * cs);
    }
    g.cell_size = (float)cs;
    g.margin = (float)margin;
    const int nx = g.dim[0], ny = g.dim[1], nz = g.dim[2];
    const size_t nc = g.ncells();

    // (cell, face) pairs, one vector per thread over ascending face ranges
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint64_t>> pairs(nt);
    auto emit = [&](size_t b, size_t e, unsigned tid) {
        std::vector<uint64_t>& out = pairs[tid];
        const double half[3] = {0.5 * cs + margin, 0.5 * cs + margin, 0.5 * cs + margin};
        for (size_t f = b; f < e; ++f) {
            double v[3][3];
            double tmin[3], tmax[3];
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) v[j][k] = xyz[3 * (size_t)faces[3 * f + j] + k];
            int c0[3], c1[3];
            for (int k = 0; k < 3; ++k) {
                tmin[k] = std::min(v[0][k], std::min(v[1][k], v[2][k])) - margin;
                tmax[k] = std::max(v[0][k], std::max(v[1][k], v[2][k])) + margin;
                c0[k] = std::max(0, (int)std::floor((tmin[k] - g.lo[k]) / cs));
                c1[k] = std::min(g.dim[k] - 1, (int)std::floor((tmax[k] - g.lo[k]) / cs));
            }
            for (int z = c0[2]; z <= c1[2]; ++z)
                for (int y = c0[1]; y <= c1[1]; ++y)
                    for (int x = c0[0]; x <= c1[0]; ++x) {
                        double c[3] = {g.lo[0] + (x + 0.5) * cs, g.lo[1] + (y + 0.5) * cs,
                                       g.lo[2] + (z + 0.5) * cs};
                        if (tri_box_overlap(c, half, v[0], v[1], v[2])) {
                            uint64_t cell = ((uint64_t)z * ny + y) * nx + x;
                            out.push_back((cell << 32) | (uint64_t)f);
                        }
                    }
        }
    };
    if (nf < 4096 || nt == 1) {
        emit(0, nf, 0);
    } else {
        std::vector<std::thread> th;
        size_t chunk = (nf + nt - 1) / nt;
        for (unsigned t = 0; t < nt; ++t) {
            size_t b = t * chunk, e = std::min<size_t>(nf, b + chunk);
            if (b >= e) break;
            th.emplace_back(emit, b, e, t);
        }
        for (auto& x : th) x.join();
    }
    // counting sort by cell, stable in (thread, face) order = ascending faces per cell
    g.cell_start.assign(nc + 1, 0u);
    size_t total = 0;
    for (auto& v : pairs) {
        total += v.size();
        for (uint64_t pr : v) g.cell_start[(pr >> 32) + 1]++;
    }
    if (total > 0xFFFFFFFFull) {
        err = "grid has more than 2^32 face references; lower the resolution";
        return false;
    }
    for (size_t c = 0; c < nc; ++c) g.cell_start[c + 1] += g.cell_start[c];
    g.cell_faces.assign(total, 0u);
    {
        std::vector<uint32_t> fill(g.cell_start.begin(), g.cell_start.end() - 1);
        for (auto& v : pairs)
            for (uint64_t pr : v) g.cell_faces[fill[pr >> 32]++] = (uint32_t)(pr & 0xFFFFFFFFu);
    }
    pairs.clear();

    // Chebyshev distance (in cells, capped at 255) from every cell to the nearest occupied one
    std::vector<uint8_t> a(nc), b(nc);
    parallel_for((size_t)ny * nz, [&](size_t lb, size_t le, unsigned) {
        for (size_t line = lb; line < le; ++line) {
            size_t base = line * nx;
            int d = 255;
            for (int x = 0; x < nx; ++x) {
                bool occ = g.cell_start[base + x + 1] > g.cell_start[base + x];
                d = occ ? 0 : std::min(255, d + 1);
                a[base + x] = (uint8_t)d;
            }
            d = 255;
            for (int x = nx - 1; x >= 0; --x) {
                d = (a[base + x] == 0) ? 0 : std::min(255, d + 1);
                a[base + x] = (uint8_t)std::min<int>(a[base + x], d);
            }
        }
    });
    parallel_for((size_t)nz * nx, [&](size_t lb, size_t le, unsigned) {
        for (size_t line = lb; line < le; ++line) {
            size_t z = line / nx, x = line % nx;
            size_t base = z * (size_t)nx * ny + x;
            minmax_line(a.data() + base, b.data() + base, ny, (size_t)nx);
        }
    });
    parallel_for((size_t)ny * nx, [&](size_t lb, size_t le, unsigned) {
        for (size_t line = lb; line < le; ++line) {
            minmax_line(b.data() + line, a.data() + line, nz, (size_t)nx * ny);
        }
    });
    g.dist.swap(a);
    return true;
}

}  // namespace linon
