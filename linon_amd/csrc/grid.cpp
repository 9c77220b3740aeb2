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

bool grid_geometry(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                   GridGeom& geo, std::string& err) {
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
        geo.dim[k] = n;
        geo.lo[k] = (float)l;
        geo.hi[k] = (float)(l + n * cs);
    }
    geo.cs = cs;
    geo.margin = margin;
    return true;
}

bool build_grid(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                GridBuild& g, std::string& err) {
    g = GridBuild();
    GridGeom geo;
    if (!grid_geometry(xyz, nv, faces, nf, resolution, geo, err)) return false;
    for (int k = 0; k < 3; ++k) {
        g.dim[k] = geo.dim[k];
        g.lo[k] = geo.lo[k];
        g.hi[k] = geo.hi[k];
    }
    g.cell_size = (float)geo.cs;
    g.margin = (float)geo.margin;
    const int nx = g.dim[0], ny = g.dim[1], nz = g.dim[2];
    const size_t nc = g.ncells();

    // (cell, face) pairs, one vector per thread over ascending face ranges
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::vector<uint64_t>> pairs(nt);
    auto emit = [&](size_t b, size_t e, unsigned tid) {
        std::vector<uint64_t>& out = pairs[tid];
        for (size_t f = b; f < e; ++f) {
            double v[3][3];
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 3; ++k) v[j][k] = xyz[3 * (size_t)faces[3 * f + j] + k];
            int c0[3], c1[3];
            face_cell_range(geo, v, c0, c1);
            for (int z = c0[2]; z <= c1[2]; ++z)
                for (int y = c0[1]; y <= c1[1]; ++y)
                    for (int x = c0[0]; x <= c1[0]; ++x)
                        if (face_in_cell(geo, v, x, y, z)) {
                            uint64_t cell = ((uint64_t)z * ny + y) * nx + x;
                            out.push_back((cell << 32) | (uint64_t)f);
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
