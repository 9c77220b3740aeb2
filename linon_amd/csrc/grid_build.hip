// Device build of the uniform grid (the host build in grid.cpp computes the same arrays):
//   1. count: one lane per face counts, per overlapped cell, one reference (atomic add);
//   2. exclusive scan of the per-cell counts (64-bit, block scans + one pass over the block
//      sums) -> cell_start;
//   3. fill: the same overlap decisions again, each reference placed at cell_start[cell] plus
//      an atomic cursor;
//   4. each cell's list sorted by face id (the host build's ascending order);
//   5. Chebyshev distance field: a 1-D distance along x, then min-max passes along y and z.
// The overlap test is the shared double-precision separating-axis test with contraction off,
// so every decision, and therefore every array, equals the host build's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "device/grid_sat.hpp"
#include "grid_build.hpp"

namespace linon {

namespace {

constexpr int kScanBlock = 1024;  // elements per block of the local scan (256 lanes x 4)

__device__ __forceinline__ void load_face(const float* xyz, const uint32_t* faces, uint32_t f, double v[3][3]) {
    for (int j = 0; j < 3; ++j) {
        const float* p = xyz + 3 * (size_t)faces[3 * (size_t)f + j];
        for (int k = 0; k < 3; ++k) v[j][k] = (double)p[k];
    }
}

// Faces whose cell range holds more than this many cells are tested cell-parallel by a whole
// block (grid_cells_big_kernel) instead of by one lane: a wall of the Cornell box spans 2 x 256^2
// cells of the 256^3 grid, against a handful for an icosphere face.
constexpr size_t kBigFaceCells = 256;

__device__ __forceinline__ size_t range_cells(const int c0[3], const int c1[3]) {
    return (size_t)(c1[0] - c0[0] + 1) * (size_t)(c1[1] - c0[1] + 1) * (size_t)(c1[2] - c0[2] + 1);
}

// One reference of face f to cell (x, y, z): count it (FILL = false) or place it (FILL = true).
template <bool FILL>
__device__ __forceinline__ void add_reference(const GridGeom& g, size_t f, int x, int y, int z, unsigned* counter,
                                              const unsigned* cell_start, unsigned* cell_faces, unsigned* overflow) {
    const size_t cell = ((size_t)z * (size_t)g.dim[1] + (size_t)y) * (size_t)g.dim[0] + (size_t)x;
    if (!FILL) {
        atomicAdd(counter + cell, 1u);
    } else {
        const unsigned slot = atomicAdd(counter + cell, 1u);
        const unsigned b = cell_start[cell], e = cell_start[cell + 1];
        if (slot < e - b)
            cell_faces[b + slot] = (unsigned)f;
        else
            atomicOr(overflow, 1u);  // counts and placements disagree: reported, never written
    }
}

// Count (FILL = false) or place (FILL = true) the references of one face per lane.  Faces with
// more than kBigFaceCells cells in range are left to grid_cells_big_kernel; the counting pass
// lists them in big[0 .. *nbig).
template <bool FILL>
__global__ void __launch_bounds__(256) grid_cells_kernel(GridGeom g, const float* xyz, const uint32_t* faces,
                                                         uint32_t nf, unsigned* counter, const unsigned* cell_start,
                                                         unsigned* cell_faces, unsigned* overflow, unsigned* big,
                                                         unsigned* nbig) {
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    double v[3][3];
    load_face(xyz, faces, (uint32_t)f, v);
    int c0[3], c1[3];
    face_cell_range(g, v, c0, c1);
    if (c1[0] < c0[0] || c1[1] < c0[1] || c1[2] < c0[2]) return;  // outside the grid (cannot happen)
    if (range_cells(c0, c1) > kBigFaceCells) {
        if (!FILL) big[atomicAdd(nbig, 1u)] = (unsigned)f;  // at most nf entries
        return;
    }
    for (int z = c0[2]; z <= c1[2]; ++z)
        for (int y = c0[1]; y <= c1[1]; ++y)
            for (int x = c0[0]; x <= c1[0]; ++x)
                if (face_in_cell(g, v, x, y, z))
                    add_reference<FILL>(g, f, x, y, z, counter, cell_start, cell_faces, overflow);
}

// The faces listed by the counting pass, one block per face at a time (grid-stride over the
// list), its lanes striding over the face's cell range.
template <bool FILL>
__global__ void __launch_bounds__(256) grid_cells_big_kernel(GridGeom g, const float* xyz, const uint32_t* faces,
                                                             const unsigned* big, const unsigned* nbig,
                                                             unsigned* counter, const unsigned* cell_start,
                                                             unsigned* cell_faces, unsigned* overflow) {
    const unsigned n = *nbig;
    for (unsigned i = blockIdx.x; i < n; i += gridDim.x) {
        const size_t f = big[i];
        double v[3][3];
        load_face(xyz, faces, (uint32_t)f, v);
        int c0[3], c1[3];
        face_cell_range(g, v, c0, c1);
        const size_t sx = (size_t)(c1[0] - c0[0] + 1), sy = (size_t)(c1[1] - c0[1] + 1);
        const size_t cells = range_cells(c0, c1);
        for (size_t k = threadIdx.x; k < cells; k += blockDim.x) {
            const int x = c0[0] + (int)(k % sx);
            const int y = c0[1] + (int)((k / sx) % sy);
            const int z = c0[2] + (int)(k / (sx * sy));
            if (face_in_cell(g, v, x, y, z))
                add_reference<FILL>(g, f, x, y, z, counter, cell_start, cell_faces, overflow);
        }
    }
}

__device__ __forceinline__ unsigned long long block_inclusive_scan(unsigned long long v, unsigned long long* s) {
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < (int)blockDim.x; off <<= 1) {
        const unsigned long long x = t >= off ? s[t - off] : 0ull;
        __syncthreads();
        s[t] += x;
        __syncthreads();
    }
    return s[t];
}

// Exclusive scan inside each block of kScanBlock counts; block totals to bsum.
__global__ void __launch_bounds__(256) scan_local_kernel(const unsigned* in, size_t n, unsigned long long* out,
                                                         unsigned long long* bsum) {
    __shared__ unsigned long long s[256];
    const size_t base = (size_t)blockIdx.x * kScanBlock + (size_t)threadIdx.x * 4;
    unsigned long long v[4], t = 0;
    for (int k = 0; k < 4; ++k) {
        v[k] = base + k < n ? in[base + k] : 0u;
        t += v[k];
    }
    const unsigned long long incl = block_inclusive_scan(t, s);
    unsigned long long run = incl - t;
    for (int k = 0; k < 4; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == blockDim.x - 1) bsum[blockIdx.x] = incl;
}

// Exclusive scan of the nb block totals in place (one block, each lane a contiguous chunk);
// the grand total to *total.
__global__ void __launch_bounds__(1024) scan_sums_kernel(unsigned long long* bsum, size_t nb,
                                                         unsigned long long* total) {
    __shared__ unsigned long long s[1024];
    const size_t chunk = (nb + blockDim.x - 1) / blockDim.x;
    const size_t b0 = (size_t)threadIdx.x * chunk, b1 = b0 + chunk < nb ? b0 + chunk : nb;
    unsigned long long t = 0;
    for (size_t i = b0; i < b1; ++i) t += bsum[i];
    const unsigned long long incl = block_inclusive_scan(t, s);
    unsigned long long run = incl - t;
    for (size_t i = b0; i < b1; ++i) {
        const unsigned long long x = bsum[i];
        bsum[i] = run;
        run += x;
    }
    if (threadIdx.x == blockDim.x - 1) *total = incl;
}

__global__ void __launch_bounds__(256) scan_finish_kernel(const unsigned long long* local,
                                                          const unsigned long long* bsum, size_t n, unsigned* out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (unsigned)(local[i] + bsum[i / kScanBlock]);
}

__global__ void __launch_bounds__(256) fill_u32_kernel(unsigned* p, size_t n, unsigned value) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = value;
}

// Ascending face ids inside every cell (lists are short; insertion sort per lane).
__global__ void __launch_bounds__(256) sort_cells_kernel(const unsigned* cell_start, size_t nc, unsigned* faces) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const unsigned b = cell_start[c], e = cell_start[c + 1];
    for (unsigned i = b + 1; i < e; ++i) {
        const unsigned key = faces[i];
        unsigned j = i;
        while (j > b && faces[j - 1] > key) {
            faces[j] = faces[j - 1];
            --j;
        }
        faces[j] = key;
    }
}

// Distance along x to the nearest occupied cell of the row (capped at 255): a forward and a
// backward sweep per row.
__global__ void __launch_bounds__(256) dist_x_kernel(const unsigned* cell_start, int nx, size_t rows,
                                                     unsigned char* a) {
    const size_t row = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const size_t base = row * (size_t)nx;
    int d = 255;
    for (int x = 0; x < nx; ++x) {
        const bool occ = cell_start[base + x + 1] > cell_start[base + x];
        d = occ ? 0 : (d + 1 < 255 ? d + 1 : 255);
        a[base + x] = (unsigned char)d;
    }
    d = 255;
    for (int x = nx - 1; x >= 0; --x) {
        d = a[base + x] == 0 ? 0 : (d + 1 < 255 ? d + 1 : 255);
        if (d < a[base + x]) a[base + x] = (unsigned char)d;
    }
}

// out[i] = min over j of max(|i - j|, in[j]) along one line (n samples, stride s), searched
// outwards from i until the offset alone reaches the best value.  y lines: line = (z, x);
// z lines: line = (y, x).
__global__ void __launch_bounds__(256) dist_minmax_kernel(const unsigned char* in, unsigned char* out, int nx,
                                                          int ny, int nz, int axis) {
    const size_t line = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t base, s;
    int n;
    if (axis == 1) {
        if (line >= (size_t)nz * nx) return;
        base = (line / nx) * (size_t)nx * ny + line % nx;
        s = (size_t)nx;
        n = ny;
    } else {
        if (line >= (size_t)ny * nx) return;
        base = line;
        s = (size_t)nx * ny;
        n = nz;
    }
    const unsigned char* src = in + base;
    unsigned char* dst = out + base;
    for (int i = 0; i < n; ++i) {
        int best = src[(size_t)i * s];
        for (int dlt = 1; dlt < best; ++dlt) {
            int g = 255;
            if (i - dlt >= 0) g = min(g, (int)src[(size_t)(i - dlt) * s]);
            if (i + dlt < n) g = min(g, (int)src[(size_t)(i + dlt) * s]);
            best = min(best, max(dlt, g));
        }
        dst[(size_t)i * s] = (unsigned char)best;
    }
}

// Per-face (v0, v1 - v0, v2 - v0) as three float4 (f32 subtraction, as the shader's edges).
__global__ void __launch_bounds__(256) pack_faces_kernel(const float* xyz, const uint32_t* faces, uint32_t nf,
                                                         float4* tris) {
    const size_t f = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const float* a = xyz + 3 * (size_t)faces[3 * f];
    const float* b = xyz + 3 * (size_t)faces[3 * f + 1];
    const float* c = xyz + 3 * (size_t)faces[3 * f + 2];
    tris[3 * f] = make_float4(a[0], a[1], a[2], 0.0f);
    tris[3 * f + 1] = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.0f);
    tris[3 * f + 2] = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.0f);
}

unsigned blocks_for(size_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void DeviceGrid::release() {
    for (void** p : {(void**)&cell_start, (void**)&cell_faces, (void**)&dist, (void**)&tris}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    refs = 0;
}

bool build_grid_device(const GridGeom& g, const float* h_xyz, uint32_t nv, const uint32_t* h_faces, uint32_t nf,
                       hipStream_t s, DeviceGrid& out, std::string& err) {
    out.release();
    const size_t nx = (size_t)g.dim[0], ny = (size_t)g.dim[1], nz = (size_t)g.dim[2];
    const size_t nc = nx * ny * nz;
    const size_t nb = (nc + kScanBlock - 1) / kScanBlock;
    float* d_xyz = nullptr;
    uint32_t* d_faces = nullptr;
    unsigned* d_count = nullptr;
    unsigned long long *d_local = nullptr, *d_bsum = nullptr, *d_total = nullptr;
    unsigned* d_flag = nullptr;
    unsigned* d_big = nullptr;   // faces left to the block-per-face kernel, and their count
    unsigned* d_nbig = nullptr;
    unsigned char* d_tmp = nullptr;
    bool ok = false;
    auto check = [&](hipError_t e, const char* what) {
        if (e != hipSuccess) err = std::string("device grid build: ") + what + ": " + hipGetErrorString(e);
        return e == hipSuccess;
    };
    do {
        if (!check(hipMalloc(&d_xyz, std::max<size_t>(1, 3 * (size_t)nv) * 4), "alloc") ||
            !check(hipMalloc(&d_faces, std::max<size_t>(1, 3 * (size_t)nf) * 4), "alloc") ||
            !check(hipMalloc(&d_count, nc * 4), "alloc") || !check(hipMalloc(&d_local, nc * 8), "alloc") ||
            !check(hipMalloc(&d_bsum, nb * 8), "alloc") || !check(hipMalloc(&d_total, 8), "alloc") ||
            !check(hipMalloc(&d_flag, 4), "alloc") || !check(hipMalloc(&d_tmp, nc), "alloc") ||
            !check(hipMalloc(&d_big, std::max<size_t>(1, (size_t)nf) * 4), "alloc") ||
            !check(hipMalloc(&d_nbig, 4), "alloc") ||
            !check(hipMalloc(&out.cell_start, (nc + 1) * 4), "alloc") || !check(hipMalloc(&out.dist, nc), "alloc") ||
            !check(hipMalloc(&out.tris, std::max<size_t>(1, (size_t)nf) * 48), "alloc"))
            break;
        if (nv && !check(hipMemcpyAsync(d_xyz, h_xyz, 3 * (size_t)nv * 4, hipMemcpyHostToDevice, s), "upload")) break;
        if (nf && !check(hipMemcpyAsync(d_faces, h_faces, 3 * (size_t)nf * 4, hipMemcpyHostToDevice, s), "upload"))
            break;
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nc, 256)), dim3(256), 0, s, d_count, nc, 0u);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, s, d_flag, (size_t)1, 0u);
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, s, d_nbig, (size_t)1, 0u);
        const unsigned big_blocks = (unsigned)std::max<size_t>(1, std::min<size_t>(nf, 8192));
        if (nf) {
            hipLaunchKernelGGL(grid_cells_kernel<false>, dim3(blocks_for(nf, 256)), dim3(256), 0, s, g, d_xyz,
                               d_faces, nf, d_count, (const unsigned*)nullptr, (unsigned*)nullptr, d_flag, d_big,
                               d_nbig);
            hipLaunchKernelGGL(grid_cells_big_kernel<false>, dim3(big_blocks), dim3(256), 0, s, g, d_xyz, d_faces,
                               (const unsigned*)d_big, (const unsigned*)d_nbig, d_count, (const unsigned*)nullptr,
                               (unsigned*)nullptr, d_flag);
            hipLaunchKernelGGL(pack_faces_kernel, dim3(blocks_for(nf, 256)), dim3(256), 0, s, d_xyz, d_faces, nf,
                               reinterpret_cast<float4*>(out.tris));
        }
        hipLaunchKernelGGL(scan_local_kernel, dim3((unsigned)nb), dim3(256), 0, s, d_count, nc, d_local, d_bsum);
        hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(1024), 0, s, d_bsum, nb, d_total);
        hipLaunchKernelGGL(scan_finish_kernel, dim3(blocks_for(nc, 256)), dim3(256), 0, s, d_local, d_bsum, nc,
                           out.cell_start);
        if (!check(hipGetLastError(), "count/scan launch")) break;
        unsigned long long total = 0;
        if (!check(hipMemcpyAsync(&total, d_total, 8, hipMemcpyDeviceToHost, s), "total") ||
            !check(hipStreamSynchronize(s), "count/scan"))
            break;
        if (total > 0xFFFFFFFFull) {
            err = "grid has more than 2^32 face references; lower the resolution";
            break;
        }
        hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(256), 0, s, out.cell_start + nc, (size_t)1,
                           (unsigned)total);
        if (!check(hipMalloc(&out.cell_faces, std::max<size_t>(1, (size_t)total) * 4), "alloc")) break;
        out.refs = (size_t)total;
        hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks_for(nc, 256)), dim3(256), 0, s, d_count, nc, 0u);
        if (nf) {
            hipLaunchKernelGGL(grid_cells_kernel<true>, dim3(blocks_for(nf, 256)), dim3(256), 0, s, g, d_xyz,
                               d_faces, nf, d_count, (const unsigned*)out.cell_start, out.cell_faces, d_flag, d_big,
                               d_nbig);
            hipLaunchKernelGGL(grid_cells_big_kernel<true>, dim3(big_blocks), dim3(256), 0, s, g, d_xyz, d_faces,
                               (const unsigned*)d_big, (const unsigned*)d_nbig, d_count,
                               (const unsigned*)out.cell_start, out.cell_faces, d_flag);
        }
        hipLaunchKernelGGL(sort_cells_kernel, dim3(blocks_for(nc, 256)), dim3(256), 0, s,
                           (const unsigned*)out.cell_start, nc, out.cell_faces);
        hipLaunchKernelGGL(dist_x_kernel, dim3(blocks_for(ny * nz, 256)), dim3(256), 0, s,
                           (const unsigned*)out.cell_start, (int)nx, ny * nz, out.dist);
        hipLaunchKernelGGL(dist_minmax_kernel, dim3(blocks_for(nz * nx, 256)), dim3(256), 0, s,
                           (const unsigned char*)out.dist, d_tmp, (int)nx, (int)ny, (int)nz, 1);
        hipLaunchKernelGGL(dist_minmax_kernel, dim3(blocks_for(ny * nx, 256)), dim3(256), 0, s,
                           (const unsigned char*)d_tmp, out.dist, (int)nx, (int)ny, (int)nz, 2);
        if (!check(hipGetLastError(), "fill/distance launch")) break;
        unsigned flag = 0;
        if (!check(hipMemcpyAsync(&flag, d_flag, 4, hipMemcpyDeviceToHost, s), "flag") ||
            !check(hipStreamSynchronize(s), "fill/distance"))
            break;
        if (flag) {
            err = "device grid build: count and fill passes disagree";
            break;
        }
        ok = true;
    } while (false);
    for (void* p : {(void*)d_xyz, (void*)d_faces, (void*)d_count, (void*)d_local, (void*)d_bsum, (void*)d_total,
                    (void*)d_flag, (void*)d_tmp, (void*)d_big, (void*)d_nbig})
        if (p) (void)hipFree(p);
    if (!ok) out.release();
    return ok;
}

}  // namespace linon
