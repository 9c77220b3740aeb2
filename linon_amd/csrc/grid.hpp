// Uniform-grid acceleration for short ray segments (host-side build).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace linon {

struct GridBuild {
    float lo[3] = {0, 0, 0};
    float hi[3] = {0, 0, 0};
    float cell_size = 1.0f;
    float margin = 0.0f;     // cells list every face within `margin` of the cell box
    int dim[3] = {1, 1, 1};
    std::vector<uint32_t> cell_start;  // ncells + 1
    std::vector<uint32_t> cell_faces;  // ascending face ids per cell
    std::vector<uint8_t> dist;         // Chebyshev distance (cells) to the nearest occupied cell
    size_t ncells() const { return (size_t)dim[0] * dim[1] * dim[2]; }
};

// Per-face (v0, v1 - v0, v2 - v0) as three float4, the layout the kernels read.
void pack_faces(const float* xyz, const uint32_t* faces, uint32_t nf, std::vector<float>& out);

// Builds the grid with at most `resolution` cells along the longest axis.  Returns false and
// sets `err` for invalid input (an index out of range, non-finite coordinates).
bool build_grid(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                GridBuild& g, std::string& err);

// Exposed for tests: Akenine-Moeller separating-axis triangle/box overlap (double precision).
bool tri_box_overlap(const double c[3], const double half[3], const double v0[3], const double v1[3],
                     const double v2[3]);

}  // namespace linon
