// Uniform-grid acceleration for short ray segments (host-side build).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "device/grid_sat.hpp"

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

// Grid geometry for a mesh: bounds of the referenced vertices padded by a margin, cubic cells
// with at most `resolution` along the longest axis.  Validates the input (indices in range,
// finite coordinates); returns false and sets `err` otherwise.
bool grid_geometry(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                   GridGeom& geo, std::string& err);

// Builds the grid on the host (threads).  Returns false and sets `err` for invalid input.
bool build_grid(const float* xyz, uint32_t nv, const uint32_t* faces, uint32_t nf, int resolution,
                GridBuild& g, std::string& err);

}  // namespace linon
