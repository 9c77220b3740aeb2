// Device build of the uniform grid (grid_build.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "device/grid_sat.hpp"

namespace linon {

// Device arrays of a built grid (owned by the caller once returned).
struct DeviceGrid {
    unsigned* cell_start = nullptr;  // ncells + 1
    unsigned* cell_faces = nullptr;  // refs, ascending face ids per cell
    unsigned char* dist = nullptr;   // ncells, Chebyshev cell distance to occupancy, capped at 255
    float* tris = nullptr;           // 12 floats per face: v0, v1 - v0, v2 - v0 (float4 each)
    size_t refs = 0;
    void release();
};

// Builds the grid of geometry `g` (from grid_geometry, which also validated the mesh) on the
// device: uploads the mesh, counts, scans, fills, sorts and computes the distance field on
// stream s, synchronising twice.  On failure returns false with `err` set and nothing allocated.
bool build_grid_device(const GridGeom& g, const float* h_xyz, uint32_t nv, const uint32_t* h_faces, uint32_t nf,
                       hipStream_t s, DeviceGrid& out, std::string& err);

}  // namespace linon
