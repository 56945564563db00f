// nifti.hpp -- NIfTI-1 single-file images (nifti.cpp).
#pragma once

#include <cstdint>
#include <string>

namespace claire_amd {
namespace nifti {

// nx receives the grid size (n1, n2, n3) = (nz, ny, nx) of the image; data
// (capacity floats, may be null) receives the voxels in file order, which is
// the grid's flat order; header (may be null) receives the 348 header bytes.
void read(const std::string& path, int nx[3], float* data, int64_t capacity, unsigned char* header);
// FLOAT32 volume of grid size (n1, n2, n3); header and spacing may be null.
void write(const std::string& path, const int nx[3], const float* data, const float* spacing,
           const unsigned char* header);

}  // namespace nifti
}  // namespace claire_amd
