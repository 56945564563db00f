// petscbin.hpp -- PETSc binary vector files (petscbin.cpp).
#pragma once

#include <cstdint>
#include <string>

namespace claire_amd {
namespace petscbin {

// n receives the vector length; data (capacity floats, may be null to query
// the length) receives the values in file order.
void read(const std::string& path, int64_t* n, float* data, int64_t capacity);
// Writes n float values with a 64-bit-index header.
void write(const std::string& path, int64_t n, const float* data);

}  // namespace petscbin
}  // namespace claire_amd
