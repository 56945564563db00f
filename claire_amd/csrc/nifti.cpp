// nifti.cpp -- NIfTI-1 image files (.nii, .nii.gz) for the C ABI.
//
// Replaces ReadWriteReg::ReadNII / WriteNII (src/ReadWriteReg.cpp:794-1475),
// which go through nifticlib, with a reader and writer of the NIfTI-1
// single-file format written from the format's published layout: a 348-byte
// header (sizeof_hdr = 348, dim[8] at byte 40, datatype at 70, bitpix at 72,
// pixdim[8] at 76, vox_offset at 108, scl_slope/scl_inter at 112/116,
// magic "n+1" at 344), then the voxels at vox_offset, x fastest.  Either
// byte order is accepted (a swapped sizeof_hdr marks the other one).  Files
// ending in ".gz" go through zlib.
//
// Layout conventions of the reference: the image's (nx, ny, nz) become the
// grid's (n3, n2, n1) (src/ReadWriteReg.cpp:817-819), so the voxel order is
// the grid's flat order (x3 fastest) and the data are copied in file order,
// converted to float without applying scl_slope/scl_inter, like
// nifti_image_load + static_cast<ScalarType> (src/ReadWriteReg.cpp:1083-1110).
// Written images are FLOAT32 with the header of a previously read image
// when one is given (nifti_copy_nim_info, src/ReadWriteReg.cpp:1157-1170).
#include <zlib.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "nifti.hpp"

namespace claire_amd {
namespace nifti {

constexpr int kHeaderBytes = 348;
constexpr int kVoxOffset = 352;  // header + 4-byte extension flag

enum DataType {
    kUInt8 = 2, kInt16 = 4, kInt32 = 8, kFloat32 = 16, kFloat64 = 64, kInt8 = 256, kUInt16 = 512, kUInt32 = 768,
    kInt64 = 1024, kUInt64 = 1280
};

bool ends_with(const std::string& s, const std::string& suffix) {
    return s.size() >= suffix.size() && s.compare(s.size() - suffix.size(), suffix.size(), suffix) == 0;
}

// Whole-file read through zlib (gzread also reads uncompressed files).
std::vector<unsigned char> read_file(const std::string& path) {
    gzFile f = gzopen(path.c_str(), "rb");
    CLR_REQUIRE(f != nullptr, kErrInvalidArgument, "cannot open " + path);
    std::vector<unsigned char> buf;
    unsigned char chunk[1 << 16];
    for (;;) {
        const int n = gzread(f, chunk, sizeof(chunk));
        if (n < 0) {
            gzclose(f);
            throw Error(kErrInvalidArgument, "read error in " + path);
        }
        if (n == 0) break;
        buf.insert(buf.end(), chunk, chunk + n);
    }
    gzclose(f);
    return buf;
}

void write_file(const std::string& path, const std::vector<unsigned char>& data) {
    if (ends_with(path, ".gz")) {
        gzFile f = gzopen(path.c_str(), "wb6");
        CLR_REQUIRE(f != nullptr, kErrInvalidArgument, "cannot create " + path);
        size_t off = 0;
        while (off < data.size()) {
            const unsigned n = static_cast<unsigned>(std::min<size_t>(data.size() - off, 1u << 30));
            if (gzwrite(f, data.data() + off, n) != static_cast<int>(n)) {
                gzclose(f);
                throw Error(kErrInvalidArgument, "write error in " + path);
            }
            off += n;
        }
        CLR_REQUIRE(gzclose(f) == Z_OK, kErrInvalidArgument, "write error in " + path);
    } else {
        FILE* f = std::fopen(path.c_str(), "wb");
        CLR_REQUIRE(f != nullptr, kErrInvalidArgument, "cannot create " + path);
        const size_t n = std::fwrite(data.data(), 1, data.size(), f);
        const int rc = std::fclose(f);
        CLR_REQUIRE(n == data.size() && rc == 0, kErrInvalidArgument, "write error in " + path);
    }
}

template <class T>
T get(const unsigned char* p, bool swap) {
    unsigned char b[sizeof(T)];
    std::memcpy(b, p, sizeof(T));
    if (swap)
        for (size_t i = 0; i < sizeof(T) / 2; ++i) std::swap(b[i], b[sizeof(T) - 1 - i]);
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

template <class T>
void put(unsigned char* p, T v) {
    std::memcpy(p, &v, sizeof(T));
}

struct Header {
    bool swap = false;
    int dim[8] = {0};
    int datatype = 0, bitpix = 0;
    float vox_offset = 0.f;
};

Header parse(const std::vector<unsigned char>& buf, const std::string& path) {
    CLR_REQUIRE(buf.size() >= static_cast<size_t>(kHeaderBytes), kErrInvalidArgument, path + ": file too short");
    Header h;
    const int32_t sz = get<int32_t>(buf.data(), false);
    if (sz != kHeaderBytes) {
        CLR_REQUIRE(get<int32_t>(buf.data(), true) == kHeaderBytes, kErrInvalidArgument,
                    path + ": not a NIfTI-1 file (sizeof_hdr != 348)");
        h.swap = true;
    }
    CLR_REQUIRE(std::memcmp(buf.data() + 344, "n+1", 4) == 0, kErrInvalidArgument,
                path + ": only single-file NIfTI-1 (magic \"n+1\") is supported");
    for (int i = 0; i < 8; ++i) h.dim[i] = get<int16_t>(buf.data() + 40 + 2 * i, h.swap);
    h.datatype = get<int16_t>(buf.data() + 70, h.swap);
    h.bitpix = get<int16_t>(buf.data() + 72, h.swap);
    h.vox_offset = get<float>(buf.data() + 108, h.swap);
    CLR_REQUIRE(h.dim[0] >= 3 && h.dim[0] <= 7, kErrInvalidArgument, path + ": dim[0] must be 3..7");
    for (int i = 4; i <= h.dim[0]; ++i)
        CLR_REQUIRE(h.dim[i] == 1, kErrInvalidArgument, path + ": only single 3D volumes are supported");
    for (int i = 1; i <= 3; ++i) CLR_REQUIRE(h.dim[i] > 0, kErrInvalidArgument, path + ": non-positive dimension");
    return h;
}

template <class T>
void convert(const unsigned char* src, int64_t n, bool swap, float* out) {
    for (int64_t i = 0; i < n; ++i) out[i] = static_cast<float>(get<T>(src + i * sizeof(T), swap));
}

// Reads a 3D volume; nx receives (n1, n2, n3) of the grid.  With data null
// only the size (and header) is returned.
void read(const std::string& p, int nx[3], float* data, int64_t capacity, unsigned char* header) {
    const std::vector<unsigned char> buf = read_file(p);
    const Header h = parse(buf, p);
    // (nx, ny, nz) of the image -> (n3, n2, n1) of the grid
    nx[0] = h.dim[3];
    nx[1] = h.dim[2];
    nx[2] = h.dim[1];
    if (header) std::memcpy(header, buf.data(), kHeaderBytes);
    if (!data) return;
    const int64_t n = static_cast<int64_t>(nx[0]) * nx[1] * nx[2];
    CLR_REQUIRE(capacity >= n, kErrInvalidArgument, "output buffer too small");
    const int64_t off = static_cast<int64_t>(h.vox_offset);
    CLR_REQUIRE(off >= kHeaderBytes, kErrInvalidArgument, p + ": vox_offset before the end of the header");
    int bytes = 0;
    switch (h.datatype) {
        case kUInt8: case kInt8: bytes = 1; break;
        case kInt16: case kUInt16: bytes = 2; break;
        case kInt32: case kUInt32: case kFloat32: bytes = 4; break;
        case kInt64: case kUInt64: case kFloat64: bytes = 8; break;
        default: throw Error(kErrInvalidArgument, p + ": unsupported NIfTI datatype " + std::to_string(h.datatype));
    }
    CLR_REQUIRE(h.bitpix == 8 * bytes, kErrInvalidArgument, p + ": bitpix does not match the datatype");
    CLR_REQUIRE(static_cast<int64_t>(buf.size()) >= off + n * bytes, kErrInvalidArgument,
                p + ": file shorter than its voxel data");
    const unsigned char* src = buf.data() + off;
    switch (h.datatype) {
        case kUInt8: convert<uint8_t>(src, n, h.swap, data); break;
        case kInt8: convert<int8_t>(src, n, h.swap, data); break;
        case kInt16: convert<int16_t>(src, n, h.swap, data); break;
        case kUInt16: convert<uint16_t>(src, n, h.swap, data); break;
        case kInt32: convert<int32_t>(src, n, h.swap, data); break;
        case kUInt32: convert<uint32_t>(src, n, h.swap, data); break;
        case kInt64: convert<int64_t>(src, n, h.swap, data); break;
        case kUInt64: convert<uint64_t>(src, n, h.swap, data); break;
        case kFloat32: convert<float>(src, n, h.swap, data); break;
        case kFloat64: convert<double>(src, n, h.swap, data); break;
        default: throw Error(kErrInvalidArgument, p + ": unsupported NIfTI datatype " + std::to_string(h.datatype));
    }
}

// Writes a FLOAT32 volume of grid size nx = (n1, n2, n3).  header: 348 bytes
// of a previously read image (kept except for the fields set here) or null;
// spacing: grid spacing of (x1, x2, x3) for pixdim, or null.
void write(const std::string& path, const int nx[3], const float* data, const float* spacing,
           const unsigned char* header) {
    for (int i = 0; i < 3; ++i)
        CLR_REQUIRE(nx[i] > 0 && nx[i] < 32768, kErrInvalidArgument, "dimension out of the NIfTI-1 range");
    const int64_t n = static_cast<int64_t>(nx[0]) * nx[1] * nx[2];
    std::vector<unsigned char> buf(kVoxOffset + n * sizeof(float), 0);
    unsigned char* h = buf.data();
    if (header && get<int32_t>(header, false) == kHeaderBytes) {
        std::memcpy(h, header, kHeaderBytes);
    } else if (header) {
        // template in the other byte order: keep only pixdim
        for (int i = 0; i < 8; ++i) put<float>(h + 76 + 4 * i, get<float>(header + 76 + 4 * i, true));
    } else {
        for (int i = 0; i < 4; ++i) put<float>(h + 76 + 4 * i, 1.0f);  // qfac and unit spacing
    }
    put<int32_t>(h + 0, kHeaderBytes);
    const int16_t dim[8] = {3, static_cast<int16_t>(nx[2]), static_cast<int16_t>(nx[1]), static_cast<int16_t>(nx[0]),
                            1, 1, 1, 1};
    for (int i = 0; i < 8; ++i) put<int16_t>(h + 40 + 2 * i, dim[i]);
    put<int16_t>(h + 70, static_cast<int16_t>(kFloat32));
    put<int16_t>(h + 72, 32);
    if (spacing)  // pixdim of (x, y, z) = spacing of (x3, x2, x1)
        for (int i = 0; i < 3; ++i) put<float>(h + 76 + 4 * (i + 1), spacing[2 - i]);
    put<float>(h + 108, static_cast<float>(kVoxOffset));
    put<float>(h + 112, 0.0f);  // scl_slope: no scaling
    put<float>(h + 116, 0.0f);  // scl_inter
    std::memcpy(h + 344, "n+1", 4);
    std::memset(h + kHeaderBytes, 0, kVoxOffset - kHeaderBytes);  // no extensions
    std::memcpy(buf.data() + kVoxOffset, data, n * sizeof(float));
    write_file(path, buf);
}

}  // namespace nifti
}  // namespace claire_amd
