// petscbin.cpp -- PETSc binary vector files (.bin) for the C ABI.
//
// Replaces ReadWriteReg::ReadBIN / WriteBIN (src/ReadWriteReg.cpp:1476-1537),
// which call VecLoad / VecView on a PetscViewerBinary.  The file layout is
// PETSc's published binary vector format, all big-endian: the class id
// VEC_FILE_CLASSID = 1211214 and the global length as PetscInt, then the
// values as PetscScalar.  The reference's PETSc is configured
// --with-precision=single --with-64-bit-indices (deps/build_libs.sh:176,392),
// so its files carry 8-byte header words and 4-byte floats; files with
// 4-byte header words (a 32-bit-index PETSc) are read as well: the high half
// of a 64-bit class id is zero, the first word of a 32-bit one is not.  The
// values are in the grid's flat order (x3 fastest) on one rank.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"
#include "petscbin.hpp"

namespace claire_amd {
namespace petscbin {

constexpr int64_t kVecClassId = 1211214;

uint64_t load_be(const unsigned char* p, int bytes) {
    uint64_t v = 0;
    for (int i = 0; i < bytes; ++i) v = (v << 8) | p[i];
    return v;
}

void store_be(unsigned char* p, uint64_t v, int bytes) {
    for (int i = bytes - 1; i >= 0; --i, v >>= 8) p[i] = static_cast<unsigned char>(v & 0xFF);
}

struct File {
    FILE* f;
    explicit File(FILE* f_) : f(f_) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

void read(const std::string& path, int64_t* n, float* data, int64_t capacity) {
    File file(std::fopen(path.c_str(), "rb"));
    CLR_REQUIRE(file.f != nullptr, kErrInvalidArgument, "cannot open " + path);
    unsigned char head[16];
    const size_t got = std::fread(head, 1, sizeof(head), file.f);
    CLR_REQUIRE(got >= 8, kErrInvalidArgument, path + ": file too short");
    int word = 0;
    if (load_be(head, 4) == static_cast<uint64_t>(kVecClassId)) {
        word = 4;
    } else {
        CLR_REQUIRE(got == 16 && load_be(head, 8) == static_cast<uint64_t>(kVecClassId), kErrInvalidArgument,
                    path + ": not a PETSc binary vector (class id 1211214)");
        word = 8;
    }
    const uint64_t len = load_be(head + word, word);
    CLR_REQUIRE(len < (uint64_t(1) << 40), kErrInvalidArgument, path + ": implausible vector length");
    *n = static_cast<int64_t>(len);
    if (!data) return;
    CLR_REQUIRE(capacity >= *n, kErrInvalidArgument, "output buffer too small");
    CLR_REQUIRE(std::fseek(file.f, 2 * word, SEEK_SET) == 0, kErrInvalidArgument, path + ": seek failed");
    std::vector<unsigned char> buf(static_cast<size_t>(1) << 20);
    int64_t done = 0;
    while (done < *n) {
        const int64_t chunk = std::min<int64_t>(*n - done, static_cast<int64_t>(buf.size() / 4));
        const size_t r = std::fread(buf.data(), 4, static_cast<size_t>(chunk), file.f);
        CLR_REQUIRE(static_cast<int64_t>(r) == chunk, kErrInvalidArgument, path + ": file shorter than its length");
        for (int64_t i = 0; i < chunk; ++i) {
            const uint32_t bits = static_cast<uint32_t>(load_be(buf.data() + 4 * i, 4));
            std::memcpy(data + done + i, &bits, 4);
        }
        done += chunk;
    }
}

void write(const std::string& path, int64_t n, const float* data) {
    CLR_REQUIRE(n >= 0, kErrInvalidArgument, "negative vector length");
    File file(std::fopen(path.c_str(), "wb"));
    CLR_REQUIRE(file.f != nullptr, kErrInvalidArgument, "cannot create " + path);
    unsigned char head[16];
    store_be(head, static_cast<uint64_t>(kVecClassId), 8);
    store_be(head + 8, static_cast<uint64_t>(n), 8);
    bool ok = std::fwrite(head, 1, sizeof(head), file.f) == sizeof(head);
    std::vector<unsigned char> buf(static_cast<size_t>(1) << 20);
    int64_t done = 0;
    while (ok && done < n) {
        const int64_t chunk = std::min<int64_t>(n - done, static_cast<int64_t>(buf.size() / 4));
        for (int64_t i = 0; i < chunk; ++i) {
            uint32_t bits;
            std::memcpy(&bits, data + done + i, 4);
            store_be(buf.data() + 4 * i, bits, 4);
        }
        ok = std::fwrite(buf.data(), 4, static_cast<size_t>(chunk), file.f) == static_cast<size_t>(chunk);
        done += chunk;
    }
    const int rc = std::fclose(file.f);
    file.f = nullptr;
    CLR_REQUIRE(ok && rc == 0, kErrInvalidArgument, "write error in " + path);
}

}  // namespace petscbin
}  // namespace claire_amd
