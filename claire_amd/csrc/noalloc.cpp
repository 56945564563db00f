// noalloc.cpp -- the spectral gradient without device allocation.
//
// Callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc) get the context-free
// gradient grad_noalloc: the same five FFT launches as SpectralOps::Grad,
// with the three half spectra and the twiddle tables in the file-scope
// device array `scratch` instead of buffers the library allocates.
//
// Layout of `scratch` on each device:
//   [0, kTwiddleBytes)        twiddle tables of every supported axis length,
//                             written once per device at set-up and never
//                             changed afterwards;
//   [kTwiddleBytes, ...)      three half spectra of the current call, each
//                             starting on a 256-byte boundary.
// Every call on a device shares that device's spectra, so calls that
// overlap on one device (two host threads, or two streams without an
// ordering between them) race.  No part of the spectra must start at zero:
// the first pass overwrites the whole of spectrum 0 and the fused x1 pass
// the whole of spectra 0..2.
#include <mutex>

#include "kernels.hpp"

namespace claire_amd {

namespace {

// Axis lengths the FFT supports, and the twiddle-table region that holds one
// table per length (8 + 16 + ... + 512 + 48 + 96 + 150 + 192 = 1502 values).
constexpr int kSizes[] = {8, 16, 32, 64, 128, 256, 512, 48, 96, 150, 192};
constexpr int kNumSizes = sizeof(kSizes) / sizeof(kSizes[0]);
constexpr int kTwiddleValues = 8 + 16 + 32 + 64 + 128 + 256 + 512 + 48 + 96 + 150 + 192;
constexpr size_t kTwiddleBytes = 12288;
static_assert(kTwiddleValues * sizeof(c32) <= kTwiddleBytes, "twiddle region too small");

constexpr size_t kSpectrumAlign = 256;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_setup_mutex;
unsigned char* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch`, per device

size_t round_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Offset of the table of length n inside the twiddle region, in values.
int twiddle_offset(int n) {
    int off = 0;
    for (int i = 0; i < kNumSizes; ++i) {
        if (kSizes[i] == n) return off;
        off += kSizes[i];
    }
    return -1;
}

// Looks up the device address of `scratch` on the current device and writes
// the twiddle tables, once per device.  Host code uses only the returned
// address, never the array's own name.
unsigned char* scratch_on_current_device() {
    int dev = -1;
    CLR_HIP_CHECK(hipGetDevice(&dev));
    CLR_REQUIRE(dev >= 0 && dev < kMaxDevices, kErrInvalidArgument, "device index out of range");
    std::lock_guard<std::mutex> lock(g_setup_mutex);
    if (g_scratch_ptr[dev]) return g_scratch_ptr[dev];
    void* p = nullptr;
    CLR_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    CLR_REQUIRE(p != nullptr, kErrHip, "no device address for the scratch array");
    c32 tables[kTwiddleValues];
    for (int i = 0, off = 0; i < kNumSizes; off += kSizes[i], ++i) fft_twiddles(kSizes[i], tables + off);
    CLR_HIP_CHECK(hipMemcpy(p, tables, sizeof(tables), hipMemcpyHostToDevice));
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
    return g_scratch_ptr[dev];
}

}  // namespace

size_t noalloc_grad_bytes(const int nx[3]) {
    const int64_t nlc = static_cast<int64_t>(nx[0]) * nx[1] * (nx[2] / 2 + 1);
    return kTwiddleBytes + 3 * round_up(static_cast<size_t>(nlc) * sizeof(c32), kSpectrumAlign);
}

void grad_noalloc(const int nx[3], const float* f, float* g, hipStream_t st) {
    for (int a = 0; a < 3; ++a)
        CLR_REQUIRE(nx[a] >= 4 && nx[a] <= 512, kErrInvalidArgument, "grid size must be 4..512 per axis");
    const size_t need = noalloc_grad_bytes(nx);
    CLR_REQUIRE(need <= kNoAllocScratchBytes, kErrUnsupportedSize,
                "grid " + std::to_string(nx[0]) + "x" + std::to_string(nx[1]) + "x" + std::to_string(nx[2]) +
                    " needs " + std::to_string(need) + " bytes of scratch; the allocation-free gradient holds " +
                    std::to_string(kNoAllocScratchBytes) + " (three half spectra plus the twiddle tables)");
    for (int a = 0; a < 3; ++a)
        CLR_REQUIRE(fft_size_supported(nx[a]), kErrUnsupportedSize,
                    "grid size " + std::to_string(nx[a]) +
                        " not supported (powers of two 8..512, and 48, 96, 150, 192)");
    const Dims d = make_dims(nx);
    unsigned char* base = scratch_on_current_device();
    c32* tw_base = reinterpret_cast<c32*>(base);
    c32* const tw[3] = {tw_base + twiddle_offset(nx[0]), tw_base + twiddle_offset(nx[1]),
                        tw_base + twiddle_offset(nx[2])};
    const size_t spec_bytes = round_up(static_cast<size_t>(d.nlc) * sizeof(c32), kSpectrumAlign);
    c32* s[3];
    for (int i = 0; i < 3; ++i) s[i] = reinterpret_cast<c32*>(base + kTwiddleBytes + i * spec_bytes);
    const FFTPlan3D plan(d, tw);
    // the launches of SpectralOps::Grad, in the same order
    const int64_t nl = d.nl;
    float* out[3] = {g, g + nl, g + 2 * nl};
    plan.x3_r2c(f, s[0], st);
    plan.x2(s[0], -1, st);
    plan.x1_grad(s[0], s, st);
    plan.x2_n(s, 3, +1, st);
    plan.x3_c2r_n(s, out, 3, st);
}

}  // namespace claire_amd
