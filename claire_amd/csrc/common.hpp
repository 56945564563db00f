// common.hpp -- shared definitions of the claire_amd HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace claire_amd {

// Error raised by host code; the C-ABI layer converts it into a status code
// plus a message (reg::ThrowError / CHKERRQ play this role in the reference,
// src/CLAIREUtils.cpp:260-309).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum Status : int {
    kOk = 0,
    kErrInvalidArgument = 1,
    kErrUnsupportedSize = 2,
    kErrHip = 3,
    kErrState = 4,
    kErrOutOfMemory = 5,
};

#define CLR_HIP_CHECK(expr)                                                                                  \
    do {                                                                                                   \
        hipError_t _e = (expr);                                                                            \
        if (_e != hipSuccess)                                                                              \
            throw ::claire_amd::Error(::claire_amd::kErrHip,                                               \
                                      std::string(#expr) + ": " + hipGetErrorString(_e));                  \
    } while (0)

#define CLR_REQUIRE(cond, code, msg)                                                                       \
    do {                                                                                                   \
        if (!(cond)) throw ::claire_amd::Error((code), (msg));                                             \
    } while (0)

// Grid sizes and the reference's derived constants (computed in FP32 exactly
// as the single-precision build does: hx = 2π/n src/RegOpt.cpp:935,
// hd = hx0·hx1·hx2 include/RegOpt.hpp:485-489, scale = 1/(n0·n1·n2)
// src/RegOpt.cpp:2505-2512).
struct Dims {
    int n[3];
    int nh;        // n[2]/2 + 1
    int64_t nl;    // real points
    int64_t nlc;   // complex half-spectrum coefficients
    float hx[3];
    float hd;
    float scale;
};

Dims make_dims(const int nx[3]);

// Regularization norms (RegNorm in the reference, include/RegOpt.hpp).
enum RegNormKind : int { kRegH1 = 0, kRegH2 = 1, kRegH1SN = 2, kRegH2SN = 3 };

// k-space operators applied in the fused x1 pass.
enum SpecOp : int {
    kOpRegApply = 0,     // first/second variation multiplier
    kOpRegInverse = 1,   // inverse
    kOpRegInvSqrt = 2,   // inverse square root
    kOpH2Functional = 3, // scale·√β0·(lapik + √β1)
    kOpScaledLaplace = 4,// scale·lapik
    kOpProjStokes = 5,   // CLAIREStokes::ApplyProjection
    kOpProjRelaxed = 6   // CLAIREDivReg::ApplyProjection
};

// Registration models (RegOpt m_RegModel).
enum RegModelKind : int { kModelCompressible = 0, kModelStokes = 1, kModelRelaxedStokes = 2 };

struct SpecParams {
    int op;
    int norm;
    float beta0, beta1;
    float hd, scale;
    float beta2;  // divergence weight (projection of the relaxed model)
};

}  // namespace claire_amd
