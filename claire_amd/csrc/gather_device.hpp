// gather_device.hpp -- device helpers shared by the semi-Lagrangian gathers.
//
// Query-point handling restated from the reference:
//   * periodic wrap into (−1/N, 1) (Interp3_Plan::fast_scatter, src/Interp3_Plan.cpp:340-362);
//   * index space u = q·N + 3 with ghost width 3 (rescale_xyz, src/interp3.cpp:44-47);
//   * stencil base floor(u) − 1, weights c_k·Π_{l≠k}(t − l) in the FP32 product
//     order of optimized_interp3_ghost_xyz_p (src/interp3.cpp:2508-2533);
//   * ghost index j ↔ grid index (j − 3) mod N (the ghost layer of
//     accfft_get_ghost_xyz, src/ghost.cpp:555-612, as modular indexing).
#pragma once

#include "kernels.hpp"

namespace claire_amd {
namespace gather {

typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

struct Axis {
    int idx[4];
    float w[4];
};

// hh = float(1.0/N): the wrap threshold h of fast_scatter (src/Interp3_Plan.cpp:328-331).
__device__ __forceinline__ void setup_axis(float q, int n, float hh, Axis& A) {
#pragma unroll 1
    for (int it = 0; it < 64 && q <= -hh; ++it) q += 1.0f;
#pragma unroll 1
    for (int it = 0; it < 64 && q >= 1.0f; ++it) q -= 1.0f;
    const float u = q * static_cast<float>(n) + 3.0f;
    const int gi = __float2int_rd(u) - 1;
    const float x = u - static_cast<float>(gi);
    const float c0 = static_cast<float>(-1.0 / 6.0), c1 = 0.5f, c2 = -0.5f, c3 = static_cast<float>(1.0 / 6.0);
    const float x0 = x, x1 = x - 1.0f, x2 = x - 2.0f, x3 = x - 3.0f;
    A.w[0] = ((c0 * x1) * x2) * x3;
    A.w[1] = ((c1 * x0) * x2) * x3;
    A.w[2] = ((c2 * x0) * x1) * x3;
    A.w[3] = ((c3 * x0) * x1) * x2;
    int i0 = gi - 3;
    if (i0 < -2 || i0 > n) {  // only for non-finite coordinates
        i0 %= n;
        if (i0 < 0) i0 += n;
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        int i = i0 + a;
        i += (i < 0) ? n : 0;
        i -= (i >= n) ? n : 0;
        A.idx[a] = i;
    }
}

struct Stencil {
    Axis a1, a2, a3;
};

__device__ __forceinline__ void setup_stencil(const float q[3], const GatherDims& g, const float hh[3], Stencil& S) {
    setup_axis(q[0], g.n1, hh[0], S.a1);
    setup_axis(q[1], g.n2, hh[1], S.a2);
    setup_axis(q[2], g.n3, hh[2], S.a3);
}

// Σ_a w1[a] Σ_b w2[b] Σ_c w3[c] f[i1_a, i2_b, i3_c] for ND fields sharing the stencil.
template <int ND>
__device__ __forceinline__ void eval_stencil(const float* const f[ND], const Stencil& S, const GatherDims& g,
                                             float out[ND]) {
    const bool contig = S.a3.idx[0] <= g.n3 - 4;
#pragma unroll
    for (int d = 0; d < ND; ++d) out[d] = 0.0f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        float sa[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) sa[d] = 0.0f;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int off = (S.a1.idx[a] * g.n2 + S.a2.idx[b]) * g.n3;
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                float r;
                if (contig) {
                    const f4u q = *reinterpret_cast<const f4u*>(f[d] + off + S.a3.idx[0]);
                    r = S.a3.w[0] * q.x + S.a3.w[1] * q.y + S.a3.w[2] * q.z + S.a3.w[3] * q.w;
                } else {
                    const float* row = f[d] + off;
                    r = S.a3.w[0] * row[S.a3.idx[0]] + S.a3.w[1] * row[S.a3.idx[1]] + S.a3.w[2] * row[S.a3.idx[2]] +
                        S.a3.w[3] * row[S.a3.idx[3]];
                }
                sa[d] += S.a2.w[b] * r;
            }
        }
#pragma unroll
        for (int d = 0; d < ND; ++d) out[d] += S.a1.w[a] * sa[d];
    }
}

struct HH {
    float h[3];
};

__host__ __device__ inline HH make_hh(const GatherDims& g) {
    HH h;
    h.h[0] = static_cast<float>(1.0 / g.n1);
    h.h[1] = static_cast<float>(1.0 / g.n2);
    h.h[2] = static_cast<float>(1.0 / g.n3);
    return h;
}

}  // namespace gather
}  // namespace claire_amd
