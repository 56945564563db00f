// gather_device.hpp -- device helpers shared by the semi-Lagrangian gathers.
//
// Query-point handling restated from the reference:
//   * periodic wrap into (−1/N, 1) (Interp3_Plan::fast_scatter, src/Interp3_Plan.cpp:340-362);
//   * index space u = q·N + 3 with ghost width 3 (rescale_xyz, src/interp3.cpp:44-47);
//     like Interp3_Plan, which wraps and rescales the query points once in
//     scatter() and reuses them in every interpolate(), a trajectory is turned
//     once into a "plan" of rescaled points u (one float4 per point);
//   * stencil base floor(u) − 1, weights c_k·Π_{l≠k}(t − l) in the FP32 product
//     order of optimized_interp3_ghost_xyz_p (src/interp3.cpp:2508-2533);
//   * ghost index j ↔ grid index (j − 3) mod N (the ghost layer of
//     accfft_get_ghost_xyz, src/ghost.cpp:555-612, as modular indexing).
#pragma once

#include "kernels.hpp"

namespace claire_amd {
namespace gather {

typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

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

struct Axis {
    int idx[4];
    float w[4];
};

// Periodic wrap into (−h, 1), hh = float(1.0/N) (src/Interp3_Plan.cpp:328-362).
__device__ __forceinline__ float wrap_q(float q, float hh) {
#pragma unroll 1
    for (int it = 0; it < 64 && q <= -hh; ++it) q += 1.0f;
#pragma unroll 1
    for (int it = 0; it < 64 && q >= 1.0f; ++it) q -= 1.0f;
    return q;
}

// Rescaled index-space coordinate u = q·N + 3 (rescale_xyz, src/interp3.cpp:44-47).
__device__ __forceinline__ float index_coord(float q, int n, float hh) {
    return wrap_q(q, hh) * static_cast<float>(n) + 3.0f;
}

// Grid index of the first stencil node before the periodic wrap: floor(u) − 4.
__device__ __forceinline__ int stencil_start(float u) { return __float2int_rd(u) - 4; }

__device__ __forceinline__ void setup_axis_u(float u, int n, Axis& A) {
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

__device__ __forceinline__ void setup_axis(float q, int n, float hh, Axis& A) {
    setup_axis_u(index_coord(q, n, hh), n, A);
}

struct Stencil {
    Axis a1, a2, a3;
};

__device__ __forceinline__ void setup_stencil(const float q[3], const GatherDims& g, const float hh[3], Stencil& S) {
    setup_axis(q[0], g.n1, hh[0], S.a1);
    setup_axis(q[1], g.n2, hh[1], S.a2);
    setup_axis(q[2], g.n3, hh[2], S.a3);
}

__device__ __forceinline__ void setup_stencil_u(const float u[3], const GatherDims& g, Stencil& S) {
    setup_axis_u(u[0], g.n1, S.a1);
    setup_axis_u(u[1], g.n2, S.a2);
    setup_axis_u(u[2], g.n3, S.a3);
}

// Σ_a w1[a] Σ_b w2[b] Σ_c w3[c] f[i1_a, i2_b, i3_c] for ND fields sharing the
// stencil, read straight from global memory.
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

// Block → contiguous range of points, spread so that each XCD (blocks are
// dispatched round-robin over 8 XCDs) owns one contiguous slab.
__device__ __forceinline__ int xcd_remap(int b, int nblocks) {
    return (nblocks % 8 == 0) ? ((b & 7) * (nblocks >> 3) + (b >> 3)) : b;
}

// ---------------------------------------------------------------------------
// Epilogues of the trajectory-based gathers: the query comes from a plan
// (u1, u2, u3, aux) and the epilogue decides what to do with f(X).
// ---------------------------------------------------------------------------
struct QPre {
    float u[3];
    float aux;   // per-point coefficient from the plan (adjoint update c)
    float aux2;  // per-point operand loaded with the query (inc-state s_j)
};

struct EpiPlain {  // out = f(X)
    const float4* U;
    float* out;
    int64_t ozstride;
    __device__ void select(int z) { out += z * ozstride; }
    __device__ QPre load(int64_t i) const {
        const float4 a = U[i];
        return {{a.x, a.y, a.z}, 0.f, 0.f};
    }
    __device__ void store(int64_t i, float v, const QPre&) const { out[i] = v; }
};

struct EpiIncStep {  // out = sign·(f(X) − s)   (src/CLAIRE.cpp:3228-3273)
    const float4* U;
    const float* s;
    float* out;
    float sign;
    __device__ void select(int) {}
    __device__ QPre load(int64_t i) const {
        const float4 a = U[i];
        return {{a.x, a.y, a.z}, 0.f, s[i]};
    }
    __device__ void store(int64_t i, float v, const QPre& p) const { out[i] = sign * (v - p.aux2); }
};

struct EpiAdjStep {  // out = c·f(X)   (src/CLAIRE.cpp:2668-2673, 3955-3962)
    const float4* U;
    float* out;
    __device__ void select(int) {}
    __device__ QPre load(int64_t i) const {
        const float4 a = U[i];
        return {{a.x, a.y, a.z}, a.w, 0.f};
    }
    __device__ void store(int64_t i, float v, const QPre& p) const { out[i] = p.aux * v; }
};

}  // namespace gather
}  // namespace claire_amd
