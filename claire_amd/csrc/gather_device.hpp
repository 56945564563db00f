// gather_device.hpp -- device helpers shared by the semi-Lagrangian gathers.
//
// Query-point handling restated from the reference:
//   * periodic wrap into (−1/N, 1) (Interp3_Plan::fast_scatter, src/Interp3_Plan.cpp:340-362);
//   * index space u = q·N + 3 with ghost width 3 (rescale_xyz, src/interp3.cpp:44-47);
//     a trajectory is turned once into a packed plan (below);
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

// Tricubic Lagrange weights at t in the FP32 product order of
// optimized_interp3_ghost_xyz_p (src/interp3.cpp:2508-2533).
__device__ __forceinline__ void lagrange_weights(float x, float w[4]) {
    const float c0 = static_cast<float>(-1.0 / 6.0), c1 = 0.5f, c2 = -0.5f, c3 = static_cast<float>(1.0 / 6.0);
    const float x0 = x, x1 = x - 1.0f, x2 = x - 2.0f, x3 = x - 3.0f;
    w[0] = ((c0 * x1) * x2) * x3;
    w[1] = ((c1 * x0) * x2) * x3;
    w[2] = ((c2 * x0) * x1) * x3;
    w[3] = ((c3 * x0) * x1) * x2;
}

__device__ __forceinline__ void setup_axis_u(float u, int n, Axis& A) {
    const int gi = __float2int_rd(u) - 1;
    lagrange_weights(u - static_cast<float>(gi), A.w);
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
// Packed plan.  Like Interp3_Plan, which wraps and rescales the query points
// once in scatter() and reuses them in every interpolate(), a trajectory is
// turned once into a plan; here one 32-bit word per axis and point:
//   bits  0..22  mantissa of t = u − (floor(u) − 1) ∈ [1, 2), the local
//                coordinate the reference's weights are evaluated at
//                (src/interp3.cpp:2508-2533).  t is exact: u − floor(u) + 1
//                is a float subtraction without rounding for u ≥ 2;
//   bits 23..31  signed offset o = (floor(u) − 4) − i of the first stencil
//                node from the point's own grid index i, taken modulo N
//                into [−N/2, N/2); nodes are (i + o + a) mod N, a = 0..3.
// Decoding costs two integer instructions per axis and reproduces the
// reference's floor/weights bit for bit.  N <= 512 keeps o in 9 bits.
// ---------------------------------------------------------------------------
constexpr int kPlanMaxN = 512;

struct __attribute__((packed, aligned(4))) PlanPt {
    uint32_t w[3];
};

__device__ __forceinline__ float plan_t(uint32_t w) { return __uint_as_float((w & 0x7FFFFFu) | 0x3F800000u); }
__device__ __forceinline__ int plan_off(uint32_t w) { return static_cast<int>(w) >> 23; }

// Encodes one axis; returns false when u is not a finite, wrapped coordinate.
__device__ __forceinline__ bool plan_encode(float u, int idx, int n, uint32_t& w, int& off) {
    if (!(u >= 0.0f && u < static_cast<float>(n + 8))) return false;
    const int fu = __float2int_rd(u);
    const float t = u - static_cast<float>(fu - 1);
    // o ∈ [−n − 3, n + 3] (0 <= fu <= n + 7, 0 <= idx < n): two rounds of a
    // conditional shift by n bring it into [−⌊n/2⌋, n − ⌊n/2⌋) for every
    // n >= 4, exactly as ((o + ⌊n/2⌋) mod n) − ⌊n/2⌋, without an integer
    // division
    int o = (fu - 4) - idx;
    const int h = n / 2;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (o < -h) o += n;
        if (o >= n - h) o -= n;
    }
    off = o;
    w = (__float_as_uint(t) & 0x7FFFFFu) | (static_cast<uint32_t>(o) << 23);
    return true;
}

// ---------------------------------------------------------------------------
// Epilogues of the trajectory-based gathers.  Direct kernels use aux(i) (the
// per-point operand of the update) and store(i, f(X), aux); the LDS kernel
// reads auxp() and writes apply(f(X), aux) to outp() itself, and kHasAux
// says whether there is an operand at all.  Epilogues with kCoord need the
// point's grid indices and provide apply_at(f(X), aux, i1, i2, i3) instead;
// kOut2 adds a second output, apply2(f(X), aux) into outp2() (LDS kernel
// only).
// ---------------------------------------------------------------------------
struct EpiPlain {  // out = f(X)
    static constexpr bool kHasAux = false;
    static constexpr bool kCoord = false;
    static constexpr bool kOut2 = false;
    float* out;
    int64_t ozstride;
    __device__ void select(int z) { out += z * ozstride; }
    __device__ const float* auxp() const { return nullptr; }
    __device__ float* outp() const { return out; }
    __device__ float apply(float v, float) const { return v; }
    __device__ float aux(int64_t) const { return 0.0f; }
    __device__ void store(int64_t i, float v, float) const { out[i] = v; }
};

struct EpiIncStep {  // out = sign·(f(X) − s)   (src/CLAIRE.cpp:3228-3273)
    static constexpr bool kHasAux = true;
    static constexpr bool kCoord = false;
    static constexpr bool kOut2 = false;
    const float* s;
    float* out;
    float sign;
    __device__ void select(int) {}
    __device__ const float* auxp() const { return s; }
    __device__ float* outp() const { return out; }
    __device__ float apply(float v, float sv) const { return sign * (v - sv); }
    __device__ float aux(int64_t i) const { return s[i]; }
    __device__ void store(int64_t i, float v, float sv) const { out[i] = sign * (v - sv); }
};

struct EpiAdjStep {  // out = c·f(X)   (src/CLAIRE.cpp:2668-2673, 3955-3962)
    static constexpr bool kHasAux = true;
    static constexpr bool kCoord = false;
    static constexpr bool kOut2 = false;
    const float* c;
    float* out;
    __device__ void select(int) {}
    __device__ const float* auxp() const { return c; }
    __device__ float* outp() const { return out; }
    __device__ float apply(float v, float cv) const { return cv * v; }
    __device__ float aux(int64_t i) const { return c[i]; }
    __device__ void store(int64_t i, float v, float cv) const { out[i] = cv * v; }
};

// Update coefficient c = 1 + ht/2·(α·dX + α·(1 + ht·α·dX)·d) from the
// gathered dX = d(X) and d = ∇·v at the point (the expression of
// k_adjoint_coeff: the SL adjoint step for α = 1 on the adjoint trajectory,
// det(∇y^{±1}) and the continuity equation for α = ±1 on the state one).
struct EpiAdjCoeff {
    static constexpr bool kHasAux = true;
    static constexpr bool kCoord = false;
    static constexpr bool kOut2 = false;
    const float* d;
    float* out;
    float ht, alpha;
    __device__ void select(int) {}
    __device__ const float* auxp() const { return d; }
    __device__ float* outp() const { return out; }
    __device__ float apply(float dx, float dv) const {
        const float hthalf = 0.5f * ht;
        const float adx = alpha * dx;
        return 1.0f + hthalf * (adx + alpha * (1.0f + ht * adx) * dv);
    }
    __device__ float aux(int64_t i) const { return d[i]; }
    __device__ void store(int64_t i, float v, float dv) const { out[i] = apply(v, dv); }
};

// Corrector of the RK2 characteristic (src/SemiLagrangian.cpp:235-278) for
// component z = blockIdx.y, gathered along the Euler predictor's plan:
// X_z = (x_z − σ·ht/2·(v_z(X̃) + v_z(x)))/2π, stored per component (out +
// z·zstride); aux is v_z at the point.  Same expressions as k_traj_rk2.
struct EpiTrajRk2 {
    static constexpr bool kHasAux = true;
    static constexpr bool kCoord = true;
    static constexpr bool kOut2 = false;
    const float* v;
    float* out;
    int64_t zstride;
    float hx0, hx1, hx2;  // grid spacings
    float shth;  // σ·ht/2
    int z;
    float hz;  // spacing of axis z, set by select
    __device__ void select(int zz) {
        z = zz;
        // a blend, not a conditional: the compiler turns a conditional choice of
        // members into an indexed load, which keeps the struct in scratch memory
        hz = hx0 * static_cast<float>(zz == 0) + hx1 * static_cast<float>(zz == 1) + hx2 * static_cast<float>(zz == 2);
        v += zz * zstride;
        out += zz * zstride;
    }
    __device__ const float* auxp() const { return v; }
    __device__ float* outp() const { return out; }
    __device__ float apply_at(float vx, float vl, int i1, int i2, int i3) const {
        const int idx = z == 0 ? i1 : (z == 1 ? i2 : i3);
        const float x = hz * static_cast<float>(idx);
        const float t = x - shth * (vx + vl);
        return static_cast<float>(static_cast<double>(t) / (2.0 * 3.14159265358979323846));
    }
};

// Stages of the RK4 characteristic (src/SemiLagrangian.cpp:297-459) for
// component z, gathered along the plan of the previous stage point, with the
// float/double steps of k_traj_rk4:
//   STAGE 1 (aux v_z):  f = v + 2·v(X¹) (double), out X² = (x − σ·ht/2·v(X¹))/2π;
//   STAGE 2 (aux f_z):  f' = f + 2·v(X²) (double), out X³ = (x − σ·ht·v(X²))/2π;
//   STAGE 3 (aux f'_z): out X = (x − σ·(ht/6)·(f' + v(X³)))/2π (double).
// out holds the next point per component (out + z·zstride), out2 = the new
// stage sum, in a buffer other than aux (stage 3 has no out2).
template <int STAGE>
struct EpiTrajRk4 {
    static constexpr bool kHasAux = true;
    static constexpr bool kCoord = true;
    static constexpr bool kOut2 = STAGE < 3;
    const float* a;  // v (stage 1) or f
    float* out;
    float* f;
    int64_t zstride;
    float hx0, hx1, hx2;  // grid spacings
    float step;  // σ·ht/2 (stage 1), σ·ht (stage 2)
    double s6;   // σ·(ht/6) (stage 3)
    int z;
    float hz;  // spacing of axis z, set by select
    __device__ void select(int zz) {
        z = zz;
        // a blend, not a conditional: the compiler turns a conditional choice of
        // members into an indexed load, which keeps the struct in scratch memory
        hz = hx0 * static_cast<float>(zz == 0) + hx1 * static_cast<float>(zz == 1) + hx2 * static_cast<float>(zz == 2);
        a += zz * zstride;
        out += zz * zstride;
        f += zz * zstride;
    }
    __device__ const float* auxp() const { return a; }
    __device__ float* outp() const { return out; }
    __device__ float* outp2() const { return f; }
    __device__ float apply_at(float vx, float av, int i1, int i2, int i3) const {
        const int idx = z == 0 ? i1 : (z == 1 ? i2 : i3);
        const float x = hz * static_cast<float>(idx);
        constexpr double twopi = 2.0 * 3.14159265358979323846;
        if constexpr (STAGE < 3) {
            const float t = x - step * vx;
            return static_cast<float>(static_cast<double>(t) / twopi);
        } else {
            const float fs = av + vx;
            return static_cast<float>((static_cast<double>(x) - s6 * static_cast<double>(fs)) / twopi);
        }
    }
    __device__ float apply2(float vx, float av) const {
        return static_cast<float>(static_cast<double>(av) + 2.0 * static_cast<double>(vx));
    }
};

}  // namespace gather
}  // namespace claire_amd
