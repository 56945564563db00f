// interp.hip -- semi-Lagrangian gathers for gfx950.
//
// Replaces Interp3_Plan::scatter/interpolate + optimized_interp3_ghost_xyz_p
// (src/Interp3_Plan.cpp:294-845, src/interp3.cpp:2492-2719) and the ghost
// exchange accfft_get_ghost_xyz (src/ghost.cpp:555-612) on one GPU:
//   * the periodic wrap of the query point is the reference's
//     (src/Interp3_Plan.cpp:340-362): q ∈ (−1/N, 1);
//   * index space u = q·N + 3 with the ghost width 3 of the reference
//     (rescale_xyz, src/interp3.cpp:44-47), stencil base floor(u) − 1 and
//     weights c_k·Π_{l≠k}(t − l) in the reference's FP32 product order
//     (src/interp3.cpp:2508-2533);
//   * the ghost copy is replaced by modular indexing: ghost index j is grid
//     index (j − 3) mod N, so no padded copy of the field is ever written.
// One thread computes one output point; a block is 256 consecutive points
// (whole or partial x3 rows).  The four x3-consecutive stencil values of a
// row are one unaligned 16-byte load unless the stencil wraps around the row
// end.  Blocks are remapped so that each of the 8 XCDs works on one
// contiguous slab of the grid and its L2 holds that slab's stencil rows.
// The transport updates of the solves are fused into the gathers.
#include "gather_device.hpp"

namespace claire_amd {

namespace {

using namespace gather;

constexpr int kThreads = 256;

// Block → contiguous range of points, spread so that each XCD (blocks are
// dispatched round-robin over 8 XCDs) owns one contiguous slab.
__device__ __forceinline__ int64_t point_index(int nblocks) {
    const int b = blockIdx.x;
    const int L = (nblocks % 8 == 0) ? ((b & 7) * (nblocks >> 3) + (b >> 3)) : b;
    return static_cast<int64_t>(L) * kThreads + threadIdx.x;
}

__device__ __forceinline__ void load_q3(const float* X, int64_t i, float q[3]) {
    q[0] = X[3 * i + 0];
    q[1] = X[3 * i + 1];
    q[2] = X[3 * i + 2];
}

__global__ __launch_bounds__(kThreads) void k_gather1(GatherDims g, HH hh, const float* __restrict__ f,
                                                      const float* __restrict__ X, float* __restrict__ out, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float q[3];
    load_q3(X, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {f};
    float r[1];
    eval_stencil<1>(fs, S, g, r);
    out[i] = r[0];
}

struct P3 {
    const float* in[3];
    float* out[3];
};

__global__ __launch_bounds__(kThreads) void k_gather3(GatherDims g, HH hh, P3 p, const float* __restrict__ X,
                                                      int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float q[3];
    load_q3(X, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    float r[3];
    eval_stencil<3>(p.in, S, g, r);
    p.out[0][i] = r[0];
    p.out[1][i] = r[1];
    p.out[2][i] = r[2];
}

// SemiLagrangian::ComputeTrajectoryRK2 (src/SemiLagrangian.cpp:192-288)
// with the Euler predictor kept in registers.
struct TrajParams {
    float hx[3];
    float sht;    // σ·ht
    float shth;   // σ·ht/2
};

__global__ __launch_bounds__(kThreads) void k_traj_rk2(GatherDims g, HH hh, TrajParams tp, P3 v,
                                                       float* __restrict__ X, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    const int64_t plane = static_cast<int64_t>(g.n2) * g.n3;
    const int i1 = static_cast<int>(i / plane);
    const int i2 = static_cast<int>((i / g.n3) % g.n2);
    const int i3 = static_cast<int>(i % g.n3);
    const float x[3] = {tp.hx[0] * static_cast<float>(i1), tp.hx[1] * static_cast<float>(i2),
                        tp.hx[2] * static_cast<float>(i3)};
    const double twopi = 2.0 * 3.14159265358979323846;
    float vl[3], q[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        vl[d] = v.in[d][i];
        const float t = x[d] - tp.sht * vl[d];
        q[d] = static_cast<float>(static_cast<double>(t) / twopi);
    }
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    float vx[3];
    eval_stencil<3>(v.in, S, g, vx);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float t = x[d] - tp.shth * (vx[d] + vl[d]);
        X[3 * i + d] = static_cast<float>(static_cast<double>(t) / twopi);
    }
}

__global__ __launch_bounds__(kThreads) void k_adjoint_coeff(GatherDims g, HH hh, const float* __restrict__ Xa,
                                                            const float* __restrict__ divv, float ht,
                                                            float4* __restrict__ X4, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float q[3];
    load_q3(Xa, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {divv};
    float dx[1];
    eval_stencil<1>(fs, S, g, dx);
    const float d = divv[i];
    const float hthalf = 0.5f * ht;
    const float c = 1.0f + hthalf * (dx[0] + (1.0f + ht * dx[0]) * d);
    X4[i] = make_float4(q[0], q[1], q[2], c);
}

__global__ __launch_bounds__(kThreads) void k_inc_state_step(GatherDims g, HH hh, const float* __restrict__ Xs,
                                                             const float* __restrict__ mt, const float* __restrict__ sj,
                                                             float* __restrict__ out, float sign, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float q[3];
    load_q3(Xs, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {mt};
    float r[1];
    eval_stencil<1>(fs, S, g, r);
    out[i] = sign * (r[0] - sj[i]);
}

__global__ __launch_bounds__(kThreads) void k_adjoint_step(GatherDims g, HH hh, const float4* __restrict__ X4,
                                                           const float* __restrict__ lam, float* __restrict__ out,
                                                           int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    const float4 a = X4[i];
    const float q[3] = {a.x, a.y, a.z};
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {lam};
    float r[1];
    eval_stencil<1>(fs, S, g, r);
    out[i] = a.w * r[0];
}

struct FinalP {
    const float* lam;  // nt fields λ^0..λ^{nt−1}
    const float* gm;   // (nt+1)·3 fields
    float* b[3];
    float* lam_out;    // optional: λ^nt
};

__global__ __launch_bounds__(kThreads) void k_adjoint_final(GatherDims g, HH hh, int nt, float ht, float hd,
                                                            const float4* __restrict__ X4, FinalP p, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    const int64_t nl = g.nl;
    const float4 a = X4[i];
    const float q[3] = {a.x, a.y, a.z};
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {p.lam + static_cast<int64_t>(nt - 1) * nl};
    float r[1];
    eval_stencil<1>(fs, S, g, r);
    const float lam_nt = a.w * r[0];
    if (p.lam_out) p.lam_out[i] = lam_nt;
    // body force, accumulated in the reference's order (src/CLAIRE.cpp:3940-3999)
    float b0 = 0.f, b1 = 0.f, b2 = 0.f;
    for (int j = 0; j < nt; ++j) {
        const float w = (j == 0) ? 0.5f * ht : ht;
        const float l = p.lam[static_cast<int64_t>(j) * nl + i];
        const float* gmj = p.gm + static_cast<int64_t>(3 * (nt - j)) * nl;
        b0 += w * gmj[i] * l;
        b1 += w * gmj[nl + i] * l;
        b2 += w * gmj[2 * nl + i] * l;
    }
    const float wl = 0.5f * ht;
    b0 += wl * p.gm[i] * lam_nt;
    b1 += wl * p.gm[nl + i] * lam_nt;
    b2 += wl * p.gm[2 * nl + i] * lam_nt;
    p.b[0][i] = b0 * hd;
    p.b[1][i] = b1 * hd;
    p.b[2][i] = b2 * hd;
}

__global__ __launch_bounds__(kThreads) void k_body_force_static(int64_t nl, float hd, const float* __restrict__ lam,
                                                                const float* __restrict__ gm, P3 b) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i >= nl) return;
    const float l = lam[i];
    b.out[0][i] = (l * gm[i]) * hd;
    b.out[1][i] = (l * gm[nl + i]) * hd;
    b.out[2][i] = (l * gm[2 * nl + i]) * hd;
}


int nblocks_for(int64_t nl) { return static_cast<int>((nl + kThreads - 1) / kThreads); }

void check_gather_dims(const GatherDims& g) {
    CLR_REQUIRE(g.n1 >= 4 && g.n2 >= 4 && g.n3 >= 4, kErrInvalidArgument, "gather needs n >= 4 per axis");
    CLR_REQUIRE(static_cast<int64_t>(g.n1) * g.n2 * g.n3 == g.nl, kErrInvalidArgument, "inconsistent gather dims");
    CLR_REQUIRE(g.nl < (int64_t(1) << 31), kErrInvalidArgument, "grid too large for 32-bit stencil offsets");
}

}  // namespace

GatherDims gather_dims(const Dims& d) { return GatherDims{d.n[0], d.n[1], d.n[2], d.nl}; }

void launch_gather1(const GatherDims& g, const float* f, const float* X, float* out, hipStream_t st) {
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL(k_gather1, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), f, X, out, nb);
}

void launch_gather3(const GatherDims& g, const float* const f[3], const float* X, float* const out[3], hipStream_t st) {
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    P3 p{{f[0], f[1], f[2]}, {out[0], out[1], out[2]}};
    hipLaunchKernelGGL(k_gather3, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), p, X, nb);
}

void launch_trajectory_rk2(const Dims& d, const float* const v[3], float sigma, float ht, float* X, hipStream_t st) {
    const GatherDims g = gather_dims(d);
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    TrajParams tp;
    for (int a = 0; a < 3; ++a) tp.hx[a] = d.hx[a];
    const float hthalf = 0.5f * ht;
    tp.sht = sigma * ht;
    tp.shth = sigma * hthalf;
    P3 p{{v[0], v[1], v[2]}, {nullptr, nullptr, nullptr}};
    hipLaunchKernelGGL(k_traj_rk2, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), tp, p, X, nb);
}

void launch_adjoint_coeff(const GatherDims& g, const float* Xa, const float* divv, float ht, float* X4,
                          hipStream_t st) {
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL(k_adjoint_coeff, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), Xa, divv, ht,
                       reinterpret_cast<float4*>(X4), nb);
}

void launch_inc_state_step(const GatherDims& g, const float* Xs, const float* mt, const float* sj, float* out,
                           int negate, hipStream_t st) {
    check_gather_dims(g);
    CLR_REQUIRE(mt != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL(k_inc_state_step, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), Xs, mt, sj, out,
                       negate ? -1.0f : 1.0f, nb);
}

void launch_adjoint_step(const GatherDims& g, const float* X4, const float* lam, float* out, hipStream_t st) {
    check_gather_dims(g);
    CLR_REQUIRE(lam != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL(k_adjoint_step, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g),
                       reinterpret_cast<const float4*>(X4), lam, out, nb);
}

void launch_adjoint_final(const GatherDims& g, int nt, float ht, float hd, const float* X4, const float* lam,
                          const float* gm, float* const b[3], float* lam_out, hipStream_t st) {
    check_gather_dims(g);
    CLR_REQUIRE(nt >= 1, kErrInvalidArgument, "nt must be >= 1");
    const int nb = nblocks_for(g.nl);
    FinalP p{lam, gm, {b[0], b[1], b[2]}, lam_out};
    hipLaunchKernelGGL(k_adjoint_final, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), nt, ht, hd,
                       reinterpret_cast<const float4*>(X4), p, nb);
}

void launch_body_force_static(int64_t nl, float hd, const float* lam, const float* gm, float* const b[3],
                              hipStream_t st) {
    const int nb = nblocks_for(nl);
    P3 p{{nullptr, nullptr, nullptr}, {b[0], b[1], b[2]}};
    hipLaunchKernelGGL(k_body_force_static, dim3(nb), dim3(kThreads), 0, st, nl, hd, lam, gm, p);
}

}  // namespace claire_amd
