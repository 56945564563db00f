// rk2.hip -- element-wise stages of the explicit RK2 transport solvers
// (-pdesolver rk2): the state, adjoint and GN incremental solves of
// src/CLAIRE.cpp:1995-2105, 2365-2548, 2962-3135 and 3504-3683.
//
// Each RK2 step of the reference is two spectral derivatives (grad or div,
// done by SpectralOps) with an element-wise pass after each; the passes below
// are those element-wise loops, with the reference's FP32 expressions in the
// same order, fused so that every field is read once per pass.  nc = 1.
#include "kernels.hpp"

namespace claire_amd {

namespace {

constexpr int kThreads = 256;

int blocks_for(int64_t n) {
    const int64_t b = (n + kThreads - 1) / kThreads;
    return static_cast<int>(b < 1 ? 1 : b);
}

__device__ __forceinline__ int64_t gid() { return static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; }

// −g·v, summed left to right as −g1·v1 − g2·v2 − g3·v3.
__device__ __forceinline__ float neg_dot(const Vec3Ptr& g, const Vec3Ptr& v, int64_t i) {
    return -g.p[0][i] * v.p[0][i] - g.p[1][i] * v.p[1][i] - g.p[2][i] * v.p[2][i];
}

// State, first stage (src/CLAIRE.cpp:2048-2058): rhs0 = −∇m^j·v,
// m̄ = m^j + ht·rhs0.
__global__ __launch_bounds__(kThreads) void k_state_stage1(int64_t nl, float ht, Vec3Ptr g, Vec3Ptr v,
                                                           const float* __restrict__ m, float* __restrict__ rhs0,
                                                           float* __restrict__ mbar) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float r = neg_dot(g, v, i);
    rhs0[i] = r;
    mbar[i] = m[i] + ht * r;
}

// State, second stage (src/CLAIRE.cpp:2066-2077):
// m^{j+1} = m^j + ht/2·(rhs0 + rhs1), rhs1 = −∇m̄·v.
__global__ __launch_bounds__(kThreads) void k_state_stage2(int64_t nl, float hthalf, Vec3Ptr g, Vec3Ptr v,
                                                           const float* __restrict__ m, const float* __restrict__ rhs0,
                                                           float* __restrict__ mnext) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float r1 = neg_dot(g, v, i);
    mnext[i] = m[i] + hthalf * (rhs0[i] + r1);
}

// Incremental state, first stage (src/CLAIRE.cpp:3086-3095):
// rhs0 = −∇m̃^j·v − ∇m^j·ṽ, m̃̄ = m̃^j + ht·rhs0.  With has_gmt = 0 the
// incremental state is zero (j = 0, m̃⁰ = 0) and its gradient terms vanish.
__global__ __launch_bounds__(kThreads) void k_inc_stage1(int64_t nl, float ht, int has_gmt, Vec3Ptr gmt, Vec3Ptr v,
                                                         Vec3Ptr gm, Vec3Ptr vt, const float* __restrict__ mt,
                                                         float* __restrict__ rhs0, float* __restrict__ mtbar) {
    const int64_t i = gid();
    if (i >= nl) return;
    float r;
    if (has_gmt)
        r = neg_dot(gmt, v, i) - gm.p[0][i] * vt.p[0][i] - gm.p[1][i] * vt.p[1][i] - gm.p[2][i] * vt.p[2][i];
    else
        r = neg_dot(gm, vt, i);
    rhs0[i] = r;
    mtbar[i] = (has_gmt ? mt[i] : 0.0f) + ht * r;
}

// Incremental state, second stage (src/CLAIRE.cpp:3111-3121):
// m̃^{j+1} = m̃^j + ht/2·(rhs1 + rhs0), rhs1 = −∇m̃̄·v − ∇m^{j+1}·ṽ.
// negate: store −m̃^{j+1} (the incremental adjoint's final condition).
__global__ __launch_bounds__(kThreads) void k_inc_stage2(int64_t nl, float hthalf, int has_mt, Vec3Ptr gmt, Vec3Ptr v,
                                                         Vec3Ptr gm, Vec3Ptr vt, const float* __restrict__ mt,
                                                         const float* __restrict__ rhs0, float* __restrict__ out,
                                                         int negate) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float r1 =
        neg_dot(gmt, v, i) - gm.p[0][i] * vt.p[0][i] - gm.p[1][i] * vt.p[1][i] - gm.p[2][i] * vt.p[2][i];
    const float r = (has_mt ? mt[i] : 0.0f) + hthalf * (r1 + rhs0[i]);
    out[i] = negate ? -r : r;
}

// Incremental state for v = 0 (src/CLAIRE.cpp:3036-3052): nt explicit steps
// m̃ ← m̃ − ht·(∇m⁰·ṽ) with a constant right-hand side.
__global__ __launch_bounds__(kThreads) void k_inc_static(int64_t nl, int nt, float ht, Vec3Ptr gm, Vec3Ptr vt,
                                                         float* __restrict__ out, int negate) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float s = gm.p[0][i] * vt.p[0][i] + gm.p[1][i] * vt.p[1][i] + gm.p[2][i] * vt.p[2][i];
    float mt = 0.0f;
    for (int j = 0; j < nt; ++j) mt = mt - ht * s;
    out[i] = negate ? -mt : mt;
}

// Adjoint (and GN incremental adjoint) stages, in the flux form
// −∂t λ − ∇·(λv) = 0 integrated backward (src/CLAIRE.cpp:2436-2475,
// 3567-3609): w = λ·v before the first divergence.
__global__ __launch_bounds__(kThreads) void k_adj_flux(int64_t nl, const float* __restrict__ lam, Vec3Ptr v, Vec3Out w) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float l = lam[i];
    w.p[0][i] = l * v.p[0][i];
    w.p[1][i] = l * v.p[1][i];
    w.p[2][i] = l * v.p[2][i];
}

// λ̄ = λ^j + ht·rhs0, w = v·λ̄ (before the second divergence).
__global__ __launch_bounds__(kThreads) void k_adj_mid(int64_t nl, float ht, const float* __restrict__ lam,
                                                      const float* __restrict__ rhs0, Vec3Ptr v, Vec3Out w) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float lb = lam[i] + ht * rhs0[i];
    w.p[0][i] = v.p[0][i] * lb;
    w.p[1][i] = v.p[1][i] * lb;
    w.p[2][i] = v.p[2][i] * lb;
}

// λ^{j+1} = λ^j + ht/2·(rhs0 + rhs1); with next = 1 also w = λ^{j+1}·v for
// the following step.
__global__ __launch_bounds__(kThreads) void k_adj_end(int64_t nl, float hthalf, const float* __restrict__ lam,
                                                      const float* __restrict__ rhs0, const float* __restrict__ rhs1,
                                                      float* __restrict__ lnext, int next, Vec3Ptr v, Vec3Out w) {
    const int64_t i = gid();
    if (i >= nl) return;
    const float l = lam[i] + hthalf * (rhs0[i] + rhs1[i]);
    lnext[i] = l;
    if (next) {
        w.p[0][i] = l * v.p[0][i];
        w.p[1][i] = l * v.p[1][i];
        w.p[2][i] = l * v.p[2][i];
    }
}

}  // namespace

void launch_rk2_state_stage1(int64_t nl, float ht, const Vec3Ptr& g, const Vec3Ptr& v, const float* m, float* rhs0,
                             float* mbar, hipStream_t st) {
    hipLaunchKernelGGL(k_state_stage1, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, ht, g, v, m, rhs0, mbar);
}
void launch_rk2_state_stage2(int64_t nl, float hthalf, const Vec3Ptr& g, const Vec3Ptr& v, const float* m,
                             const float* rhs0, float* mnext, hipStream_t st) {
    hipLaunchKernelGGL(k_state_stage2, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, hthalf, g, v, m, rhs0, mnext);
}
void launch_rk2_inc_stage1(int64_t nl, float ht, bool has_gmt, const Vec3Ptr& gmt, const Vec3Ptr& v,
                           const Vec3Ptr& gm, const Vec3Ptr& vt, const float* mt, float* rhs0, float* mtbar,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_inc_stage1, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, ht, has_gmt ? 1 : 0, gmt, v, gm,
                       vt, mt, rhs0, mtbar);
}
void launch_rk2_inc_stage2(int64_t nl, float hthalf, bool has_mt, const Vec3Ptr& gmt, const Vec3Ptr& v,
                           const Vec3Ptr& gm, const Vec3Ptr& vt, const float* mt, const float* rhs0, float* out,
                           bool negate, hipStream_t st) {
    hipLaunchKernelGGL(k_inc_stage2, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, hthalf, has_mt ? 1 : 0, gmt, v,
                       gm, vt, mt, rhs0, out, negate ? 1 : 0);
}
void launch_rk2_inc_static(int64_t nl, int nt, float ht, const Vec3Ptr& gm, const Vec3Ptr& vt, float* out, bool negate,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_inc_static, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, nt, ht, gm, vt, out,
                       negate ? 1 : 0);
}
void launch_rk2_adj_flux(int64_t nl, const float* lam, const Vec3Ptr& v, const Vec3Out& w, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_flux, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, lam, v, w);
}
void launch_rk2_adj_mid(int64_t nl, float ht, const float* lam, const float* rhs0, const Vec3Ptr& v, const Vec3Out& w,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_adj_mid, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, ht, lam, rhs0, v, w);
}
void launch_rk2_adj_end(int64_t nl, float hthalf, const float* lam, const float* rhs0, const float* rhs1, float* lnext,
                        bool next, const Vec3Ptr& v, const Vec3Out& w, hipStream_t st) {
    hipLaunchKernelGGL(k_adj_end, dim3(blocks_for(nl)), dim3(kThreads), 0, st, nl, hthalf, lam, rhs0, rhs1, lnext,
                       next ? 1 : 0, v, w);
}

}  // namespace claire_amd
