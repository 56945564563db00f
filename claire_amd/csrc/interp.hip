// interp.hip -- semi-Lagrangian trajectory, plans and direct gathers (gfx950).
//
// Replaces Interp3_Plan::scatter/interpolate + optimized_interp3_ghost_xyz_p
// (src/Interp3_Plan.cpp:294-845, src/interp3.cpp:2492-2719) and the ghost
// exchange accfft_get_ghost_xyz (src/ghost.cpp:555-612) on one GPU.  The
// arithmetic per point is in gather_device.hpp.  Kernels here:
//   * the fused RK2 characteristic (Euler predictor kept in registers);
//   * the packed plan of a trajectory (12 B per point, gather_device.hpp),
//     together with the exact per-axis displacement bound that decides
//     whether the LDS-staged gathers can be used;
//   * direct gathers (one thread per point, stencil rows read from global
//     memory as unaligned 16-byte loads) for arbitrary query points and for
//     trajectories whose displacement exceeds the LDS window.
#include <algorithm>

#include "gather_device.hpp"

namespace claire_amd {

namespace {

using namespace gather;

constexpr int kThreads = 256;

__device__ __forceinline__ int64_t point_index(int nblocks) {
    return static_cast<int64_t>(xcd_remap(blockIdx.x, nblocks)) * kThreads + threadIdx.x;
}

// Grid indices (i1, i2, i3) of flat point i in 32-bit arithmetic (the
// gathers require nl < 2^31; 64-bit division is a long emulated sequence).
__device__ __forceinline__ void grid_index(const GatherDims& g, int64_t i, int idx[3]) {
    const uint32_t ui = static_cast<uint32_t>(i);
    const uint32_t plane = static_cast<uint32_t>(g.n2) * static_cast<uint32_t>(g.n3);
    const uint32_t i1 = ui / plane;
    const uint32_t r = ui - i1 * plane;
    const uint32_t i2 = r / static_cast<uint32_t>(g.n3);
    idx[0] = static_cast<int>(i1);
    idx[1] = static_cast<int>(i2);
    idx[2] = static_cast<int>(r - i2 * static_cast<uint32_t>(g.n3));
}

// Grid indices of a point index that advances by a fixed stride (the
// bounded-grid loops of the plan kernels): the divisions happen once per
// thread, and each step adds the stride's (i1, i2, i3) parts with carries.
struct GridWalk {
    int idx[3];
    int s[3];
    __device__ GridWalk(const GatherDims& g, int64_t i0, int64_t stride) {
        grid_index(g, i0, idx);
        grid_index(g, stride, s);  // stride = s0·n2·n3 + s1·n3 + s2, s1 < n2, s2 < n3
    }
    __device__ void advance(const GatherDims& g) {
        idx[2] += s[2];
        int c = idx[2] >= g.n3;
        if (c) idx[2] -= g.n3;
        idx[1] += s[1] + c;
        c = idx[1] >= g.n2;
        if (c) idx[1] -= g.n2;
        idx[0] += s[0] + c;
    }
};

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
    float sht;   // σ·ht
    float shth;  // σ·ht/2
};

__global__ __launch_bounds__(kThreads) void k_traj_rk2(GatherDims g, HH hh, TrajParams tp, P3 v,
                                                       float* __restrict__ X, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    int idx[3];
    grid_index(g, i, idx);
    const float x[3] = {tp.hx[0] * static_cast<float>(idx[0]), tp.hx[1] * static_cast<float>(idx[1]),
                        tp.hx[2] * static_cast<float>(idx[2])};
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

// SemiLagrangian::ComputeTrajectoryRK4 (src/SemiLagrangian.cpp:297-459), in
// three launches, each one gather of v at the previous stage point:
//   STAGE 1: X¹ = (x − σ·ht/2·v)/2π in registers, f = v + 2·v(X¹),
//            X² = (x − σ·ht/2·v(X¹))/2π;
//   STAGE 2: f += 2·v(X²), X³ = (x − σ·ht·v(X²))/2π;
//   STAGE 3: f += v(X³), X = (x − σ·(ht/6)·f)/2π.
// f (3 floats per point, interleaved like X) and the stage points live in F
// and X between launches.  Precision follows the reference: f += 2.0·vX in
// double, the last update and every division by 2π in double.
struct Rk4Params {
    float hx[3];
    float shth;   // σ·ht/2
    float sht;    // σ·ht
    double s6;    // σ·(ht/6.0)
};

__device__ __forceinline__ void grid_point(const GatherDims& g, const float hx[3], int64_t i, float x[3]) {
    int idx[3];
    grid_index(g, i, idx);
    x[0] = hx[0] * static_cast<float>(idx[0]);
    x[1] = hx[1] * static_cast<float>(idx[1]);
    x[2] = hx[2] * static_cast<float>(idx[2]);
}

template <int STAGE>
__global__ __launch_bounds__(kThreads) void k_traj_rk4(GatherDims g, HH hh, Rk4Params tp, P3 v,
                                                       float* __restrict__ X, float* __restrict__ F, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float x[3];
    grid_point(g, tp.hx, i, x);
    const double twopi = 2.0 * 3.14159265358979323846;
    float vl[3], q[3];
    if constexpr (STAGE == 1) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            vl[d] = v.in[d][i];
            const float t = x[d] - tp.shth * vl[d];
            q[d] = static_cast<float>(static_cast<double>(t) / twopi);
        }
    } else {
        load_q3(X, i, q);
    }
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    float vx[3];
    eval_stencil<3>(v.in, S, g, vx);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if constexpr (STAGE == 1) {
            F[3 * i + d] = static_cast<float>(static_cast<double>(vl[d]) + 2.0 * static_cast<double>(vx[d]));
            const float t = x[d] - tp.shth * vx[d];
            X[3 * i + d] = static_cast<float>(static_cast<double>(t) / twopi);
        } else if constexpr (STAGE == 2) {
            F[3 * i + d] = static_cast<float>(static_cast<double>(F[3 * i + d]) + 2.0 * static_cast<double>(vx[d]));
            const float t = x[d] - tp.sht * vx[d];
            X[3 * i + d] = static_cast<float>(static_cast<double>(t) / twopi);
        } else {
            const float f = F[3 * i + d] + vx[d];
            X[3 * i + d] = static_cast<float>((static_cast<double>(x[d]) - tp.s6 * static_cast<double>(f)) / twopi);
        }
    }
}

// Block maximum of `need` into *dmax (one atomic per block).
__device__ __forceinline__ void block_max_to(int need, int* sh, int* dmax) {
    sh[threadIdx.x] = need;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(dmax, sh[0]);
}

// Encodes the three axes of foot point q at grid index idx; returns the
// displacement bound this point needs (1 << 20 when not encodable).
__device__ __forceinline__ int encode_point(const GatherDims& g, const HH& hh, const float q[3], const int idx[3],
                                            PlanPt& pt) {
    const int n[3] = {g.n1, g.n2, g.n3};
    int need = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float u = index_coord(q[a], n[a], hh.h[a]);
        int lo = 0;
        if (!plan_encode(u, idx[a], n[a], pt.w[a], lo)) {
            need = 1 << 20;
            pt.w[a] = 0u;  // t = 1, o = 0: never read
            continue;
        }
        need = max(need, max(-lo - 1, lo + 1));
    }
    return need;
}

// Packed plan of a trajectory (gather_device.hpp) and the smallest D such
// that every stencil lies within [i − D − 1, i + D + 2] per axis (the window
// of the LDS-staged gathers).  Points that cannot be encoded (non-finite
// coordinates) set D to 1 << 20, which routes every gather along this
// trajectory to the direct kernels on the raw foot points.  A bounded grid
// strides over the points, so that the blocks' maxima meet in few atomics
// (one per block: tens of thousands on one address serialize).
constexpr int kPlanBlocks = 2048;

__global__ __launch_bounds__(kThreads) void k_make_plan(GatherDims g, HH hh, const float* __restrict__ X,
                                                        PlanPt* __restrict__ P, int* __restrict__ dmax) {
    __shared__ int sh[kThreads];
    int need = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    GridWalk gw(g, i0, stride);
    for (int64_t i = i0; i < g.nl; i += stride, gw.advance(g)) {
        float q[3];
        load_q3(X, i, q);
        PlanPt pt;
        need = max(need, encode_point(g, hh, q, gw.idx, pt));
        P[i] = pt;
    }
    block_max_to(need, sh, dmax);
}

__global__ void k_zero_int(int* p) { *p = 0; }

// Plan of the RK2 Euler predictor X̃ = (x − σ·ht·v)/2π, with the same float
// and double steps as k_traj_rk2 (bounded grid, like k_make_plan).
__global__ __launch_bounds__(kThreads) void k_predictor_plan(GatherDims g, HH hh, TrajParams tp, P3 v,
                                                             PlanPt* __restrict__ P, int* __restrict__ dmax) {
    __shared__ int sh[kThreads];
    int need = 0;
    const double twopi = 2.0 * 3.14159265358979323846;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    GridWalk gw(g, i0, stride);
    for (int64_t i = i0; i < g.nl; i += stride, gw.advance(g)) {
        float q[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float x = tp.hx[d] * static_cast<float>(gw.idx[d]);
            const float t = x - tp.sht * v.in[d][i];
            q[d] = static_cast<float>(static_cast<double>(t) / twopi);
        }
        PlanPt pt;
        need = max(need, encode_point(g, hh, q, gw.idx, pt));
        P[i] = pt;
    }
    block_max_to(need, sh, dmax);
}

// Plan of foot points given per component (W + d·nl), which are also
// written interleaved into X (3 floats per point) unless X is null.
__global__ __launch_bounds__(kThreads) void k_make_plan_soa(GatherDims g, HH hh, const float* __restrict__ W,
                                                            float* __restrict__ X, PlanPt* __restrict__ P,
                                                            int* __restrict__ dmax) {
    __shared__ int sh[kThreads];
    int need = 0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    GridWalk gw(g, i0, stride);
    for (int64_t i = i0; i < g.nl; i += stride, gw.advance(g)) {
        const float q[3] = {W[i], W[g.nl + i], W[2 * g.nl + i]};
        if (X) {
            X[3 * i + 0] = q[0];
            X[3 * i + 1] = q[1];
            X[3 * i + 2] = q[2];
        }
        PlanPt pt;
        need = max(need, encode_point(g, hh, q, gw.idx, pt));
        P[i] = pt;
    }
    block_max_to(need, sh, dmax);
}

// Update coefficient c = 1 + ht/2·(α·dX + α·(1 + ht·α·dX)·d), d = ∇·v,
// dX = d(X).  With α = 1 on the adjoint trajectory, the SL adjoint step
// λ⁺ = λX + ht/2·(λX·dX + (λX + ht·λX·dX)·d) (src/CLAIRE.cpp:2668-2673,
// 3955-3962) is λ⁺ = c·λX; with α = ±1 on the state trajectory, the SL
// step of det(∇y^{±1}) (src/DeformationFields.cpp:714-721) is J⁺ = c·J(X).
__global__ __launch_bounds__(kThreads) void k_adjoint_coeff(GatherDims g, HH hh, const float* __restrict__ X,
                                                            const float* __restrict__ divv, float ht, float alpha,
                                                            float* __restrict__ C, int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    float q[3];
    load_q3(X, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {divv};
    float dx[1];
    eval_stencil<1>(fs, S, g, dx);
    const float d = divv[i];
    const float hthalf = 0.5f * ht;
    const float adx = alpha * dx[0];
    C[i] = 1.0f + hthalf * (adx + alpha * (1.0f + ht * adx) * d);
}

// Gather along a trajectory straight from the raw foot points (any
// displacement).
template <class Epi>
__global__ __launch_bounds__(kThreads) void k_gather_direct(GatherDims g, HH hh, const float* __restrict__ X,
                                                            const float* __restrict__ fbase, int64_t fzstride, Epi epi,
                                                            int nblocks) {
    const int64_t i = point_index(nblocks);
    if (i >= g.nl) return;
    const float* f = fbase + static_cast<int64_t>(blockIdx.y) * fzstride;
    epi.select(blockIdx.y);
    const float aux = epi.aux(i);
    float q[3];
    load_q3(X, i, q);
    Stencil S;
    setup_stencil(q, g, hh.h, S);
    const float* fs[1] = {f};
    float r[1];
    eval_stencil<1>(fs, S, g, r);
    epi.store(i, r[0], aux);
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

template <class Epi>
void launch_direct(const GatherDims& g, const float* X, const float* f, int64_t fz, int nz, const Epi& e,
                   hipStream_t st) {
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL((k_gather_direct<Epi>), dim3(nb, nz), dim3(kThreads), 0, st, g, make_hh(g), X, f, fz, e, nb);
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

void launch_trajectory_rk4(const Dims& d, const float* const v[3], float sigma, float ht, float* X, float* F,
                           hipStream_t st) {
    const GatherDims g = gather_dims(d);
    check_gather_dims(g);
    CLR_REQUIRE(F != nullptr && F != X, kErrInvalidArgument, "RK4 trajectory needs a separate 3-field work buffer");
    const int nb = nblocks_for(g.nl);
    Rk4Params tp;
    for (int a = 0; a < 3; ++a) tp.hx[a] = d.hx[a];
    const float hthalf = 0.5f * ht;
    tp.shth = sigma * hthalf;
    tp.sht = sigma * ht;
    tp.s6 = static_cast<double>(sigma) * (static_cast<double>(ht) / 6.0);
    P3 p{{v[0], v[1], v[2]}, {nullptr, nullptr, nullptr}};
    const HH hh = make_hh(g);
    hipLaunchKernelGGL(k_traj_rk4<1>, dim3(nb), dim3(kThreads), 0, st, g, hh, tp, p, X, F, nb);
    hipLaunchKernelGGL(k_traj_rk4<2>, dim3(nb), dim3(kThreads), 0, st, g, hh, tp, p, X, F, nb);
    hipLaunchKernelGGL(k_traj_rk4<3>, dim3(nb), dim3(kThreads), 0, st, g, hh, tp, p, X, F, nb);
}

int launch_make_plan(const GatherDims& g, const float* X, uint32_t* P, int* scratch, hipStream_t st) {
    check_gather_dims(g);
    CLR_REQUIRE(g.n1 <= kPlanMaxN && g.n2 <= kPlanMaxN && g.n3 <= kPlanMaxN, kErrUnsupportedSize,
                "trajectory plans support at most 512 points per axis");
    const int nb = std::min(nblocks_for(g.nl), kPlanBlocks);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, scratch);
    hipLaunchKernelGGL(k_make_plan, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), X, reinterpret_cast<PlanPt*>(P),
                       scratch);
    int h = 0;
    CLR_HIP_CHECK(hipMemcpyAsync(&h, scratch, sizeof(int), hipMemcpyDeviceToHost, st));
    CLR_HIP_CHECK(hipStreamSynchronize(st));
    return h;
}

int launch_rk2_predictor_plan(const Dims& d, const float* const v[3], float sigma, float ht, uint32_t* P,
                              int* scratch, hipStream_t st) {
    const GatherDims g = gather_dims(d);
    check_gather_dims(g);
    CLR_REQUIRE(g.n1 <= kPlanMaxN && g.n2 <= kPlanMaxN && g.n3 <= kPlanMaxN, kErrUnsupportedSize,
                "trajectory plans support at most 512 points per axis");
    TrajParams tp;
    for (int a = 0; a < 3; ++a) tp.hx[a] = d.hx[a];
    tp.sht = sigma * ht;
    tp.shth = sigma * (0.5f * ht);
    P3 p{{v[0], v[1], v[2]}, {nullptr, nullptr, nullptr}};
    const int nb = std::min(nblocks_for(g.nl), kPlanBlocks);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, scratch);
    hipLaunchKernelGGL(k_predictor_plan, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), tp, p,
                       reinterpret_cast<PlanPt*>(P), scratch);
    int h = 0;
    CLR_HIP_CHECK(hipMemcpyAsync(&h, scratch, sizeof(int), hipMemcpyDeviceToHost, st));
    CLR_HIP_CHECK(hipStreamSynchronize(st));
    return h;
}

int launch_make_plan_soa(const GatherDims& g, const float* W, float* X, uint32_t* P, int* scratch, hipStream_t st) {
    check_gather_dims(g);
    CLR_REQUIRE(g.n1 <= kPlanMaxN && g.n2 <= kPlanMaxN && g.n3 <= kPlanMaxN, kErrUnsupportedSize,
                "trajectory plans support at most 512 points per axis");
    const int nb = std::min(nblocks_for(g.nl), kPlanBlocks);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, scratch);
    hipLaunchKernelGGL(k_make_plan_soa, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), W, X,
                       reinterpret_cast<PlanPt*>(P), scratch);
    int h = 0;
    CLR_HIP_CHECK(hipMemcpyAsync(&h, scratch, sizeof(int), hipMemcpyDeviceToHost, st));
    CLR_HIP_CHECK(hipStreamSynchronize(st));
    return h;
}

void launch_adjoint_coeff(const GatherDims& g, const float* X, const float* divv, float ht, float alpha, float* C,
                          hipStream_t st) {
    check_gather_dims(g);
    const int nb = nblocks_for(g.nl);
    hipLaunchKernelGGL(k_adjoint_coeff, dim3(nb), dim3(kThreads), 0, st, g, make_hh(g), X, divv, ht, alpha, C, nb);
}

void launch_direct_gather(const GatherDims& g, const float* X, const float* const f[], int nf, float* const out[],
                          hipStream_t st) {
    const int64_t fz = nf > 1 ? (f[1] - f[0]) : 0;
    const int64_t oz = nf > 1 ? (out[1] - out[0]) : 0;
    for (int d = 2; d < nf; ++d)
        CLR_REQUIRE(f[d] - f[d - 1] == fz && out[d] - out[d - 1] == oz, kErrInvalidArgument,
                    "vector gather needs equally spaced components");
    launch_direct(g, X, f[0], fz, nf, EpiPlain{out[0], oz}, st);
}

void launch_direct_inc_state_step(const GatherDims& g, const float* X, const float* mt, const float* sj, float* out,
                                  int negate, hipStream_t st) {
    CLR_REQUIRE(mt != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    launch_direct(g, X, mt, 0, 1, EpiIncStep{sj, out, negate ? -1.0f : 1.0f}, st);
}

void launch_direct_adjoint_step(const GatherDims& g, const float* X, const float* C, const float* lam, float* out,
                                hipStream_t st) {
    CLR_REQUIRE(lam != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    launch_direct(g, X, lam, 0, 1, EpiAdjStep{C, out}, st);
}

void launch_body_force_static(int64_t nl, float hd, const float* lam, const float* gm, float* const b[3],
                              hipStream_t st) {
    const int nb = nblocks_for(nl);
    P3 p{{nullptr, nullptr, nullptr}, {b[0], b[1], b[2]}};
    hipLaunchKernelGGL(k_body_force_static, dim3(nb), dim3(kThreads), 0, st, nl, hd, lam, gm, p);
}

}  // namespace claire_amd
