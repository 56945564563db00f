// pointwise.hip -- element-wise updates and FP64 reductions.
//
// Replaces the PETSc Vec operations on the path (VecAXPY, VecWAXPY, VecScale,
// VecSet, VecCopy, VecTDot, VecNorm; e.g. src/VecField.cpp:577-692,
// src/RegularizationH1.cpp:105-143) and the SL2 mismatch reduction
// (src/DistanceMeasureSL2.cpp:124-147).  Reductions accumulate in FP64 with a
// fixed block count and a fixed combination order, so results do not depend
// on scheduling.
#include "kernels.hpp"

namespace claire_amd {

namespace {

constexpr int kThreads = 256;

int blocks_for(int64_t n, int per_thread = 1) {
    const int64_t b = (n + static_cast<int64_t>(kThreads) * per_thread - 1) / (static_cast<int64_t>(kThreads) * per_thread);
    return static_cast<int>(b < 1 ? 1 : b);
}

__global__ __launch_bounds__(kThreads) void k_fill(float* __restrict__ x, float v, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) x[i] = v;
}
__global__ __launch_bounds__(kThreads) void k_copy(const float* __restrict__ x, float* __restrict__ y, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) y[i] = x[i];
}
__global__ __launch_bounds__(kThreads) void k_scale(float* __restrict__ x, float a, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) x[i] *= a;
}
__global__ __launch_bounds__(kThreads) void k_axpy(float a, const float* __restrict__ x, float* __restrict__ y, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) y[i] = a * x[i] + y[i];
}
__global__ __launch_bounds__(kThreads) void k_waxpy(float a, const float* __restrict__ x, const float* __restrict__ y,
                                                    float* __restrict__ w, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) w[i] = a * x[i] + y[i];
}
__global__ __launch_bounds__(kThreads) void k_sub(const float* __restrict__ x, const float* __restrict__ y,
                                                  float* __restrict__ z, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) z[i] = x[i] - y[i];
}

// Conjugate-gradient scalars on the device (PcgState).  One thread; the
// arithmetic is the host's: double division and square root, then the
// rounding to float that the host's update launches apply.
__global__ void k_pcg_init(PcgState* s, const double* __restrict__ slots, double tol, double bnorm, int kmax) {
    s->rz = slots[2];
    s->tol = tol;
    s->bnorm = bnorm;
    s->rnorm = bnorm;
    s->a = s->na = s->beta = 0.0f;
    s->k = 0;
    s->kmax = kmax;
    s->done = 0;
}
__device__ __forceinline__ void pcg_alpha(PcgState* s, double pw) {
    if (s->done) return;
    s->k += 1;
    if (!(pw > 0.0)) {  // indefinite operator or breakdown: stop before the update
        s->done = 1;
        return;
    }
    const double a = s->rz / pw;
    s->a = static_cast<float>(a);
    s->na = static_cast<float>(-a);
}
__device__ __forceinline__ void pcg_beta(PcgState* s, double rr, double rz_new) {
    if (s->done) return;
    const double rnorm = sqrt(rr);
    s->rnorm = rnorm;
    if (rnorm <= s->tol || !(rnorm < 1e6 * s->bnorm) || s->k >= s->kmax) {
        s->done = 1;
        return;
    }
    s->beta = static_cast<float>(rz_new / s->rz);
    s->rz = rz_new;
}
__global__ void k_pcg_alpha(PcgState* s, const double* __restrict__ slots) { pcg_alpha(s, slots[0]); }
__global__ void k_pcg_beta(PcgState* s, const double* __restrict__ slots) { pcg_beta(s, slots[1], slots[2]); }
__global__ __launch_bounds__(kThreads) void k_axpy_pcg(const PcgState* __restrict__ s, int sign,
                                                       const float* __restrict__ x, float* __restrict__ y,
                                                       int64_t n) {
    if (s->done) return;
    const float a = sign > 0 ? s->a : s->na;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) y[i] = a * x[i] + y[i];
}
// d += α·p and r −= α·w in one pass (the two updates of a CG step)
__global__ __launch_bounds__(kThreads) void k_axpy2_pcg(const PcgState* __restrict__ s, const float* __restrict__ p,
                                                        float* __restrict__ d, const float* __restrict__ w,
                                                        float* __restrict__ r, int64_t n) {
    if (s->done) return;
    const float a = s->a, na = s->na;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) {
        d[i] = a * p[i] + d[i];
        r[i] = na * w[i] + r[i];
    }
}
__global__ __launch_bounds__(kThreads) void k_xpay_pcg(const PcgState* __restrict__ s, float* __restrict__ p,
                                                       const float* __restrict__ z, int64_t n) {
    if (s->done) return;
    const float b = s->beta;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if (i < n) p[i] = b * p[i] + z[i];
}

enum RedOp { kRedDot = 0, kRedAbsMax = 1, kRedSqDist = 2, kRedMin = 3, kRedMax = 4, kRedSum = 5 };

template <int OP>
__device__ __forceinline__ double red_init() {
    if constexpr (OP == kRedMin) return INFINITY;
    if constexpr (OP == kRedMax) return -INFINITY;
    return 0.0;
}

template <int OP>
__device__ __forceinline__ double red_combine(double x, double y) {
    if constexpr (OP == kRedMin) return fmin(x, y);
    if constexpr (OP == kRedMax || OP == kRedAbsMax) return fmax(x, y);
    return x + y;
}

template <int OP>
__device__ double block_reduce(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = red_combine<OP>(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    return sh[0];
}

template <int OP>
__device__ __forceinline__ double red_step(double acc, float x, float y) {
    if constexpr (OP == kRedDot) {
        return acc + static_cast<double>(x) * static_cast<double>(y);
    } else if constexpr (OP == kRedAbsMax) {
        return fmax(acc, fabs(static_cast<double>(x)));
    } else if constexpr (OP == kRedSqDist) {
        const double d = static_cast<double>(x - y);
        return acc + d * d;
    } else {
        return red_combine<OP>(acc, static_cast<double>(x));
    }
}

// Each thread folds its elements in index order into one FP64 accumulator;
// the grid and the order are fixed, so the result does not depend on
// scheduling.  VEC = true reads 16-byte vectors (both operands 16-byte
// aligned), two of them in flight per thread; the n mod 4 tail elements go
// to the first threads of block 0.
template <int OP, bool VEC>
__global__ __launch_bounds__(kThreads) void k_reduce_partial(const float* __restrict__ a, const float* __restrict__ b,
                                                             int64_t n, double* __restrict__ partial) {
    __shared__ double sh[kThreads];
    double acc = red_init<OP>();
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    constexpr bool kB = (OP == kRedDot || OP == kRedSqDist);
    if constexpr (VEC) {
        const int64_t n4 = n >> 2;
        const float4* a4 = reinterpret_cast<const float4*>(a);
        const float4* b4 = reinterpret_cast<const float4*>(kB ? b : a);
        int64_t i = first;
        for (; i + stride < n4; i += 2 * stride) {
            const float4 x0 = a4[i], x1 = a4[i + stride];
            const float4 y0 = kB ? b4[i] : x0, y1 = kB ? b4[i + stride] : x1;
            acc = red_step<OP>(acc, x0.x, y0.x);
            acc = red_step<OP>(acc, x0.y, y0.y);
            acc = red_step<OP>(acc, x0.z, y0.z);
            acc = red_step<OP>(acc, x0.w, y0.w);
            acc = red_step<OP>(acc, x1.x, y1.x);
            acc = red_step<OP>(acc, x1.y, y1.y);
            acc = red_step<OP>(acc, x1.z, y1.z);
            acc = red_step<OP>(acc, x1.w, y1.w);
        }
        if (i < n4) {
            const float4 x0 = a4[i];
            const float4 y0 = kB ? b4[i] : x0;
            acc = red_step<OP>(acc, x0.x, y0.x);
            acc = red_step<OP>(acc, x0.y, y0.y);
            acc = red_step<OP>(acc, x0.z, y0.z);
            acc = red_step<OP>(acc, x0.w, y0.w);
        }
        const int64_t tail = 4 * n4 + threadIdx.x;
        if (blockIdx.x == 0 && tail < n) acc = red_step<OP>(acc, a[tail], kB ? b[tail] : 0.0f);
    } else {
        for (int64_t i = first; i < n; i += stride) acc = red_step<OP>(acc, a[i], kB ? b[i] : 0.0f);
    }
    const double r = block_reduce<OP>(acc, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

template <int OP>
__global__ __launch_bounds__(kThreads) void k_reduce_final(const double* __restrict__ partial, int np,
                                                           double* __restrict__ result) {
    __shared__ double sh[kThreads];
    double acc = red_init<OP>();
    for (int i = threadIdx.x; i < np; i += kThreads) acc = red_combine<OP>(acc, partial[i]);
    const double r = block_reduce<OP>(acc, sh);
    if (threadIdx.x == 0) result[0] = r;
}

// Two dot products a·b and a·c in one pass, each folded exactly as
// k_reduce_partial<kRedDot> folds it (same grid, same element order): the
// partials of a·b go to partial[0, nblocks), those of a·c after them.
template <bool VEC>
__global__ __launch_bounds__(kThreads) void k_reduce_partial_dot2(const float* __restrict__ a,
                                                                  const float* __restrict__ b,
                                                                  const float* __restrict__ c, int64_t n,
                                                                  double* __restrict__ partial) {
    __shared__ double sh[kThreads];
    double ab = 0.0, ac = 0.0;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x;
    if constexpr (VEC) {
        const int64_t n4 = n >> 2;
        const float4* a4 = reinterpret_cast<const float4*>(a);
        const float4* b4 = reinterpret_cast<const float4*>(b);
        const float4* c4 = reinterpret_cast<const float4*>(c);
        auto fold4 = [&](const float4& x, const float4& y, const float4& z) {
            ab = red_step<kRedDot>(ab, x.x, y.x);
            ab = red_step<kRedDot>(ab, x.y, y.y);
            ab = red_step<kRedDot>(ab, x.z, y.z);
            ab = red_step<kRedDot>(ab, x.w, y.w);
            ac = red_step<kRedDot>(ac, x.x, z.x);
            ac = red_step<kRedDot>(ac, x.y, z.y);
            ac = red_step<kRedDot>(ac, x.z, z.z);
            ac = red_step<kRedDot>(ac, x.w, z.w);
        };
        int64_t i = first;
        for (; i + stride < n4; i += 2 * stride) {
            const float4 x0 = a4[i], x1 = a4[i + stride];
            const float4 y0 = b4[i], y1 = b4[i + stride];
            const float4 z0 = c4[i], z1 = c4[i + stride];
            fold4(x0, y0, z0);
            fold4(x1, y1, z1);
        }
        if (i < n4) fold4(a4[i], b4[i], c4[i]);
        const int64_t tail = 4 * n4 + threadIdx.x;
        if (blockIdx.x == 0 && tail < n) {
            ab = red_step<kRedDot>(ab, a[tail], b[tail]);
            ac = red_step<kRedDot>(ac, a[tail], c[tail]);
        }
    } else {
        for (int64_t i = first; i < n; i += stride) {
            ab = red_step<kRedDot>(ab, a[i], b[i]);
            ac = red_step<kRedDot>(ac, a[i], c[i]);
        }
    }
    const double rab = block_reduce<kRedDot>(ab, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = rab;
    __syncthreads();
    const double rac = block_reduce<kRedDot>(ac, sh);
    if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = rac;
}

// Final sums of the CG products, then the CG scalar step on thread 0:
//   NSUM = 1: slots[0] = pᵀw, then α (as k_pcg_alpha);
//   NSUM = 2: slots[1] = rᵀr, slots[2] = rᵀz, then the stopping test and β
//             (as k_pcg_beta).
template <int NSUM>
__global__ __launch_bounds__(kThreads) void k_reduce_final_pcg(const double* __restrict__ partial, int np,
                                                               double* __restrict__ slots, PcgState* s) {
    __shared__ double sh[kThreads];
    double sums[NSUM];
#pragma unroll
    for (int k = 0; k < NSUM; ++k) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < np; i += kThreads) acc = red_combine<kRedDot>(acc, partial[k * np + i]);
        sums[k] = block_reduce<kRedDot>(acc, sh);
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    if constexpr (NSUM == 1) {
        slots[0] = sums[0];
        pcg_alpha(s, sums[0]);
    } else {
        slots[1] = sums[0];
        slots[2] = sums[1];
        pcg_beta(s, sums[0], sums[1]);
    }
}

template <int OP>
void reduce(const float* a, const float* b, int64_t n, double* partial, double* result, hipStream_t st) {
    const uintptr_t align = reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b ? b : a);
    if ((align & 15u) == 0)
        hipLaunchKernelGGL((k_reduce_partial<OP, true>), dim3(kReduceBlocks), dim3(kThreads), 0, st, a, b, n, partial);
    else
        hipLaunchKernelGGL((k_reduce_partial<OP, false>), dim3(kReduceBlocks), dim3(kThreads), 0, st, a, b, n,
                           partial);
    hipLaunchKernelGGL(k_reduce_final<OP>, dim3(1), dim3(kThreads), 0, st, partial, kReduceBlocks, result);
}

}  // namespace

void launch_fill(float* x, float value, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_fill, dim3(blocks_for(n)), dim3(kThreads), 0, st, x, value, n);
}
void launch_copy(const float* x, float* y, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_copy, dim3(blocks_for(n)), dim3(kThreads), 0, st, x, y, n);
}
void launch_scale(float* x, float alpha, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_scale, dim3(blocks_for(n)), dim3(kThreads), 0, st, x, alpha, n);
}
void launch_axpy(float alpha, const float* x, float* y, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_axpy, dim3(blocks_for(n)), dim3(kThreads), 0, st, alpha, x, y, n);
}
void launch_waxpy(float alpha, const float* x, const float* y, float* w, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_waxpy, dim3(blocks_for(n)), dim3(kThreads), 0, st, alpha, x, y, w, n);
}
void launch_sub(const float* x, const float* y, float* z, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_sub, dim3(blocks_for(n)), dim3(kThreads), 0, st, x, y, z, n);
}
void launch_pcg_init(PcgState* s, const double* slots, double tol, double bnorm, int kmax, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg_init, dim3(1), dim3(1), 0, st, s, slots, tol, bnorm, kmax);
}
void launch_pcg_alpha(PcgState* s, const double* slots, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg_alpha, dim3(1), dim3(1), 0, st, s, slots);
}
void launch_pcg_beta(PcgState* s, const double* slots, hipStream_t st) {
    hipLaunchKernelGGL(k_pcg_beta, dim3(1), dim3(1), 0, st, s, slots);
}
void launch_axpy_pcg(const PcgState* s, int sign, const float* x, float* y, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_axpy_pcg, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, sign, x, y, n);
}
void launch_axpy2_pcg(const PcgState* s, const float* p, float* d, const float* w, float* r, int64_t n,
                      hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_axpy2_pcg, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, p, d, w, r, n);
}
void launch_dot_pcg_alpha(const float* p, const float* w, int64_t n, double* partial, double* slots, PcgState* s,
                          hipStream_t st) {
    const uintptr_t align = reinterpret_cast<uintptr_t>(p) | reinterpret_cast<uintptr_t>(w);
    if ((align & 15u) == 0)
        hipLaunchKernelGGL((k_reduce_partial<kRedDot, true>), dim3(kReduceBlocks), dim3(kThreads), 0, st, p, w, n,
                           partial);
    else
        hipLaunchKernelGGL((k_reduce_partial<kRedDot, false>), dim3(kReduceBlocks), dim3(kThreads), 0, st, p, w, n,
                           partial);
    hipLaunchKernelGGL(k_reduce_final_pcg<1>, dim3(1), dim3(kThreads), 0, st, partial, kReduceBlocks, slots, s);
}
void launch_dots_pcg_beta(const float* r, const float* z, int64_t n, double* partial, double* slots, PcgState* s,
                          hipStream_t st) {
    const uintptr_t align = reinterpret_cast<uintptr_t>(r) | reinterpret_cast<uintptr_t>(z);
    if ((align & 15u) == 0)
        hipLaunchKernelGGL(k_reduce_partial_dot2<true>, dim3(kReduceBlocks), dim3(kThreads), 0, st, r, r, z, n,
                           partial);
    else
        hipLaunchKernelGGL(k_reduce_partial_dot2<false>, dim3(kReduceBlocks), dim3(kThreads), 0, st, r, r, z, n,
                           partial);
    hipLaunchKernelGGL(k_reduce_final_pcg<2>, dim3(1), dim3(kThreads), 0, st, partial, kReduceBlocks, slots, s);
}
void launch_xpay_pcg(const PcgState* s, float* p, const float* z, int64_t n, hipStream_t st) {
    if (n > 0) hipLaunchKernelGGL(k_xpay_pcg, dim3(blocks_for(n)), dim3(kThreads), 0, st, s, p, z, n);
}
void launch_dot(const float* a, const float* b, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedDot>(a, b, n, partial, result, st);
}
void launch_absmax(const float* a, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedAbsMax>(a, nullptr, n, partial, result, st);
}
void launch_sqdist(const float* a, const float* b, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedSqDist>(a, b, n, partial, result, st);
}
void launch_min(const float* a, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedMin>(a, nullptr, n, partial, result, st);
}
void launch_max(const float* a, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedMax>(a, nullptr, n, partial, result, st);
}
void launch_sum(const float* a, int64_t n, double* partial, double* result, hipStream_t st) {
    reduce<kRedSum>(a, nullptr, n, partial, result, st);
}

}  // namespace claire_amd
