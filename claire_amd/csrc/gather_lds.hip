// gather_lds.hip -- LDS-staged semi-Lagrangian gathers for gfx950.
//
// Same arithmetic as the direct gather (query wrap, u = q·N + 3, tricubic
// Lagrange weights of src/interp3.cpp:2508-2533, modular ghost indexing), but
// the field is read through LDS:
//
//   * a block owns a 16 (x2) × 64 (x3) tile of output points and marches
//     along x1 over a chunk of 32 planes ("2.5D" streaming);
//   * LDS holds a ring of input planes, each the tile plus a halo of D+1
//     (low side) and D+2 (high side) cells in x2 and x3, where D bounds the
//     per-axis characteristic displacement in cells (D = 2 covers CFL 2);
//   * the ring has one spare slot: while the block evaluates plane x1 the
//     global loads of plane x1+D+3 are in flight and land in the spare slot
//     afterwards, so each plane costs one barrier;
//   * every input element is read from HBM/L2 once per (tile, chunk), so the
//     64 stencil reads per point hit LDS instead of the L1/TA path;
//   * a point whose stencil leaves the tile (displacement above D) falls
//     back to the direct global gather, so D only affects speed.
// Blocks are remapped so that each XCD streams one contiguous range of
// tiles and keeps the shared halo rows in its own L2.
#include "gather_device.hpp"

namespace claire_amd {

namespace {

using namespace gather;

constexpr int kLT2 = 16;        // tile extent in x2
constexpr int kLT3 = 64;        // tile extent in x3 (one wave per tile row)
constexpr int kLC1 = 32;        // planes per block
constexpr int kLThreads = 256;  // 4 waves; each thread evaluates 4 points per plane
constexpr int kPtsPerThread = kLT2 / (kLThreads / kLT3);

template <int D>
struct LdsGeo {
    static constexpr int HLO = D + 1, HHI = D + 2;
    static constexpr int E2 = kLT2 + HLO + HHI, E3 = kLT3 + HLO + HHI;
    static constexpr int P = E2 * E3;       // floats per LDS plane
    static constexpr int WIN = HLO + HHI + 1;  // planes read while evaluating one output plane
    static constexpr int W = WIN + 1;          // ring slots (one spare for the prefetch)
    static constexpr int FILL = (P + kLThreads - 1) / kLThreads;
};

__device__ __forceinline__ int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

__device__ __forceinline__ int wrap_once(int a, int n) {
    a += (a < 0) ? n : 0;
    a -= (a >= n) ? n : 0;
    return a;
}

// ---------------------------------------------------------------------------
// Epilogues: where the query point comes from and what is done with f(X).
// ---------------------------------------------------------------------------
struct EpiPlain {  // out = f(X)
    const float* X;
    float* out;
    int64_t ozstride;
    __device__ void select(int z) { out += z * ozstride; }
    __device__ float load_q(int64_t i, float q[3]) const {
        q[0] = X[3 * i];
        q[1] = X[3 * i + 1];
        q[2] = X[3 * i + 2];
        return 0.f;
    }
    __device__ void store(int64_t i, float v, float) const { out[i] = v; }
};

struct EpiIncStep {  // out = sign·(f(X) − s)
    const float* X;
    const float* s;
    float* out;
    float sign;
    __device__ void select(int) {}
    __device__ float load_q(int64_t i, float q[3]) const {
        q[0] = X[3 * i];
        q[1] = X[3 * i + 1];
        q[2] = X[3 * i + 2];
        return 0.f;
    }
    __device__ void store(int64_t i, float v, float) const { out[i] = sign * (v - s[i]); }
};

struct EpiAdjStep {  // out = c·f(X), (X, c) packed per point
    const float4* X4;
    float* out;
    __device__ void select(int) {}
    __device__ float load_q(int64_t i, float q[3]) const {
        const float4 a = X4[i];
        q[0] = a.x;
        q[1] = a.y;
        q[2] = a.z;
        return a.w;
    }
    __device__ void store(int64_t i, float v, float c) const { out[i] = c * v; }
};

// λ^nt = c·λ^{nt−1}(X) and the body force b = hd·Σ_j w_j·λ^j·∇m^{nt−j}
// (trapezoid weights w_0 = w_nt = ht/2), accumulated in the reference's order
// (src/CLAIRE.cpp:3940-3999, 2668-2709).
struct EpiAdjFinal {
    const float4* X4;
    const float* lam;  // λ^0..λ^{nt−1}
    const float* gm;   // ∇m^j, (nt+1)·3 fields
    float* b0;
    float* b1;
    float* b2;
    float* lam_out;
    int nt;
    float ht, hd;
    int64_t nl;
    __device__ void select(int) {}
    __device__ float load_q(int64_t i, float q[3]) const {
        const float4 a = X4[i];
        q[0] = a.x;
        q[1] = a.y;
        q[2] = a.z;
        return a.w;
    }
    __device__ void store(int64_t i, float v, float c) const {
        const float lam_nt = c * v;
        if (lam_out) lam_out[i] = lam_nt;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f;
        for (int j = 0; j < nt; ++j) {
            const float w = (j == 0) ? 0.5f * ht : ht;
            const float l = lam[static_cast<int64_t>(j) * nl + i];
            const float* g = gm + static_cast<int64_t>(3 * (nt - j)) * nl;
            s0 += w * g[i] * l;
            s1 += w * g[nl + i] * l;
            s2 += w * g[2 * nl + i] * l;
        }
        const float wl = 0.5f * ht;
        s0 += wl * gm[i] * lam_nt;
        s1 += wl * gm[nl + i] * lam_nt;
        s2 += wl * gm[2 * nl + i] * lam_nt;
        b0[i] = s0 * hd;
        b1[i] = s1 * hd;
        b2[i] = s2 * hd;
    }
};

template <int D, class Epi>
__global__ __launch_bounds__(kLThreads) void k_gather_lds(GatherDims g, HH hh, const float* __restrict__ fbase,
                                                          int64_t fzstride, Epi epi, int nblocks) {
    using G = LdsGeo<D>;
    __shared__ float lds[G::W * G::P];
    const int tid = threadIdx.x;
    const int nt3 = (g.n3 + kLT3 - 1) / kLT3;
    const int nt2 = (g.n2 + kLT2 - 1) / kLT2;
    const int b = blockIdx.x;
    const int L = (nblocks % 8 == 0) ? ((b & 7) * (nblocks >> 3) + (b >> 3)) : b;
    const int t3 = L % nt3;
    const int t2 = (L / nt3) % nt2;
    const int ch = L / (nt3 * nt2);
    const float* f = fbase + static_cast<int64_t>(blockIdx.y) * fzstride;
    epi.select(blockIdx.y);
    const int x1s = ch * kLC1;
    const int x1e = min(x1s + kLC1, g.n1);
    const int r2 = t2 * kLT2 - G::HLO;
    const int r3 = t3 * kLT3 - G::HLO;
    const int64_t plane = static_cast<int64_t>(g.n2) * g.n3;

    // Offsets (within a global plane) of the tile elements this thread loads.
    int foff[G::FILL];
#pragma unroll
    for (int k = 0; k < G::FILL; ++k) {
        const int e = tid + kLThreads * k;
        if (e < G::P) {
            const int e2 = e / G::E3;
            const int e3 = e - e2 * G::E3;
            foff[k] = pmod(r2 + e2, g.n2) * g.n3 + pmod(r3 + e3, g.n3);
        } else {
            foff[k] = -1;
        }
    }
    const int base_u = x1s - G::HLO;  // unwrapped index of the plane in ring slot 0

    // prologue: planes x1s−HLO … x1s+HHI
    for (int k = 0; k < G::WIN; ++k) {
        const float* src = f + pmod(base_u + k, g.n1) * plane;
        float* dst = lds + k * G::P;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q)
            if (foff[q] >= 0) dst[tid + kLThreads * q] = src[foff[q]];
    }
    __syncthreads();

    const int x3l = tid % kLT3;
    const int x2l0 = tid / kLT3;
    for (int x1 = x1s; x1 < x1e; ++x1) {
        // prefetch the plane needed by the next step
        float pre[G::FILL];
        const bool more = x1 + 1 < x1e;
        if (more) {
            const float* src = f + pmod(x1 + G::HHI + 1, g.n1) * plane;
#pragma unroll
            for (int q = 0; q < G::FILL; ++q) pre[q] = (foff[q] >= 0) ? src[foff[q]] : 0.f;
        }
        const int cs = (x1 - x1s) % G::W;  // ring slot of plane x1 − HLO
#pragma unroll
        for (int k = 0; k < kPtsPerThread; ++k) {
            const int x2 = t2 * kLT2 + x2l0 + (kLThreads / kLT3) * k;
            const int x3 = t3 * kLT3 + x3l;
            if (x2 >= g.n2 || x3 >= g.n3) continue;
            const int64_t i = (static_cast<int64_t>(x1) * g.n2 + x2) * g.n3 + x3;
            float q[3];
            const float aux = epi.load_q(i, q);
            Stencil S;
            setup_stencil(q, g, hh.h, S);
            const int d1 = wrap_once(S.a1.idx[0] - (x1 - G::HLO), g.n1);
            const int l2 = wrap_once(S.a2.idx[0] - r2, g.n2);
            const int l3 = wrap_once(S.a3.idx[0] - r3, g.n3);
            float val;
            if (d1 <= G::WIN - 4 && l2 <= G::E2 - 4 && l3 <= G::E3 - 4) {
                val = 0.f;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    int slot = cs + d1 + a;
                    slot -= (slot >= G::W) ? G::W : 0;
                    const float* p = lds + slot * G::P + l2 * G::E3 + l3;
                    float sa = 0.f;
#pragma unroll
                    for (int bb = 0; bb < 4; ++bb) {
                        const float* r = p + bb * G::E3;
                        const float rv = S.a3.w[0] * r[0] + S.a3.w[1] * r[1] + S.a3.w[2] * r[2] + S.a3.w[3] * r[3];
                        sa += S.a2.w[bb] * rv;
                    }
                    val += S.a1.w[a] * sa;
                }
            } else {
                const float* fs[1] = {f};
                float r[1];
                eval_stencil<1>(fs, S, g, r);
                val = r[0];
            }
            epi.store(i, val, aux);
        }
        if (more) {
            float* dst = lds + ((x1 - x1s + G::WIN) % G::W) * G::P;
#pragma unroll
            for (int q = 0; q < G::FILL; ++q)
                if (foff[q] >= 0) dst[tid + kLThreads * q] = pre[q];
        }
        __syncthreads();
    }
}

__global__ void k_zero_int(int* p) { *p = 0; }

// Largest |floor(δ)| over all points and axes, δ = q·N − i the characteristic
// displacement in cells (wrapped to the nearest periodic image).
__global__ __launch_bounds__(256) void k_max_displacement(GatherDims g, const float* __restrict__ X, int stride,
                                                          int* __restrict__ out) {
    __shared__ int sh[256];
    int m = 0;
    const int64_t step = static_cast<int64_t>(gridDim.x) * 256;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < g.nl; i += step) {
        const int64_t plane = static_cast<int64_t>(g.n2) * g.n3;
        const int idx[3] = {static_cast<int>(i / plane), static_cast<int>((i / g.n3) % g.n2), static_cast<int>(i % g.n3)};
        const int n[3] = {g.n1, g.n2, g.n3};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float dlt = X[stride * i + a] * static_cast<float>(n[a]) - static_cast<float>(idx[a]);
            dlt -= static_cast<float>(n[a]) * rintf(dlt / static_cast<float>(n[a]));
            int k = static_cast<int>(floorf(dlt));
            if (!(dlt == dlt) || fabsf(dlt) > 1e6f) k = 1 << 20;  // non-finite: force the fallback
            k = k < 0 ? -k : k;
            m = max(m, k);
        }
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] = max(sh[threadIdx.x], sh[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(out, sh[0]);
}

// Incremental state right-hand sides, streamed with 16-byte accesses:
//   s_j = ht/2·[(∇m^j)(X_s)·ṽ(X_s) + ∇m^{j+1}·ṽ],  j = 0..nt−1;
// writes m̃¹ = −s_0 (λ̃⁰ = s_0 when nt = 1) and s_1..s_{nt−1}.
struct RhsP {
    const float* vx[3];
    const float* vt[3];
    const float* gX;
    const float* gm;
    float* s;
    float* mt1;
};

__global__ __launch_bounds__(256) void k_inc_state_rhs(int64_t nl4, int nt, float hthalf, RhsP p) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= nl4) return;
    const int64_t nl = nl4 * 4;
    auto ld = [&](const float* a) { return reinterpret_cast<const float4*>(a)[i]; };
    const float4 x0 = ld(p.vx[0]), x1 = ld(p.vx[1]), x2 = ld(p.vx[2]);
    const float4 v0 = ld(p.vt[0]), v1 = ld(p.vt[1]), v2 = ld(p.vt[2]);
    for (int j = 0; j < nt; ++j) {
        const float* gx = p.gX + static_cast<int64_t>(3 * j) * nl;
        const float* gn = p.gm + static_cast<int64_t>(3 * (j + 1)) * nl;
        const float4 a0 = ld(gx), a1 = ld(gx + nl), a2 = ld(gx + 2 * nl);
        const float4 c0 = ld(gn), c1 = ld(gn + nl), c2 = ld(gn + 2 * nl);
        float4 s;
        s.x = hthalf * (a0.x * x0.x + a1.x * x1.x + a2.x * x2.x) + hthalf * (c0.x * v0.x + c1.x * v1.x + c2.x * v2.x);
        s.y = hthalf * (a0.y * x0.y + a1.y * x1.y + a2.y * x2.y) + hthalf * (c0.y * v0.y + c1.y * v1.y + c2.y * v2.y);
        s.z = hthalf * (a0.z * x0.z + a1.z * x1.z + a2.z * x2.z) + hthalf * (c0.z * v0.z + c1.z * v1.z + c2.z * v2.z);
        s.w = hthalf * (a0.w * x0.w + a1.w * x1.w + a2.w * x2.w) + hthalf * (c0.w * v0.w + c1.w * v1.w + c2.w * v2.w);
        if (j == 0) {
            if (nt != 1) s = make_float4(-s.x, -s.y, -s.z, -s.w);
            reinterpret_cast<float4*>(p.mt1)[i] = s;
        } else {
            reinterpret_cast<float4*>(p.s + static_cast<int64_t>(j) * nl)[i] = s;
        }
    }
}

template <int D, class Epi>
void launch_lds(const GatherDims& g, const float* f, int64_t fzstride, int nz, const Epi& epi, hipStream_t st) {
    const int nt3 = (g.n3 + kLT3 - 1) / kLT3;
    const int nt2 = (g.n2 + kLT2 - 1) / kLT2;
    const int nch = (g.n1 + kLC1 - 1) / kLC1;
    const int nb = nt3 * nt2 * nch;
    hipLaunchKernelGGL((k_gather_lds<D, Epi>), dim3(nb, nz), dim3(kLThreads), 0, st, g, make_hh(g), f, fzstride, epi, nb);
}

template <class Epi>
void dispatch(int dbound, const GatherDims& g, const float* f, int64_t fzstride, int nz, const Epi& epi,
              hipStream_t st) {
    CLR_REQUIRE(g.n1 >= 4 && g.n2 >= 4 && g.n3 >= 4, kErrInvalidArgument, "gather needs n >= 4 per axis");
    CLR_REQUIRE(g.nl < (int64_t(1) << 31), kErrInvalidArgument, "grid too large for 32-bit stencil offsets");
    if (dbound <= 2)
        launch_lds<2>(g, f, fzstride, nz, epi, st);
    else
        launch_lds<4>(g, f, fzstride, nz, epi, st);
}

}  // namespace

int lds_gather_max_displacement() { return 4; }

void launch_lds_gather(int dbound, const GatherDims& g, const float* const f[], int nf, const float* X, float* const out[],
                       hipStream_t st) {
    // components must be equally spaced (one field array with a stride)
    const int64_t fz = nf > 1 ? (f[1] - f[0]) : 0;
    const int64_t oz = nf > 1 ? (out[1] - out[0]) : 0;
    for (int d = 2; d < nf; ++d)
        CLR_REQUIRE(f[d] - f[d - 1] == fz && out[d] - out[d - 1] == oz, kErrInvalidArgument,
                    "vector gather needs equally spaced components");
    EpiPlain e{X, out[0], oz};
    dispatch(dbound, g, f[0], fz, nf, e, st);
}

void launch_lds_inc_state_step(int dbound, const GatherDims& g, const float* Xs, const float* mt, const float* sj,
                               float* out, int negate, hipStream_t st) {
    CLR_REQUIRE(mt != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    EpiIncStep e{Xs, sj, out, negate ? -1.0f : 1.0f};
    dispatch(dbound, g, mt, 0, 1, e, st);
}

void launch_lds_adjoint_step(int dbound, const GatherDims& g, const float* X4, const float* lam, float* out,
                             hipStream_t st) {
    CLR_REQUIRE(lam != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    EpiAdjStep e{reinterpret_cast<const float4*>(X4), out};
    dispatch(dbound, g, lam, 0, 1, e, st);
}

void launch_lds_adjoint_final(int dbound, const GatherDims& g, int nt, float ht, float hd, const float* X4,
                              const float* lam, const float* gm, float* const b[3], float* lam_out, hipStream_t st) {
    CLR_REQUIRE(nt >= 1, kErrInvalidArgument, "nt must be >= 1");
    EpiAdjFinal e{reinterpret_cast<const float4*>(X4), lam, gm, b[0], b[1], b[2], lam_out, nt, ht, hd, g.nl};
    dispatch(dbound, g, lam + static_cast<int64_t>(nt - 1) * g.nl, 0, 1, e, st);
}

int max_displacement(const GatherDims& g, const float* X, int stride, int* scratch, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, scratch);
    hipLaunchKernelGGL(k_max_displacement, dim3(512), dim3(256), 0, st, g, X, stride, scratch);
    int h = 0;
    CLR_HIP_CHECK(hipMemcpyAsync(&h, scratch, sizeof(int), hipMemcpyDeviceToHost, st));
    CLR_HIP_CHECK(hipStreamSynchronize(st));
    return h;
}

void launch_inc_state_rhs(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3],
                          const float* gX, const float* gm, float* s, float* mt1, hipStream_t st) {
    CLR_REQUIRE(nl % 4 == 0, kErrInvalidArgument, "grid size must be a multiple of 4 points");
    CLR_REQUIRE(nt >= 1, kErrInvalidArgument, "nt must be >= 1");
    const int64_t nl4 = nl / 4;
    RhsP p{{vx[0], vx[1], vx[2]}, {vt[0], vt[1], vt[2]}, gX, gm, s, mt1};
    const int nb = static_cast<int>((nl4 + 255) / 256);
    hipLaunchKernelGGL(k_inc_state_rhs, dim3(nb), dim3(256), 0, st, nl4, nt, 0.5f * ht, p);
}

}  // namespace claire_amd
