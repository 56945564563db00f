// gather_lds.hip -- LDS-staged semi-Lagrangian gathers and the streaming
// transport kernels of the Hessian matvec (gfx950).
//
// Same arithmetic as the direct gather (tricubic Lagrange weights of
// src/interp3.cpp:2508-2533 on the rescaled plan points, modular ghost
// indexing), but the field is read through LDS:
//
//   * a block owns a 16 (x2) × 32 or 64 (x3) tile of output points and
//     marches along x1 over a chunk of planes ("2.5D" streaming), the chunk
//     length chosen per launch so that the blocks fill whole rounds of the
//     GPU;
//   * LDS holds a ring of input planes, each the tile plus a halo of D+1
//     (low side) and D+2 (high side) cells in x2 and x3, where D is the exact
//     per-axis stencil displacement bound of the plan (D <= 2 covers CFL 2),
//     so every stencil of the tile lies inside the ring by construction;
//   * the ring has one spare slot: while the block evaluates plane x1 the
//     global loads of plane x1+D+3 and of the next plane's query points are
//     in flight and land in the spare slot / registers afterwards, so each
//     plane costs one barrier and no exposed HBM latency;
//   * every input element is read from HBM/L2 once per (tile, chunk); the 64
//     stencil values per point come from LDS (ds_read2_b32 pairs).
// Blocks are remapped so that each XCD streams one contiguous range of
// tiles and keeps the shared halo rows in its own L2.
#include <algorithm>
#include <cstdlib>
#include <limits>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "gather_device.hpp"

namespace claire_amd {

namespace {

using namespace gather;

constexpr int kLT2 = 16;        // tile extent in x2
constexpr int kLThreads = 512;  // 8 waves; wave w covers tile rows 2w and 2w+1

// LDS plane geometry for displacement bound D.  A plane holds E2 rows of E3
// valid columns at a row pitch that is a multiple of 32 dwords: a wave reads
// two tile rows of 32 points each (one row per half-wave), so lanes whose
// stencils start on different rows or planes still hit different banks
// unless their columns collide mod 32.
//   D <= 2: a 16 × 32 tile, one point per thread, 53 KB of LDS -- three
//           blocks (24 waves) per CU;
//   D <= 4: a 16 × 64 tile, two points per thread 32 columns apart, 135 KB.
template <int D>
struct LdsGeo {
    static constexpr int LT3 = (D <= 2) ? 32 : 64;  // tile extent in x3
    static constexpr int PTS = LT3 / 32;             // points per thread and plane
    static constexpr int WAVES_PER_EU = (D <= 2) ? 6 : 2;
    static constexpr int HLO = D + 1, HHI = D + 2;
    static constexpr int E2 = kLT2 + HLO + HHI, E3 = LT3 + HLO + HHI;
    static constexpr int PITCH = (E3 + 31) / 32 * 32;
    static constexpr int P = E2 * PITCH;       // dwords per LDS plane
    static constexpr int NV = E2 * E3;         // valid elements per plane
    static constexpr int WIN = HLO + HHI + 1;  // planes read while evaluating one output plane
    static constexpr int W = WIN + 1;          // ring slots (one spare for the prefetch)
    static constexpr int FILL = (NV + kLThreads - 1) / kLThreads;
};

__device__ __forceinline__ int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}

// Raw buffer accesses: a 32-bit per-lane byte offset plus a uniform byte
// offset (the plane), against a resource that bounds the array (reads past
// it return 0, writes past it are dropped).  Word 3 of a gfx9 raw buffer
// resource: 32-bit data, no swizzle.
using rsrc_t = __amdgpu_buffer_rsrc_t;
constexpr int kBufferWord3 = 0x00020000;

__device__ __forceinline__ rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, static_cast<int>(bytes), kBufferWord3);
}
// Cache policy of the buffer loads: 0 for the field planes (neighbouring
// tiles re-read their halo rows from L2), kNonTemporal for the plan words and
// the per-point coefficients, which are read once.
constexpr int kNonTemporal = 2;
__device__ __forceinline__ float bload(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}
__device__ __forceinline__ float bload_once(rsrc_t r, uint32_t voff, uint32_t soff) {
    return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, kNonTemporal));
}
__device__ __forceinline__ void bstore(rsrc_t r, float v, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, voff, soff, 0);
}
__device__ __forceinline__ PlanPt bload_plan(rsrc_t r, uint32_t voff, uint32_t soff) {
    typedef unsigned int u3 __attribute__((ext_vector_type(3)));
    const u3 w = __builtin_amdgcn_raw_buffer_load_b96(r, voff, soff, kNonTemporal);
    PlanPt p;
    p.w[0] = w.x;
    p.w[1] = w.y;
    p.w[2] = w.z;
    return p;
}

// Keeps the compiler from splitting an LDS base address into a register part
// plus a constant too large for the ds_read2 offset fields; the stencil's row
// and column offsets then all fit the instruction's immediates.
__device__ __forceinline__ int opaque(int x) {
    asm volatile("" : "+v"(x));
    return x;
}

template <class Epi>
__device__ __forceinline__ float* out2_of(const Epi& e) {
    if constexpr (Epi::kOut2)
        return e.outp2();
    else
        return e.outp();
}

template <int D, class Epi>
__global__ __launch_bounds__(kLThreads) __attribute__((amdgpu_waves_per_eu(LdsGeo<D>::WAVES_PER_EU))) void k_gather_lds(
    GatherDims g, const PlanPt* __restrict__ plan, const float* __restrict__ fbase, int64_t fzstride, Epi epi,
    int nblocks, int chunk) {
    using G = LdsGeo<D>;
    __shared__ float lds[G::W * G::P];
    const int tid = threadIdx.x;
    const int nt3 = (g.n3 + G::LT3 - 1) / G::LT3;
    const int nt2 = (g.n2 + kLT2 - 1) / kLT2;
    // one block per (tile, chunk, field), a tile's fields adjacent (as in
    // k_gather_lds_pair)
    const int nz = static_cast<int>(gridDim.x) / nblocks;
    const int Lz = xcd_remap(blockIdx.x, nblocks * nz);
    const int z = Lz % nz;
    const int L = Lz / nz;
    const int t3 = L % nt3;
    const int t2 = (L / nt3) % nt2;
    const int ch = L / (nt3 * nt2);
    const float* f = fbase + static_cast<int64_t>(z) * fzstride;
    epi.select(z);
    // Byte sizes fit 32 bits: the host requires nl < 2^28 (12 B per point).
    const uint32_t fbytes = static_cast<uint32_t>(g.nl) * 4u;
    const rsrc_t rf = make_rsrc(f, fbytes);
    const rsrc_t rp = make_rsrc(plan, 3u * fbytes);
    const rsrc_t ra = make_rsrc(Epi::kHasAux ? epi.auxp() : f, fbytes);
    const rsrc_t ro = make_rsrc(epi.outp(), fbytes);
    const rsrc_t ro2 = make_rsrc(out2_of(epi), fbytes);
    const uint32_t plane4 = static_cast<uint32_t>(g.n2) * static_cast<uint32_t>(g.n3) * 4u;  // bytes per plane
    const int x1s = ch * chunk;
    const int x1e = min(x1s + chunk, g.n1);
    const int r2 = t2 * kLT2 - G::HLO;
    const int r3 = t3 * G::LT3 - G::HLO;

    // Offsets (within a global plane and within an LDS plane) of the tile
    // elements this thread loads; elements past the tile reload element
    // NV−1 (never stored).
    uint32_t foff[G::FILL];
    int loff[G::FILL];
#pragma unroll
    for (int k = 0; k < G::FILL; ++k) {
        const int e = min(tid + kLThreads * k, G::NV - 1);
        const int e2 = e / G::E3;
        const int e3 = e - e2 * G::E3;
        foff[k] = 4u * static_cast<uint32_t>(pmod(r2 + e2, g.n2) * g.n3 + pmod(r3 + e3, g.n3));  // bytes
        loff[k] = e2 * G::PITCH + e3;
    }

    // This thread's two output points within a plane: wave w covers tile rows
    // 2w and 2w+1, 32 columns each, and the two points are 32 columns apart.
    // Invalid points load point 0 of the plane and skip the store.  rc4 is
    // the byte offset of the point's own cell in an LDS plane.
    const int lane = tid & 63;
    const int lx2 = 2 * (tid >> 6) + (lane >> 5);
    const int x2 = t2 * kLT2 + lx2;
    uint32_t off[G::PTS];
    int rc4[G::PTS];
    bool valid[G::PTS];
    int x3s[G::PTS];
#pragma unroll
    for (int k = 0; k < G::PTS; ++k) {
        const int lx3 = (lane & 31) + 32 * k;
        const int x3 = t3 * G::LT3 + lx3;
        x3s[k] = x3;
        valid[k] = x2 < g.n2 && x3 < g.n3;
        off[k] = valid[k] ? 4u * static_cast<uint32_t>(x2 * g.n3 + x3) : 0u;  // bytes of a float
        rc4[k] = 4 * ((lx2 + G::HLO) * G::PITCH + lx3 + G::HLO);
    }

    // prologue: planes x1s−HLO … x1s+HHI into ring slots 0 … WIN−1, and the
    // query points of plane x1s
    for (int k = 0; k < G::WIN; ++k) {
        const uint32_t src = static_cast<uint32_t>(pmod(x1s - G::HLO + k, g.n1)) * plane4;
        float* dst = lds + k * G::P;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q)
            if (tid + kLThreads * q < G::NV) dst[loff[q]] = bload(rf, foff[q], src);
    }
    // Query points of plane x1 live in (qa, aa) or (qb, ab); the loop below
    // alternates the two so that the prefetch of the next plane's points
    // needs no register copies.
    PlanPt qa[G::PTS], qb[G::PTS];
    float aa[G::PTS], ab[G::PTS];
    {
        const uint32_t pb = static_cast<uint32_t>(x1s) * plane4;
#pragma unroll
        for (int k = 0; k < G::PTS; ++k) {
            qa[k] = bload_plan(rp, 3u * off[k], 3u * pb);
            aa[k] = Epi::kHasAux ? bload_once(ra, off[k], pb) : 0.0f;
        }
    }
    __syncthreads();

    constexpr int P4 = 4 * G::P;  // bytes per ring slot
    const char* ldsb = reinterpret_cast<const char*>(lds);
    int cs = 0;  // ring slot of plane x1 − HLO
    // Input planes are loaded two steps before they are landed: step x1
    // lands plane x1 + HHI + 1 (in pa or pb, loaded by step x1 − 1) and
    // loads plane x1 + HHI + 2 into the other set.
    float pa[G::FILL], pb2[G::FILL];
    {
        const uint32_t src = static_cast<uint32_t>(pmod(x1s + G::HHI + 1, g.n1)) * plane4;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q) pa[q] = bload(rf, foff[q], src);
    }
    // periodic index of plane x1 + HHI + 2, advanced by one per step (a
    // scalar modulo per step costs about 20 scalar instructions)
    int xf = pmod(x1s + G::HHI + 2, g.n1);
    auto step = [&](int x1, const PlanPt (&cur)[G::PTS], const float (&caux)[G::PTS], PlanPt (&nxt)[G::PTS],
                    float (&naux)[G::PTS], const float (&pre)[G::FILL], float (&pnext)[G::FILL]) {
        // Prefetch for the next steps, unconditionally (past the chunk end it
        // reloads a plane of this chunk; the values are never used).
        const int xn = (x1 + 1 < x1e) ? x1 + 1 : x1;
        {
            const uint32_t src = static_cast<uint32_t>(xf) * plane4;
            xf = (xf + 1 == g.n1) ? 0 : xf + 1;
#pragma unroll
            for (int q = 0; q < G::FILL; ++q) pnext[q] = bload(rf, foff[q], src);
        }
        {
            const uint32_t pb = static_cast<uint32_t>(xn) * plane4;
#pragma unroll
            for (int k = 0; k < G::PTS; ++k) {
                nxt[k] = bload_plan(rp, 3u * off[k], 3u * pb);
                naux[k] = Epi::kHasAux ? bload_once(ra, off[k], pb) : 0.0f;
            }
        }

        // Stencils: weights and the LDS byte addresses of rows 0 and 3 of the
        // four planes of each point.  The plan's bound keeps every offset
        // within the ring; the clamps only pin LDS addresses inside the
        // allocation.
        float w1[G::PTS][4], w2[G::PTS][4], w3[G::PTS][4];
        int base[G::PTS][4], base3[G::PTS][4];
#pragma unroll
        for (int k = 0; k < G::PTS; ++k) {
            lagrange_weights(plan_t(cur[k].w[0]), w1[k]);
            lagrange_weights(plan_t(cur[k].w[1]), w2[k]);
            lagrange_weights(plan_t(cur[k].w[2]), w3[k]);
            const int lb4 = min(max(rc4[k] + plan_off(cur[k].w[1]) * (4 * G::PITCH) + 4 * plan_off(cur[k].w[2]), 0),
                                4 * (G::P - 3 * G::PITCH - 4));
            int sw = min(max(cs + G::HLO + plan_off(cur[k].w[0]), 0), 2 * G::W - 4);
            sw -= (sw >= G::W) ? G::W : 0;  // ring slot of the first stencil plane
            base[k][0] = opaque(sw * P4 + lb4);
#pragma unroll
            for (int a = 1; a < 4; ++a) base[k][a] = opaque(base[k][a - 1] + ((sw + a == G::W) ? P4 - G::W * P4 : P4));
            // rows 0-2 fit the 8-bit dword offsets of ds_read2_b32 from base;
            // row 3 (3·PITCH + 3 dwords away) needs its own base
#pragma unroll
            for (int a = 0; a < 4; ++a) base3[k][a] = opaque(base[k][a] + 3 * 4 * G::PITCH);
        }
        float v[G::PTS];
#pragma unroll
        for (int k = 0; k < G::PTS; ++k) v[k] = 0.f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float sa[G::PTS];
#pragma unroll
            for (int k = 0; k < G::PTS; ++k) sa[k] = 0.f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
#pragma unroll
                for (int k = 0; k < G::PTS; ++k) {
                    const float* p = (b < 3) ? reinterpret_cast<const float*>(ldsb + base[k][a]) + b * G::PITCH
                                             : reinterpret_cast<const float*>(ldsb + base3[k][a]);
                    const float r = w3[k][0] * p[0] + w3[k][1] * p[1] + w3[k][2] * p[2] + w3[k][3] * p[3];
                    sa[k] += w2[k][b] * r;
                }
            }
#pragma unroll
            for (int k = 0; k < G::PTS; ++k) v[k] += w1[k][a] * sa[k];
        }
        {
            const uint32_t pb = static_cast<uint32_t>(x1) * plane4;
#pragma unroll
            for (int k = 0; k < G::PTS; ++k) {
                if (!valid[k]) continue;
                if constexpr (Epi::kCoord)
                    bstore(ro, epi.apply_at(v[k], caux[k], x1, x2, x3s[k]), off[k], pb);
                else
                    bstore(ro, epi.apply(v[k], caux[k]), off[k], pb);
                if constexpr (Epi::kOut2) bstore(ro2, epi.apply2(v[k], caux[k]), off[k], pb);
            }
        }

        // land the prefetched plane (x1 + HHI + 1) in the spare ring slot
        const int spare = (cs == 0) ? G::W - 1 : cs - 1;
        float* dst = lds + spare * G::P;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q)
            if (tid + kLThreads * q < G::NV) dst[loff[q]] = pre[q];
        cs = (cs + 1 == G::W) ? 0 : cs + 1;
        __syncthreads();
    };

    int x1 = x1s;
    for (; x1 + 1 < x1e; x1 += 2) {
        step(x1, qa, aa, qb, ab, pa, pb2);
        step(x1 + 1, qb, ab, qa, aa, pb2, pa);
    }
    if (x1 < x1e) step(x1, qa, aa, qb, ab, pa, pb2);
}

// Pair layout for D <= 2.  Entry c of an LDS row holds the two consecutive
// values (f[c], f[c+1]) of the tile row, so the four values of a stencil row
// starting at column c are the 8-byte entries c and c + 2: one ds_read2_b64
// per stencil row instead of two ds_read2_b32, with no alignment condition
// on c.  The price is twice the LDS per plane.  A block keeps a ring of
// W = WIN + 2·PERIOD − 1 planes and synchronizes once per PERIOD = 2 output
// planes (the planes of two consecutive stencil windows stay resident while
// the next two are landed in slots neither window reads): 11 planes of a
// 16 × 32 tile with its halo at a row pitch of 80 dwords take 79.1 KB, two
// blocks (16 waves) per CU.  The pitch of 96 dwords that a one-plane period
// affords has a third of the bank conflicts (8 k against 22 k cycles per CU
// and launch at 256³), but the halved barriers gain more (adjoint step
// 107.2 → 105.1 µs; at one plane per barrier the pitch of 80 alone costs
// 6 µs).
struct PairGeo {
    static constexpr int D = 2;
    static constexpr int LT3 = 32;
    static constexpr int WAVES_PER_EU = 4;
    static constexpr int HLO = D + 1, HHI = D + 2;
    static constexpr int E2 = kLT2 + HLO + HHI, E3 = LT3 + HLO + HHI;
    static constexpr int NE3 = E3 - 1;                   // entries per row (the last column only fills entry E3 − 2)
    static constexpr int PITCH = (2 * NE3 + 15) / 16 * 16;  // dwords per row
    static constexpr int P = E2 * PITCH;                 // dwords per LDS plane
    static constexpr int NV = E2 * NE3;                  // entries per plane
    static constexpr int WIN = HLO + HHI + 1;            // planes of one output plane's stencils
    static constexpr int PERIOD = 2;                     // output planes per barrier
    static constexpr int W = WIN + 2 * PERIOD - 1;       // ring slots
    static constexpr int FILL = (NV + kLThreads - 1) / kLThreads;
    // every ds_read2_b64 of a stencil (rows 0-3, entries c and c + 2) is
    // within the instruction's 8-bit offsets (in 8-byte units) of the base
    static_assert((3 * 4 * PITCH + 16) / 8 <= 255, "row 3 out of ds_read2_b64 offset range");
    // the kernel's steps are unrolled by four and barriers close periods
    static_assert(4 % PERIOD == 0, "the period must divide the unroll of four");
    // two blocks (16 waves) per CU in the 160 KB of LDS
    static_assert(2 * W * P * 4 <= 160 * 1024, "ring too large for two blocks per CU");
};

template <class Epi>
__global__ __launch_bounds__(kLThreads) __attribute__((amdgpu_waves_per_eu(PairGeo::WAVES_PER_EU))) void
k_gather_lds_pair(GatherDims g, const PlanPt* __restrict__ plan, const float* __restrict__ fbase, int64_t fzstride,
                  Epi epi, int nblocks, int chunk) {
    using G = PairGeo;
    __shared__ __align__(16) float lds[G::W * G::P];
    const int tid = threadIdx.x;
    const int nt3 = (g.n3 + G::LT3 - 1) / G::LT3;
    const int nt2 = (g.n2 + kLT2 - 1) / kLT2;
    // One block per (tile, chunk, field) on a 1D grid.  The fields of a tile
    // are neighbours in each XCD's range, so the plan words one of them loads
    // are L2 hits for the others (three-field gather at 256³ 288.0 -> 285.4
    // µs, outputs identical).
    const int nz = static_cast<int>(gridDim.x) / nblocks;
    const int Lz = xcd_remap(blockIdx.x, nblocks * nz);
    const int z = Lz % nz;
    const int L = Lz / nz;
    const int t3 = L % nt3;
    const int t2 = (L / nt3) % nt2;
    const int ch = L / (nt3 * nt2);
    const float* f = fbase + static_cast<int64_t>(z) * fzstride;
    epi.select(z);
    const uint32_t fbytes = static_cast<uint32_t>(g.nl) * 4u;
    const rsrc_t rf = make_rsrc(f, fbytes);
    const rsrc_t rp = make_rsrc(plan, 3u * fbytes);
    const rsrc_t ra = make_rsrc(Epi::kHasAux ? epi.auxp() : f, fbytes);
    const rsrc_t ro = make_rsrc(epi.outp(), fbytes);
    const rsrc_t ro2 = make_rsrc(out2_of(epi), fbytes);
    const uint32_t plane4 = static_cast<uint32_t>(g.n2) * static_cast<uint32_t>(g.n3) * 4u;
    const int x1s = ch * chunk;
    const int x1e = min(x1s + chunk, g.n1);
    const int r2 = t2 * kLT2 - G::HLO;
    const int r3 = t3 * G::LT3 - G::HLO;

    // Entries this thread fills: the byte offsets (within a global plane) of
    // the entry's two values and the entry's dword offset in an LDS plane.
    // Entries past the plane reload entry NV−1 (never stored).
    uint32_t foff[G::FILL][2];
    int loff[G::FILL];
#pragma unroll
    for (int k = 0; k < G::FILL; ++k) {
        const int e = min(tid + kLThreads * k, G::NV - 1);
        const int e2 = e / G::NE3;
        const int e3 = e - e2 * G::NE3;
        const int row = pmod(r2 + e2, g.n2) * g.n3;
        foff[k][0] = 4u * static_cast<uint32_t>(row + pmod(r3 + e3, g.n3));
        foff[k][1] = 4u * static_cast<uint32_t>(row + pmod(r3 + e3 + 1, g.n3));
        loff[k] = e2 * G::PITCH + 2 * e3;
    }

    const int lane = tid & 63;
    const int lx2 = 2 * (tid >> 6) + (lane >> 5);
    const int x2 = t2 * kLT2 + lx2;
    const int lx3 = lane & 31;
    const int x3 = t3 * G::LT3 + lx3;
    const bool valid = x2 < g.n2 && x3 < g.n3;
    const uint32_t off = valid ? 4u * static_cast<uint32_t>(x2 * g.n3 + x3) : 0u;
    // byte offset of the point's own cell (its entry) in an LDS plane
    const int rc = 4 * ((lx2 + G::HLO) * G::PITCH + 2 * (lx3 + G::HLO));

    auto store2 = [&](float* dst, int q, float a, float b) {
        *reinterpret_cast<float2*>(dst + loff[q]) = make_float2(a, b);
    };
    // prologue: input planes x1s − HLO … x1s + HHI + PERIOD − 1 (the stencils
    // of the first PERIOD output planes) into ring slots 0 … WIN + PERIOD − 2
    for (int k = 0; k < G::WIN + G::PERIOD - 1; ++k) {
        const uint32_t src = static_cast<uint32_t>(pmod(x1s - G::HLO + k, g.n1)) * plane4;
        float* dst = lds + k * G::P;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q)
            if (tid + kLThreads * q < G::NV) store2(dst, q, bload(rf, foff[q][0], src), bload(rf, foff[q][1], src));
    }
    // Loads run ahead of their use: while plane x1 is evaluated, the plan
    // words (and the epilogue operand) of plane x1 + LA_PT and the input
    // plane x1 + HHI + PERIOD + LA_PL are issued.  Q[k & 3] / A[k & 3] hold
    // the plan words and operand of step k; PRE[k % NPRE] the input plane
    // that step k lands in a free ring slot.  Measured at 256³ (outputs
    // identical): plane loads two steps ahead instead of one take the steps
    // from 104.5 to 103.7 µs for 8 more VGPRs (123); plan words three steps
    // ahead instead of two are slower (105 µs).
    constexpr int LA_PT = 2;  // steps between a plan word load and its use
    constexpr int LA_PL = 2;  // steps between an input plane load and its landing
    constexpr int NPRE = (LA_PL == 1) ? 2 : 4;
    PlanPt Q[4];
    float A[4];
    float PRE[NPRE][G::FILL][2];
    auto load_point = [&](int x, PlanPt& q, float& a) {
        const int xc = min(x, x1e - 1);  // past the chunk end: a plane of this chunk, never used
        const uint32_t pb = static_cast<uint32_t>(xc) * plane4;
        q = bload_plan(rp, 3u * off, 3u * pb);
        a = Epi::kHasAux ? bload_once(ra, off, pb) : 0.0f;
    };
    auto load_plane = [&](int xw, float (&pre)[G::FILL][2]) {  // xw: plane index in [0, n1)
        const uint32_t src = static_cast<uint32_t>(xw) * plane4;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q) {
            pre[q][0] = bload(rf, foff[q][0], src);
            pre[q][1] = bload(rf, foff[q][1], src);
        }
    };
    for (int k = 0; k < LA_PT; ++k) load_point(x1s + k, Q[k], A[k]);
    for (int k = 0; k < LA_PL; ++k) load_plane(pmod(x1s + G::HHI + G::PERIOD + k, g.n1), PRE[k]);
    // periodic index of the input plane the next step loads, advanced by one
    // per step (a scalar modulo per step costs about 20 scalar instructions)
    int xf = pmod(x1s + G::HHI + G::PERIOD + LA_PL, g.n1);
    __syncthreads();

    constexpr int P4 = 4 * G::P;
    const char* ldsb = reinterpret_cast<const char*>(lds);
    int cs = 0;
    // Stencil rows kept in registers across output planes.  A point's
    // stencil spans input planes x1 + o1 … x1 + o1 + 3; when the next plane's
    // point has the same offsets (o1, o2, o3) -- about 95% of points on a
    // CFL-2 trajectory -- its planes 0..2 are this point's planes 1..3, so
    // only its plane 3 is read from LDS.  S[k & 3] holds the rows of the
    // plane that step k reads as its plane 3; step k's plane a lives in
    // S[(k + a) & 3], which is the set step k + 1 uses for plane a − 1.
    // Lanes whose offsets changed (and every lane on a chunk's first plane)
    // read all four planes.  The values and their order of summation are
    // those of the LDS reads, so results are unchanged bit for bit.
    float2 S[4][4][2];
    int prev_o1 = 0, prev_lb = -1;
    auto step = [&](auto kconst, int x1, bool first) {
        constexpr int K = decltype(kconst)::value;
        const PlanPt& cur = Q[K & 3];
        const float caux = A[K & 3];
        load_point(x1 + LA_PT, Q[(K + LA_PT) & 3], A[(K + LA_PT) & 3]);
        load_plane(xf, PRE[(K + LA_PL) % NPRE]);  // plane x1 + HHI + PERIOD + LA_PL
        xf = (xf + 1 == g.n1) ? 0 : xf + 1;
        // land input plane x1 + HHI + PERIOD (loaded LA_PL steps earlier) in
        // the slot of plane x1 + HHI + PERIOD − W, which no step between
        // this period's barriers reads.  Landing it before the point's
        // stencil is read, not after the point is stored, lets the writes
        // drain while the point is evaluated, and a wave reaches the
        // period's barrier with no landing pending (measured at 256³,
        // outputs identical: adjoint step 103.9 -> 103.1 µs, three-field
        // gather 289.4 -> 287.3 µs).
        int spare = cs + G::WIN + G::PERIOD - 1;
        spare -= (spare >= G::W) ? G::W : 0;
        float* dst = lds + spare * G::P;
#pragma unroll
        for (int q = 0; q < G::FILL; ++q)
            if (tid + kLThreads * q < G::NV) store2(dst, q, PRE[K % NPRE][q][0], PRE[K % NPRE][q][1]);
        float w1[4], w2[4], w3[4];
        lagrange_weights(plan_t(cur.w[0]), w1);
        lagrange_weights(plan_t(cur.w[1]), w2);
        lagrange_weights(plan_t(cur.w[2]), w3);
        // the plan's bound keeps every stencil inside the ring; the clamps
        // only pin LDS addresses inside the allocation
        const int lb = min(max(rc + plan_off(cur.w[1]) * (4 * G::PITCH) + 8 * plan_off(cur.w[2]), 0),
                           4 * (G::P - 3 * G::PITCH - 6));
        int sw = min(max(cs + G::HLO + plan_off(cur.w[0]), 0), 2 * G::W - 4);
        sw -= (sw >= G::W) ? G::W : 0;
        int base[4];
        base[0] = opaque(sw * P4 + lb);
#pragma unroll
        for (int a = 1; a < 4; ++a) base[a] = opaque(base[a - 1] + ((sw + a == G::W) ? P4 - G::W * P4 : P4));
        // same offsets as the previous plane's point: the same x1 offset and
        // the same in-plane stencil address (which fixes the x2 and x3 offsets)
        const int o1 = plan_off(cur.w[0]);
        const bool reuse = !first && o1 == prev_o1 && lb == prev_lb;
        prev_o1 = o1;
        prev_lb = lb;
        auto read_plane = [&](int a, float2 (&R)[4][2]) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const char* row = ldsb + base[a] + b * 4 * G::PITCH;
                R[b][0] = *reinterpret_cast<const float2*>(row);
                R[b][1] = *reinterpret_cast<const float2*>(row + 16);
            }
        };
        // The stencil reads are issued at raised wave priority, so that a
        // wave's reads go out ahead of the other waves' contractions and
        // their LDS latency overlaps that work (outputs identical; at 256³
        // adjoint step 103.4 -> 101.1 µs, inc-state step 104.1 -> 101.0 µs,
        // three-field gather 285.6 -> 287.2 µs, matvec +0.9%; priority 3
        // instead of 1, or raised from the landing on, gained no more).
        __builtin_amdgcn_s_setprio(1);
        if (!reuse) {
            read_plane(0, S[K & 3]);
            read_plane(1, S[(K + 1) & 3]);
            read_plane(2, S[(K + 2) & 3]);
        }
        read_plane(3, S[(K + 3) & 3]);
        __builtin_amdgcn_s_setprio(0);
        float v = 0.f;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            float sa = 0.f;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float2 lo = S[(K + a) & 3][b][0];
                const float2 hi = S[(K + a) & 3][b][1];
                const float r = w3[0] * lo.x + w3[1] * lo.y + w3[2] * hi.x + w3[3] * hi.y;
                sa += w2[b] * r;
            }
            v += w1[a] * sa;
        }
        if (valid) {
            const uint32_t pb = static_cast<uint32_t>(x1) * plane4;
            if constexpr (Epi::kCoord)
                bstore(ro, epi.apply_at(v, caux, x1, x2, x3), off, pb);
            else
                bstore(ro, epi.apply(v, caux), off, pb);
            if constexpr (Epi::kOut2) bstore(ro2, epi.apply2(v, caux), off, pb);
        }
        cs = (cs + 1 == G::W) ? 0 : cs + 1;
        if constexpr (K % G::PERIOD == G::PERIOD - 1) __syncthreads();
    };
    using K0 = std::integral_constant<int, 0>;
    using K1 = std::integral_constant<int, 1>;
    using K2 = std::integral_constant<int, 2>;
    using K3 = std::integral_constant<int, 3>;

    // four planes per trip (the register sets rotate once)
    int x1 = x1s;
    bool first = true;
    for (; x1 + 3 < x1e; x1 += 4) {
        step(K0{}, x1, first);
        step(K1{}, x1 + 1, false);
        step(K2{}, x1 + 2, false);
        step(K3{}, x1 + 3, false);
        first = false;
    }
    // the last 0-3 planes continue the rotation
    if (x1 < x1e) step(K0{}, x1, first);
    if (x1 + 1 < x1e) step(K1{}, x1 + 1, false);
    if (x1 + 2 < x1e) step(K2{}, x1 + 2, false);
}

// Incremental state right-hand sides, streamed with 16-byte accesses:
//   s_j = ht/2·[(∇m^j)(X_s)·ṽ(X_s) + ∇m^{j+1}·ṽ],  j = 0..nt−1
// (the two decrements of src/CLAIRE.cpp:3253-3273);
// writes m̃¹ = −s_0 (λ̃⁰ = s_0 when nt = 1) and s_1..s_{nt−1}.
// 4 consecutive values as float4.
// 16-byte (FP32) or 8-byte (16-bit) streaming loads with the non-temporal
// hint: every element of these arrays is read once per launch, and the hint
// keeps the streams from displacing reusable lines (measured on a read-8,
// write-1 stream at 256³: 5.84 TB/s with the hint, 5.19 TB/s without).
using f4v = float __attribute__((ext_vector_type(4)));
using u2v = unsigned __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float4 ld4(const float* a, int64_t i) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(a) + i);
    return make_float4(v.x, v.y, v.z, v.w);
}

// Sources of the Newton-invariant gradient fields, field k at offset k·nl:
//   SrcF32  FP32 (the plain form);
//   SrcBfp  16-bit block floating point (the mixed form): int16 mantissas
//           with one power-of-two scale per block of 256 consecutive values
//           of a field (kBfpBlock4 float4 groups, one wave of the streaming
//           kernels); value = q·scale exactly.
constexpr int kBfpBlock4 = 64;
struct SrcF32 {
    const float* p;
    __device__ float4 load(int k, int64_t nl, int64_t i) const { return ld4(p + k * nl, i); }
};
struct SrcBfp {
    const int16_t* q;
    const float* s;
    int64_t nblk;  // scale blocks per field
    __device__ float4 load(int k, int64_t nl, int64_t i) const {
        const u2v bits = __builtin_nontemporal_load(reinterpret_cast<const u2v*>(q + k * nl) + i);
        const float sc = s[k * nblk + i / kBfpBlock4];
        const int16_t a = static_cast<int16_t>(bits.x & 0xFFFFu), b = static_cast<int16_t>(bits.x >> 16);
        const int16_t c = static_cast<int16_t>(bits.y & 0xFFFFu), d = static_cast<int16_t>(bits.y >> 16);
        return make_float4(static_cast<float>(a) * sc, static_cast<float>(b) * sc, static_cast<float>(c) * sc,
                           static_cast<float>(d) * sc);
    }
};

template <class Src>
struct RhsP {
    const float* vx[3];
    const float* vt[3];
    Src gX;
    Src gm;
    float* s;
    float* mt1;
};

// Src = SrcF32 (plain) or SrcBfp (mixed form: the Newton-invariant gradient
// fields stored as 16-bit block floating point, accumulated in FP32).
// Each wave of the streaming kernels covers kStreamK consecutive runs of 64
// float4 (1 KB of FP32) of every stream, lane l taking float4 l of each run,
// so each of the 30-odd streams is read in 2 KB bursts.  Measured at 256³
// (outputs identical): k_inc_state_rhs 399 → 367 µs and k_body_force 256 →
// 243 µs with two runs per wave; four runs (227 VGPRs for the right-hand
// sides) were slower than two.  A microbenchmark of 30 read and 4 write
// streams gave 5.6, 6.2 and 6.6 TB/s for one, two and four runs.
constexpr int kStreamK = 2;
__device__ __forceinline__ int64_t stream_base(int k) {
    return static_cast<int64_t>(blockIdx.x) * 256 * k + (threadIdx.x >> 6) * 64 * k + (threadIdx.x & 63);
}

template <class Src>
__global__ __launch_bounds__(256) void k_inc_state_rhs(int64_t nl4, int nt, float hthalf, RhsP<Src> p) {
    const int64_t i0 = stream_base(kStreamK);
    const int64_t nl = nl4 * 4;
    float4 x0[kStreamK], x1[kStreamK], x2[kStreamK], v0[kStreamK], v1[kStreamK], v2[kStreamK];
#pragma unroll
    for (int k = 0; k < kStreamK; ++k) {
        const int64_t i = min(i0 + 64 * k, nl4 - 1);  // past the end: reload the last group, never stored
        x0[k] = ld4(p.vx[0], i); x1[k] = ld4(p.vx[1], i); x2[k] = ld4(p.vx[2], i);
        v0[k] = ld4(p.vt[0], i); v1[k] = ld4(p.vt[1], i); v2[k] = ld4(p.vt[2], i);
    }
    for (int j = 0; j < nt; ++j) {
        float4 a0[kStreamK], a1[kStreamK], a2[kStreamK], c0[kStreamK], c1[kStreamK], c2[kStreamK];
#pragma unroll
        for (int k = 0; k < kStreamK; ++k) {
            const int64_t i = min(i0 + 64 * k, nl4 - 1);
            a0[k] = p.gX.load(3 * j, nl, i); a1[k] = p.gX.load(3 * j + 1, nl, i); a2[k] = p.gX.load(3 * j + 2, nl, i);
            c0[k] = p.gm.load(3 * (j + 1), nl, i); c1[k] = p.gm.load(3 * (j + 1) + 1, nl, i);
            c2[k] = p.gm.load(3 * (j + 1) + 2, nl, i);
        }
#pragma unroll
        for (int k = 0; k < kStreamK; ++k) {
            const int64_t i = i0 + 64 * k;
            if (i >= nl4) continue;
            float4 s;
            s.x = hthalf * (a0[k].x * x0[k].x + a1[k].x * x1[k].x + a2[k].x * x2[k].x) +
                  hthalf * (c0[k].x * v0[k].x + c1[k].x * v1[k].x + c2[k].x * v2[k].x);
            s.y = hthalf * (a0[k].y * x0[k].y + a1[k].y * x1[k].y + a2[k].y * x2[k].y) +
                  hthalf * (c0[k].y * v0[k].y + c1[k].y * v1[k].y + c2[k].y * v2[k].y);
            s.z = hthalf * (a0[k].z * x0[k].z + a1[k].z * x1[k].z + a2[k].z * x2[k].z) +
                  hthalf * (c0[k].z * v0[k].z + c1[k].z * v1[k].z + c2[k].z * v2[k].z);
            s.w = hthalf * (a0[k].w * x0[k].w + a1[k].w * x1[k].w + a2[k].w * x2[k].w) +
                  hthalf * (c0[k].w * v0[k].w + c1[k].w * v1[k].w + c2[k].w * v2[k].w);
            if (j == 0) {
                if (nt != 1) s = make_float4(-s.x, -s.y, -s.z, -s.w);
                reinterpret_cast<float4*>(p.mt1)[i] = s;
            } else {
                reinterpret_cast<float4*>(p.s + static_cast<int64_t>(j) * nl)[i] = s;
            }
        }
    }
}

// Body force b_d = hd·Σ_{j=0..nt} w_j·λ^j·∂_d m^{nt−j} with trapezoid weights
// w_0 = w_nt = ht/2, w_j = ht, accumulated in the reference's order
// (src/CLAIRE.cpp:2668-2709 and 3940-3999, scaled by hd at :2341/:3471).
// Streamed with 16-byte accesses; lam holds λ^0..λ^nt.
template <class Src>
struct BodyP {
    const float* lam;
    Src gm;
    float* b[3];
};

template <class Src>
__global__ __launch_bounds__(256) void k_body_force(int64_t nl4, int nt, float ht, float hd, BodyP<Src> p) {
    const int64_t i0 = stream_base(kStreamK);
    const int64_t nl = nl4 * 4;
    float4 s[kStreamK][3];
#pragma unroll
    for (int k = 0; k < kStreamK; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) s[k][d] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int j = 0; j <= nt; ++j) {
        const float w = (j == 0 || j == nt) ? 0.5f * ht : ht;
        float4 l[kStreamK], g[kStreamK][3];
#pragma unroll
        for (int k = 0; k < kStreamK; ++k) {
            const int64_t i = min(i0 + 64 * k, nl4 - 1);  // past the end: reload the last group, never stored
            l[k] = ld4(p.lam + static_cast<int64_t>(j) * nl, i);
#pragma unroll
            for (int d = 0; d < 3; ++d) g[k][d] = p.gm.load(3 * (nt - j) + d, nl, i);
        }
#pragma unroll
        for (int k = 0; k < kStreamK; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                s[k][d].x += w * g[k][d].x * l[k].x;
                s[k][d].y += w * g[k][d].y * l[k].y;
                s[k][d].z += w * g[k][d].z * l[k].z;
                s[k][d].w += w * g[k][d].w * l[k].w;
            }
    }
#pragma unroll
    for (int k = 0; k < kStreamK; ++k) {
        const int64_t i = i0 + 64 * k;
        if (i >= nl4) continue;
#pragma unroll
        for (int d = 0; d < 3; ++d)
            reinterpret_cast<float4*>(p.b[d])[i] =
                make_float4(s[k][d].x * hd, s[k][d].y * hd, s[k][d].z * hd, s[k][d].w * hd);
    }
}

// FP32 → 16-bit block floating point, field blockIdx.y, 4 values per thread.
// A wave (64 threads, 256 values) shares one scale 2^(E−15), E the binary
// exponent of its largest magnitude (max |v| < 2^E), so every mantissa
// q = round(v / 2^(E−15)) fits int16 (clamped to ±32767 at the top); the
// rounding error is at most 2^−16 of the block's largest magnitude.  Blocks
// whose largest magnitude is below 2^−100 are stored as zero.
__global__ __launch_bounds__(256) void k_to_bfp16(const float* __restrict__ in, int16_t* __restrict__ q,
                                                  float* __restrict__ s, int64_t nl4, int64_t nblk) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t field = blockIdx.y;
    const int64_t nl = nl4 * 4;
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    if (i < nl4) a = reinterpret_cast<const float4*>(in + field * nl)[i];
    float m = fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(a.w)));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    float scale = 1.0f, inv = 0.0f;
    if (m >= 0x1p-100f) {
        int e = 0;
        (void)frexpf(m, &e);  // m = f·2^e, f ∈ [0.5, 1)
        scale = ldexpf(1.0f, e - 15);
        inv = ldexpf(1.0f, 15 - e);
    }
    auto quant = [&](float v) {
        const float r = rintf(v * inv);
        return static_cast<int16_t>(fminf(fmaxf(r, -32767.0f), 32767.0f));
    };
    const int64_t blk = i / kBfpBlock4;
    if (i < nl4) {
        const uint32_t lo = static_cast<uint16_t>(quant(a.x)) | (static_cast<uint32_t>(static_cast<uint16_t>(quant(a.y))) << 16);
        const uint32_t hi = static_cast<uint16_t>(quant(a.z)) | (static_cast<uint32_t>(static_cast<uint16_t>(quant(a.w))) << 16);
        u2v bits;
        bits.x = lo;
        bits.y = hi;
        reinterpret_cast<u2v*>(q + field * nl)[i] = bits;
        if (i % kBfpBlock4 == 0) s[field * nblk + blk] = scale;
    }
}

// Planes per block: at most 48, and the chunk count that fills whole rounds
// of resident blocks (LDS allows 3 blocks per CU for D <= 2, 1 for D <= 4),
// weighed against the per-block prologue of WIN planes.
int num_cus() {
    // one device per process; the first call (any thread) queries it
    static const int cus = [] {
        int dev = 0, n = 0;
        CLR_HIP_CHECK(hipGetDevice(&dev));
        CLR_HIP_CHECK(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
        return n > 0 ? n : 1;
    }();
    return cus;
}

// CLAIRE_AMD_GATHER_CHUNK (planes, >= 4) overrides the choice, for tuning.
int chunk_override() {
    // read at every launch (not cached) so that a process can switch it
    const char* e = std::getenv("CLAIRE_AMD_GATHER_CHUNK");
    return e ? std::atoi(e) : 0;
}

// CLAIRE_AMD_GATHER_PAIR=0 selects the one-value-per-entry layout for single-field D <= 2 gathers
// (three blocks per CU) instead of the pair layout (two), for comparison.
bool pair_layout() {
    static const bool on = [] {
        const char* e = std::getenv("CLAIRE_AMD_GATHER_PAIR");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

// per_cu: resident blocks per CU (the LDS of one block decides it).
template <int D>
int choose_chunk(int n1, int64_t tiles, int per_cu) {
    using G = LdsGeo<D>;
    if (chunk_override() >= 4) return std::min(chunk_override(), n1);
    const int64_t slots = static_cast<int64_t>(num_cus()) * per_cu;
    // Chunks longer than 48 planes measured slower at 256³ (fewer, longer
    // blocks start and stream in lockstep); 32-43 planes were best.
    const int cmax = std::min(n1, 48);
    int best = cmax;
    double best_score = -1.0;
    for (int chunk = cmax; chunk >= 4; --chunk) {
        const int64_t nch = (n1 + chunk - 1) / chunk;
        if ((n1 + (chunk - 1) - 1) / (chunk - 1) == nch && chunk > 4) continue;  // same count, larger chunk kept
        const int64_t blocks = tiles * nch;
        const int64_t rounds = (blocks + slots - 1) / slots;
        const double fill = static_cast<double>(blocks) / static_cast<double>(rounds * slots);
        const double score = fill * chunk / static_cast<double>(chunk + G::WIN / 4);
        if (score > best_score + 1e-9) {
            best_score = score;
            best = chunk;
        }
    }
    return best;
}

// Chunk-length autotuning.  The first launch of a gather shape (bound D,
// grid, number of fields) times the heuristic's choice and a few other
// lengths on the launch stream and keeps the fastest for every later launch
// of that shape.  The chunk length only changes how the planes are split
// among blocks, not the arithmetic of any point, so every candidate writes
// the same output; repeating a launch is harmless because no gather writes a
// buffer it reads (the launchers check it: the steps refuse to run in place,
// and the RK4 stage sums ping-pong between two buffers).  Tuning is skipped while the stream is being captured into
// a graph; CLAIRE_AMD_GATHER_TUNE=0 turns it off (the heuristic is used).
struct ShapeKey {
    int d, n1, n2, n3, nz;
    bool operator<(const ShapeKey& o) const {
        return std::tie(d, n1, n2, n3, nz) < std::tie(o.d, o.n1, o.n2, o.n3, o.nz);
    }
};
std::mutex g_tune_mu;
std::map<ShapeKey, int> g_tuned;

bool tuning_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("CLAIRE_AMD_GATHER_TUNE");
        return !(e && std::atoi(e) == 0);
    }();
    return on;
}

// Launches through `launch(chunk)` with the tuned chunk length of `key`,
// tuning it on the first launch of that shape (see ShapeKey).
template <class Launch, class Heuristic>
void tuned_launch(const ShapeKey& key, int n1, hipStream_t st, Launch&& launch, Heuristic&& heuristic) {
    if (chunk_override() >= 4) return launch(std::min(chunk_override(), n1));
    {
        std::lock_guard<std::mutex> lock(g_tune_mu);
        const auto it = g_tuned.find(key);
        if (it != g_tuned.end()) return launch(it->second);
    }
    const int guess = heuristic();
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    if (!tuning_enabled() || hipStreamIsCapturing(st, &cap) != hipSuccess || cap != hipStreamCaptureStatusNone)
        return launch(guess);
    std::vector<int> cand{guess};
    for (int c : {16, 24, 32, 40, 48})
        if (c <= n1 && std::find(cand.begin(), cand.end(), c) == cand.end()) cand.push_back(c);
    launch(guess);  // warm-up; its output is the result
    int best = guess;
    if (cand.size() > 1) {
        hipEvent_t ev[2];
        CLR_HIP_CHECK(hipEventCreate(&ev[0]));
        CLR_HIP_CHECK(hipEventCreate(&ev[1]));
        float best_ms = std::numeric_limits<float>::max();
        for (int c : cand) {
            CLR_HIP_CHECK(hipEventRecord(ev[0], st));
            launch(c);
            CLR_HIP_CHECK(hipEventRecord(ev[1], st));
            CLR_HIP_CHECK(hipEventSynchronize(ev[1]));
            float ms = 0.f;
            CLR_HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    std::lock_guard<std::mutex> lock(g_tune_mu);
    g_tuned[key] = best;
}

template <int D, bool PAIR, class Epi>
void launch_lds(const GatherDims& g, const uint32_t* P, const float* f, int64_t fzstride, int nz, const Epi& epi,
                hipStream_t st) {
    static_assert(!PAIR || D == 2, "the pair layout is for D <= 2");
    const int nt3 = (g.n3 + LdsGeo<D>::LT3 - 1) / LdsGeo<D>::LT3;
    const int nt2 = (g.n2 + kLT2 - 1) / kLT2;
    auto launch = [&](int chunk) {
        const int nch = (g.n1 + chunk - 1) / chunk;
        const int nb = nt3 * nt2 * nch;
        if constexpr (PAIR)
            hipLaunchKernelGGL((k_gather_lds_pair<Epi>), dim3(nb * nz), dim3(kLThreads), 0, st, g,
                               reinterpret_cast<const PlanPt*>(P), f, fzstride, epi, nb, chunk);
        else
            hipLaunchKernelGGL((k_gather_lds<D, Epi>), dim3(nb * nz), dim3(kLThreads), 0, st, g,
                               reinterpret_cast<const PlanPt*>(P), f, fzstride, epi, nb, chunk);
    };
    const ShapeKey key{PAIR ? 100 + D : D, g.n1, g.n2, g.n3, nz};
    tuned_launch(key, g.n1, st, launch, [&] {
        return choose_chunk<D>(g.n1, static_cast<int64_t>(nt3) * nt2 * nz, PAIR ? 2 : (D <= 2 ? 3 : 1));
    });
}

template <class Epi>
void dispatch(int dbound, const GatherDims& g, const uint32_t* P, const float* f, int64_t fzstride, int nz,
              const Epi& epi, hipStream_t st) {
    CLR_REQUIRE(g.n1 >= 4 && g.n2 >= 4 && g.n3 >= 4, kErrInvalidArgument, "gather needs n >= 4 per axis");
    CLR_REQUIRE(g.nl < (int64_t(1) << 28), kErrInvalidArgument, "grid too large for 32-bit buffer byte offsets");
    CLR_REQUIRE(dbound >= 0 && dbound <= 4, kErrInvalidArgument, "displacement bound too large for the LDS gather");
    // The pair layout is faster for one field per launch (the solves' steps:
    // 110 µs against 115 at 256³, before later changes to both kernels) and
    // for plain three-field gathers (295 against 300 µs); the trajectory
    // stages were 6-12% slower with it (two blocks per CU instead of three)
    // and now only match the one-value layout, so they keep that one.
    constexpr bool kPlain = std::is_same<Epi, EpiPlain>::value;
    if (dbound <= 2 && (nz == 1 || kPlain) && pair_layout())
        launch_lds<2, true>(g, P, f, fzstride, nz, epi, st);
    else if (dbound <= 2)
        launch_lds<2, false>(g, P, f, fzstride, nz, epi, st);
    else
        launch_lds<4, false>(g, P, f, fzstride, nz, epi, st);
}

bool use_lds(const TrajectoryPlan& tp) { return tp.dbound >= 0 && tp.dbound <= lds_gather_max_displacement(); }

}  // namespace

int lds_gather_max_displacement() { return 4; }

void launch_plan_gather(const TrajectoryPlan& tp, const GatherDims& g, const float* const f[], int nf,
                        float* const out[], hipStream_t st) {
    if (!use_lds(tp)) return launch_direct_gather(g, tp.X, f, nf, out, st);
    const int64_t fz = nf > 1 ? (f[1] - f[0]) : 0;
    const int64_t oz = nf > 1 ? (out[1] - out[0]) : 0;
    for (int d = 2; d < nf; ++d)
        CLR_REQUIRE(f[d] - f[d - 1] == fz && out[d] - out[d - 1] == oz, kErrInvalidArgument,
                    "vector gather needs equally spaced components");
    dispatch(tp.dbound, g, tp.P, f[0], fz, nf, EpiPlain{out[0], oz}, st);
}

void launch_plan_inc_state_step(const TrajectoryPlan& tp, const GatherDims& g, const float* mt, const float* sj,
                                float* out, int negate, hipStream_t st) {
    if (!use_lds(tp)) return launch_direct_inc_state_step(g, tp.X, mt, sj, out, negate, st);
    CLR_REQUIRE(mt != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    dispatch(tp.dbound, g, tp.P, mt, 0, 1, EpiIncStep{sj, out, negate ? -1.0f : 1.0f}, st);
}

void launch_plan_adjoint_step(const TrajectoryPlan& tp, const GatherDims& g, const float* C, const float* lam,
                              float* out, hipStream_t st) {
    if (!use_lds(tp)) return launch_direct_adjoint_step(g, tp.X, C, lam, out, st);
    CLR_REQUIRE(lam != out, kErrInvalidArgument, "semi-Lagrangian step cannot run in place");
    dispatch(tp.dbound, g, tp.P, lam, 0, 1, EpiAdjStep{C, out}, st);
}

void launch_plan_adjoint_coeff(const TrajectoryPlan& tp, const GatherDims& g, const float* divv, float ht, float alpha,
                               float* C, hipStream_t st) {
    if (!use_lds(tp)) return launch_adjoint_coeff(g, tp.X, divv, ht, alpha, C, st);
    CLR_REQUIRE(divv != C, kErrInvalidArgument, "coefficient cannot overwrite ∇·v");
    dispatch(tp.dbound, g, tp.P, divv, 0, 1, EpiAdjCoeff{divv, C, ht, alpha}, st);
}

void launch_lds_traj_rk2_corrector(const Dims& d, const uint32_t* P, int dbound, const float* const v[3], float sigma,
                                   float ht, float* W, hipStream_t st) {
    const GatherDims g = gather_dims(d);
    const int64_t vz = v[1] - v[0];
    CLR_REQUIRE(v[2] - v[1] == vz && vz == g.nl, kErrInvalidArgument, "velocity components must be contiguous");
    EpiTrajRk2 e{v[0], W, g.nl, d.hx[0], d.hx[1], d.hx[2], sigma * (0.5f * ht), 0};
    dispatch(dbound, g, P, v[0], vz, 3, e, st);
}

void launch_lds_traj_rk4_stage(int stage, const Dims& d, const uint32_t* P, int dbound, const float* const v[3],
                               float sigma, float ht, const float* Fin, float* Fout, float* W, hipStream_t st) {
    const GatherDims g = gather_dims(d);
    const int64_t vz = v[1] - v[0];
    CLR_REQUIRE(v[2] - v[1] == vz && vz == g.nl, kErrInvalidArgument, "velocity components must be contiguous");
    // the tuner may repeat a launch: no stage may overwrite what it reads
    CLR_REQUIRE(stage == 1 || Fin != nullptr, kErrInvalidArgument, "RK4 stage needs the previous stage sum");
    CLR_REQUIRE(stage == 3 || (Fout != nullptr && Fout != Fin), kErrInvalidArgument,
                "RK4 stage sum must go to a buffer distinct from its input");
    const float hx[3] = {d.hx[0], d.hx[1], d.hx[2]};
    const float hthalf = 0.5f * ht;
    const double s6 = static_cast<double>(sigma) * (static_cast<double>(ht) / 6.0);
    switch (stage) {
        case 1: {
            EpiTrajRk4<1> e{v[0], W, Fout, g.nl, hx[0], hx[1], hx[2], sigma * hthalf, s6, 0};
            return dispatch(dbound, g, P, v[0], vz, 3, e, st);
        }
        case 2: {
            EpiTrajRk4<2> e{Fin, W, Fout, g.nl, hx[0], hx[1], hx[2], sigma * ht, s6, 0};
            return dispatch(dbound, g, P, v[0], vz, 3, e, st);
        }
        case 3: {
            EpiTrajRk4<3> e{Fin, W, W, g.nl, hx[0], hx[1], hx[2], 0.0f, s6, 0};  // no second output
            return dispatch(dbound, g, P, v[0], vz, 3, e, st);
        }
        default:
            throw Error(kErrInvalidArgument, "RK4 stage must be 1, 2 or 3");
    }
}

template <class Src>
void body_force(int64_t nl, int nt, float ht, float hd, const float* lam, const Src& gm, float* const b[3],
                hipStream_t st) {
    CLR_REQUIRE(nl % 4 == 0, kErrInvalidArgument, "grid size must be a multiple of 4 points");
    const int64_t nl4 = nl / 4;
    BodyP<Src> p{lam, gm, {b[0], b[1], b[2]}};
    const int nb = static_cast<int>((nl4 + 256 * kStreamK - 1) / (256 * kStreamK));
    hipLaunchKernelGGL((k_body_force<Src>), dim3(nb), dim3(256), 0, st, nl4, nt, ht, hd, p);
}

template <class Src>
void inc_state_rhs(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3], const Src& gX,
                   const Src& gm, float* s, float* mt1, hipStream_t st) {
    CLR_REQUIRE(nl % 4 == 0, kErrInvalidArgument, "grid size must be a multiple of 4 points");
    CLR_REQUIRE(nt >= 1, kErrInvalidArgument, "nt must be >= 1");
    const int64_t nl4 = nl / 4;
    RhsP<Src> p{{vx[0], vx[1], vx[2]}, {vt[0], vt[1], vt[2]}, gX, gm, s, mt1};
    const int nb = static_cast<int>((nl4 + 256 * kStreamK - 1) / (256 * kStreamK));
    hipLaunchKernelGGL((k_inc_state_rhs<Src>), dim3(nb), dim3(256), 0, st, nl4, nt, 0.5f * ht, p);
}

int64_t bfp16_blocks(int64_t nl) { return (nl / 4 + kBfpBlock4 - 1) / kBfpBlock4; }

namespace {

SrcBfp bfp_src(const Bfp16Fields& f, int64_t nl) { return SrcBfp{f.q, f.s, bfp16_blocks(nl)}; }

}  // namespace

void launch_body_force(int64_t nl, int nt, float ht, float hd, const float* lam, const float* gm, float* const b[3],
                       hipStream_t st) {
    body_force(nl, nt, ht, hd, lam, SrcF32{gm}, b, st);
}

void launch_body_force_bfp(int64_t nl, int nt, float ht, float hd, const float* lam, const Bfp16Fields& gm,
                           float* const b[3], hipStream_t st) {
    body_force(nl, nt, ht, hd, lam, bfp_src(gm, nl), b, st);
}

void launch_inc_state_rhs(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3],
                          const float* gX, const float* gm, float* s, float* mt1, hipStream_t st) {
    inc_state_rhs(nl, nt, ht, vx, vt, SrcF32{gX}, SrcF32{gm}, s, mt1, st);
}

void launch_inc_state_rhs_bfp(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3],
                              const Bfp16Fields& gX, const Bfp16Fields& gm, float* s, float* mt1, hipStream_t st) {
    inc_state_rhs(nl, nt, ht, vx, vt, bfp_src(gX, nl), bfp_src(gm, nl), s, mt1, st);
}

void launch_to_bfp16(const float* in, int64_t nl, int nfields, const Bfp16Fields& out, hipStream_t st) {
    CLR_REQUIRE(nl % 4 == 0, kErrInvalidArgument, "field length must be a multiple of 4");
    CLR_REQUIRE(nfields >= 1 && nfields <= 65535, kErrInvalidArgument, "1 to 65535 fields");
    const int64_t nl4 = nl / 4;
    const int nb = static_cast<int>((nl4 + 255) / 256);
    hipLaunchKernelGGL(k_to_bfp16, dim3(nb, nfields), dim3(256), 0, st, in, out.q, out.s, nl4, bfp16_blocks(nl));
}

}  // namespace claire_amd
