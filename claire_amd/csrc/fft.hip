// fft.hip -- hand-written 3D real FFT for gfx950 with fused k-space operators.
//
// Replaces AccFFT's accfft_execute_r2c_t / accfft_execute_c2r_t /
// accfft_grad_t / accfft_divergence_t (called e.g. at src/CLAIRE.cpp:2622,
// 3243, src/RegularizationH1.cpp:198-200,253-255; not vendored) on one GPU.
//
// Layout (same as the reference at one rank): real field [x1][x2][x3], x3
// fastest; half spectrum [k1][k2][k3], k3 in [0, n3/2], complex interleaved.
//
// A 3D transform is three passes over HBM:
//   x3 pass  rows of n3 reals, two rows packed as one complex row
//            (z = a + i·b), one complex FFT, split by Hermitian symmetry;
//   x2 pass  batched strided complex FFTs, 16 adjacent columns per block so
//            that each row access is a 128-byte segment;
//   x1 pass  same, and this is where k-space operators are fused: the block
//            holds complete x1 columns, so forward FFT, per-mode multiplier
//            and inverse FFT happen in one read and one write.
// Inside a block a length-N FFT is a Stockham autosort sequence of radix-16/8/
// 5/4/3/2 stages; each thread keeps V values in registers (V a multiple of
// every radix of the length) and the block swaps them through LDS between
// stages.  The inverse uses the forward radices in
// reverse order, so the forward output layout in registers is exactly the
// inverse input layout and fused forward·multiply·inverse needs no exchange.
// Twiddles come from a table generated in double precision on the host.
#include "kernels.hpp"

namespace claire_amd {

namespace {

__device__ __forceinline__ c32 cadd(c32 a, c32 b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ c32 csub(c32 a, c32 b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ c32 cmul(c32 a, c32 b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// a·(DIR·i)
template <int DIR>
__device__ __forceinline__ c32 mul_i(c32 a) { return {-DIR * a.y, DIR * a.x}; }

constexpr float kC8 = 0.70710678118654752440f;   // cos(π/4)
constexpr float kC16 = 0.92387953251128675613f;  // cos(π/8)
constexpr float kS16 = 0.38268343236508977173f;  // sin(π/8)

template <int DIR>
__device__ __forceinline__ void dft2(c32* v) {
    c32 t = v[0];
    v[0] = cadd(t, v[1]);
    v[1] = csub(t, v[1]);
}

template <int DIR>
__device__ __forceinline__ void dft4s(c32& v0, c32& v1, c32& v2, c32& v3) {
    c32 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = mul_i<DIR>(csub(v1, v3));
    v0 = cadd(a0, a2);
    v2 = csub(a0, a2);
    v1 = cadd(a1, a3);
    v3 = csub(a1, a3);
}

template <int DIR>
__device__ __forceinline__ void dft4(c32* v) { dft4s<DIR>(v[0], v[1], v[2], v[3]); }

template <int DIR>
__device__ __forceinline__ void dft8(c32* v) {
    c32 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    c32 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4s<DIR>(e0, e1, e2, e3);
    dft4s<DIR>(o0, o1, o2, o3);
    // o_k *= w8^k, w8 = exp(DIR·2πi/8)
    o1 = {kC8 * (o1.x - DIR * o1.y), kC8 * (DIR * o1.x + o1.y)};
    o2 = mul_i<DIR>(o2);
    o3 = {kC8 * (-o3.x - DIR * o3.y), kC8 * (DIR * o3.x - o3.y)};
    v[0] = cadd(e0, o0);
    v[4] = csub(e0, o0);
    v[1] = cadd(e1, o1);
    v[5] = csub(e1, o1);
    v[2] = cadd(e2, o2);
    v[6] = csub(e2, o2);
    v[3] = cadd(e3, o3);
    v[7] = csub(e3, o3);
}

template <int DIR>
__device__ __forceinline__ void dft16(c32* v) {
    c32 e[8], o[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        e[i] = v[2 * i];
        o[i] = v[2 * i + 1];
    }
    dft8<DIR>(e);
    dft8<DIR>(o);
    // w16^k = (cos(πk/8), DIR·sin(πk/8))
    const float cs[8] = {1.f, kC16, kC8, kS16, 0.f, -kS16, -kC8, -kC16};
    const float sn[8] = {0.f, kS16, kC8, kC16, 1.f, kC16, kC8, kS16};
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        c32 w = {cs[k], DIR * sn[k]};
        c32 ow = (k == 0) ? o[k] : (k == 4 ? mul_i<DIR>(o[k]) : cmul(o[k], w));
        v[k] = cadd(e[k], ow);
        v[k + 8] = csub(e[k], ow);
    }
}

constexpr float kS3 = 0.86602540378443864676f;   // sin(2π/3)
constexpr float kC51 = 0.30901699437494742410f;  // cos(2π/5)
constexpr float kC52 = -0.80901699437494742410f; // cos(4π/5)
constexpr float kS51 = 0.95105651629515357212f;  // sin(2π/5)
constexpr float kS52 = 0.58778525229247312917f;  // sin(4π/5)

// X_k = Σ_m x_m·w^{mk}, w = exp(DIR·2πi/3)
template <int DIR>
__device__ __forceinline__ void dft3(c32* v) {
    const c32 t1 = cadd(v[1], v[2]);
    const c32 t2 = {v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y};
    const c32 d = csub(v[1], v[2]);
    const c32 t3 = mul_i<DIR>(c32{kS3 * d.x, kS3 * d.y});
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, t3);
    v[2] = csub(t2, t3);
}

// X_k = Σ_m x_m·w^{mk}, w = exp(DIR·2πi/5)
template <int DIR>
__device__ __forceinline__ void dft5(c32* v) {
    const c32 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
    const c32 t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    const c32 a1 = {v[0].x + kC51 * t1.x + kC52 * t2.x, v[0].y + kC51 * t1.y + kC52 * t2.y};
    const c32 a2 = {v[0].x + kC52 * t1.x + kC51 * t2.x, v[0].y + kC52 * t1.y + kC51 * t2.y};
    const c32 b1 = mul_i<DIR>(c32{kS51 * t3.x + kS52 * t4.x, kS51 * t3.y + kS52 * t4.y});
    const c32 b2 = mul_i<DIR>(c32{kS52 * t3.x - kS51 * t4.x, kS52 * t3.y - kS51 * t4.y});
    v[0] = cadd(v[0], cadd(t1, t2));
    v[1] = cadd(a1, b1);
    v[4] = csub(a1, b1);
    v[2] = cadd(a2, b2);
    v[3] = csub(a2, b2);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(c32* v) {
    if constexpr (R == 2) dft2<DIR>(v);
    else if constexpr (R == 3) dft3<DIR>(v);
    else if constexpr (R == 4) dft4<DIR>(v);
    else if constexpr (R == 5) dft5<DIR>(v);
    else if constexpr (R == 8) dft8<DIR>(v);
    else if constexpr (R == 16) dft16<DIR>(v);
    else static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16, "unsupported radix");
}

// ---------------------------------------------------------------------------
// Stockham stage machinery.  A length-N transform is run by T = N/V threads
// per column; "virtual" butterfly j of a radix-R stage is done by thread
// j mod T.  Register convention at stage entry: v[q·R + r] = x[j + r·N/R],
// j = t + T·q.  After the last stage (span NS = N/R): v[q·R + r] = X[j + r·N/R].
// ---------------------------------------------------------------------------
template <int... R>
struct Radices {};

template <int N>
struct Plan;
// V = values per thread; Fwd radices; Inv = reverse order.
template <> struct Plan<8>   { static constexpr int V = 8;  using Fwd = Radices<8>;         using Inv = Radices<8>; };
template <> struct Plan<16>  { static constexpr int V = 16; using Fwd = Radices<16>;        using Inv = Radices<16>; };
template <> struct Plan<32>  { static constexpr int V = 16; using Fwd = Radices<16, 2>;     using Inv = Radices<2, 16>; };
template <> struct Plan<64>  { static constexpr int V = 16; using Fwd = Radices<16, 4>;     using Inv = Radices<4, 16>; };
template <> struct Plan<128> { static constexpr int V = 16; using Fwd = Radices<16, 8>;     using Inv = Radices<8, 16>; };
template <> struct Plan<256> { static constexpr int V = 16; using Fwd = Radices<16, 16>;    using Inv = Radices<16, 16>; };
template <> struct Plan<512> { static constexpr int V = 16; using Fwd = Radices<16, 16, 2>; using Inv = Radices<2, 16, 16>; };
// lengths with factors 3 and 5 (every radix divides V)
template <> struct Plan<48>  { static constexpr int V = 12; using Fwd = Radices<4, 4, 3>;     using Inv = Radices<3, 4, 4>; };
template <> struct Plan<96>  { static constexpr int V = 24; using Fwd = Radices<8, 4, 3>;     using Inv = Radices<3, 4, 8>; };
template <> struct Plan<150> { static constexpr int V = 30; using Fwd = Radices<2, 3, 5, 5>;  using Inv = Radices<5, 5, 3, 2>; };
template <> struct Plan<192> { static constexpr int V = 24; using Fwd = Radices<8, 8, 3>;     using Inv = Radices<3, 8, 8>; };

template <class L> struct First;
template <int R, int... RS> struct First<Radices<R, RS...>> { static constexpr int value = R; };
template <class L> struct Last;
template <int R> struct Last<Radices<R>> { static constexpr int value = R; };
template <int R, int R2, int... RS> struct Last<Radices<R, R2, RS...>> { static constexpr int value = Last<Radices<R2, RS...>>::value; };

template <int N, int V, int DIR, int NS, int R>
__device__ __forceinline__ void twiddle_butterfly(c32 (&v)[V], int t, const c32* __restrict__ tw) {
    constexpr int T = N / V;
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
        const int j = t + T * q;
        if constexpr (NS > 1) {
            const int jm = j % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                c32 w = tw[jm * r * (N / (NS * R))];
                if constexpr (DIR > 0) w.y = -w.y;
                v[q * R + r] = cmul(v[q * R + r], w);
            }
        }
        dft<R, DIR>(&v[q * R]);
    }
}

// LDS layouts of the exchanges: lay(n, c) is the slot of element n of the
// block's column (or row pair) c.
//   ColLayout: columns interleaved with an odd pitch; the lanes of a wave run
//              over c, so any row access hits distinct banks.
//   RowLayout: each row contiguous with one pad slot per 16 elements; the
//              lanes of a wave run over n, and the radix-16 strides n = 16t + r
//              and n = t + 16r both map the 16 lanes of a row to 16 distinct
//              bank pairs.
struct ColLayout {
    int LD;
    __device__ __forceinline__ int operator()(int n, int c) const { return n * LD + c; }
};

template <int N>
struct RowLayout {
    static constexpr int CP = N + N / 16;
    __device__ __forceinline__ int operator()(int n, int c) const { return c * CP + n + (n >> 4); }
};

// Write stage outputs (span NS, radix R) to LDS, then load the inputs of the
// next radix-R2 stage.
template <int N, int V, int NS, int R, int R2, class Lay>
__device__ __forceinline__ void exchange(c32 (&v)[V], c32* s, const Lay& lay, int c, int t) {
    constexpr int T = N / V;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < V / R; ++q) {
        const int j = t + T * q;
        const int base = (j / NS) * NS * R + (j % NS);
#pragma unroll
        for (int r = 0; r < R; ++r) s[lay(base + r * NS, c)] = v[q * R + r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < V / R2; ++q) {
        const int j = t + T * q;
#pragma unroll
        for (int r = 0; r < R2; ++r) v[q * R2 + r] = s[lay(j + r * (N / R2), c)];
    }
}

template <int N, int V, int DIR, int NS, class L>
struct Stages;

template <int N, int V, int DIR, int NS, int R>
struct Stages<N, V, DIR, NS, Radices<R>> {
    template <class Lay>
    __device__ __forceinline__ static void run(c32 (&v)[V], c32*, const Lay&, int, int t, const c32* tw) {
        twiddle_butterfly<N, V, DIR, NS, R>(v, t, tw);
    }
};

template <int N, int V, int DIR, int NS, int R, int R2, int... RS>
struct Stages<N, V, DIR, NS, Radices<R, R2, RS...>> {
    template <class Lay>
    __device__ __forceinline__ static void run(c32 (&v)[V], c32* s, const Lay& lay, int c, int t, const c32* tw) {
        twiddle_butterfly<N, V, DIR, NS, R>(v, t, tw);
        exchange<N, V, NS, R, R2>(v, s, lay, c, t);
        Stages<N, V, DIR, NS * R, Radices<R2, RS...>>::run(v, s, lay, c, t, tw);
    }
};

template <int N, int DIR, class Lay>
__device__ __forceinline__ void fft_regs(c32 (&v)[Plan<N>::V], c32* s, const Lay& lay, int c, int t,
                                         const c32* tw) {
    if constexpr (DIR < 0)
        Stages<N, Plan<N>::V, DIR, 1, typename Plan<N>::Fwd>::run(v, s, lay, c, t, tw);
    else
        Stages<N, Plan<N>::V, DIR, 1, typename Plan<N>::Inv>::run(v, s, lay, c, t, tw);
}

// Block geometry: B columns (or row pairs) per block, T threads per column.
template <int N>
struct Geo {
    static constexpr int V = Plan<N>::V;
    static constexpr int T = N / V;
    static constexpr int B = (T >= 32) ? 8 : ((T >= 16) ? 16 : (256 / T));
    static constexpr int THREADS = B * T;
    static constexpr int LD = B + 1;  // ColLayout pitch (odd: spreads banks)
    static constexpr int ROW_ELEMS = B * RowLayout<N>::CP;
    static constexpr int LDS_ELEMS = (N * LD > ROW_ELEMS) ? N * LD : ROW_ELEMS;
};

// Geometry of the plain column passes (k_col, the x2 passes): Geo<N> with
// twice the columns per block at N = 256 and 512 (256-byte row segments per
// block instead of 128; measured: three-field x2 pass at 256³ 82.8 → 79.8
// µs, the 512³ H2 matvec 29.10 → 28.95 ms, outputs identical).  The fused
// x1 pass and the row passes keep Geo<N> (wider blocks made them slower).
template <int N>
struct ColGeo {
    static constexpr int V = Plan<N>::V;
    static constexpr int T = N / V;
    static constexpr int B = (N == 256 || N == 512) ? 2 * Geo<N>::B : Geo<N>::B;
    static constexpr int THREADS = B * T;
    static constexpr int LD = B + 1;
    static constexpr int LDS_ELEMS = N * LD;
};

__device__ __forceinline__ int wavenum(int k, int n) { return k > n / 2 ? k - n : (k == n / 2 ? 0 : k); }

// Per-mode real multiplier, FP32 expressions in the reference's order
// (src/RegularizationH1.cpp:222-229,373-382; src/RegularizationH2.cpp:123-128,
// 236-239,375-381; src/RegularizationH1SN.cpp:210-217,352-359;
// src/RegularizationH2SN.cpp:125-130,241-244,385-390).
__device__ __forceinline__ float spec_multiplier(const SpecParams& p, int w1, int w2, int w3) {
    const float lapik = -static_cast<float>(w1 * w1 + w2 * w2 + w3 * w3);
    const float b0 = p.beta0, b1 = p.beta1;
    switch (p.op) {
        case kOpRegApply:
            switch (p.norm) {
                case kRegH1: return p.hd * p.scale * b0 * (-lapik + b1);
                case kRegH2: return p.scale * b0 * (lapik * lapik + b1);
                case kRegH1SN: return -p.hd * p.scale * b0 * lapik;
                default: return p.hd * p.scale * b0 * (lapik * lapik);
            }
        case kOpRegInverse:
        case kOpRegInvSqrt: {
            float regop;
            switch (p.norm) {
                case kRegH1: regop = b0 * (-lapik + b1); break;
                case kRegH2: regop = b0 * (lapik * lapik + b1); break;
                case kRegH1SN: regop = (fabsf(lapik) == 0.0f) ? b0 : -b0 * lapik; break;
                default: regop = (fabsf(lapik) == 0.0f) ? b0 : b0 * (lapik * lapik); break;
            }
            if (p.op == kOpRegInvSqrt) regop = sqrtf(regop);
            return p.scale / regop;
        }
        case kOpH2Functional: return p.scale * b0 * (lapik + b1);  // b0, b1 hold √β0, √β1
        default: return p.scale * lapik;
    }
}

// ---------------------------------------------------------------------------
// x3 pass, real -> half complex.  Rows 2p and 2p+1 are packed as a + i·b.
// ---------------------------------------------------------------------------
// Up to three fields per launch: blockIdx.y selects the field (a vector
// field's components go through one launch instead of three).
struct R2CPtrs {
    const float* in[3];
    c32* out[3];
};
struct C2RPtrs {
    const c32* in[3];
    float* out[3];
    const float* add[3];
};
struct ColPtrs {
    const c32* in[3];
    c32* out[3];
};

template <int N>
__global__ __launch_bounds__(Geo<N>::THREADS) void k_row_r2c(R2CPtrs ptr, int64_t nrows, const c32* __restrict__ tw) {
    const float* __restrict__ in = ptr.in[blockIdx.y];
    c32* __restrict__ out = ptr.out[blockIdx.y];
    using G = Geo<N>;
    __shared__ c32 s[G::LDS_ELEMS];
    constexpr int V = G::V, T = G::T, R0 = First<typename Plan<N>::Fwd>::value,
                  RL = Last<typename Plan<N>::Fwd>::value;
    constexpr int NH = N / 2 + 1;
    const int t = threadIdx.x % T;
    const int c = threadIdx.x / T;
    const int64_t ra = 2 * (static_cast<int64_t>(blockIdx.x) * G::B + c);
    const int64_t rb = ra + 1;
    const bool va = ra < nrows, vb = rb < nrows;
    const float* pa = in + ra * N;
    const float* pb = in + rb * N;
    c32 v[V];
#pragma unroll
    for (int q = 0; q < V / R0; ++q) {
        const int j = t + T * q;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int n = j + r * (N / R0);
            v[q * R0 + r] = {va ? pa[n] : 0.f, vb ? pb[n] : 0.f};
        }
    }
    const RowLayout<N> lay;
    fft_regs<N, -1>(v, s, lay, c, t, tw);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < V / RL; ++q) {
        const int j = t + T * q;
#pragma unroll
        for (int r = 0; r < RL; ++r) s[lay(j + r * (N / RL), c)] = v[q * RL + r];
    }
    __syncthreads();
    c32* oa = out + ra * NH;
    c32* ob = out + rb * NH;
    for (int k = t; k < NH; k += T) {
        const c32 zk = s[lay(k, c)];
        const c32 zm = s[lay(k == 0 ? 0 : N - k, c)];
        if (va) oa[k] = {0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)};
        if (vb) ob[k] = {0.5f * (zk.y + zm.y), 0.5f * (zm.x - zk.x)};
    }
}

// ---------------------------------------------------------------------------
// x3 pass, half complex -> real, with epilogue out = alpha·x + beta·add.
// The imaginary parts of the k3 = 0 and k3 = n3/2 coefficients are ignored
// (FFTW c2r convention).
// ---------------------------------------------------------------------------
// ADD: some field has an added term (its values are loaded before the
// transform, in registers the plain pass does not need).
template <int N, bool ADD>
__global__ __launch_bounds__(Geo<N>::THREADS) void k_row_c2r(C2RPtrs ptr, int64_t nrows, const c32* __restrict__ tw,
                                                             float alpha, float beta) {
    const c32* __restrict__ in = ptr.in[blockIdx.y];
    float* __restrict__ out = ptr.out[blockIdx.y];
    const float* __restrict__ add = ADD ? ptr.add[blockIdx.y] : nullptr;
    using G = Geo<N>;
    __shared__ c32 s[G::LDS_ELEMS];
    constexpr int V = G::V, T = G::T, R0 = First<typename Plan<N>::Inv>::value,
                  RL = Last<typename Plan<N>::Inv>::value;
    constexpr int NH = N / 2 + 1;
    const int t = threadIdx.x % T;
    const int c = threadIdx.x / T;
    const int64_t ra = 2 * (static_cast<int64_t>(blockIdx.x) * G::B + c);
    const int64_t rb = ra + 1;
    const bool va = ra < nrows, vb = rb < nrows;
    const c32* ha = in + ra * NH;
    const c32* hb = in + rb * NH;
    c32 v[V];
    // the epilogue's added field, loaded before the transform so that its
    // latency overlaps the butterflies (three-field pass with the add at
    // 256³: 115.7 -> 113.0 µs, outputs identical)
    constexpr int NQ = (N % 4 == 0) ? (N / 4 + T - 1) / T : 1;
    float4 qa[NQ], qb[NQ];
    if constexpr (N % 4 == 0) {
        const float* aa0 = add ? add + ra * N : nullptr;
        const float* ab0 = add ? add + rb * N : nullptr;
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int m = t + T * i;
            if (aa0 && va && m < N / 4) qa[i] = reinterpret_cast<const float4*>(aa0)[m];
            if (ab0 && vb && m < N / 4) qb[i] = reinterpret_cast<const float4*>(ab0)[m];
        }
    }
#pragma unroll
    for (int q = 0; q < V / R0; ++q) {
        const int j = t + T * q;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const int k = j + r * (N / R0);
            const bool mirror = k > N / 2;
            const int kk = mirror ? N - k : k;
            c32 A = va ? ha[kk] : c32{0.f, 0.f};
            c32 Bv = vb ? hb[kk] : c32{0.f, 0.f};
            if (kk == 0 || kk == N / 2) {
                A.y = 0.f;
                Bv.y = 0.f;
            }
            if (mirror) {
                A.y = -A.y;
                Bv.y = -Bv.y;
            }
            v[q * R0 + r] = {A.x - Bv.y, A.y + Bv.x};
        }
    }
    const RowLayout<N> lay;
    fft_regs<N, +1>(v, s, lay, c, t, tw);
    float* oa = out + ra * N;
    float* ob = out + rb * N;
    const float* aa = add ? add + ra * N : nullptr;
    const float* ab = add ? add + rb * N : nullptr;
    if constexpr (N % 4 == 0) {
        // Stage the result through LDS so that each lane writes 4 consecutive
        // reals (16-byte stores) of both rows.
        __syncthreads();
#pragma unroll
        for (int q = 0; q < V / RL; ++q) {
            const int j = t + T * q;
#pragma unroll
            for (int r = 0; r < RL; ++r) s[lay(j + r * (N / RL), c)] = v[q * RL + r];
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NQ; ++i) {
            const int m = t + T * i;
            if (m >= N / 4) break;
            const c32 z0 = s[lay(4 * m, c)], z1 = s[lay(4 * m + 1, c)], z2 = s[lay(4 * m + 2, c)],
                      z3 = s[lay(4 * m + 3, c)];
            if (va) {
                float4 o = make_float4(alpha * z0.x, alpha * z1.x, alpha * z2.x, alpha * z3.x);
                if (aa) {
                    const float4 x = qa[i];
                    o = make_float4(o.x + beta * x.x, o.y + beta * x.y, o.z + beta * x.z, o.w + beta * x.w);
                }
                reinterpret_cast<float4*>(oa)[m] = o;
            }
            if (vb) {
                float4 o = make_float4(alpha * z0.y, alpha * z1.y, alpha * z2.y, alpha * z3.y);
                if (ab) {
                    const float4 x = qb[i];
                    o = make_float4(o.x + beta * x.x, o.y + beta * x.y, o.z + beta * x.z, o.w + beta * x.w);
                }
                reinterpret_cast<float4*>(ob)[m] = o;
            }
        }
    } else {
#pragma unroll
        for (int q = 0; q < V / RL; ++q) {
            const int j = t + T * q;
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int n = j + r * (N / RL);
                const c32 z = v[q * RL + r];
                if (va) oa[n] = aa ? alpha * z.x + beta * aa[n] : alpha * z.x;
                if (vb) ob[n] = ab ? alpha * z.y + beta * ab[n] : alpha * z.y;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Strided column pass (x2 or x1), plain forward or inverse, in place allowed.
// Column cg starts at (cg / inner)·outer + cg % inner; elements are estride
// apart.
// ---------------------------------------------------------------------------
struct ColGeom {
    int64_t ncols, inner, outer, estride;
};

// Offset of column cg's first element; column counts and the inner extent
// fit 32 bits (64-bit division is a long emulated sequence).
template <int N>
__device__ __forceinline__ int64_t col_base(const ColGeom& g, int64_t cg) {
    const uint32_t c = static_cast<uint32_t>(cg), in = static_cast<uint32_t>(g.inner);
    const uint32_t q = c / in;
    return static_cast<int64_t>(q) * g.outer + static_cast<int64_t>(c - q * in);
}

// Raw buffer access to one field of a complex spectrum of `elems` values: a
// per-lane byte offset (the lane's column and row) plus a uniform one (the
// element's stride multiple), both 32-bit, so a column pass keeps one address
// register per lane instead of a 64-bit address per element.  The launchers
// require fields under 2 GiB.
using spec_rsrc_t = __amdgpu_buffer_rsrc_t;
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ spec_rsrc_t spec_rsrc(const c32* p, int64_t elems) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<c32*>(p), 0, static_cast<int>(elems * 8), 0x00020000);
}
__device__ __forceinline__ c32 bload_c(spec_rsrc_t r, uint32_t voff, uint32_t soff) {
    const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0);
    return {__uint_as_float(w.x), __uint_as_float(w.y)};
}
__device__ __forceinline__ void bstore_c(spec_rsrc_t r, c32 v, uint32_t voff, uint32_t soff) {
    u32x2 w;
    w.x = __float_as_uint(v.x);
    w.y = __float_as_uint(v.y);
    __builtin_amdgcn_raw_buffer_store_b64(w, r, voff, soff, 0);
}

template <int N, int DIR>
__global__ __launch_bounds__(ColGeo<N>::THREADS) void k_col(ColPtrs ptr, ColGeom geo, const c32* __restrict__ tw) {
    const c32* in = ptr.in[blockIdx.y];  // may equal out (in place): no __restrict__
    c32* out = ptr.out[blockIdx.y];
    using G = ColGeo<N>;
    __shared__ c32 s[G::LDS_ELEMS];
    using L = typename std::conditional<(DIR < 0), typename Plan<N>::Fwd, typename Plan<N>::Inv>::type;
    constexpr int V = G::V, T = G::T, R0 = First<L>::value, RL = Last<L>::value;
    const int c = threadIdx.x % G::B;
    const int t = threadIdx.x / G::B;
    // Blocks are dealt round-robin over the 8 XCDs; remap them so each XCD
    // owns one contiguous range of columns and the cache lines shared by
    // neighbouring column groups (rows are not 128-byte aligned) stay in one L2.
    const unsigned nb = gridDim.x;
    const unsigned b = (nb % 8 == 0) ? (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int64_t cg = static_cast<int64_t>(b) * G::B + c;
    const bool valid = cg < geo.ncols;
    const int64_t base = valid ? col_base<N>(geo, cg) : 0;
    c32 v[V];
    const spec_rsrc_t rin = spec_rsrc(in, geo.ncols * N), rout = spec_rsrc(out, geo.ncols * N);
    const uint32_t es8 = static_cast<uint32_t>(geo.estride) * 8u;
    const uint32_t lane = static_cast<uint32_t>(base) * 8u + static_cast<uint32_t>(t) * es8;
#pragma unroll
    for (int q = 0; q < V / R0; ++q)
#pragma unroll
        for (int r = 0; r < R0; ++r) {
            const c32 x = bload_c(rin, lane, static_cast<uint32_t>(T * q + r * (N / R0)) * es8);
            v[q * R0 + r] = valid ? x : c32{0.f, 0.f};
        }
    fft_regs<N, DIR>(v, s, ColLayout{G::LD}, c, t, tw);
    if (valid) {
#pragma unroll
        for (int q = 0; q < V / RL; ++q)
#pragma unroll
            for (int r = 0; r < RL; ++r)
                bstore_c(rout, v[q * RL + r], lane, static_cast<uint32_t>(T * q + r * (N / RL)) * es8);
    }
}

// ---------------------------------------------------------------------------
// Fused x1 pass: forward x1 FFT, k-space operator, inverse x1 FFT.
//   kFuseMult  out[d] = F1⁻¹(m(k)·F1(in[d]))         d = blockIdx.y
//   kFuseGrad  out[d] = F1⁻¹(i·w_d·scale·F1(in[0]))   d = 0..2
//   kFuseDiv   out[0] = F1⁻¹(Σ_d i·w_d·scale·F1(in[d]))
// Columns of the x1 pass are (k2, k3) with cg = k2·nh + k3.
// ---------------------------------------------------------------------------
enum FuseMode { kFuseMult = 0, kFuseGrad = 1, kFuseDiv = 2, kFuseProject = 3 };

// Per-mode factor m of the models' projections, FP32/FP64 exactly as the
// reference writes it: ĉ = scale·m·(g·b̂), x̂_i = g_i·ĉ, b ← b + F⁻¹x̂.
//   Stokes (src/CLAIREStokes.cpp:457-462): m = −1/|w|², the Nyquist index
//   kept at +n/2 in |w|², −1 at w = 0.
//   Relaxed (src/CLAIREDivReg.cpp:362-404): lapik = −|g|² (Nyquist zeroed),
//   m = opik·(1/lapik), opik = 1/(β0·(1/(β2·(−lapik + 1))) + 1).
__device__ __forceinline__ float proj_factor(const SpecParams& p, int k1, int k2, int k3, int n1, int n2, int n3) {
    const int x1 = k1 > n1 / 2 ? k1 - n1 : k1, x2 = k2 > n2 / 2 ? k2 - n2 : k2, x3 = k3 > n3 / 2 ? k3 - n3 : k3;
    if (p.op == kOpProjStokes) {
        const float lap = static_cast<float>(static_cast<long>(x1) * x1 + static_cast<long>(x2) * x2 +
                                             static_cast<long>(x3) * x3);
        return lap == 0.0f ? -1.0f : static_cast<float>(-1.0 / static_cast<double>(lap));
    }
    const int z1 = k1 == n1 / 2 ? 0 : x1, z2 = k2 == n2 / 2 ? 0 : x2, z3 = k3 == n3 / 2 ? 0 : x3;
    const float lapik = -static_cast<float>(static_cast<long>(z1) * z1 + static_cast<long>(z2) * z2 +
                                            static_cast<long>(z3) * z3);
    const float lapinvik = lapik == 0.0f ? -1.0f : static_cast<float>(1.0 / static_cast<double>(lapik));
    float opik = static_cast<float>(1.0 / (static_cast<double>(p.beta2) * (static_cast<double>(-lapik) + 1.0)));
    opik = static_cast<float>(1.0 / (static_cast<double>(p.beta0) * static_cast<double>(opik) + 1.0));
    return opik * lapinvik;
}

struct Ptr3 {
    const c32* in[3];
    c32* out[3];
};

template <int N, int MODE>
__global__ __launch_bounds__(Geo<N>::THREADS) void k_x1_fused(Ptr3 ptr, ColGeom geo, const c32* __restrict__ tw,
                                                              SpecParams sp, int n2, int n3, int nh) {
    using G = Geo<N>;
    __shared__ c32 s[G::LDS_ELEMS];
    using LF = typename Plan<N>::Fwd;
    constexpr int V = G::V, T = G::T, R0 = First<LF>::value, RL = Last<LF>::value;
    const int c = threadIdx.x % G::B;
    const int t = threadIdx.x / G::B;
    const int64_t cg = static_cast<int64_t>(blockIdx.x) * G::B + c;
    const bool valid = cg < geo.ncols;
    const int64_t base = valid ? cg : 0;
    const int k2 = static_cast<int>(static_cast<uint32_t>(cg) / static_cast<uint32_t>(nh));
    const int k3 = static_cast<int>(cg) - k2 * nh;
    const int w2 = wavenum(k2, n2), w3 = wavenum(k3, n3);

    const uint32_t es8 = static_cast<uint32_t>(geo.estride) * 8u;
    const uint32_t lane = static_cast<uint32_t>(base) * 8u + static_cast<uint32_t>(t) * es8;
    auto load_fwd = [&](const c32* src, c32 (&v)[V]) {
        const spec_rsrc_t r_in = spec_rsrc(src, geo.ncols * N);
#pragma unroll
        for (int q = 0; q < V / R0; ++q)
#pragma unroll
            for (int r = 0; r < R0; ++r) {
                const c32 x = bload_c(r_in, lane, static_cast<uint32_t>(T * q + r * (N / R0)) * es8);
                v[q * R0 + r] = valid ? x : c32{0.f, 0.f};
            }
        fft_regs<N, -1>(v, s, ColLayout{G::LD}, c, t, tw);
    };
    // After the forward pass v[q·RL + r] = X[k1], k1 = j + r·N/RL; the
    // inverse plan starts with radix RL and expects the same layout.
    auto inv_store = [&](c32 (&v)[V], c32* dst) {
        fft_regs<N, +1>(v, s, ColLayout{G::LD}, c, t, tw);
        constexpr int RI = Last<typename Plan<N>::Inv>::value;
        if (valid) {
            const spec_rsrc_t r_out = spec_rsrc(dst, geo.ncols * N);
#pragma unroll
            for (int q = 0; q < V / RI; ++q)
#pragma unroll
                for (int r = 0; r < RI; ++r)
                    bstore_c(r_out, v[q * RI + r], lane, static_cast<uint32_t>(T * q + r * (N / RI)) * es8);
        }
    };

    if constexpr (MODE == kFuseMult) {
        const int d = blockIdx.y;
        c32 v[V];
        load_fwd(ptr.in[d], v);
#pragma unroll
        for (int q = 0; q < V / RL; ++q) {
            const int j = t + T * q;
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int w1 = wavenum(j + r * (N / RL), N);
                const float m = spec_multiplier(sp, w1, w2, w3);
                v[q * RL + r] = {v[q * RL + r].x * m, v[q * RL + r].y * m};
            }
        }
        inv_store(v, ptr.out[d]);
    } else if constexpr (MODE == kFuseGrad) {
        c32 x[V];
        load_fwd(ptr.in[0], x);
        const float sc = sp.scale;
#pragma unroll 1
        for (int d = 0; d < 3; ++d) {
            c32 v[V];
#pragma unroll
            for (int q = 0; q < V / RL; ++q) {
                const int j = t + T * q;
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int w = (d == 0) ? wavenum(j + r * (N / RL), N) : (d == 1 ? w2 : w3);
                    const float wd = static_cast<float>(w);
                    const c32 a = x[q * RL + r];
                    v[q * RL + r] = {-wd * a.y * sc, wd * a.x * sc};
                }
            }
            inv_store(v, ptr.out[d]);
        }
    } else if constexpr (MODE == kFuseDiv) {
        c32 acc[V];
#pragma unroll
        for (int i = 0; i < V; ++i) acc[i] = {0.f, 0.f};
#pragma unroll 1
        for (int d = 0; d < 3; ++d) {
            c32 v[V];
            load_fwd(ptr.in[d], v);
#pragma unroll
            for (int q = 0; q < V / RL; ++q) {
                const int j = t + T * q;
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    const int w = (d == 0) ? wavenum(j + r * (N / RL), N) : (d == 1 ? w2 : w3);
                    const float wd = static_cast<float>(w);
                    const c32 a = v[q * RL + r];
                    // Σ_d i·w_d·a_d, accumulated as the oracle and AccFFT order the terms
                    acc[q * RL + r].x += -wd * a.y;
                    acc[q * RL + r].y += wd * a.x;
                }
            }
        }
        const float sc = sp.scale;
#pragma unroll
        for (int i = 0; i < V; ++i) acc[i] = {acc[i].x * sc, acc[i].y * sc};
        inv_store(acc, ptr.out[0]);
    } else {
        // projection of a vector field: all three components at every mode
        c32 a[V], b[V], e[V];
        load_fwd(ptr.in[0], a);
        load_fwd(ptr.in[1], b);
        load_fwd(ptr.in[2], e);
        const float sc = sp.scale;
        const float g2 = static_cast<float>(w2), g3 = static_cast<float>(w3);
#pragma unroll
        for (int q = 0; q < V / RL; ++q) {
            const int j = t + T * q;
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const int k1 = j + r * (N / RL);
                const float g1 = static_cast<float>(wavenum(k1, N));
                const float m = proj_factor(sp, k1, k2, k3, N, n2, n3);
                const int i = q * RL + r;
                float cre = -sc * (g1 * a[i].x + g2 * b[i].x + g3 * e[i].x);
                float cim = sc * (g1 * a[i].y + g2 * b[i].y + g3 * e[i].y);
                cre *= m;
                cim *= m;
                b[i] = {-g2 * cre, g2 * cim};
                e[i] = {-g3 * cre, g3 * cim};
                a[i] = {cre * -g1, cim * g1};
            }
        }
        inv_store(a, ptr.out[0]);
        inv_store(b, ptr.out[1]);
        inv_store(e, ptr.out[2]);
    }
}

// ---------------------------------------------------------------------------
// Host-side dispatch on the (runtime) transform length.
// ---------------------------------------------------------------------------
template <template <int> class F, class... A>
void dispatch_n(int n, A&&... a) {
    switch (n) {
        case 8: F<8>::go(a...); break;
        case 16: F<16>::go(a...); break;
        case 32: F<32>::go(a...); break;
        case 64: F<64>::go(a...); break;
        case 128: F<128>::go(a...); break;
        case 256: F<256>::go(a...); break;
        case 512: F<512>::go(a...); break;
        case 48: F<48>::go(a...); break;
        case 96: F<96>::go(a...); break;
        case 150: F<150>::go(a...); break;
        case 192: F<192>::go(a...); break;
        default: throw Error(kErrUnsupportedSize, "FFT length " + std::to_string(n) + " not supported");
    }
}

template <int N>
struct RowR2C {
    static void go(const R2CPtrs& p, int ncomp, int64_t nrows, const c32* tw, hipStream_t st) {
        using G = Geo<N>;
        const int64_t npairs = (nrows + 1) / 2;
        const int64_t blocks = (npairs + G::B - 1) / G::B;
        hipLaunchKernelGGL(k_row_r2c<N>, dim3(static_cast<unsigned>(blocks), ncomp), dim3(G::THREADS), 0, st, p, nrows,
                           tw);
    }
};

template <int N>
struct RowC2R {
    static void go(const C2RPtrs& p, int ncomp, int64_t nrows, const c32* tw, float alpha, float beta, hipStream_t st) {
        using G = Geo<N>;
        const int64_t npairs = (nrows + 1) / 2;
        const int64_t blocks = (npairs + G::B - 1) / G::B;
        const dim3 grid(static_cast<unsigned>(blocks), ncomp);
        if (p.add[0] || p.add[1] || p.add[2])
            hipLaunchKernelGGL((k_row_c2r<N, true>), grid, dim3(G::THREADS), 0, st, p, nrows, tw, alpha, beta);
        else
            hipLaunchKernelGGL((k_row_c2r<N, false>), grid, dim3(G::THREADS), 0, st, p, nrows, tw, alpha, beta);
    }
};

template <int N>
struct ColPass {
    static void go(const ColPtrs& p, int ncomp, ColGeom geo, int dir, const c32* tw, hipStream_t st) {
        using G = ColGeo<N>;
        CLR_REQUIRE(geo.ncols * N * 8 < (int64_t(1) << 31), kErrInvalidArgument,
                    "spectrum field too large for 32-bit buffer offsets");
        const dim3 grid(static_cast<unsigned>((geo.ncols + G::B - 1) / G::B), ncomp);
        if (dir < 0)
            hipLaunchKernelGGL((k_col<N, -1>), grid, dim3(G::THREADS), 0, st, p, geo, tw);
        else
            hipLaunchKernelGGL((k_col<N, +1>), grid, dim3(G::THREADS), 0, st, p, geo, tw);
    }
};

ColPtrs col_ptrs(c32* const data[], int ncomp) {
    CLR_REQUIRE(ncomp >= 1 && ncomp <= 3, kErrInvalidArgument, "1 to 3 fields per FFT pass");
    ColPtrs p{};
    for (int c = 0; c < ncomp; ++c) p.in[c] = p.out[c] = data[c];
    return p;
}

template <int N>
struct X1Fused {
    static void go(int mode, Ptr3 p, ColGeom geo, const c32* tw, SpecParams sp, int n2, int n3, int nh, int ncomp,
                   hipStream_t st) {
        using G = Geo<N>;
        CLR_REQUIRE(geo.ncols * N * 8 < (int64_t(1) << 31), kErrInvalidArgument,
                    "spectrum field too large for 32-bit buffer offsets");
        const int64_t blocks = (geo.ncols + G::B - 1) / G::B;
        dim3 grid(static_cast<unsigned>(blocks), mode == kFuseMult ? ncomp : 1);
        switch (mode) {
            case kFuseMult:
                hipLaunchKernelGGL((k_x1_fused<N, kFuseMult>), grid, dim3(G::THREADS), 0, st, p, geo, tw, sp, n2, n3, nh);
                break;
            case kFuseGrad:
                hipLaunchKernelGGL((k_x1_fused<N, kFuseGrad>), grid, dim3(G::THREADS), 0, st, p, geo, tw, sp, n2, n3, nh);
                break;
            case kFuseProject:
                hipLaunchKernelGGL((k_x1_fused<N, kFuseProject>), grid, dim3(G::THREADS), 0, st, p, geo, tw, sp, n2, n3,
                                   nh);
                break;
            default:
                hipLaunchKernelGGL((k_x1_fused<N, kFuseDiv>), grid, dim3(G::THREADS), 0, st, p, geo, tw, sp, n2, n3, nh);
                break;
        }
    }
};

ColGeom geom_x2(const Dims& d) {
    // columns (i1, k3): cg = i1·nh + k3, start i1·n2·nh + k3, stride nh
    return ColGeom{static_cast<int64_t>(d.n[0]) * d.nh, d.nh, static_cast<int64_t>(d.n[1]) * d.nh, d.nh};
}
ColGeom geom_x1(const Dims& d) {
    // columns (k2, k3): cg = k2·nh + k3, start cg, stride n2·nh
    const int64_t plane = static_cast<int64_t>(d.n[1]) * d.nh;
    return ColGeom{plane, plane, 0, plane};
}

}  // namespace

bool fft_size_supported(int n) {
    return n == 8 || n == 16 || n == 32 || n == 64 || n == 128 || n == 256 || n == 512 || n == 48 || n == 96 ||
           n == 150 || n == 192;
}

void fft_twiddles(int n, c32* h) {
    for (int k = 0; k < n; ++k) {
        const double ang = -2.0 * 3.14159265358979323846 * static_cast<double>(k) / static_cast<double>(n);
        h[k] = {static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))};
    }
}

namespace {
void require_fft_sizes(const Dims& d) {
    for (int a = 0; a < 3; ++a)
        CLR_REQUIRE(fft_size_supported(d.n[a]), kErrUnsupportedSize,
                    "grid size " + std::to_string(d.n[a]) +
                        " not supported (powers of two 8..512, and 48, 96, 150, 192)");
}
}  // namespace

FFTPlan3D::FFTPlan3D(const Dims& d) : dims(d) {
    require_fft_sizes(d);
    for (int a = 0; a < 3; ++a) {
        std::vector<c32> h(d.n[a]);
        fft_twiddles(d.n[a], h.data());
        CLR_HIP_CHECK(hipMalloc(&tw[a], sizeof(c32) * d.n[a]));
        CLR_HIP_CHECK(hipMemcpy(tw[a], h.data(), sizeof(c32) * d.n[a], hipMemcpyHostToDevice));
    }
}

FFTPlan3D::FFTPlan3D(const Dims& d, c32* const tw_ext[3]) : dims(d), owns_tw(false) {
    require_fft_sizes(d);
    for (int a = 0; a < 3; ++a) {
        CLR_REQUIRE(tw_ext[a] != nullptr, kErrInvalidArgument, "missing twiddle table");
        tw[a] = tw_ext[a];
    }
}

FFTPlan3D::~FFTPlan3D() {
    if (!owns_tw) return;
    for (int a = 0; a < 3; ++a)
        if (tw[a]) (void)hipFree(tw[a]);
}

void FFTPlan3D::r2c(const float* in, c32* out, hipStream_t st) const {
    x3_r2c(in, out, st);
    x2(out, -1, st);
    x1(out, -1, st);
}

void FFTPlan3D::c2r(c32* in, float* out, hipStream_t st, float alpha, const float* add, float beta) const {
    x1(in, +1, st);
    x2(in, +1, st);
    x3_c2r(in, out, st, alpha, add, beta);
}

void FFTPlan3D::x3_r2c_n(const float* const in[], c32* const out[], int ncomp, hipStream_t st) const {
    CLR_REQUIRE(ncomp >= 1 && ncomp <= 3, kErrInvalidArgument, "1 to 3 fields per FFT pass");
    R2CPtrs p{};
    for (int c = 0; c < ncomp; ++c) {
        p.in[c] = in[c];
        p.out[c] = out[c];
    }
    dispatch_n<RowR2C>(dims.n[2], p, ncomp, static_cast<int64_t>(dims.n[0]) * dims.n[1], tw[2], st);
}
void FFTPlan3D::x3_c2r_n(const c32* const in[], float* const out[], int ncomp, hipStream_t st, float alpha,
                         const float* const add[], float beta) const {
    CLR_REQUIRE(ncomp >= 1 && ncomp <= 3, kErrInvalidArgument, "1 to 3 fields per FFT pass");
    C2RPtrs p{};
    for (int c = 0; c < ncomp; ++c) {
        p.in[c] = in[c];
        p.out[c] = out[c];
        p.add[c] = add ? add[c] : nullptr;
    }
    dispatch_n<RowC2R>(dims.n[2], p, ncomp, static_cast<int64_t>(dims.n[0]) * dims.n[1], tw[2], alpha, beta, st);
}
void FFTPlan3D::x2_n(c32* const data[], int ncomp, int dir, hipStream_t st) const {
    dispatch_n<ColPass>(dims.n[1], col_ptrs(data, ncomp), ncomp, geom_x2(dims), dir, tw[1], st);
}

void FFTPlan3D::x3_r2c(const float* in, c32* out, hipStream_t st) const {
    const float* i[1] = {in};
    c32* o[1] = {out};
    x3_r2c_n(i, o, 1, st);
}
void FFTPlan3D::x3_c2r(const c32* in, float* out, hipStream_t st, float alpha, const float* add, float beta) const {
    const c32* i[1] = {in};
    float* o[1] = {out};
    const float* a[1] = {add};
    x3_c2r_n(i, o, 1, st, alpha, a, beta);
}
void FFTPlan3D::x2(c32* data, int dir, hipStream_t st) const {
    c32* d[1] = {data};
    x2_n(d, 1, dir, st);
}
void FFTPlan3D::x1(c32* data, int dir, hipStream_t st) const {
    c32* d[1] = {data};
    dispatch_n<ColPass>(dims.n[0], col_ptrs(d, 1), 1, geom_x1(dims), dir, tw[0], st);
}

void FFTPlan3D::x1_multiplier(const c32* const in[3], c32* const out[3], int ncomp, const SpecParams& sp,
                              hipStream_t st) const {
    Ptr3 p{};
    for (int d = 0; d < ncomp; ++d) {
        p.in[d] = in[d];
        p.out[d] = out[d];
    }
    dispatch_n<X1Fused>(dims.n[0], static_cast<int>(kFuseMult), p, geom_x1(dims), tw[0], sp, dims.n[1], dims.n[2],
                        dims.nh, ncomp, st);
}

void FFTPlan3D::x1_grad(const c32* in, c32* const out[3], hipStream_t st) const {
    Ptr3 p{};
    p.in[0] = in;
    for (int d = 0; d < 3; ++d) p.out[d] = out[d];
    SpecParams sp{};
    sp.scale = dims.scale;
    dispatch_n<X1Fused>(dims.n[0], static_cast<int>(kFuseGrad), p, geom_x1(dims), tw[0], sp, dims.n[1], dims.n[2],
                        dims.nh, 3, st);
}

void FFTPlan3D::x1_project(c32* const data[3], const SpecParams& sp, hipStream_t st) const {
    CLR_REQUIRE(sp.op == kOpProjStokes || sp.op == kOpProjRelaxed, kErrInvalidArgument, "not a projection");
    Ptr3 p{};
    for (int d = 0; d < 3; ++d) p.in[d] = p.out[d] = data[d];
    dispatch_n<X1Fused>(dims.n[0], static_cast<int>(kFuseProject), p, geom_x1(dims), tw[0], sp, dims.n[1], dims.n[2],
                        dims.nh, 3, st);
}

void FFTPlan3D::x1_div(const c32* const in[3], c32* out, hipStream_t st) const {
    Ptr3 p{};
    for (int d = 0; d < 3; ++d) p.in[d] = in[d];
    p.out[0] = out;
    SpecParams sp{};
    sp.scale = dims.scale;
    dispatch_n<X1Fused>(dims.n[0], static_cast<int>(kFuseDiv), p, geom_x1(dims), tw[0], sp, dims.n[1], dims.n[2],
                        dims.nh, 3, st);
}

}  // namespace claire_amd
