// claire_oracle.hpp -- CPU oracle for the CLAIRE registration hot path.
//
// TEST INFRASTRUCTURE ONLY.  This is a plain C++17 restatement of the
// reference's algorithm, templated on the real type: instantiated with
// double it is the accuracy oracle, instantiated with float it is the
// single-precision "reference CPU path" (the reference's FP32 PETSc build
// restated, OpenMP over grid points, with the reference's OpenMP races on
// shared temporaries privatized).  Only tests, the smoke check and the CPU
// baseline leg of the benchmark may load it, and only as the checker.  The
// product library never links it.
//
// Every function cites the reference lines it restates (paths relative to
// the reference repository root).
//
// Parity pins (see tests/test_oracle_*.py):
//   fft / r2c / c2r ..... numpy.fft (pocketfft, an independent library) and
//                         brute-force DFT on tiny grids
//   grad / div .......... closed forms on trigonometric polynomials,
//                         Nyquist mode mapped to zero
//   reg operators ....... Fourier-mode eigenpairs, inverse∘apply = hd·I / I
//   interpolation ....... partition of unity, node reproduction, exactness
//                         for per-axis cubics, integer shifts, brute force
//   trajectory .......... constant velocity closed form
//   transport solves .... v = 0 closed forms, constant-v translation
//   Hessian matvec ...... v = 0 closed form, symmetry, linearity, Taylor
//                         test against the gradient
#pragma once

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>
#include <algorithm>
#include <array>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace oracle {

constexpr double kPi = 3.14159265358979323846264338327950288;

// ---------------------------------------------------------------------------
// Grid conventions.
//   hx[i] = 2π/nx[i]                         src/RegOpt.cpp:935
//   hd    = hx0·hx1·hx2                      include/RegOpt.hpp:485-489
//   ht    = 1/nt (horizon [0,1])             include/RegOpt.hpp:497-500
//   scale = 1/(n0·n1·n2)                     src/RegOpt.cpp:2505-2512
//   flat index l = i1·n2·n3 + i2·n3 + i3     include/CLAIREUtils.hpp:198-212
// ---------------------------------------------------------------------------
template <class T>
struct Grid {
    int n[3];
    int64_t nl;      // number of real grid points
    int nh;          // n[2]/2 + 1 (half-spectrum length along x3)
    int64_t nlc;     // number of complex coefficients n0·n1·nh
    T hx[3];
    T hd;
    T scale;
    explicit Grid(const int nx[3]) {
        for (int i = 0; i < 3; ++i) {
            n[i] = nx[i];
            if (nx[i] < 2) throw std::invalid_argument("grid size must be >= 2");
            hx[i] = static_cast<T>(kPi * 2.0 / static_cast<double>(static_cast<T>(nx[i])));
        }
        nl = static_cast<int64_t>(n[0]) * n[1] * n[2];
        nh = n[2] / 2 + 1;
        nlc = static_cast<int64_t>(n[0]) * n[1] * nh;
        hd = hx[0] * hx[1] * hx[2];
        T s = T(1);
        for (int i = 0; i < 3; ++i) s *= static_cast<T>(n[i]);
        scale = T(1) / s;
    }
};

// Wave number with the Nyquist mode mapped to zero.
// include/CLAIREUtils.hpp:248-255 (ComputeWaveNumber)
inline long wavenumber(long k, long n) {
    if (k > n / 2) return k - n;
    if (k == n / 2) return 0;
    return k;
}

// ---------------------------------------------------------------------------
// FFT.  The reference calls AccFFT (github.com/amirgholami/accfft, master,
// not vendored; deps/get_libs.sh:8) on top of FFTW 3.3.6-pl2
// (deps/get_libs.sh:16).  Both compute the unnormalized DFT
//   X[k] = Σ_n x[n]·exp(-2πi·k·n/N)  (forward),  exp(+2πi…) (backward).
// This is an own recursive mixed-radix decimation-in-time Cooley–Tukey
// transform with twiddles generated in double precision; it is pinned against
// numpy.fft in the tests.
// ---------------------------------------------------------------------------
template <class T>
class FFT1D {
  public:
    using C = std::complex<T>;
    explicit FFT1D(int n) : n_(n) {
        int m = n;
        for (int p : {4, 2, 3, 5}) {
            while (m % p == 0) {
                factors_.push_back(p);
                m /= p;
            }
        }
        for (int p = 7; m > 1; p += 2) {
            while (m % p == 0) { factors_.push_back(p); m /= p; }
        }
        tw_.resize(n);
        for (int k = 0; k < n; ++k) {
            double a = -2.0 * kPi * static_cast<double>(k) / static_cast<double>(n);
            tw_[k] = C(static_cast<T>(std::cos(a)), static_cast<T>(std::sin(a)));
        }
    }
    int size() const { return n_; }
    // out[k] = Σ in[j·is]·w^{jk}; sign -1 forward, +1 backward.
    void exec(const C* in, int64_t is, C* out, int sign) const {
        rec(in, is, out, n_, 0, 1, sign);
    }

  private:
    C twid(int64_t idx, int sign) const {
        C w = tw_[static_cast<size_t>(idx % n_)];
        return sign < 0 ? w : std::conj(w);
    }
    void rec(const C* in, int64_t is, C* out, int n, size_t fi, int64_t twstep, int sign) const {
        if (n == 1) { out[0] = in[0]; return; }
        const int p = factors_[fi];
        const int m = n / p;
        for (int q = 0; q < p; ++q) rec(in + q * is, is * p, out + q * m, m, fi + 1, twstep * p, sign);
        C t[16];
        C* tp = t;
        std::vector<C> big;
        if (p > 16) { big.resize(p); tp = big.data(); }
        for (int k = 0; k < m; ++k) {
            for (int q = 0; q < p; ++q) {
                C y = out[q * m + k];
                tp[q] = (q == 0 || k == 0) ? y : y * twid(static_cast<int64_t>(q) * k * twstep, sign);
            }
            const int64_t pstep = static_cast<int64_t>(n_) / p;
            for (int s = 0; s < p; ++s) {
                C acc = tp[0];
                for (int q = 1; q < p; ++q) acc += tp[q] * twid(static_cast<int64_t>((q * s) % p) * pstep, sign);
                out[k + m * s] = acc;
            }
        }
    }
    int n_;
    std::vector<int> factors_;
    std::vector<C> tw_;
};

// 3D real-to-complex transform, unnormalized; output layout [k1][k2][k3],
// k3 in [0, n3/2].  (AccFFT single-rank layout, src/RegularizationH1.cpp:212-233)
template <class T>
void r2c3d(const Grid<T>& g, const T* in, std::complex<T>* out) {
    using C = std::complex<T>;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2], nh = g.nh;
    FFT1D<T> f0(n0), f1(n1), f2(n2);
#pragma omp parallel
    {
        std::vector<C> a(std::max({n0, n1, n2})), b(std::max({n0, n1, n2}));
#pragma omp for collapse(2) schedule(static)
        for (int i0 = 0; i0 < n0; ++i0)
            for (int i1 = 0; i1 < n1; ++i1) {
                const T* row = in + (static_cast<int64_t>(i0) * n1 + i1) * n2;
                for (int i = 0; i < n2; ++i) a[i] = C(row[i], T(0));
                f2.exec(a.data(), 1, b.data(), -1);
                C* o = out + (static_cast<int64_t>(i0) * n1 + i1) * nh;
                for (int k = 0; k < nh; ++k) o[k] = b[k];
            }
#pragma omp for collapse(2) schedule(static)
        for (int i0 = 0; i0 < n0; ++i0)
            for (int k = 0; k < nh; ++k) {
                C* base = out + static_cast<int64_t>(i0) * n1 * nh + k;
                for (int i = 0; i < n1; ++i) a[i] = base[static_cast<int64_t>(i) * nh];
                f1.exec(a.data(), 1, b.data(), -1);
                for (int i = 0; i < n1; ++i) base[static_cast<int64_t>(i) * nh] = b[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int k = 0; k < nh; ++k) {
                C* base = out + static_cast<int64_t>(i1) * nh + k;
                const int64_t st = static_cast<int64_t>(n1) * nh;
                for (int i = 0; i < n0; ++i) a[i] = base[i * st];
                f0.exec(a.data(), 1, b.data(), -1);
                for (int i = 0; i < n0; ++i) base[i * st] = b[i];
            }
    }
}

// 3D complex-to-real transform, unnormalized.  Like FFTW's c2r, the
// imaginary parts of the k3 = 0 and k3 = n3/2 coefficients are discarded in
// the last (x3) stage.  `in` is not modified.
template <class T>
void c2r3d(const Grid<T>& g, const std::complex<T>* in, T* out) {
    using C = std::complex<T>;
    const int n0 = g.n[0], n1 = g.n[1], n2 = g.n[2], nh = g.nh;
    FFT1D<T> f0(n0), f1(n1), f2(n2);
    std::vector<C> w(in, in + g.nlc);
#pragma omp parallel
    {
        std::vector<C> a(std::max({n0, n1, n2})), b(std::max({n0, n1, n2}));
#pragma omp for collapse(2) schedule(static)
        for (int i1 = 0; i1 < n1; ++i1)
            for (int k = 0; k < nh; ++k) {
                C* base = w.data() + static_cast<int64_t>(i1) * nh + k;
                const int64_t st = static_cast<int64_t>(n1) * nh;
                for (int i = 0; i < n0; ++i) a[i] = base[i * st];
                f0.exec(a.data(), 1, b.data(), +1);
                for (int i = 0; i < n0; ++i) base[i * st] = b[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (int i0 = 0; i0 < n0; ++i0)
            for (int k = 0; k < nh; ++k) {
                C* base = w.data() + static_cast<int64_t>(i0) * n1 * nh + k;
                for (int i = 0; i < n1; ++i) a[i] = base[static_cast<int64_t>(i) * nh];
                f1.exec(a.data(), 1, b.data(), +1);
                for (int i = 0; i < n1; ++i) base[static_cast<int64_t>(i) * nh] = b[i];
            }
#pragma omp for collapse(2) schedule(static)
        for (int i0 = 0; i0 < n0; ++i0)
            for (int i1 = 0; i1 < n1; ++i1) {
                const C* h = w.data() + (static_cast<int64_t>(i0) * n1 + i1) * nh;
                for (int k = 0; k < nh; ++k) a[k] = h[k];
                a[0] = C(a[0].real(), T(0));
                if (n2 % 2 == 0) a[n2 / 2] = C(a[n2 / 2].real(), T(0));
                for (int k = nh; k < n2; ++k) a[k] = std::conj(a[n2 - k]);
                f2.exec(a.data(), 1, b.data(), +1);
                T* row = out + (static_cast<int64_t>(i0) * n1 + i1) * n2;
                for (int i = 0; i < n2; ++i) row[i] = b[i].real();
            }
    }
}

// Visit every half-spectrum coefficient with its (Nyquist-zeroed) wave numbers.
template <class T, class F>
void for_each_mode(const Grid<T>& g, F&& fn) {
    const int n0 = g.n[0], n1 = g.n[1], nh = g.nh;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k0 = 0; k0 < n0; ++k0)
        for (int k1 = 0; k1 < n1; ++k1)
            for (int k2 = 0; k2 < nh; ++k2) {
                long w[3] = {wavenumber(k0, g.n[0]), wavenumber(k1, g.n[1]), wavenumber(k2, g.n[2])};
                int64_t i = (static_cast<int64_t>(k0) * n1 + k1) * nh + k2;
                fn(i, w);
            }
}

// Spectral gradient ∇f = F⁻¹(i·w·f̂)·scale (AccFFT accfft_grad_t; call sites
// e.g. src/CLAIRE.cpp:2622,3243; one r2c and three c2r, include/RegOpt.hpp:183-186).
template <class T>
void grad(const Grid<T>& g, const T* f, T* g1, T* g2, T* g3) {
    using C = std::complex<T>;
    std::vector<C> fh(g.nlc), gh(g.nlc);
    r2c3d(g, f, fh.data());
    T* outs[3] = {g1, g2, g3};
    for (int d = 0; d < 3; ++d) {
        const T s = g.scale;
        for_each_mode(g, [&](int64_t i, const long* w) {
            const T wd = static_cast<T>(w[d]);
            const C v = fh[i];
            gh[i] = C(-wd * v.imag() * s, wd * v.real() * s);
        });
        c2r3d(g, gh.data(), outs[d]);
    }
}

// Spectral divergence ∇·v = F⁻¹(Σ_j i·w_j·v̂_j)·scale (AccFFT
// accfft_divergence_t; e.g. src/CLAIRE.cpp:2622,3911; three r2c and one c2r).
template <class T>
void div(const Grid<T>& g, const T* v1, const T* v2, const T* v3, T* d) {
    using C = std::complex<T>;
    std::vector<C> a(g.nlc), b(g.nlc), c(g.nlc), acc(g.nlc);
    r2c3d(g, v1, a.data());
    r2c3d(g, v2, b.data());
    r2c3d(g, v3, c.data());
    const T s = g.scale;
    for_each_mode(g, [&](int64_t i, const long* w) {
        const T w0 = static_cast<T>(w[0]), w1 = static_cast<T>(w[1]), w2 = static_cast<T>(w[2]);
        const T re = -w0 * a[i].imag() - w1 * b[i].imag() - w2 * c[i].imag();
        const T im = w0 * a[i].real() + w1 * b[i].real() + w2 * c[i].real();
        acc[i] = C(re * s, im * s);
    });
    c2r3d(g, acc.data(), d);
}

// ---------------------------------------------------------------------------
// Registration models (RegOpt m_RegModel; CLAIREInterface picks the class,
// src/CLAIREInterface.cpp:685-690): 0 COMPRESSIBLE (CLAIRE), 1 STOKES
// (CLAIREStokes: pure-transport adjoints, projected body forces), 2
// RELAXEDSTOKES (CLAIREDivReg: compressible adjoints, body forces projected
// with a β-weighted multiplier, divergence penalty in the objective).  Test
// infrastructure sets it per call.
// ---------------------------------------------------------------------------
enum RegModel { kCompressible = 0, kStokes = 1, kRelaxedStokes = 2 };

inline int& reg_model() {
    static int model = kCompressible;
    return model;
}

// Hessian operator (m_KrylovMethod.matvectype): 0 DEFAULTMATVEC (HessMatVec),
// 1 PRECONDMATVEC (PrecondHessMatVec), 2 PRECONDMATVECSYM
// (PrecondHessMatVecSym).  Test infrastructure sets it per call.
inline int& matvec_type() {
    static int t = 0;
    return t;
}

// β of the model's projection (m_RegNorm.beta: β0 and, for RELAXEDSTOKES,
// the divergence weight β2); the same values as the regularization's.
inline double* reg_model_beta() {
    static double b[3] = {0.0, 0.0, 0.0};
    return b;
}

// ApplyProjection, in place: b ← b + F⁻¹[x̂] with, per mode,
//   ĉ = scale·m·(g·b̂),  x̂_i = g_i·ĉ,
// where g is the Nyquist-zeroed wave vector and
//   STOKES (src/CLAIREStokes.cpp:446-513)        m = −1/|w|² (w with the
//          Nyquist index kept at +n/2), m = −1 at w = 0;
//   RELAXEDSTOKES (src/CLAIREDivReg.cpp:354-430) m = opik·lapinvik with
//          lapik = −|g|², lapinvik = 1/lapik (−1 at 0),
//          opik = 1/(β0·(1/(β2·(|g|² + 1))) + 1).
// The statement order and the double literals follow the reference: the
// real part is formed as −scale·Σ g_i·Re b̂_i and each output real part as
// −g_i·ĉ_re, which is g_i·scale·m·Σ … .
template <class T>
void project(const Grid<T>& g, int model, const T* beta, T* const b[3]) {
    if (model == kCompressible) return;
    using C = std::complex<T>;
    std::vector<C> h0(g.nlc), h1(g.nlc), h2(g.nlc);
    r2c3d(g, b[0], h0.data());
    r2c3d(g, b[1], h1.data());
    r2c3d(g, b[2], h2.data());
    const T scale = g.scale;
    const int n0 = g.n[0], n1 = g.n[1], nh = g.nh;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k0 = 0; k0 < n0; ++k0)
        for (int k1 = 0; k1 < n1; ++k1)
            for (int k2 = 0; k2 < nh; ++k2) {
                const long kk[3] = {k0, k1, k2};
                long wx[3];
                for (int a = 0; a < 3; ++a) wx[a] = kk[a] > g.n[a] / 2 ? kk[a] - g.n[a] : kk[a];
                T m;
                if (model == kStokes) {
                    T lapinvik = static_cast<T>(wx[0] * wx[0] + wx[1] * wx[1] + wx[2] * wx[2]);
                    lapinvik = lapinvik == T(0) ? T(-1) : static_cast<T>(-1.0 / lapinvik);
                    m = lapinvik;
                } else {
                    long wz[3];
                    for (int a = 0; a < 3; ++a) wz[a] = kk[a] == g.n[a] / 2 ? 0 : wx[a];
                    const T lapik = -static_cast<T>(wz[0] * wz[0] + wz[1] * wz[1] + wz[2] * wz[2]);
                    const T lapinvik = lapik == T(0) ? T(-1) : static_cast<T>(1.0 / lapik);
                    T opik = static_cast<T>(1.0 / (beta[2] * (-lapik + 1.0)));
                    opik = static_cast<T>(1.0 / (beta[0] * opik + 1.0));
                    m = opik * lapinvik;
                }
                T gr[3];
                for (int a = 0; a < 3; ++a) gr[a] = static_cast<T>(kk[a] == g.n[a] / 2 ? 0 : wx[a]);
                const int64_t i = (static_cast<int64_t>(k0) * n1 + k1) * nh + k2;
                const C x1 = h0[i], x2 = h1[i], x3 = h2[i];
                T cre = -scale * (gr[0] * x1.real() + gr[1] * x2.real() + gr[2] * x3.real());
                T cim = scale * (gr[0] * x1.imag() + gr[1] * x2.imag() + gr[2] * x3.imag());
                cre *= m;
                cim *= m;
                h1[i] = C(-gr[1] * cre, gr[1] * cim);
                h2[i] = C(-gr[2] * cre, gr[2] * cim);
                h0[i] = C(cre * -gr[0], cim * gr[0]);
            }
    std::vector<T> x(g.nl);
    C* hs[3] = {h0.data(), h1.data(), h2.data()};
    for (int d = 0; d < 3; ++d) {
        c2r3d(g, hs[d], x.data());
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < g.nl; ++l) b[d][l] += x[l];
    }
}

// K[b]·hd, the end of SolveAdjointEquation / SolveIncAdjointEquation
// (src/CLAIRE.cpp:2337-2341, 3467-3471): the model's projection, then the
// scaling by hd.
template <class T>
void project_and_scale(const Grid<T>& g, T* const b[3]) {
    const T beta[3] = {static_cast<T>(reg_model_beta()[0]), static_cast<T>(reg_model_beta()[1]),
                       static_cast<T>(reg_model_beta()[2])};
    project(g, reg_model(), beta, b);
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < g.nl; ++i) b[d][i] *= g.hd;
}

// ---------------------------------------------------------------------------
// Regularization operators (per-mode multipliers on each velocity component).
// ---------------------------------------------------------------------------
enum RegNorm { kH1 = 0, kH2 = 1, kH1SN = 2, kH2SN = 3 };

// First variation / second variation (identical; HessianMatVec calls
// EvaluateGradient, src/RegularizationH1.cpp:277-303).
//   H1   src/RegularizationH1.cpp:222-229     regop = hd·scale·β0·(−lapik + β1)
//   H2   src/RegularizationH2.cpp:236-239     regop = scale·β0·(lapik² + β1)   (no hd)
//   H1SN src/RegularizationH1SN.cpp:210-217   regop = −hd·scale·β·lapik
//   H2SN src/RegularizationH2SN.cpp:241-244   regop = hd·scale·β·lapik²
template <class T>
T reg_apply_op(int norm, const long* w, T hd, T scale, const T* beta) {
    const T lapik = -static_cast<T>(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    switch (norm) {
        case kH1:   return hd * scale * beta[0] * (-lapik + beta[1]);
        case kH2:   return scale * beta[0] * (lapik * lapik + beta[1]);
        case kH1SN: return -hd * scale * beta[0] * lapik;
        case kH2SN: return hd * scale * beta[0] * (lapik * lapik);
    }
    throw std::invalid_argument("unknown regularization norm");
}

// Inverse (sqrt = true: inverse square root).
//   H1   src/RegularizationH1.cpp:373-382
//   H2   src/RegularizationH2.cpp:375-381
//   H1SN src/RegularizationH1SN.cpp:352-359   (β at w = 0)
//   H2SN src/RegularizationH2SN.cpp:385-390   (β at w = 0)
template <class T>
T reg_inverse_op(int norm, const long* w, T scale, const T* beta, bool applysqrt) {
    const T lapik = -static_cast<T>(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    T regop;
    switch (norm) {
        case kH1:   regop = beta[0] * (-lapik + beta[1]); break;
        case kH2:   regop = beta[0] * (lapik * lapik + beta[1]); break;
        case kH1SN: regop = (std::fabs(lapik) == T(0)) ? beta[0] : -beta[0] * lapik; break;
        case kH2SN: regop = (std::fabs(lapik) == T(0)) ? beta[0] : beta[0] * (lapik * lapik); break;
        default: throw std::invalid_argument("unknown regularization norm");
    }
    if (applysqrt) regop = std::sqrt(regop);
    return scale / regop;
}

// Apply a per-mode real multiplier to the three components of a vector field.
template <class T, class Op>
void apply_multiplier3(const Grid<T>& g, const T* const in[3], T* const out[3], Op&& op) {
    using C = std::complex<T>;
    std::vector<C> h(g.nlc);
    std::vector<T> mult(g.nlc);
    for_each_mode(g, [&](int64_t i, const long* w) { mult[i] = op(w); });
    for (int d = 0; d < 3; ++d) {
        r2c3d(g, in[d], h.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < g.nlc; ++i) h[i] *= mult[i];
        c2r3d(g, h.data(), out[d]);
    }
}

// Regularization::EvaluateGradient / HessianMatVec.  β0 == 0 gives zero
// (src/RegularizationH1.cpp:188-190).
template <class T>
void reg_apply(const Grid<T>& g, int norm, const T* beta, const T* const in[3], T* const out[3]) {
    if (beta[0] == T(0)) {
        for (int d = 0; d < 3; ++d) std::fill(out[d], out[d] + g.nl, T(0));
        return;
    }
    apply_multiplier3(g, in, out, [&](const long* w) { return reg_apply_op<T>(norm, w, g.hd, g.scale, beta); });
}

// Regularization::ApplyInverse.  β0 == 0 copies (src/RegularizationH1.cpp:336-340).
template <class T>
void reg_inverse(const Grid<T>& g, int norm, const T* beta, const T* const in[3], T* const out[3], bool applysqrt) {
    if (beta[0] == T(0)) {
        for (int d = 0; d < 3; ++d) std::copy(in[d], in[d] + g.nl, out[d]);
        return;
    }
    apply_multiplier3(g, in, out, [&](const long* w) { return reg_inverse_op<T>(norm, w, g.scale, beta, applysqrt); });
}

template <class T>
double dot(const T* a, const T* b, int64_t n) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < n; ++i) s += static_cast<double>(a[i]) * static_cast<double>(b[i]);
    return s;
}

// Regularization::EvaluateFunctional.
//   H1   src/RegularizationH1.cpp:64-156   ½·hd·β0·(Σ‖∂_j v_i‖² + β1·‖v‖²)
//   H2   src/RegularizationH2.cpp:64-170   ½·hd·Σ‖F⁻¹(scale·√β0·(lapik+√β1)·v̂)‖²
//   H1SN src/RegularizationH1SN.cpp:64-150 ½·hd·β·Σ‖∂_j v_i‖²
//   H2SN src/RegularizationH2SN.cpp:66-170 ½·β·hd·‖Δv‖²
template <class T>
double reg_functional(const Grid<T>& g, int norm, const T* beta, const T* const v[3]) {
    if (beta[0] == T(0)) return 0.0;
    std::vector<T> a(g.nl), b(g.nl), c(g.nl);
    double acc = 0.0;
    if (norm == kH1 || norm == kH1SN) {
        for (int d = 0; d < 3; ++d) {
            grad(g, v[d], a.data(), b.data(), c.data());
            acc += dot(a.data(), a.data(), g.nl) + dot(b.data(), b.data(), g.nl) + dot(c.data(), c.data(), g.nl);
        }
        if (norm == kH1) {
            double l2 = 0.0;
            for (int d = 0; d < 3; ++d) l2 += dot(v[d], v[d], g.nl);
            return 0.5 * g.hd * beta[0] * (acc + beta[1] * l2);
        }
        return 0.5 * g.hd * beta[0] * acc;
    }
    T* outs[3] = {a.data(), b.data(), c.data()};
    if (norm == kH2) {
        const T sb0 = std::sqrt(beta[0]), sb1 = std::sqrt(beta[1]);
        apply_multiplier3(g, v, outs, [&](const long* w) {
            const T lapik = -static_cast<T>(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            return g.scale * sb0 * (lapik + sb1);
        });
        for (int d = 0; d < 3; ++d) acc += dot(outs[d], outs[d], g.nl);
        return 0.5 * g.hd * acc;
    }
    apply_multiplier3(g, v, outs, [&](const long* w) {
        const T lapik = -static_cast<T>(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
        return g.scale * lapik;
    });
    for (int d = 0; d < 3; ++d) acc += dot(outs[d], outs[d], g.nl);
    return 0.5 * beta[0] * g.hd * acc;
}

// Regularization::GetExtremeEigValsInvOp.
//   H1   src/RegularizationH1.cpp:427-453, H2 src/RegularizationH2.cpp:426-450,
//   H1SN src/RegularizationH1SN.cpp:406-430, H2SN src/RegularizationH2SN.cpp:435-460
template <class T>
void reg_eigvals_inv(const Grid<T>& g, int norm, const T* beta, T* emin, T* emax) {
    T w[3];
    for (int i = 0; i < 3; ++i) w[i] = static_cast<T>(g.n[i]) / T(2);
    T regop = -(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    switch (norm) {
        case kH1:   regop = beta[0] * (-regop + beta[1]); *emin = T(1) / regop; *emax = T(1) / beta[1]; break;
        case kH2:   regop = beta[0] * (regop * regop) + beta[1]; *emin = T(1) / regop; *emax = T(1) / beta[1]; break;
        case kH1SN: regop = -beta[0] * regop + beta[1]; *emin = T(1) / regop; *emax = T(1) / beta[1]; break;
        // emax = 1 for H2SN: the operator has no constant term ("by definition;
        // it's 1/0", src/RegularizationH2SN.cpp:453)
        case kH2SN: regop = beta[0] * (regop * regop); *emin = T(1) / regop; *emax = T(1); break;
        default: throw std::invalid_argument("unknown regularization norm");
    }
}

// ---------------------------------------------------------------------------
// Semi-Lagrangian interpolation.
// ---------------------------------------------------------------------------
// Periodic wrap of normalized query coordinates into (−1/N, 1)
// (Interp3_Plan::fast_scatter, src/Interp3_Plan.cpp:328-362).
template <class T>
inline T wrap_coord(T q, T h) {
    int guard = 0;
    while (q <= -h && guard++ < 1024) q += T(1);
    guard = 0;
    while (q >= T(1) && guard++ < 1024) q -= T(1);
    return q;
}

// Tricubic Lagrange interpolation of a periodic field at one normalized
// query point.  Index space u = q·N + g with g = 3 ghost layers
// (rescale_xyz, src/interp3.cpp:24-60); stencil base floor(u)−1 and weights
// L_k(t) = c_k·Π_{l≠k}(t−l), c = (−1/6, 1/2, −1/2, 1/6)
// (optimized_interp3_ghost_xyz_p, src/interp3.cpp:2508-2533); the ghost copy
// of accfft_get_ghost_xyz (src/ghost.cpp:555-612) is replaced by modular
// indexing (ghost index j maps to (j − g) mod N).
template <class T>
T interp_point(const T* f, const int* n, const T q[3]) {
    const T lagr[4] = {static_cast<T>(-1.0 / 6.0), static_cast<T>(0.5), static_cast<T>(-0.5), static_cast<T>(1.0 / 6.0)};
    int idx[3][4];
    T M[3][4];
    for (int j = 0; j < 3; ++j) {
        const T hh = static_cast<T>(1.0 / n[j]);
        const T qq = wrap_coord(q[j], hh);
        T u = qq * static_cast<T>(n[j]) + static_cast<T>(3);
        const T fu = std::floor(u);
        int gi = std::isfinite(fu) ? static_cast<int>(fu) - 1 : 0;
        const T x = u - static_cast<T>(gi);
        for (int k = 0; k < 4; ++k) {
            T m = lagr[k];
            for (int l = 0; l < 4; ++l)
                if (k != l) m *= (x - static_cast<T>(l));
            M[j][k] = m;
            long ii = static_cast<long>(gi) - 3 + k;
            ii %= n[j];
            if (ii < 0) ii += n[j];
            idx[j][k] = static_cast<int>(ii);
        }
    }
    T val = T(0);
    for (int a = 0; a < 4; ++a) {
        T sa = T(0);
        for (int b = 0; b < 4; ++b) {
            const T* row = f + (static_cast<int64_t>(idx[0][a]) * n[1] + idx[1][b]) * n[2];
            T r = T(0);
            for (int c = 0; c < 4; ++c) r += M[2][c] * row[idx[2][c]];
            sa += M[1][b] * r;
        }
        val += M[0][a] * sa;
    }
    return val;
}

// Interpolate a scalar field at interleaved normalized query points X
// (SemiLagrangian::Interpolate, src/SemiLagrangian.cpp:536-594).
template <class T>
void interp(const Grid<T>& g, const T* f, const T* X, T* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < g.nl; ++i) out[i] = interp_point(f, g.n, X + 3 * i);
}

// RK2 characteristic (SemiLagrangian::ComputeTrajectoryRK2,
// src/SemiLagrangian.cpp:192-288).  sigma = +1 state, −1 adjoint.  Output
// X interleaved, normalized by 2π.
template <class T>
void trajectory_rk2(const Grid<T>& g, const T* const v[3], T ht, int sigma, T* X) {
    const T s = static_cast<T>(sigma);
    const T hthalf = static_cast<T>(0.5) * ht;
    const int n1 = g.n[1], n2 = g.n[2];
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < g.nl; ++l) {
        const int i1 = static_cast<int>(l / (static_cast<int64_t>(n1) * n2));
        const int i2 = static_cast<int>((l / n2) % n1);
        const int i3 = static_cast<int>(l % n2);
        const T x[3] = {g.hx[0] * static_cast<T>(i1), g.hx[1] * static_cast<T>(i2), g.hx[2] * static_cast<T>(i3)};
        for (int d = 0; d < 3; ++d)
            X[3 * l + d] = static_cast<T>(static_cast<double>(x[d] - s * ht * v[d][l]) / (2.0 * kPi));
    }
    std::vector<T> vx(g.nl);
    std::vector<T> out(3 * g.nl);
    for (int d = 0; d < 3; ++d) {
        interp(g, v[d], X, vx.data());
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < g.nl; ++l) out[3 * l + d] = vx[l];
    }
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < g.nl; ++l) {
        const int i1 = static_cast<int>(l / (static_cast<int64_t>(n1) * n2));
        const int i2 = static_cast<int>((l / n2) % n1);
        const int i3 = static_cast<int>(l % n2);
        const T x[3] = {g.hx[0] * static_cast<T>(i1), g.hx[1] * static_cast<T>(i2), g.hx[2] * static_cast<T>(i3)};
        for (int d = 0; d < 3; ++d)
            X[3 * l + d] = static_cast<T>(static_cast<double>(x[d] - s * hthalf * (out[3 * l + d] + v[d][l])) / (2.0 * kPi));
    }
}

// SemiLagrangian::ComputeTrajectoryRK4 (src/SemiLagrangian.cpp:297-459),
// selected by the -rkorder option.  Four stages of the stationary backward
// characteristic, σ = +1 state / −1 adjoint, with the reference's mixed
// precision: f += 2.0·vX in double, hthalf = 0.5·ht and ht in float, ht/6.0
// in double, and the division by 2π in double.
template <class T>
void trajectory_rk4(const Grid<T>& g, const T* const v[3], T ht, int sigma, T* X) {
    const T s = static_cast<T>(sigma);
    const T hthalf = static_cast<T>(0.5) * ht;
    const int n1 = g.n[1], n2 = g.n[2];
    const int64_t nl = g.nl;
    auto coords = [&](int64_t l, T x[3]) {
        const int i1 = static_cast<int>(l / (static_cast<int64_t>(n1) * n2));
        const int i2 = static_cast<int>((l / n2) % n1);
        const int i3 = static_cast<int>(l % n2);
        x[0] = g.hx[0] * static_cast<T>(i1);
        x[1] = g.hx[1] * static_cast<T>(i2);
        x[2] = g.hx[2] * static_cast<T>(i3);
    };
    std::vector<T> f(3 * nl), vx(3 * nl), tmp(nl);
    auto interp_v = [&]() {
        for (int d = 0; d < 3; ++d) {
            interp(g, v[d], X, tmp.data());
#pragma omp parallel for schedule(static)
            for (int64_t l = 0; l < nl; ++l) vx[3 * l + d] = tmp[l];
        }
    };
    // stage 1: f = v, X = x − σ·ht/2·v
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < nl; ++l) {
        T x[3];
        coords(l, x);
        for (int d = 0; d < 3; ++d) {
            f[3 * l + d] = v[d][l];
            X[3 * l + d] = static_cast<T>(static_cast<double>(x[d] - s * hthalf * v[d][l]) / (2.0 * kPi));
        }
    }
    // stages 2 and 3: f += 2·v(X), X = x − σ·(ht/2 | ht)·v(X)
    for (int stage = 2; stage <= 3; ++stage) {
        interp_v();
        const T step = (stage == 2) ? hthalf : ht;
#pragma omp parallel for schedule(static)
        for (int64_t l = 0; l < nl; ++l) {
            T x[3];
            coords(l, x);
            for (int d = 0; d < 3; ++d) {
                const T w = vx[3 * l + d];
                f[3 * l + d] = static_cast<T>(f[3 * l + d] + 2.0 * w);
                X[3 * l + d] = static_cast<T>(static_cast<double>(x[d] - s * step * w) / (2.0 * kPi));
            }
        }
    }
    // stage 4: f += v(X), X = x − σ·(ht/6)·f
    interp_v();
#pragma omp parallel for schedule(static)
    for (int64_t l = 0; l < nl; ++l) {
        T x[3];
        coords(l, x);
        for (int d = 0; d < 3; ++d) {
            f[3 * l + d] += vx[3 * l + d];
            X[3 * l + d] = static_cast<T>((x[d] - s * (ht / 6.0) * f[3 * l + d]) / (2.0 * kPi));
        }
    }
}

// The characteristic scheme of the semi-Lagrangian solves (RegOpt
// m_PDESolver.rkorder, default 2).  Test infrastructure sets it per call.
inline int& rk_order() {
    static int order = 2;
    return order;
}

template <class T>
void trajectory(const Grid<T>& g, const T* const v[3], T ht, int sigma, T* X) {
    if (rk_order() == 4)
        trajectory_rk4(g, v, ht, sigma, X);
    else
        trajectory_rk2(g, v, ht, sigma, X);
}

// ---------------------------------------------------------------------------
// Transport solves (GN, SL, nc = 1).
// ---------------------------------------------------------------------------
template <class T>
bool velocity_is_zero(const Grid<T>& g, const T* const v[3]) {
    // CLAIREBase::IsVelocityZero, src/CLAIREBase.cpp:578-605 (∞-norm == 0)
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < g.nl; ++i)
            if (v[d][i] != T(0)) return false;
    return true;
}

// State equation m^{j+1} = m^j(X_s), m^0 = mT; v = 0 copies
// (CLAIRE::SolveStateEquation src/CLAIRE.cpp:1896-1920,
//  SolveStateEquationSL src/CLAIRE.cpp:2116-2170).  m has (nt+1)·nl entries.
// Hyperbolic constraint of the state solve (m_PDESolver.pdetype,
// include/RegOpt.hpp:42-46): 0 the continuity equation ∂t m + ∇·(m v) = 0,
// 1 the transport equation ∂t m + v·∇m = 0 (default, src/RegOpt.cpp:1089).
// Only the state solve reads it (CLAIRE::SolveStateEquation,
// src/CLAIRE.cpp:1923-1926).  Test infrastructure sets it per call.
enum PDEType { kContinuityEq = 0, kTransportEq = 1 };

inline int& pde_type() {
    static int t = kTransportEq;
    return t;
}

// SL solve of the continuity equation, CLAIRE::SolveContinuityEquationSL
// (src/CLAIRE.cpp:2735-2854), on the state trajectory X: for j = 0..nt−1
//   mx = m^j(X), rhs0 = −mx·(∇·v)(X), rhs1 = −(mx + ht·rhs0)·∇·v,
//   m^{j+1} = mx + ht/2·(rhs0 + rhs1);   m^0 is already in place.
template <class T>
void solve_continuity(const Grid<T>& g, int nt, const T* const v[3], const T* X, T* m) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    std::vector<T> divv(nl), divvx(nl), mx(nl);
    div(g, v[0], v[1], v[2], divv.data());
    interp(g, divv.data(), X, divvx.data());
    for (int j = 0; j < nt; ++j) {
        interp(g, m + j * nl, X, mx.data());
        T* next = m + (j + 1) * nl;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T x = mx[i];
            const T rhs0 = -x * divvx[i];
            const T rhs1 = -(x + ht * rhs0) * divv[i];
            next[i] = x + hthalf * (rhs0 + rhs1);
        }
    }
}

// Time integrator of the transport solves (m_PDESolver.type,
// include/RegOpt.hpp:31-35): 0 explicit RK2 ("-pdesolver rk2"), 2 SL
// (default, src/RegOpt.cpp:1083).  Test infrastructure sets it per call.
enum PDESolverType { kRK2 = 0, kRK2A = 1, kSL = 2 };

inline int& pde_solver() {
    static int s = kSL;
    return s;
}

// −a·b summed as −a1·b1 − a2·b2 − a3·b3.
template <class T>
inline T neg_dot3(const std::vector<T>* a, const T* const b[3], int64_t i) {
    return -a[0][i] * b[0][i] - a[1][i] * b[1][i] - a[2][i] * b[2][i];
}

// SolveStateEquationRK2 (src/CLAIRE.cpp:1995-2105): for j = 0..nt−1
//   rhs0 = −∇m^j·v, m̄ = m^j + ht·rhs0, rhs1 = −∇m̄·v,
//   m^{j+1} = m^j + ht/2·(rhs0 + rhs1);   m^0 is already in place.
template <class T>
void solve_state_rk2(const Grid<T>& g, int nt, const T* const v[3], T* m) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    std::vector<T> gr[3] = {std::vector<T>(nl), std::vector<T>(nl), std::vector<T>(nl)};
    std::vector<T> mbar(nl), rhs0(nl);
    for (int j = 0; j < nt; ++j) {
        const T* mj = m + static_cast<int64_t>(j) * nl;
        T* mnext = m + static_cast<int64_t>(j + 1) * nl;
        grad(g, mj, gr[0].data(), gr[1].data(), gr[2].data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            rhs0[i] = neg_dot3(gr, v, i);
            mbar[i] = mj[i] + ht * rhs0[i];
        }
        grad(g, mbar.data(), gr[0].data(), gr[1].data(), gr[2].data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T rhs1 = neg_dot3(gr, v, i);
            mnext[i] = mj[i] + hthalf * (rhs0[i] + rhs1);
        }
    }
}

// SolveAdjointEquationRK2 (src/CLAIRE.cpp:2365-2548) and
// SolveIncAdjointEquationGNRK2 (src/CLAIRE.cpp:3504-3683), GN (one slot),
// nc = 1: λ holds the final condition; for j = 0..nt−1
//   rhs0 = ∇·(λv), λ̄ = λ + ht·rhs0, rhs1 = ∇·(λ̄v),
//   b += w_j·∇m^{nt−j}·λ (λ before the update), λ ← λ + ht/2·(rhs0 + rhs1),
// then b += ht/2·∇m^0·λ (in double, as written); w_0 = ht/2, else ht.
// b is not projected or scaled here.
template <class T>
void adjoint_rk2(const Grid<T>& g, int nt, const T* const v[3], const T* m, T* lambda, T* const b[3]) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    std::vector<T> w[3] = {std::vector<T>(nl), std::vector<T>(nl), std::vector<T>(nl)};
    std::vector<T> rhs0(nl), rhs1(nl);
    T scale = ht;
    for (int j = 0; j < nt; ++j) {
        const T* mj = m + static_cast<int64_t>(nt - j) * nl;
        if (j == 0) scale *= static_cast<T>(0.5);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T l = lambda[i];
            for (int d = 0; d < 3; ++d) w[d][i] = l * v[d][i];
        }
        div(g, w[0].data(), w[1].data(), w[2].data(), rhs0.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T lbar = lambda[i] + ht * rhs0[i];
            for (int d = 0; d < 3; ++d) w[d][i] = v[d][i] * lbar;
        }
        div(g, w[0].data(), w[1].data(), w[2].data(), rhs1.data());
        grad(g, mj, w[0].data(), w[1].data(), w[2].data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T l = lambda[i];
            lambda[i] = l + hthalf * (rhs0[i] + rhs1[i]);
            for (int d = 0; d < 3; ++d) b[d][i] += scale * w[d][i] * l / T(1);
        }
        if (j == 0) scale *= static_cast<T>(2.0);
    }
    grad(g, m, w[0].data(), w[1].data(), w[2].data());
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nl; ++i) {
        const T l = lambda[i];
        for (int d = 0; d < 3; ++d) b[d][i] += 0.5 * scale * w[d][i] * l / T(1);
    }
}

// SolveIncStateEquationRK2 (src/CLAIRE.cpp:2962-3135), GN (one slot),
// nc = 1: m̃⁰ = 0; for j = 0..nt−1
//   rhs0 = −∇m̃·v − ∇m^j·ṽ, m̃̄ = m̃ + ht·rhs0,
//   rhs1 = −∇m̃̄·v − ∇m^{j+1}·ṽ, m̃ ← m̃ + ht/2·(rhs1 + rhs0);
// for v = 0, nt steps m̃ ← m̃ − ht·(∇m^0·ṽ) (src/CLAIRE.cpp:3022-3052).
template <class T>
void solve_inc_state_rk2(const Grid<T>& g, int nt, const T* const v[3], const T* m, const T* const vt[3], T* mt) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    std::fill(mt, mt + nl, T(0));
    std::vector<T> gm[3] = {std::vector<T>(nl), std::vector<T>(nl), std::vector<T>(nl)};
    if (velocity_is_zero(g, v)) {
        grad(g, m, gm[0].data(), gm[1].data(), gm[2].data());
        for (int j = 0; j < nt; ++j)
            for (int64_t i = 0; i < nl; ++i)
                mt[i] = mt[i] - ht * (gm[0][i] * vt[0][i] + gm[1][i] * vt[1][i] + gm[2][i] * vt[2][i]);
        return;
    }
    std::vector<T> gt[3] = {std::vector<T>(nl), std::vector<T>(nl), std::vector<T>(nl)};
    std::vector<T> mtbar(nl), rhs0(nl);
    for (int j = 0; j < nt; ++j) {
        grad(g, m + static_cast<int64_t>(j) * nl, gm[0].data(), gm[1].data(), gm[2].data());
        grad(g, mt, gt[0].data(), gt[1].data(), gt[2].data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            rhs0[i] = neg_dot3(gt, v, i) - gm[0][i] * vt[0][i] - gm[1][i] * vt[1][i] - gm[2][i] * vt[2][i];
            mtbar[i] = mt[i] + ht * rhs0[i];
        }
        grad(g, m + static_cast<int64_t>(j + 1) * nl, gm[0].data(), gm[1].data(), gm[2].data());
        grad(g, mtbar.data(), gt[0].data(), gt[1].data(), gt[2].data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T rhs1 = neg_dot3(gt, v, i) - gm[0][i] * vt[0][i] - gm[1][i] * vt[1][i] - gm[2][i] * vt[2][i];
            mt[i] = mt[i] + hthalf * (rhs1 + rhs0[i]);
        }
    }
}

template <class T>
void solve_state(const Grid<T>& g, int nt, const T* const v[3], const T* mT, T* m) {
    std::copy(mT, mT + g.nl, m);
    if (velocity_is_zero(g, v)) {
        for (int j = 1; j <= nt; ++j) std::copy(mT, mT + g.nl, m + j * g.nl);
        return;
    }
    // the continuity equation is always solved by SL (src/CLAIRE.cpp:1923-1926)
    if (pde_type() != kContinuityEq && pde_solver() == kRK2) {
        solve_state_rk2(g, nt, v, m);
        return;
    }
    const T ht = T(1) / static_cast<T>(nt);
    std::vector<T> X(3 * g.nl);
    trajectory(g, v, ht, +1, X.data());
    if (pde_type() == kContinuityEq) {
        solve_continuity(g, nt, v, X.data(), m);
        return;
    }
    for (int j = 0; j < nt; ++j) interp(g, m + j * g.nl, X.data(), m + (j + 1) * g.nl);
}

// SL2 distance ½·hd·Σ(mR − m^nt)² (DistanceMeasureSL2::EvaluateFunctional,
// src/DistanceMeasureSL2.cpp:93-150).
template <class T>
double distance_sl2(const Grid<T>& g, int nt, const T* m, const T* mR) {
    const T* m1 = m + static_cast<int64_t>(nt) * g.nl;
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < g.nl; ++i) {
        const double dr = static_cast<double>(static_cast<T>(mR[i] - m1[i]));
        s += dr * dr;
    }
    return 0.5 * static_cast<double>(g.hd) * s;
}

// Adjoint equation with body force, scaled by hd.  λ¹ = mR − m^nt
// (SetFinalConditionAE, src/DistanceMeasureSL2.cpp:156-218);
// SolveAdjointEquationSL src/CLAIRE.cpp:2550-2725; v = 0 branch
// src/CLAIRE.cpp:2249-2312; ×hd src/CLAIRE.cpp:2341.
// Outputs: lambda (value at t = 0, GN single slot) and b (3·nl).
template <class T>
void solve_adjoint(const Grid<T>& g, int nt, const T* const v[3], const T* m, const T* mR, T* lambda, T* const b[3]) {
    const int64_t nl = g.nl;
    const T* m1 = m + static_cast<int64_t>(nt) * nl;
    for (int64_t i = 0; i < nl; ++i) lambda[i] = mR[i] - m1[i];
    for (int d = 0; d < 3; ++d) std::fill(b[d], b[d] + nl, T(0));
    std::vector<T> g1(nl), g2(nl), g3(nl);
    if (velocity_is_zero(g, v)) {
        grad(g, m, g1.data(), g2.data(), g3.data());
        for (int64_t i = 0; i < nl; ++i) {
            const T l = lambda[i];
            b[0][i] += l * g1[i];
            b[1][i] += l * g2[i];
            b[2][i] += l * g3[i];
        }
    } else if (pde_solver() == kRK2) {
        // CLAIREStokes overrides only the SL adjoint solves
        adjoint_rk2(g, nt, v, m, lambda, b);
    } else {
        const T ht = T(1) / static_cast<T>(nt);
        const bool stokes = reg_model() == kStokes;  // CLAIREStokes: pure transport (src/CLAIREStokes.cpp:169-296)
        std::vector<T> X(3 * nl), divv(nl), divvx(nl), lx(nl);
        trajectory(g, v, ht, -1, X.data());
        if (!stokes) {
            div(g, v[0], v[1], v[2], divv.data());
            interp(g, divv.data(), X.data(), divvx.data());
        }
        T scale = ht;
        for (int j = 0; j < nt; ++j) {
            const T* mj = m + static_cast<int64_t>(nt - j) * nl;
            if (j == 0) scale *= static_cast<T>(0.5);
            interp(g, lambda, X.data(), lx.data());
            grad(g, mj, g1.data(), g2.data(), g3.data());
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < nl; ++i) {
                const T lam = lambda[i];
                const T lamx = lx[i];
                if (stokes) {
                    lambda[i] = lamx;
                } else {
                    const T rhs0 = lamx * divvx[i];
                    const T rhs1 = (lamx + ht * rhs0) * divv[i];
                    lambda[i] = static_cast<T>(lamx + 0.5 * ht * (rhs0 + rhs1));
                }
                b[0][i] += scale * g1[i] * lam;
                b[1][i] += scale * g2[i] * lam;
                b[2][i] += scale * g3[i] * lam;
            }
            if (j == 0) scale *= static_cast<T>(2.0);
        }
        grad(g, m, g1.data(), g2.data(), g3.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T lam = lambda[i];
            b[0][i] += static_cast<T>(0.5 * scale * g1[i] * lam);
            b[1][i] += static_cast<T>(0.5 * scale * g2[i] * lam);
            b[2][i] += static_cast<T>(0.5 * scale * g3[i] * lam);
        }
    }
    project_and_scale(g, b);
}

// Incremental state equation, GN, SL (SolveIncStateEquationSL,
// src/CLAIRE.cpp:3147-3294): m̃⁰ = 0,
//   m̃^{j+1} = m̃^j(X_s) − ht/2·[(∇m^j)(X_s)·ṽ(X_s)] − ht/2·[∇m^{j+1}·ṽ].
// There is no v = 0 branch on the SL path (src/CLAIRE.cpp:2866-2950).
template <class T>
void solve_inc_state_x(const Grid<T>& g, int nt, const T* const v[3], const T* X, const T* m, const T* const vt[3],
                       T* mtilde) {
    if (pde_solver() == kRK2) {
        solve_inc_state_rk2(g, nt, v, m, vt, mtilde);
        return;
    }
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    std::vector<T> vtx0(nl), vtx1(nl), vtx2(nl), g1(nl), g2(nl), g3(nl), gx1(nl), gx2(nl), gx3(nl), tmp(nl);
    interp(g, vt[0], X, vtx0.data());
    interp(g, vt[1], X, vtx1.data());
    interp(g, vt[2], X, vtx2.data());
    std::fill(mtilde, mtilde + nl, T(0));
    for (int j = 0; j < nt; ++j) {
        interp(g, mtilde, X, tmp.data());
        std::copy(tmp.begin(), tmp.end(), mtilde);
        grad(g, m + static_cast<int64_t>(j) * nl, g1.data(), g2.data(), g3.data());
        interp(g, g1.data(), X, gx1.data());
        interp(g, g2.data(), X, gx2.data());
        interp(g, g3.data(), X, gx3.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i)
            mtilde[i] -= hthalf * (gx1[i] * vtx0[i] + gx2[i] * vtx1[i] + gx3[i] * vtx2[i]);
        grad(g, m + static_cast<int64_t>(j + 1) * nl, g1.data(), g2.data(), g3.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i)
            mtilde[i] -= hthalf * (g1[i] * vt[0][i] + g2[i] * vt[1][i] + g3[i] * vt[2][i]);
    }
}

template <class T>
void solve_inc_state(const Grid<T>& g, int nt, const T* const v[3], const T* m, const T* const vt[3], T* mtilde) {
    std::vector<T> X(3 * g.nl);
    if (pde_solver() != kRK2) trajectory(g, v, T(1) / static_cast<T>(nt), +1, X.data());
    solve_inc_state_x(g, nt, v, X.data(), m, vt, mtilde);
}

// Incremental adjoint equation, GN, SL, with the incremental body force
// scaled by hd (SolveIncAdjointEquation src/CLAIRE.cpp:3307-3491,
// SolveIncAdjointEquationGNSL src/CLAIRE.cpp:3849-4015, final condition
// λ̃ = −m̃ src/DistanceMeasureSL2.cpp:223-276, v = 0 branch
// src/CLAIRE.cpp:3367-3421, ×hd src/CLAIRE.cpp:3471).
template <class T>
void solve_inc_adjoint_x(const Grid<T>& g, int nt, const T* const v[3], const T* Xa, const T* m, const T* mtilde,
                         T* ltilde, T* const bt[3]) {
    const int64_t nl = g.nl;
    for (int64_t i = 0; i < nl; ++i) ltilde[i] = -mtilde[i];
    for (int d = 0; d < 3; ++d) std::fill(bt[d], bt[d] + nl, T(0));
    std::vector<T> g1(nl), g2(nl), g3(nl);
    if (velocity_is_zero(g, v)) {
        grad(g, m, g1.data(), g2.data(), g3.data());
        for (int64_t i = 0; i < nl; ++i) {
            const T l = ltilde[i];
            bt[0][i] += l * g1[i];
            bt[1][i] += l * g2[i];
            bt[2][i] += l * g3[i];
        }
    } else if (pde_solver() == kRK2) {
        adjoint_rk2(g, nt, v, m, ltilde, bt);
    } else {
        const T ht = T(1) / static_cast<T>(nt);
        const T hthalf = static_cast<T>(0.5) * ht;
        const bool stokes = reg_model() == kStokes;  // src/CLAIREStokes.cpp:297-406
        std::vector<T> divv(nl), divvx(nl), lx(nl);
        const T* X = Xa;
        if (!stokes) {
            div(g, v[0], v[1], v[2], divv.data());
            interp(g, divv.data(), X, divvx.data());
        }
        T scale = ht;
        for (int j = 0; j < nt; ++j) {
            const T* mj = m + static_cast<int64_t>(nt - j) * nl;
            if (j == 0) scale *= static_cast<T>(0.5);
            interp(g, ltilde, X, lx.data());
            grad(g, mj, g1.data(), g2.data(), g3.data());
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < nl; ++i) {
                const T lt = ltilde[i];
                const T ltx = lx[i];
                if (stokes) {
                    ltilde[i] = ltx;
                } else {
                    const T rhs0 = ltx * divvx[i];
                    const T rhs1 = (ltx + ht * rhs0) * divv[i];
                    ltilde[i] = ltx + hthalf * (rhs0 + rhs1);
                }
                bt[0][i] += scale * g1[i] * lt;
                bt[1][i] += scale * g2[i] * lt;
                bt[2][i] += scale * g3[i] * lt;
            }
            if (j == 0) scale *= static_cast<T>(2.0);
        }
        grad(g, m, g1.data(), g2.data(), g3.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T lt = ltilde[i];
            bt[0][i] += static_cast<T>(0.5 * scale * g1[i] * lt);
            bt[1][i] += static_cast<T>(0.5 * scale * g2[i] * lt);
            bt[2][i] += static_cast<T>(0.5 * scale * g3[i] * lt);
        }
    }
    project_and_scale(g, bt);
}

template <class T>
void solve_inc_adjoint(const Grid<T>& g, int nt, const T* const v[3], const T* m, const T* mtilde, T* ltilde,
                       T* const bt[3]) {
    std::vector<T> X;
    if (!velocity_is_zero(g, v) && pde_solver() != kRK2) {
        X.resize(3 * g.nl);
        trajectory(g, v, T(1) / static_cast<T>(nt), -1, X.data());
    }
    solve_inc_adjoint_x(g, nt, v, X.data(), m, mtilde, ltilde, bt);
}

// GN Hessian matvec Hṽ = R''ṽ + b̃ (CLAIRE::HessMatVec src/CLAIRE.cpp:1311-1379),
// optionally unscaled by 1/hd (CLAIRE::HessianMatVec src/CLAIRE.cpp:1283-1286).
// The trajectories X_s, X_a are computed once per velocity (by the state and
// adjoint solves in the reference, src/CLAIRE.cpp:2149,2607) and reused.
template <class T>
void hessian_matvec_x(const Grid<T>& g, int nt, int norm, const T* beta, const T* const v[3], const T* Xs,
                      const T* Xa, const T* m, const T* const vt[3], T* const hv[3], bool scale) {
    const int64_t nl = g.nl;
    std::vector<T> mt(nl), lt(nl), b0(nl), b1(nl), b2(nl);
    T* bt[3] = {b0.data(), b1.data(), b2.data()};
    if (matvec_type() != 0) {
        // PrecondHessMatVec (src/CLAIRE.cpp:1392-1451): hd·ṽ + (βA)⁻¹b̃[ṽ];
        // PrecondHessMatVecSym (src/CLAIRE.cpp:1463-1540):
        // hd·ṽ + (βA)^{-1/2}b̃[(βA)^{-1/2}ṽ]
        const bool sym = matvec_type() == 2;
        std::vector<T> w0(nl), w1(nl), w2(nl);
        T* w[3] = {w0.data(), w1.data(), w2.data()};
        const T* in[3] = {vt[0], vt[1], vt[2]};
        if (sym) {
            reg_inverse(g, norm, beta, vt, w, true);
            for (int d = 0; d < 3; ++d) in[d] = w[d];
        }
        solve_inc_state_x(g, nt, v, Xs, m, in, mt.data());
        solve_inc_adjoint_x(g, nt, v, Xa, m, mt.data(), lt.data(), bt);
        const T* btc[3] = {bt[0], bt[1], bt[2]};
        reg_inverse(g, norm, beta, btc, hv, sym);
        for (int d = 0; d < 3; ++d)
            for (int64_t i = 0; i < nl; ++i) hv[d][i] = g.hd * vt[d][i] + hv[d][i];
        if (!scale) {
            for (int d = 0; d < 3; ++d)
                for (int64_t i = 0; i < nl; ++i) hv[d][i] *= T(1) / g.hd;
        }
        return;
    }
    solve_inc_state_x(g, nt, v, Xs, m, vt, mt.data());
    solve_inc_adjoint_x(g, nt, v, Xa, m, mt.data(), lt.data(), bt);
    reg_apply(g, norm, beta, vt, hv);
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < nl; ++i) hv[d][i] += bt[d][i];
    if (!scale) {
        for (int d = 0; d < 3; ++d)
            for (int64_t i = 0; i < nl; ++i) hv[d][i] *= T(1) / g.hd;
    }
}

template <class T>
void hessian_matvec(const Grid<T>& g, int nt, int norm, const T* beta, const T* const v[3], const T* m,
                    const T* const vt[3], T* const hv[3], bool scale) {
    const T ht = T(1) / static_cast<T>(nt);
    std::vector<T> Xs(3 * g.nl), Xa(3 * g.nl);
    trajectory(g, v, ht, +1, Xs.data());
    trajectory(g, v, ht, -1, Xa.data());
    hessian_matvec_x(g, nt, norm, beta, v, Xs.data(), Xa.data(), m, vt, hv, scale);
}

// Reduced L2 gradient g = R'(v) + b (CLAIRE::EvaluateGradient
// src/CLAIRE.cpp:871-1008, EvaluateL2Gradient src/CLAIRE.cpp:1015-1042);
// with v = 0 the gradient is b (src/CLAIRE.cpp:960-965).  Runs the state and
// adjoint solves.
template <class T>
void gradient(const Grid<T>& g, int nt, int norm, const T* beta, const T* const v[3], const T* mT, const T* mR,
              T* const gout[3]) {
    const int64_t nl = g.nl;
    std::vector<T> m((nt + 1) * nl), lam(nl), b0(nl), b1(nl), b2(nl);
    T* b[3] = {b0.data(), b1.data(), b2.data()};
    solve_state(g, nt, v, mT, m.data());
    solve_adjoint(g, nt, v, m.data(), mR, lam.data(), b);
    if (velocity_is_zero(g, v)) {
        for (int d = 0; d < 3; ++d) std::copy(b[d], b[d] + nl, gout[d]);
        return;
    }
    reg_apply(g, norm, beta, v, gout);
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < nl; ++i) gout[d][i] += b[d][i];
}

// CLAIREDivReg::EvaluteRegularizationDIV (src/CLAIREDivReg.cpp:190-260):
// Rw = ½·hd·β2·(‖∂3 d‖² + ‖∂2 d‖² + ‖∂1 d‖² + ‖d‖²), d = ∇·v; the gradient
// is stored with its components reversed (grad into (x3, x2, x1)), so the
// products are summed in that order.
template <class T>
double divergence_penalty(const Grid<T>& g, const T* beta, const T* const v[3]) {
    std::vector<T> d(g.nl), g1(g.nl), g2(g.nl), g3(g.nl);
    div(g, v[0], v[1], v[2], d.data());
    grad(g, d.data(), g3.data(), g2.data(), g1.data());
    double r = 0.0;
    r += dot(g1.data(), g1.data(), g.nl);
    r += dot(g2.data(), g2.data(), g.nl);
    r += dot(g3.data(), g3.data(), g.nl);
    r += dot(d.data(), d.data(), g.nl);
    return 0.5 * static_cast<double>(g.hd) * static_cast<double>(beta[2]) * r;
}

// Objective J = D + R (CLAIRE::EvaluateObjective src/CLAIRE.cpp:793-865;
// R skipped when v = 0).
template <class T>
void objective(const Grid<T>& g, int nt, int norm, const T* beta, const T* const v[3], const T* mT, const T* mR,
               double* J, double* D, double* R) {
    std::vector<T> m((nt + 1) * g.nl);
    solve_state(g, nt, v, mT, m.data());
    *D = distance_sl2(g, nt, m.data(), mR);
    *R = velocity_is_zero(g, v) ? 0.0 : reg_functional(g, norm, beta, v);
    // CLAIREDivReg::EvaluateObjective adds the divergence penalty (src/CLAIREDivReg.cpp:109-188)
    if (reg_model() == kRelaxedStokes && !velocity_is_zero(g, v)) *R += divergence_penalty(g, beta, v);
    *J = *D + *R;
}

// Determinant of the deformation gradient by semi-Lagrangian transport,
// DeformationFields::ComputeDetDefGrad + ComputeDetDefGradSL
// (src/DeformationFields.cpp:315-388, 636-740): J⁰ = 1 and for j = 0..nt−1
//   JX = J^j(X_s), rhs0 = α·JX·(∇·v)(X_s), rhs1 = α·(JX + ht·rhs0)·∇·v,
//   J^{j+1} = JX + ht/2·(rhs0 + rhs1),   α = −1 for det(∇y⁻¹) (inverse).
// Also min, max and mean of J (VecMin/VecMax/VecSum/ng).
template <class T>
void det_def_grad(const Grid<T>& g, int nt, const T* const v[3], bool inverse, T* jac, double* jmin, double* jmax,
                  double* jmean) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    const T hthalf = static_cast<T>(0.5) * ht;
    const T alpha = inverse ? T(-1) : T(1);
    std::vector<T> X(3 * nl), divv(nl), divvx(nl), jx(nl);
    std::fill(jac, jac + nl, T(1));
    trajectory(g, v, ht, +1, X.data());
    div(g, v[0], v[1], v[2], divv.data());
    interp(g, divv.data(), X.data(), divvx.data());
    for (int j = 0; j < nt; ++j) {
        interp(g, jac, X.data(), jx.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nl; ++i) {
            const T jacX = jx[i];
            const T rhs0 = alpha * jacX * divvx[i];
            const T rhs1 = alpha * (jacX + ht * rhs0) * divv[i];
            jac[i] = jacX + hthalf * (rhs0 + rhs1);
        }
    }
    double mn = jac[0], mx = jac[0], sum = 0.0;
    for (int64_t i = 0; i < nl; ++i) {
        mn = std::min(mn, static_cast<double>(jac[i]));
        mx = std::max(mx, static_cast<double>(jac[i]));
        sum += jac[i];
    }
    *jmin = mn;
    *jmax = mx;
    *jmean = sum / static_cast<double>(nl);
}

// ---------------------------------------------------------------------------
// Gauss–Newton–Krylov solve, the reference's default optimizer:
//   TAO "nls" with a matrix-free Hessian (src/Optimizer.cpp:235-444),
//   KSPCG with the unpreconditioned residual norm, zero initial guess,
//   at most maxiter−1 iterations (src/Optimizer.cpp:298-311), preconditioned
//   by the inverse regularization operator (PCSHELL → ApplySpectralPrecond,
//   src/Preconditioner.cpp:521-573, applysqrt = false);
//   Eisenstat–Walker forcing (PreKrylovSolve, src/KrylovInterface.cpp:79-175):
//   rtol = max(1e-16, min(0.5, (‖g‖/‖g₀‖)^{1/2})) (superlinear) or the ratio
//   itself (quadratic), ‖g₀‖ the norm at the first Krylov solve;
//   Armijo backtracking (TaoLineSearch "armijo": σ = 1e-4, step halving,
//   at most 30 trials, initial step 1);
//   stopping test CheckConvergenceGrad (src/TaoInterface.cpp:390-512) with
//   ‖g₀‖ the gradient norm at v = 0 (InitializeOptimization,
//   src/CLAIRE.cpp:320-335), step < 2^-10 → stop, iter > maxit−1 → stop.
// If the Krylov direction is not a descent direction, or its line search
// fails, the step falls back to steepest descent −g (TAO NLS gradient step).
// Inner products accumulate in double.
// history: 8 doubles per iterate: iter, J, D, R, ‖g‖, ‖g‖/‖g₀‖, PCG
// iterations and step length of the step taken from this iterate.
// ---------------------------------------------------------------------------
struct GNOptions {
    int maxit = 50;            // OptPara.maxiter (src/RegOpt.cpp, default 50)
    double gttol = 5e-2;       // OptPara.tol[2]
    double gatol = 1e-6;       // OptPara.tol[0]
    int krylov_maxit = 50;     // KrylovMethod.maxiter
    double krylov_rtol = 1e-12;  // KrylovMethod.tol[0] (used without forcing)
    int forcing = 2;           // 0 none, 1 quadratic, 2 superlinear (SLFS default)
    int precond = 1;           // 0 none, 1 inverse regularization (INVREG default)
    int linesearch = 1;        // 0 unit step, 1 Armijo (ARMIJOLS default)
};

struct GNReport {
    int iterations = 0, reason = 0, matvecs = 0, pcg_iterations = 0, objective_evals = 0, gradient_evals = 0;
    double J0 = 0, D0 = 0, gnorm0 = 0, J = 0, D = 0, R = 0, gnorm = 0;
};

enum GNReason { kGNContinue = 0, kGNGttol = 1, kGNGatol = 2, kGNSteptol = 3, kGNMaxits = -1, kGNNan = -2,
                kGNLineSearch = -3 };

// CheckConvergenceGrad (src/TaoInterface.cpp:421-480), miniter = 0
// (src/RegOpt.cpp:1158): TAO_DIVERGED_NAN (:424-436) and
// TAO_CONVERGED_STEPTOL (:445-453) return at once; TAO_CONVERGED_GTTOL
// (:456-459), TAO_CONVERGED_GATOL (:467-470) and TAO_DIVERGED_MAXITS
// (:478-481) are set one after the other, the last holding one standing
// (then :491-494 stops).
// TAO's iteration limit is maxiter − 1 (src/Optimizer.cpp:431).
inline int convergence_reason(double J, double gnorm, double step, double g0, int iter, double gttol, double gatol,
                              int maxit) {
    if (std::isnan(J) || std::isinf(J) || std::isnan(gnorm) || std::isinf(gnorm)) return kGNNan;
    if (step < 1.0 / 1024.0) return kGNSteptol;
    int r = kGNContinue;
    if (gnorm < gttol * g0) r = kGNGttol;
    if (gnorm < gatol) r = kGNGatol;
    if (iter > maxit - 1) r = kGNMaxits;
    return r;
}

template <class T>
double dot3(const Grid<T>& g, const T* const a[3], const T* const b[3]) {
    double s = 0.0;
    for (int d = 0; d < 3; ++d)
        for (int64_t i = 0; i < g.nl; ++i) s += static_cast<double>(a[d][i]) * static_cast<double>(b[d][i]);
    return s;
}

template <class T>
void gn_solve(const Grid<T>& g, int nt, int norm, const T* beta, T* const v[3], const T* mT, const T* mR,
              const GNOptions& o, GNReport& rep, double* hist, int hist_len) {
    const int64_t nl = g.nl;
    const T ht = T(1) / static_cast<T>(nt);
    auto field3 = [&]() { return std::vector<T>(3 * nl); };
    auto ptrs = [&](std::vector<T>& a) { return std::array<T*, 3>{a.data(), a.data() + nl, a.data() + 2 * nl}; };
    std::vector<T> gv = field3(), d = field3(), r = field3(), z = field3(), p = field3(), w = field3(), vt = field3();
    auto G = ptrs(gv), Dd = ptrs(d), Rr = ptrs(r), Z = ptrs(z), P = ptrs(p), W = ptrs(w), VT = ptrs(vt);
    auto cst = [](const std::array<T*, 3>& a) { return std::array<const T*, 3>{a[0], a[1], a[2]}; };
    const std::array<const T*, 3> V{v[0], v[1], v[2]};

    // ‖g₀‖ at v = 0 (InitializeOptimization)
    {
        std::vector<T> zero = field3();
        auto Zr = ptrs(zero);
        gradient(g, nt, norm, beta, cst(Zr).data(), mT, mR, G.data());
        rep.gnorm0 = std::sqrt(dot3(g, cst(G).data(), cst(G).data()));
        rep.gradient_evals++;
    }
    double J, D, R;
    objective(g, nt, norm, beta, V.data(), mT, mR, &J, &D, &R);
    rep.objective_evals++;
    gradient(g, nt, norm, beta, V.data(), mT, mR, G.data());
    rep.gradient_evals++;
    double gnorm = std::sqrt(dot3(g, cst(G).data(), cst(G).data()));
    rep.J0 = J;
    rep.D0 = D;
    const double g0 = rep.gnorm0 > 0.0 ? rep.gnorm0 : 1.0;
    double kg0 = -1.0;  // ‖g‖ at the first Krylov solve
    double step = 1.0;
    int iter = 0;
    std::vector<T> m((nt + 1) * nl), Xs(3 * nl), Xa(3 * nl);
    for (;;) {
        double* h = (iter < hist_len && hist) ? hist + 8 * iter : nullptr;
        if (h) {
            h[0] = iter; h[1] = J; h[2] = D; h[3] = R; h[4] = gnorm; h[5] = gnorm / g0; h[6] = 0; h[7] = 0;
        }
        // CheckConvergenceGrad (src/TaoInterface.cpp:421-480): NaN, then the
        // step; of GTTOL, GATOL and MAXITS the last that holds wins
        rep.reason = convergence_reason(J, gnorm, step, g0, iter, o.gttol, o.gatol, o.maxit);
        if (rep.reason != kGNContinue) break;

        // Newton system H d = −g at the current iterate
        solve_state(g, nt, V.data(), mT, m.data());
        trajectory(g, V.data(), ht, +1, Xs.data());
        trajectory(g, V.data(), ht, -1, Xa.data());
        if (kg0 < 0.0) kg0 = gnorm;
        double rtol = o.krylov_rtol;
        if (o.forcing != 0 && kg0 > 0.0) {
            const double q = gnorm / kg0;
            rtol = std::max(1e-16, std::min(0.5, o.forcing == 1 ? q : std::sqrt(q)));
        }
        // PRECONDMATVECSYM: (βA)^{-1/2} on the right-hand side and the solution
        // (PreKrylovSolve / PostKrylovSolve, src/CLAIREBase.cpp:941-969), no
        // further preconditioner; PRECONDMATVEC is not symmetric (no CG)
        if (matvec_type() == 1) throw std::invalid_argument("PRECONDMATVEC needs a nonsymmetric Krylov method");
        const bool sym = matvec_type() == 2;
        auto apply_pc = [&](const std::array<const T*, 3>& in, const std::array<T*, 3>& out) {
            if (o.precond && !sym) {
                reg_inverse(g, norm, beta, in.data(), out.data(), false);
            } else {
                for (int c = 0; c < 3; ++c) std::copy(in[c], in[c] + nl, out[c]);
            }
        };
        // KSPCG, x0 = 0, b = −g
        for (size_t i = 0; i < d.size(); ++i) { d[i] = T(0); r[i] = -gv[i]; }
        if (sym) {
            std::vector<T> tmp = r;
            auto Tm = ptrs(tmp);
            reg_inverse(g, norm, beta, cst(Tm).data(), Rr.data(), true);
        }
        const double bnorm = std::sqrt(dot3(g, cst(Rr).data(), cst(Rr).data()));
        const double tol = std::max(rtol * bnorm, 1e-16);
        const int kmax = std::max(0, o.krylov_maxit - 1);
        int its = 0;
        double rnorm = bnorm;
        if (rnorm > tol) {
            apply_pc(cst(Rr), Z);
            double rz = dot3(g, cst(Rr).data(), cst(Z).data());
            for (int k = 1; k <= kmax; ++k) {
                if (k == 1) p = z;  // later directions are updated at the end of the iteration
                hessian_matvec_x(g, nt, norm, beta, V.data(), Xs.data(), Xa.data(), m.data(), cst(P).data(), W.data(),
                                 true);
                rep.matvecs++;
                its = k;
                const double pw = dot3(g, cst(P).data(), cst(W).data());
                if (!(pw > 0.0)) break;  // indefinite or breakdown
                const double a = rz / pw;
                for (size_t i = 0; i < d.size(); ++i) {
                    d[i] += static_cast<T>(a) * p[i];
                    r[i] -= static_cast<T>(a) * w[i];
                }
                rnorm = std::sqrt(dot3(g, cst(Rr).data(), cst(Rr).data()));
                if (rnorm <= tol || !(rnorm < 1e6 * bnorm)) break;
                apply_pc(cst(Rr), Z);
                const double rz_new = dot3(g, cst(Rr).data(), cst(Z).data());
                const double bta = rz_new / rz;
                rz = rz_new;
                for (size_t i = 0; i < p.size(); ++i) p[i] = z[i] + static_cast<T>(bta) * p[i];
            }
        }
        if (sym) {
            std::vector<T> tmp = d;
            auto Tm = ptrs(tmp);
            reg_inverse(g, norm, beta, cst(Tm).data(), Dd.data(), true);
        }
        rep.pcg_iterations += its;
        if (h) h[6] = its;

        // descent check, line search, gradient-step fallback
        double gd = dot3(g, cst(G).data(), cst(Dd).data());
        bool newton = true;
        if (!(gd < 0.0)) {
            for (size_t i = 0; i < d.size(); ++i) d[i] = -gv[i];
            gd = -gnorm * gnorm;
            newton = false;
        }
        double alpha = 1.0, Jt = 0.0, Dt = 0.0, Rt = 0.0;
        bool ok = false;
        for (;;) {
            alpha = 1.0;
            const int trials = o.linesearch ? 30 : 1;
            for (int k = 0; k < trials; ++k) {
                for (int c = 0; c < 3; ++c)
                    for (int64_t i = 0; i < nl; ++i) VT[c][i] = v[c][i] + static_cast<T>(alpha) * Dd[c][i];
                objective(g, nt, norm, beta, cst(VT).data(), mT, mR, &Jt, &Dt, &Rt);
                rep.objective_evals++;
                if (!o.linesearch || (std::isfinite(Jt) && Jt <= J + 1e-4 * alpha * gd)) { ok = true; break; }
                alpha *= 0.5;
            }
            if (ok || !newton) break;
            for (size_t i = 0; i < d.size(); ++i) d[i] = -gv[i];
            gd = -gnorm * gnorm;
            newton = false;
        }
        if (!ok) { rep.reason = kGNLineSearch; break; }
        for (int c = 0; c < 3; ++c) std::copy(VT[c], VT[c] + nl, v[c]);
        J = Jt; D = Dt; R = Rt;
        gradient(g, nt, norm, beta, V.data(), mT, mR, G.data());
        rep.gradient_evals++;
        gnorm = std::sqrt(dot3(g, cst(G).data(), cst(G).data()));
        step = alpha;
        if (h) h[7] = alpha;
        ++iter;
    }
    rep.iterations = iter;
    rep.J = J; rep.D = D; rep.R = R; rep.gnorm = gnorm;
}

}  // namespace oracle
