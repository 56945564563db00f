// kernels.hpp -- host-side launch interface of the claire_amd HIP kernels.
#pragma once

#include <cmath>
#include <vector>

#include "common.hpp"

namespace claire_amd {

struct __align__(8) c32 {
    float x, y;
};

// ---------------------------------------------------------------------------
// FFT (fft.hip)
// ---------------------------------------------------------------------------
bool fft_size_supported(int n);

class FFTPlan3D {
  public:
    explicit FFTPlan3D(const Dims& d);
    // A plan over twiddle tables the caller keeps on the device (tw_ext[a]
    // holds n[a] values e^{-2πik/n[a]}, as fft_twiddles writes them); the
    // plan neither allocates nor frees device memory.
    FFTPlan3D(const Dims& d, c32* const tw_ext[3]);
    ~FFTPlan3D();
    FFTPlan3D(const FFTPlan3D&) = delete;
    FFTPlan3D& operator=(const FFTPlan3D&) = delete;

    // Full transforms (unnormalized).  c2r overwrites its input.
    // c2r epilogue: out = alpha·F⁻¹(in) + beta·add (add may be null).
    void r2c(const float* in, c32* out, hipStream_t st) const;
    void c2r(c32* in, float* out, hipStream_t st, float alpha = 1.f, const float* add = nullptr, float beta = 0.f) const;

    // Individual passes.
    void x3_r2c(const float* in, c32* out, hipStream_t st) const;
    void x3_c2r(const c32* in, float* out, hipStream_t st, float alpha = 1.f, const float* add = nullptr,
                float beta = 0.f) const;
    void x2(c32* data, int dir, hipStream_t st) const;
    void x1(c32* data, int dir, hipStream_t st) const;
    // The same passes over 1 to 3 fields in one launch (add may be null;
    // add[c] may equal out[c] but no other field's out).
    void x3_r2c_n(const float* const in[], c32* const out[], int ncomp, hipStream_t st) const;
    void x3_c2r_n(const c32* const in[], float* const out[], int ncomp, hipStream_t st, float alpha = 1.f,
                  const float* const add[] = nullptr, float beta = 0.f) const;
    void x2_n(c32* const data[], int ncomp, int dir, hipStream_t st) const;
    // Fused x1 passes (forward x1, k-space operator, inverse x1).
    void x1_multiplier(const c32* const in[3], c32* const out[3], int ncomp, const SpecParams& sp, hipStream_t st) const;
    void x1_grad(const c32* in, c32* const out[3], hipStream_t st) const;
    void x1_div(const c32* const in[3], c32* out, hipStream_t st) const;
    // Projection of a vector field (sp.op kOpProjStokes or kOpProjRelaxed):
    // x̂_i = g_i·scale·m·(g·b̂) in place of the three spectra.
    void x1_project(c32* const data[3], const SpecParams& sp, hipStream_t st) const;

    Dims dims;

  private:
    c32* tw[3] = {nullptr, nullptr, nullptr};
    bool owns_tw = true;
};

// The twiddle table of an n-point transform: h[k] = e^{-2πik/n}, k < n,
// evaluated in double and rounded to FP32.
void fft_twiddles(int n, c32* h);

// Spectral gradient without device allocation (noalloc.cpp): g = ∇f with
// the launches of SpectralOps::Grad, its spectra and twiddles in a
// file-scope device array of kNoAllocScratchBytes per device.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
size_t noalloc_grad_bytes(const int nx[3]);
void grad_noalloc(const int nx[3], const float* f, float* g, hipStream_t st);

// ---------------------------------------------------------------------------
// Semi-Lagrangian gathers (interp.hip, gather_lds.hip).  X: normalized query
// points, interleaved xyz (3 floats per grid point).
// ---------------------------------------------------------------------------
struct GatherDims {
    int n1, n2, n3;
    int64_t nl;
};

GatherDims gather_dims(const Dims& d);

// out = f(X)
void launch_gather1(const GatherDims& g, const float* f, const float* X, float* out, hipStream_t st);
// out_d = f_d(X), d = 0..2 (shared stencil weights)
void launch_gather3(const GatherDims& g, const float* const f[3], const float* X, float* const out[3], hipStream_t st);
// RK2 characteristic, Euler predictor and velocity gather fused:
//   X̃ = (x − σ·ht·v)/2π,  X = (x − σ·ht/2·(v(X̃) + v))/2π
void launch_trajectory_rk2(const Dims& d, const float* const v[3], float sigma, float ht, float* X, hipStream_t st);
// RK4 characteristic (three launches); F: work buffer of 3·nl floats.
void launch_trajectory_rk4(const Dims& d, const float* const v[3], float sigma, float ht, float* X, float* F,
                           hipStream_t st);

// Packed plan of a trajectory: 3 words per point (12 B), one per axis, the
// fraction t ∈ [1,2) of the rescaled coordinate u = q·N + 3 and the offset of
// the first stencil node from the point (layout in gather_device.hpp).
// Returns the per-axis stencil displacement bound D in cells (synchronizes
// the stream); 1 << 20 when a foot point is not finite.  n <= 512 per axis.
// `scratch` is one device int.
int launch_make_plan(const GatherDims& g, const float* X, uint32_t* P, int* scratch, hipStream_t st);
// RK2 characteristic through the LDS-staged gather, in three passes:
//   1. the packed plan of the Euler predictor X̃ = (x − σ·ht·v)/2π into P
//      (launch_rk2_predictor_plan, which returns its displacement bound);
//   2. v gathered along it with the corrector as epilogue,
//      X = (x − σ·ht/2·(v(X̃) + v))/2π, one component per field of W (3·nl
//      floats; launch_lds_traj_rk2_corrector, bound <= 4);
//   3. X interleaved into X (3 floats per point) and its plan into P
//      (launch_make_plan_soa, which returns the bound of X).
// The foot points equal those of launch_trajectory_rk2.
int launch_rk2_predictor_plan(const Dims& d, const float* const v[3], float sigma, float ht, uint32_t* P,
                              int* scratch, hipStream_t st);
void launch_lds_traj_rk2_corrector(const Dims& d, const uint32_t* P, int dbound, const float* const v[3], float sigma,
                                   float ht, float* W, hipStream_t st);
int launch_make_plan_soa(const GatherDims& g, const float* W, float* X, uint32_t* P, int* scratch, hipStream_t st);
// RK4 characteristic the same way: the first stage point (x − σ·ht/2·v)/2π
// is launch_rk2_predictor_plan with ht/2; each stage gathers v along the
// previous point's plan (bound <= 4), accumulates the stage sum f and writes
// the next point per component into W (the final X at stage 3), whose plan
// launch_make_plan_soa builds (X null for the intermediate points).  Stage 1
// writes f into Fout (Fin unused), stage 2 reads Fin and writes Fout (two
// distinct buffers: no launch updates an operand in place, so a gather may
// be repeated while its chunk length is tuned), stage 3 reads Fin.
void launch_lds_traj_rk4_stage(int stage, const Dims& d, const uint32_t* P, int dbound, const float* const v[3],
                               float sigma, float ht, const float* Fin, float* Fout, float* W, hipStream_t st);
// The LDS-staged gathers need D <= lds_gather_max_displacement().
int lds_gather_max_displacement();
// Update coefficient per point: C = 1 + ht/2·(α·dX + α·(1 + ht·α·dX)·d)
// with d = ∇·v, dX = d(X).  α = 1 on the adjoint trajectory gives the SL
// adjoint update λ⁺ = C·λ(X_a); α = ±1 on the state trajectory gives the
// SL update of det(∇y^{±1}), J⁺ = C·J(X_s).
void launch_adjoint_coeff(const GatherDims& g, const float* X, const float* divv, float ht, float alpha, float* C,
                          hipStream_t st);

// A trajectory as the gathers see it: the packed plan, the raw foot points
// (3 floats per point, /2π) and the plan's displacement bound.
struct TrajectoryPlan {
    const uint32_t* P;
    const float* X;
    int dbound;
};

// Gathers along a trajectory.  Vector fields are nf equally spaced component
// pointers.  They run the LDS-staged kernels when dbound allows, else the
// direct kernels on the raw foot points.
//   gather:          out_d = f_d(X)
//   inc_state_step:  out = ±(f(X) − s)   (m̃^{j+1} = m̃^j(X_s) − s_j; negated for λ̃⁰ = −m̃^nt)
//   adjoint_step:    out = C·f(X)        (λ^{j+1} = C·λ^j(X_a))
void launch_plan_gather(const TrajectoryPlan& tp, const GatherDims& g, const float* const f[], int nf,
                        float* const out[], hipStream_t st);
void launch_plan_inc_state_step(const TrajectoryPlan& tp, const GatherDims& g, const float* mt, const float* sj,
                                float* out, int negate, hipStream_t st);
void launch_plan_adjoint_step(const TrajectoryPlan& tp, const GatherDims& g, const float* C, const float* lam,
                              float* out, hipStream_t st);
// launch_adjoint_coeff along a trajectory plan (LDS-staged when its bound
// allows); divv and C must not alias.
void launch_plan_adjoint_coeff(const TrajectoryPlan& tp, const GatherDims& g, const float* divv, float ht, float alpha,
                               float* C, hipStream_t st);
// Direct versions on the raw foot points X (any displacement):
void launch_direct_gather(const GatherDims& g, const float* X, const float* const f[], int nf, float* const out[],
                          hipStream_t st);
void launch_direct_inc_state_step(const GatherDims& g, const float* X, const float* mt, const float* sj, float* out,
                                  int negate, hipStream_t st);
void launch_direct_adjoint_step(const GatherDims& g, const float* X, const float* C, const float* lam, float* out,
                                hipStream_t st);

// b_d = hd·Σ_{j=0..nt} w_j·λ^j·∂_d m^{nt−j} (trapezoid weights); lam holds
// λ^0..λ^nt, gm holds ∇m^0..∇m^nt (3 fields each).
void launch_body_force(int64_t nl, int nt, float ht, float hd, const float* lam, const float* gm, float* const b[3],
                       hipStream_t st);
// Fields stored as 16-bit block floating point (the mixed form): int16
// mantissas q (field k at q + k·nl) and one power-of-two scale per block of
// 256 consecutive values of a field (field k's scales at s + k·bfp16_blocks(nl)).
struct Bfp16Fields {
    int16_t* q;
    float* s;
};
int64_t bfp16_blocks(int64_t nl);
// The same with the gradient fields in 16-bit block floating point,
// accumulated in FP32.
void launch_body_force_bfp(int64_t nl, int nt, float ht, float hd, const float* lam, const Bfp16Fields& gm,
                           float* const b[3], hipStream_t st);
// v = 0 body force: b_d = hd·λ·∂_d m
void launch_body_force_static(int64_t nl, float hd, const float* lam, const float* gm, float* const b[3],
                              hipStream_t st);
// s_j = ht/2·[(∇m^j)(X_s)·ṽX + ∇m^{j+1}·ṽ] for j = 0..nt−1 from a precomputed
// ṽX = ṽ(X_s); writes m̃¹ = −s_0 (λ̃⁰ = s_0 when nt = 1) and s_1..s_{nt−1}.
void launch_inc_state_rhs(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3],
                          const float* gX, const float* gm, float* s, float* mt1, hipStream_t st);

void launch_inc_state_rhs_bfp(int64_t nl, int nt, float ht, const float* const vx[3], const float* const vt[3],
                              const Bfp16Fields& gX, const Bfp16Fields& gm, float* s, float* mt1, hipStream_t st);
// nfields consecutive FP32 fields of nl values (nl a multiple of 4) into 16-bit
// block floating point: per block of 256 values, scale 2^(E−15) with
// max |v| < 2^E, mantissas rounded to nearest (error <= 2^−16 of the block's
// largest magnitude).
void launch_to_bfp16(const float* in, int64_t nl, int nfields, const Bfp16Fields& out, hipStream_t st);

// ---------------------------------------------------------------------------
// Element-wise kernels and reductions (pointwise.hip)
// ---------------------------------------------------------------------------
void launch_fill(float* x, float value, int64_t n, hipStream_t st);
void launch_copy(const float* x, float* y, int64_t n, hipStream_t st);
void launch_scale(float* x, float alpha, int64_t n, hipStream_t st);
// y = alpha·x + y
void launch_axpy(float alpha, const float* x, float* y, int64_t n, hipStream_t st);
// w = alpha·x + y
void launch_waxpy(float alpha, const float* x, const float* y, float* w, int64_t n, hipStream_t st);
// z = x − y
void launch_sub(const float* x, const float* y, float* z, int64_t n, hipStream_t st);

// FP64 reductions.  `partial` must hold kReduceBlocks doubles and `result`
// one double (device memory).
constexpr int kReduceBlocks = 1024;

// Device-resident scalars of a conjugate-gradient solve (reg::Optimizer::Pcg),
// so that a run of iterations needs no host round trip: the step lengths are
// formed on the device from the reduction slots, in double, exactly as the
// host would form them, and every update is skipped once the solve has
// stopped.  Slots: 0 = pᵀw, 1 = rᵀr, 2 = rᵀz.
struct PcgState {
    double rz;     // rᵀz of the current residual
    double tol;    // stop when ‖r‖ <= tol
    double bnorm;  // ‖b‖ (divergence test ‖r‖ >= 1e6·‖b‖)
    double rnorm;  // ‖r‖ after the last iteration
    float a, na, beta;  // float(α), float(−α), float(β) for the updates
    int k;         // iterations taken
    int kmax;      // iteration limit
    int done;      // 0 while running; 1 once stopped
};
// s.rz = slots[2], s.k = 0, s.done = 0, tolerances as given
void launch_pcg_init(PcgState* s, const double* slots, double tol, double bnorm, int kmax, hipStream_t st);
// k += 1; α = rz/pᵀw, or stop (before the updates) when pᵀw is not positive
void launch_pcg_alpha(PcgState* s, const double* slots, hipStream_t st);
// ‖r‖ test and the limit, else β = rᵀz_new/rᵀz, rz = rᵀz_new
void launch_pcg_beta(PcgState* s, const double* slots, hipStream_t st);
// y += (sign > 0 ? α : −α)·x unless stopped
void launch_axpy_pcg(const PcgState* s, int sign, const float* x, float* y, int64_t n, hipStream_t st);
// p = β·p + z unless stopped
void launch_xpay_pcg(const PcgState* s, float* p, const float* z, int64_t n, hipStream_t st);
// d += α·p and r −= α·w in one launch unless stopped
void launch_axpy2_pcg(const PcgState* s, const float* p, float* d, const float* w, float* r, int64_t n,
                      hipStream_t st);
// slots[0] = pᵀw (partial: kReduceBlocks doubles), then the α step
void launch_dot_pcg_alpha(const float* p, const float* w, int64_t n, double* partial, double* slots, PcgState* s,
                          hipStream_t st);
// slots[1] = rᵀr and slots[2] = rᵀz in one pass (partial: 2·kReduceBlocks
// doubles), then the stopping test and the β step
void launch_dots_pcg_beta(const float* r, const float* z, int64_t n, double* partial, double* slots, PcgState* s,
                          hipStream_t st);
void launch_dot(const float* a, const float* b, int64_t n, double* partial, double* result, hipStream_t st);
void launch_absmax(const float* a, int64_t n, double* partial, double* result, hipStream_t st);
void launch_sqdist(const float* a, const float* b, int64_t n, double* partial, double* result, hipStream_t st);
void launch_min(const float* a, int64_t n, double* partial, double* result, hipStream_t st);
void launch_max(const float* a, int64_t n, double* partial, double* result, hipStream_t st);
void launch_sum(const float* a, int64_t n, double* partial, double* result, hipStream_t st);

// ---------------------------------------------------------------------------
// Explicit RK2 transport stages (rk2.hip); each follows a spectral grad or
// div of the step (src/CLAIRE.cpp:1995-2105, 2365-2548, 2962-3135, 3504-3683)
// ---------------------------------------------------------------------------
struct Vec3Ptr {
    const float* p[3];
};
struct Vec3Out {
    float* p[3];
};
// rhs0 = −g·v, m̄ = m + ht·rhs0
void launch_rk2_state_stage1(int64_t nl, float ht, const Vec3Ptr& g, const Vec3Ptr& v, const float* m, float* rhs0,
                             float* mbar, hipStream_t st);
// m⁺ = m + ht/2·(rhs0 − g·v)
void launch_rk2_state_stage2(int64_t nl, float hthalf, const Vec3Ptr& g, const Vec3Ptr& v, const float* m,
                             const float* rhs0, float* mnext, hipStream_t st);
// rhs0 = −∇m̃·v − ∇m·ṽ, m̃̄ = m̃ + ht·rhs0 (has_gmt = false: m̃ = 0)
void launch_rk2_inc_stage1(int64_t nl, float ht, bool has_gmt, const Vec3Ptr& gmt, const Vec3Ptr& v,
                           const Vec3Ptr& gm, const Vec3Ptr& vt, const float* mt, float* rhs0, float* mtbar,
                           hipStream_t st);
// out = ±(m̃ + ht/2·(rhs1 + rhs0)), rhs1 = −∇m̃̄·v − ∇m⁺·ṽ (has_mt = false: m̃ = 0)
void launch_rk2_inc_stage2(int64_t nl, float hthalf, bool has_mt, const Vec3Ptr& gmt, const Vec3Ptr& v,
                           const Vec3Ptr& gm, const Vec3Ptr& vt, const float* mt, const float* rhs0, float* out,
                           bool negate, hipStream_t st);
// v = 0: out = ±m̃ after nt steps m̃ ← m̃ − ht·(∇m·ṽ) from m̃ = 0
void launch_rk2_inc_static(int64_t nl, int nt, float ht, const Vec3Ptr& gm, const Vec3Ptr& vt, float* out, bool negate,
                           hipStream_t st);
// w = λ·v
void launch_rk2_adj_flux(int64_t nl, const float* lam, const Vec3Ptr& v, const Vec3Out& w, hipStream_t st);
// w = v·(λ + ht·rhs0)
void launch_rk2_adj_mid(int64_t nl, float ht, const float* lam, const float* rhs0, const Vec3Ptr& v, const Vec3Out& w,
                        hipStream_t st);
// λ⁺ = λ + ht/2·(rhs0 + rhs1); next: w = λ⁺·v
void launch_rk2_adj_end(int64_t nl, float hthalf, const float* lam, const float* rhs0, const float* rhs1, float* lnext,
                        bool next, const Vec3Ptr& v, const Vec3Out& w, hipStream_t st);

}  // namespace claire_amd
