// claire.hpp -- host classes of the MI355X registration hot path.
//
// The class and method names mirror the reference's C++ API so that the
// code above the solver boundary reads the same:
//   RegOpt            include/RegOpt.hpp:475-638 (the hot-path subset)
//   VecField          include/VecField.hpp:35-106
//   SemiLagrangian    include/SemiLagrangian.hpp:37-94
//   Regularization*   include/Regularization.hpp:35-62 and the H1/H2/H1SN/H2SN subclasses
//   CLAIRE            include/CLAIRE.hpp:37-188 (GN, semi-Lagrangian, nc = 1)
// PETSc Vec arguments become device pointers to flat FP32 vectors; a vector
// field is [v1|v2|v3] with each component x3-fastest (VecField::
// SetComponents/GetComponents, src/VecField.cpp:396-476).  Methods return an
// int error code like PetscErrorCode (0 = OK); failures throw
// claire_amd::Error internally and are converted at the C-ABI boundary.
#pragma once

#include <functional>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"

namespace reg {

using claire_amd::c32;
using claire_amd::Dims;
using ScalarType = float;
using IntType = int64_t;

// Owning device allocation.
template <class T>
class DeviceArray {
  public:
    DeviceArray() = default;
    explicit DeviceArray(size_t n) { resize(n); }
    ~DeviceArray() { release(); }
    DeviceArray(const DeviceArray&) = delete;
    DeviceArray& operator=(const DeviceArray&) = delete;
    void resize(size_t n) {
        if (n == n_) return;
        release();
        if (n) {
            hipError_t e = hipMalloc(&p_, n * sizeof(T));
            if (e != hipSuccess) {
                p_ = nullptr;
                throw claire_amd::Error(claire_amd::kErrOutOfMemory,
                                        "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
                                            hipGetErrorString(e));
            }
        }
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }

  private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

// Per-kernel-class timing with HIP events recorded on the context stream.
class Profiler {
  public:
    struct Entry {
        double ms = 0.0;
        double bytes = 0.0;
        long count = 0;
    };
    explicit Profiler(hipStream_t s) : stream_(s) {}
    ~Profiler();
    bool enabled() const { return enabled_; }
    void enable(bool on);
    // Stream of later begin/end events (the problem's stream follows
    // CLAIRE::SetStream; events already recorded keep theirs).
    void set_stream(hipStream_t s) { stream_ = s; }
    void begin(const char* name, double bytes);
    void end();
    // Synchronize, fold the recorded events into the totals.
    void collect();
    void reset();
    const std::map<std::string, Entry>& entries() const { return totals_; }

  private:
    struct Pending {
        std::string name;
        double bytes;
        hipEvent_t a, b;
    };
    hipEvent_t get_event();
    hipStream_t stream_;
    bool enabled_ = false;
    std::vector<Pending> pending_;
    std::vector<hipEvent_t> pool_;
    std::map<std::string, Entry> totals_;
};

class KernelScope {
  public:
    KernelScope(Profiler* p, const char* name, double bytes) : p_(p && p->enabled() ? p : nullptr) {
        if (p_) p_->begin(name, bytes);
    }
    ~KernelScope() {
        if (p_) p_->end();
    }

  private:
    Profiler* p_;
};

// The hot-path subset of the reference's RegOpt.
struct RegOpt {
    Dims domain{};
    int nt = 4;
    int nc = 1;
    int regnorm = claire_amd::kRegH1SN;
    float beta[3] = {1e-2f, 1e-4f, 0.0f};
    hipStream_t stream = nullptr;
    Profiler* prof = nullptr;
    // 0: plain FP32.  1: mixed form of the Hessian matvec -- the
    // Newton-invariant gradient fields ∇m^j and (∇m^j)(X_s) read by the
    // streaming kernels are stored as 16-bit block floating point (int16
    // mantissas, one power-of-two scale per 256 values; FP32 arithmetic);
    // objective, gradient and reductions are unchanged.
    int precision = 0;
    // Replay the Hessian matvec as a captured HIP graph (CLAIRE::HessianMatVec).
    int graphs = 0;
    // Registration model (m_RegModel): compressible (CLAIRE), Stokes
    // (CLAIREStokes) or relaxed Stokes (CLAIREDivReg).
    int regmodel = claire_amd::kModelCompressible;
    // Hessian operator (m_KrylovMethod.matvectype, include/RegOpt.hpp:94-98):
    // 0 DEFAULTMATVEC, 1 PRECONDMATVEC, 2 PRECONDMATVECSYM.
    int matvectype = 0;
    // Order of the characteristic scheme, 2 or 4 (m_PDESolver.rkorder,
    // src/RegOpt.cpp:674-676, default 2 at src/RegOpt.cpp:1087).
    int rkorder = 2;
    // Constraint of the state solve (m_PDESolver.pdetype, include/RegOpt.hpp:42-46):
    // 0 continuity equation, 1 transport equation (default, src/RegOpt.cpp:1089).
    int pdetype = 1;
    // Time integrator of the transport solves (m_PDESolver.type,
    // include/RegOpt.hpp:31-35): 0 RK2, 2 SL (default, src/RegOpt.cpp:1083).
    int pdesolver = 2;
    // include/RegOpt.hpp:485-500, src/RegOpt.cpp:2505-2512
    float GetLebesgueMeasure() const { return domain.hd; }
    float GetTimeStepSize() const { return 1.0f / static_cast<float>(nt); }
    float ComputeFFTScale() const { return domain.scale; }
};

// FP64 reduction scratch.
class Reducer {
  public:
    explicit Reducer(const RegOpt* opt);
    double Dot(const float* a, const float* b, IntType n);
    double AbsMax(const float* a, IntType n);
    double SqDist(const float* a, const float* b, IntType n);
    double Min(const float* a, IntType n);
    double Max(const float* a, IntType n);
    double Sum(const float* a, IntType n);
    // Several dot products into result slots (0 <= slot < kSlots), read back
    // with one synchronization by Fetch: the host waits once instead of
    // once per product.
    static constexpr int kSlots = 16;
    void DotTo(const float* a, const float* b, IntType n, int slot);
    // The CG products with the CG scalar step fused into the final sum:
    // slot 0 = pᵀw then α; slots 1, 2 = rᵀr, rᵀz (one pass) then β.
    void DotAlpha(const float* p, const float* w, IntType n, claire_amd::PcgState* s);
    void DotsBeta(const float* r, const float* z, IntType n, claire_amd::PcgState* s);
    void Fetch(double* out, int nslots);
    // Device address of the result slots (read by kernels that consume the
    // products without a host round trip).
    const double* Slots() const { return m_Result.get(); }

  private:
    double fetch();
    const RegOpt* m_Opt;
    DeviceArray<double> m_Partial, m_Result;
};

class VecField {
  public:
    // Owning field (three contiguous components).
    explicit VecField(const RegOpt* opt);
    // Non-owning view of a flat [v1|v2|v3] device vector.
    VecField(const RegOpt* opt, float* flat);
    float* m_X1 = nullptr;
    float* m_X2 = nullptr;
    float* m_X3 = nullptr;
    float* Flat() const { return m_X1; }
    float* Comp(int d) const { return d == 0 ? m_X1 : (d == 1 ? m_X2 : m_X3); }
    int SetComponents(const float* flat);
    int GetComponents(float* flat) const;
    int SetValue(ScalarType v);
    int Copy(const VecField* x);
    int Scale(ScalarType a);
    int AXPY(ScalarType a, const VecField* x);
    int WAXPY(ScalarType a, const VecField* x, const VecField* y);

  private:
    const RegOpt* m_Opt;
    DeviceArray<float> m_Storage;
};

// Replaces AccFFT's plan and spectral operators (single GPU).
class SpectralOps {
  public:
    explicit SpectralOps(const RegOpt* opt);
    void R2C(const float* in, c32* out);
    void C2R(c32* in, float* out, float alpha = 1.f, const float* add = nullptr, float beta = 0.f);
    void Grad(float* const g[3], const float* f);
    void Div(float* d, const float* const v[3]);
    // out_d = alpha·F⁻¹(m(k)·F(in_d)) + beta·add_d, d = 0..2
    void Multiplier3(const claire_amd::SpecParams& sp, const float* const in[3], float* const out[3], float alpha,
                     const float* const add[3], float beta);
    // b ← b + F⁻¹[x̂] with the projection of sp.op (kOpProjStokes or
    // kOpProjRelaxed), in place on the three components.
    void Project(const claire_amd::SpecParams& sp, float* const b[3]);
    c32* Spectrum(int i) const { return m_Spec[i].get(); }

  private:
    const RegOpt* m_Opt;
    std::unique_ptr<claire_amd::FFTPlan3D> m_Plan;
    DeviceArray<c32> m_Spec[3];
};

class Regularization {
  public:
    Regularization(const RegOpt* opt, SpectralOps* spec, Reducer* red) : m_Opt(opt), m_Spec(spec), m_Red(red) {}
    virtual ~Regularization() = default;
    virtual int EvaluateFunctional(double* R, const VecField* v);
    virtual int EvaluateGradient(VecField* dvR, const VecField* v);
    virtual int HessianMatVec(VecField* dvvR, const VecField* vtilde) { return EvaluateGradient(dvvR, vtilde); }
    virtual int ApplyInverse(VecField* Ainvx, const VecField* x, bool applysqrt);
    virtual int GetExtremeEigValsInvOp(ScalarType& emin, ScalarType& emax);
    // Fused form used by HessianMatVec: out = alpha·R''v + beta·add.
    void ApplyFused(VecField* out, const VecField* v, float alpha, const VecField* add, float beta);
    virtual int Norm() const = 0;

  protected:
    claire_amd::SpecParams Params(int op) const;
    const RegOpt* m_Opt;
    SpectralOps* m_Spec;
    Reducer* m_Red;
    // work field of EvaluateFunctional, allocated on its first call and kept
    // (no allocation inside the Newton loop)
    std::unique_ptr<VecField> m_Work;
};

class RegularizationH1 : public Regularization {
  public:
    using Regularization::Regularization;
    int Norm() const override { return claire_amd::kRegH1; }
};
class RegularizationH2 : public Regularization {
  public:
    using Regularization::Regularization;
    int Norm() const override { return claire_amd::kRegH2; }
};
class RegularizationH1SN : public Regularization {
  public:
    using Regularization::Regularization;
    int Norm() const override { return claire_amd::kRegH1SN; }
};
class RegularizationH2SN : public Regularization {
  public:
    using Regularization::Regularization;
    int Norm() const override { return claire_amd::kRegH2SN; }
};

class SemiLagrangian {
  public:
    explicit SemiLagrangian(const RegOpt* opt);
    // flag "state" (σ = +1) or "adjoint" (σ = −1)
    int ComputeTrajectory(const VecField* v, const std::string& flag);
    int Interpolate(float* xo, const float* xi, const std::string& flag);
    int Interpolate(VecField* vo, const VecField* vi, const std::string& flag);
    // Out-of-place gathers along a trajectory (no aliasing allowed).
    void Gather(float* xo, const float* xi, const std::string& flag);
    void Gather3(const float* const in[3], float* const out[3], const std::string& flag);
    // Trajectory X (3 floats per point, normalized by 2π; the reference's m_X)
    const float* Trajectory(const std::string& flag) const;
    // Interpolation plan of the trajectory (Interp3_Plan after scatter):
    // packed foot points (3 words per point), the raw foot points and the
    // per-axis stencil displacement bound in cells.  Gathers along it use the
    // LDS-staged kernels when the bound is at most
    // lds_gather_max_displacement(), else the direct kernels.
    claire_amd::TrajectoryPlan Plan(const std::string& flag) const;
    bool HasTrajectory(const std::string& flag) const;
    int DisplacementBound(const std::string& flag) const { return m_D[Which(flag)]; }

  private:
    int Which(const std::string& flag) const;
    const RegOpt* m_Opt;
    DeviceArray<float> m_X[2];
    DeviceArray<uint32_t> m_P[2];
    DeviceArray<float> m_F;   // RK4 stage sum f (m_WorkVecField2 of the reference)
    DeviceArray<float> m_F2;  // RK4 stage sum after stage 2 (LDS path; kept apart from m_F)
    DeviceArray<float> m_W;  // RK2 corrector output, one component per field (LDS path)
    DeviceArray<float> m_Tmp;  // in-place interpolation (the reference's ghost copy), kept between calls
    DeviceArray<int> m_Scratch;
    bool m_Valid[2] = {false, false};
    int m_D[2] = {0, 0};
};

class CLAIRE {
  public:
    explicit CLAIRE(const RegOpt& opt);
    ~CLAIRE();

    int SetControlVariable(const float* v);
    int SetIncControlVariable(const float* vtilde);
    int SetTemplateImage(const float* mT);
    int SetReferenceImage(const float* mR);

    int SolveStateEquation();
    int SolveAdjointEquation();
    int SolveIncStateEquation();
    int SolveIncAdjointEquation();
    // Hṽ = R''ṽ + b̃ (scale = false: divided by hd)
    int HessianMatVec(float* Hvtilde, const float* vtilde, bool scale = true);
    int EvaluateObjective(double* J, double* D, double* R, const float* v);
    int EvaluateGradient(float* g, const float* v);
    int ApplyInvRegularizationOperator(float* Ainvx, const float* x, bool applysqrt);

    // Accessors for the C-ABI building blocks and tests.
    const float* StateVariable() const { return m_State.get(); }
    const float* AdjointVariable() const { return m_Adjoint.get(); }
    const VecField* BodyForce() const { return m_BodyForce.get(); }
    const float* IncStateVariable() const { return m_IncState[m_IncStateSlot].get(); }
    const float* IncAdjointVariable() const { return m_IncAdjointT0.get(); }
    const VecField* IncBodyForce() const { return m_IncBodyForce.get(); }
    SemiLagrangian* SL() { return m_SL.get(); }
    SpectralOps* Spectral() { return m_Spectral.get(); }
    Regularization* Reg() { return m_Regularization.get(); }
    Reducer* Red() { return m_Red.get(); }
    const RegOpt& Opt() const { return m_Opt; }
    bool VelocityIsZero() const { return m_VelocityIsZero; }
    size_t DeviceBytes() const;
    // Newton-invariant data used by the Hessian (computed on demand).
    int PrepareHessian();
    // 0 plain FP32, 1 mixed (RegOpt::precision).
    int SetPrecision(int precision);
    // Characteristic scheme of every later trajectory: 2 or 4
    // (SemiLagrangian::ComputeTrajectory, src/SemiLagrangian.cpp:172-178).
    int SetRKOrder(int order);
    // Constraint of every later state solve: 0 continuity, 1 transport
    // (CLAIRE::SolveStateEquation, src/CLAIRE.cpp:1923-1926).
    int SetPDEType(int type);
    // Time integrator of every later transport solve: 0 explicit RK2
    // ("-pdesolver rk2"), 2 semi-Lagrangian (src/RegOpt.cpp:653-658).
    int SetPDESolver(int type);
    // Operator applied by HessianMatVec: 0 the GN Hessian (HessMatVec),
    // 1 hd·ṽ + (βA)⁻¹b̃[ṽ] (PrecondHessMatVec), 2 hd·ṽ +
    // (βA)^{-1/2}b̃[(βA)^{-1/2}ṽ] (PrecondHessMatVecSym).
    int SetMatVecType(int type);
    // Registration model of every later solve (kModelCompressible,
    // kModelStokes, kModelRelaxedStokes); drops what was computed before.
    int SetRegModel(int model);
    // The model's ApplyProjection on a flat vector field, in place (a no-op
    // for the compressible model, src/CLAIRE.cpp:4044-4048).
    int ApplyProjection(float* v);
    // Stream of every later launch and of the profiler's events (work
    // already queued stays where it is).
    void SetStream(hipStream_t s) {
        m_Opt.stream = s;
        if (m_Opt.prof) m_Opt.prof->set_stream(s);
    }
    // Capture the Hessian matvec into HIP graphs and replay them (on/off).
    int SetMatVecGraphs(bool on);
    bool MatVecGraphs() const { return m_Opt.graphs != 0; }
    // What a captured graph of this problem depends on (trajectories and
    // their displacement bound, the velocity-is-zero branch, precision,
    // operator type): graphs of an older epoch must not be replayed.
    uint64_t GraphEpoch() const { return m_GraphEpoch; }
    // The Hessian matvec without graph replay, for callers that capture a
    // larger body around it (the Newton-invariant caches must be current:
    // PrepareHessian first).
    void HessianMatVecInline(float* Hvtilde, const float* vtilde, bool scale);
    // Captures body() on the problem's private capture stream (body launches
    // on Opt().stream, which points there during the capture) and
    // instantiates it into *exec.
    void CaptureGraph(const std::function<void()>& body, hipGraphExec_t* exec);
    int MatVecGraphCount() const { return static_cast<int>(m_Graphs.size()); }
    // det(∇y) (inverse = false) or det(∇y⁻¹) by SL transport
    // (DeformationFields::ComputeDetDefGrad); jac may be null; stats receive
    // min, max and mean.
    int ComputeDetDefGrad(bool inverse, float* jac, double stats[3]);
    void EnsureTrajectories();
    int SolveIncAdjointFrom(const float* mtilde);

  private:
    void RequireVelocity() const;
    void EnsureStateGradients();
    void EnsureGradientsAtStateTrajectory();
    void IncStateCore(const VecField* vt, float* lt0_out, bool half);
    void IncAdjointCore(const float* lt0, VecField* bt, bool half);
    void EnsureHalfGradients();
    // The Newton-invariant fields the incremental solves of the selected
    // integrator read (and their half copies in the mixed SL form).
    void EnsureMatVecCaches();
    bool MixedKernels() const { return m_Opt.precision == 1 && m_Opt.pdesolver == 2; }
    // Explicit RK2 integrators (src/CLAIRE.cpp:1995-2105, 2365-2548,
    // 2962-3135, 3504-3683); the adjoint form fills λ^1..λ^nt of m_Lam.
    void Rk2StateSteps();
    void Rk2IncState(const VecField* vt, float* lt0_out);
    void Rk2AdjointSteps();
    void HessianMatVecBody(float* Hvtilde, const VecField* vt, bool scale);
    void ProjectField(VecField* b);
    double DivergencePenalty();
    void CaptureMatVec(float* Hvtilde, const VecField* vt, bool scale, hipGraphExec_t* exec);
    void DropGraphs();

    // Captured Hessian matvecs, one per (input, output, scale) triple.  The
    // kernels, their launch shapes and every buffer address are fixed at
    // capture; what changes them (new trajectories with their displacement
    // bound, the velocity-is-zero branch, the precision) bumps m_GraphEpoch,
    // and graphs of an older epoch are dropped.  A triple runs eagerly on its
    // first two uses in an epoch (lazy allocations happen there, and a
    // Newton step whose Krylov solve needs only one or two matvecs pays no
    // capture), is captured on its third and replayed from then on.
    struct MatVecGraph {
        const float* in;
        float* out;
        bool scale;
        uint64_t epoch;
        int uses;
        hipGraphExec_t exec;
    };
    std::vector<MatVecGraph> m_Graphs;
    uint64_t m_GraphEpoch = 1;
    hipStream_t m_CaptureStream = nullptr;

    RegOpt m_Opt;
    std::unique_ptr<Reducer> m_Red;
    std::unique_ptr<SpectralOps> m_Spectral;
    std::unique_ptr<Regularization> m_Regularization;
    std::unique_ptr<SemiLagrangian> m_SL;

    std::unique_ptr<VecField> m_VelocityField, m_IncVelocityField, m_BodyForce, m_IncBodyForce, m_WorkVecField1;
    DeviceArray<float> m_TemplateImage, m_ReferenceImage;
    DeviceArray<float> m_State;          // (nt+1)·nl
    DeviceArray<float> m_Adjoint;        // λ at t = 0
    DeviceArray<float> m_IncState[2];    // ping-pong m̃
    int m_IncStateSlot = 0;
    DeviceArray<float> m_IncAdjointT0;   // λ̃ at t = 0
    DeviceArray<float> m_GradM;          // ∇m^j, (nt+1)·3·nl
    DeviceArray<float> m_GradMX;         // (∇m^j)(X_s), nt·3·nl
    DeviceArray<float> m_S;              // inc-state decrements, nt·nl
    DeviceArray<float> m_Lam;            // adjoint time series λ^0..λ^nt, (nt+1)·nl
    DeviceArray<float> m_Scalar;         // nl work
    DeviceArray<float> m_AdjCoeff;       // nl: adjoint update coefficient at every point
    DeviceArray<float> m_StateCoeff;     // nl: continuity-equation state coefficient (pdetype 0 only)
    DeviceArray<float> m_Rk2Work;        // 6·nl, RK2 integrators only: a vector field, rhs0, rhs1 / m̄
    DeviceArray<float> m_VtX;            // ṽ(X_s), 3·nl
    // 16-bit block-floating-point copies of m_GradM / m_GradMX (mixed form):
    // mantissas and per-256-value scales
    DeviceArray<int16_t> m_GradM16, m_GradMX16;
    DeviceArray<float> m_GradMScale, m_GradMXScale;
    claire_amd::Bfp16Fields GradMBfp() { return {m_GradM16.get(), m_GradMScale.get()}; }
    claire_amd::Bfp16Fields GradMXBfp() { return {m_GradMX16.get(), m_GradMXScale.get()}; }
    bool m_HalfValid = false;

    bool m_HaveVelocity = false, m_HaveIncVelocity = false, m_HaveTemplate = false, m_HaveReference = false;
    bool m_VelocityIsZero = false;
    bool m_TrajValid = false, m_StateValid = false, m_GradMValid = false, m_GradMXValid = false;
};

}  // namespace reg
