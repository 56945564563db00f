// claire.cpp -- host side of the MI355X registration hot path.
#include "claire.hpp"

#include <algorithm>

#include <cstring>

namespace claire_amd {

Dims make_dims(const int nx[3]) {
    Dims d{};
    for (int a = 0; a < 3; ++a) {
        CLR_REQUIRE(nx[a] >= 4, kErrInvalidArgument, "grid size must be >= 4 per axis");
        d.n[a] = nx[a];
        // hx = 2π/n with the division in double, stored in FP32 (src/RegOpt.cpp:935)
        d.hx[a] = static_cast<float>(3.14159265358979323846 * 2.0 / static_cast<double>(static_cast<float>(nx[a])));
    }
    d.nh = nx[2] / 2 + 1;
    d.nl = static_cast<int64_t>(nx[0]) * nx[1] * nx[2];
    d.nlc = static_cast<int64_t>(nx[0]) * nx[1] * d.nh;
    d.hd = d.hx[0] * d.hx[1] * d.hx[2];
    float s = 1.0f;
    for (int a = 0; a < 3; ++a) s *= static_cast<float>(nx[a]);
    d.scale = static_cast<float>(1.0 / static_cast<double>(s));
    return d;
}

}  // namespace claire_amd

namespace reg {

using namespace claire_amd;

// ---------------------------------------------------------------------------
// Profiler
// ---------------------------------------------------------------------------
Profiler::~Profiler() {
    for (auto& p : pending_) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
}

void Profiler::enable(bool on) { enabled_ = on; }

hipEvent_t Profiler::get_event() {
    if (!pool_.empty()) {
        hipEvent_t e = pool_.back();
        pool_.pop_back();
        return e;
    }
    hipEvent_t e;
    CLR_HIP_CHECK(hipEventCreate(&e));
    return e;
}

void Profiler::begin(const char* name, double bytes) {
    Pending p{name, bytes, get_event(), get_event()};
    CLR_HIP_CHECK(hipEventRecord(p.a, stream_));
    pending_.push_back(p);
}

void Profiler::end() { CLR_HIP_CHECK(hipEventRecord(pending_.back().b, stream_)); }

void Profiler::collect() {
    for (auto& p : pending_) {
        // each pair waits on its own end event: the pairs may sit on
        // different streams when the problem moved between streams
        CLR_HIP_CHECK(hipEventSynchronize(p.b));
        float ms = 0.f;
        CLR_HIP_CHECK(hipEventElapsedTime(&ms, p.a, p.b));
        Entry& e = totals_[p.name];
        e.ms += ms;
        e.bytes += p.bytes;
        e.count += 1;
        pool_.push_back(p.a);
        pool_.push_back(p.b);
    }
    pending_.clear();
}

void Profiler::reset() {
    collect();
    totals_.clear();
}

// ---------------------------------------------------------------------------
// Reducer (FP64)
// ---------------------------------------------------------------------------
Reducer::Reducer(const RegOpt* opt) : m_Opt(opt), m_Partial(2 * kReduceBlocks), m_Result(kSlots) {}

void Reducer::DotAlpha(const float* p, const float* w, IntType n, claire_amd::PcgState* s) {
    KernelScope k(m_Opt->prof, "reduce_dot", 8.0 * n);
    launch_dot_pcg_alpha(p, w, n, m_Partial.get(), m_Result.get(), s, m_Opt->stream);
}

void Reducer::DotsBeta(const float* r, const float* z, IntType n, claire_amd::PcgState* s) {
    KernelScope k(m_Opt->prof, "reduce_dot2", 8.0 * n);
    launch_dots_pcg_beta(r, z, n, m_Partial.get(), m_Result.get(), s, m_Opt->stream);
}

void Reducer::DotTo(const float* a, const float* b, IntType n, int slot) {
    CLR_REQUIRE(slot >= 0 && slot < kSlots, kErrInvalidArgument, "reduction slot out of range");
    KernelScope k(m_Opt->prof, "reduce_dot", 8.0 * n);
    launch_dot(a, b, n, m_Partial.get(), m_Result.get() + slot, m_Opt->stream);
}

void Reducer::Fetch(double* out, int nslots) {
    CLR_REQUIRE(nslots >= 1 && nslots <= kSlots, kErrInvalidArgument, "reduction slot count out of range");
    CLR_HIP_CHECK(hipMemcpyAsync(out, m_Result.get(), sizeof(double) * nslots, hipMemcpyDeviceToHost, m_Opt->stream));
    CLR_HIP_CHECK(hipStreamSynchronize(m_Opt->stream));
}

double Reducer::fetch() {
    double h = 0.0;
    CLR_HIP_CHECK(hipMemcpyAsync(&h, m_Result.get(), sizeof(double), hipMemcpyDeviceToHost, m_Opt->stream));
    CLR_HIP_CHECK(hipStreamSynchronize(m_Opt->stream));
    return h;
}

double Reducer::Dot(const float* a, const float* b, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_dot", 8.0 * n);
        launch_dot(a, b, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

double Reducer::AbsMax(const float* a, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_absmax", 4.0 * n);
        launch_absmax(a, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

double Reducer::Min(const float* a, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_min", 4.0 * n);
        launch_min(a, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

double Reducer::Max(const float* a, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_max", 4.0 * n);
        launch_max(a, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

double Reducer::Sum(const float* a, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_sum", 4.0 * n);
        launch_sum(a, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

double Reducer::SqDist(const float* a, const float* b, IntType n) {
    {
        KernelScope k(m_Opt->prof, "reduce_sqdist", 8.0 * n);
        launch_sqdist(a, b, n, m_Partial.get(), m_Result.get(), m_Opt->stream);
    }
    return fetch();
}

// ---------------------------------------------------------------------------
// VecField (src/VecField.cpp)
// ---------------------------------------------------------------------------
VecField::VecField(const RegOpt* opt) : m_Opt(opt), m_Storage(3 * static_cast<size_t>(opt->domain.nl)) {
    const int64_t nl = opt->domain.nl;
    m_X1 = m_Storage.get();
    m_X2 = m_X1 + nl;
    m_X3 = m_X2 + nl;
}

VecField::VecField(const RegOpt* opt, float* flat) : m_Opt(opt) {
    CLR_REQUIRE(flat != nullptr, kErrInvalidArgument, "null vector field");
    const int64_t nl = opt->domain.nl;
    m_X1 = flat;
    m_X2 = flat + nl;
    m_X3 = flat + 2 * nl;
}

int VecField::SetComponents(const float* flat) {
    CLR_REQUIRE(flat != nullptr, kErrInvalidArgument, "null input vector");
    if (flat != m_X1)
        CLR_HIP_CHECK(hipMemcpyAsync(m_X1, flat, sizeof(float) * 3 * m_Opt->domain.nl, hipMemcpyDeviceToDevice,
                                     m_Opt->stream));
    return 0;
}

int VecField::GetComponents(float* flat) const {
    CLR_REQUIRE(flat != nullptr, kErrInvalidArgument, "null output vector");
    if (flat != m_X1)
        CLR_HIP_CHECK(hipMemcpyAsync(flat, m_X1, sizeof(float) * 3 * m_Opt->domain.nl, hipMemcpyDeviceToDevice,
                                     m_Opt->stream));
    return 0;
}

int VecField::SetValue(ScalarType v) {
    launch_fill(m_X1, v, 3 * m_Opt->domain.nl, m_Opt->stream);
    return 0;
}

int VecField::Copy(const VecField* x) { return SetComponents(x->m_X1); }

int VecField::Scale(ScalarType a) {
    launch_scale(m_X1, a, 3 * m_Opt->domain.nl, m_Opt->stream);
    return 0;
}

int VecField::AXPY(ScalarType a, const VecField* x) {
    launch_axpy(a, x->m_X1, m_X1, 3 * m_Opt->domain.nl, m_Opt->stream);
    return 0;
}

int VecField::WAXPY(ScalarType a, const VecField* x, const VecField* y) {
    launch_waxpy(a, x->m_X1, y->m_X1, m_X1, 3 * m_Opt->domain.nl, m_Opt->stream);
    return 0;
}

// ---------------------------------------------------------------------------
// SpectralOps (AccFFT replacement)
// ---------------------------------------------------------------------------
SpectralOps::SpectralOps(const RegOpt* opt) : m_Opt(opt), m_Plan(new FFTPlan3D(opt->domain)) {
    for (auto& s : m_Spec) s.resize(static_cast<size_t>(opt->domain.nlc));
}

namespace {
double bytes_x3_r2c(const Dims& d) { return 4.0 * d.nl + 8.0 * d.nlc; }
double bytes_col(const Dims& d) { return 16.0 * d.nlc; }
double bytes_x3_c2r(const Dims& d, bool add) { return 8.0 * d.nlc + 4.0 * d.nl + (add ? 4.0 * d.nl : 0.0); }
}  // namespace

void SpectralOps::R2C(const float* in, c32* out) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    { KernelScope k(m_Opt->prof, "fft_x3_r2c", bytes_x3_r2c(d)); m_Plan->x3_r2c(in, out, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", bytes_col(d)); m_Plan->x2(out, -1, st); }
    { KernelScope k(m_Opt->prof, "fft_x1", bytes_col(d)); m_Plan->x1(out, -1, st); }
}

void SpectralOps::C2R(c32* in, float* out, float alpha, const float* add, float beta) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    { KernelScope k(m_Opt->prof, "fft_x1", bytes_col(d)); m_Plan->x1(in, +1, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", bytes_col(d)); m_Plan->x2(in, +1, st); }
    { KernelScope k(m_Opt->prof, "fft_x3_c2r", bytes_x3_c2r(d, add)); m_Plan->x3_c2r(in, out, st, alpha, add, beta); }
}

void SpectralOps::Grad(float* const g[3], const float* f) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    c32* s[3] = {m_Spec[0].get(), m_Spec[1].get(), m_Spec[2].get()};
    { KernelScope k(m_Opt->prof, "fft_x3_r2c", bytes_x3_r2c(d)); m_Plan->x3_r2c(f, s[0], st); }
    { KernelScope k(m_Opt->prof, "fft_x2", bytes_col(d)); m_Plan->x2(s[0], -1, st); }
    { KernelScope k(m_Opt->prof, "fft_x1_grad", 32.0 * d.nlc); m_Plan->x1_grad(s[0], s, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, +1, st); }
    { KernelScope k(m_Opt->prof, "fft_x3_c2r", 3.0 * bytes_x3_c2r(d, false)); m_Plan->x3_c2r_n(s, g, 3, st); }
}

void SpectralOps::Div(float* out, const float* const v[3]) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    c32* s[3] = {m_Spec[0].get(), m_Spec[1].get(), m_Spec[2].get()};
    { KernelScope k(m_Opt->prof, "fft_x3_r2c", 3.0 * bytes_x3_r2c(d)); m_Plan->x3_r2c_n(v, s, 3, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, -1, st); }
    { KernelScope k(m_Opt->prof, "fft_x1_div", 32.0 * d.nlc); m_Plan->x1_div(s, s[0], st); }
    { KernelScope k(m_Opt->prof, "fft_x2", bytes_col(d)); m_Plan->x2(s[0], +1, st); }
    { KernelScope k(m_Opt->prof, "fft_x3_c2r", bytes_x3_c2r(d, false)); m_Plan->x3_c2r(s[0], out, st); }
}

void SpectralOps::Multiplier3(const SpecParams& sp, const float* const in[3], float* const out[3], float alpha,
                              const float* const add[3], float beta) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    c32* s[3] = {m_Spec[0].get(), m_Spec[1].get(), m_Spec[2].get()};
    { KernelScope k(m_Opt->prof, "fft_x3_r2c", 3.0 * bytes_x3_r2c(d)); m_Plan->x3_r2c_n(in, s, 3, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, -1, st); }
    { KernelScope k(m_Opt->prof, "fft_x1_mult", 3.0 * bytes_col(d)); m_Plan->x1_multiplier(s, s, 3, sp, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, +1, st); }
    {
        KernelScope k(m_Opt->prof, "fft_x3_c2r", 3.0 * bytes_x3_c2r(d, add != nullptr));
        m_Plan->x3_c2r_n(s, out, 3, st, alpha, add, beta);
    }
}

void SpectralOps::Project(const SpecParams& sp, float* const b[3]) {
    const Dims& d = m_Opt->domain;
    hipStream_t st = m_Opt->stream;
    c32* s[3] = {m_Spec[0].get(), m_Spec[1].get(), m_Spec[2].get()};
    const float* in[3] = {b[0], b[1], b[2]};
    { KernelScope k(m_Opt->prof, "fft_x3_r2c", 3.0 * bytes_x3_r2c(d)); m_Plan->x3_r2c_n(in, s, 3, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, -1, st); }
    { KernelScope k(m_Opt->prof, "fft_x1_project", 3.0 * bytes_col(d)); m_Plan->x1_project(s, sp, st); }
    { KernelScope k(m_Opt->prof, "fft_x2", 3.0 * bytes_col(d)); m_Plan->x2_n(s, 3, +1, st); }
    {
        // b + F⁻¹[x̂]: the c2r epilogue adds the field in place (the AXPY of
        // src/CLAIREStokes.cpp:536, src/CLAIREDivReg.cpp:452)
        KernelScope k(m_Opt->prof, "fft_x3_c2r", 3.0 * bytes_x3_c2r(d, true));
        m_Plan->x3_c2r_n(s, b, 3, st, 1.0f, in, 1.0f);
    }
}

// ---------------------------------------------------------------------------
// Regularization (src/Regularization*.cpp)
// ---------------------------------------------------------------------------
SpecParams Regularization::Params(int op) const {
    SpecParams sp{};
    sp.op = op;
    sp.norm = Norm();
    sp.beta0 = m_Opt->beta[0];
    sp.beta1 = m_Opt->beta[1];
    sp.hd = m_Opt->GetLebesgueMeasure();
    sp.scale = m_Opt->ComputeFFTScale();
    return sp;
}

int Regularization::EvaluateGradient(VecField* dvR, const VecField* v) {
    if (m_Opt->beta[0] == 0.0f) return dvR->SetValue(0.0f);  // src/RegularizationH1.cpp:188-190
    const float* in[3] = {v->m_X1, v->m_X2, v->m_X3};
    float* out[3] = {dvR->m_X1, dvR->m_X2, dvR->m_X3};
    m_Spec->Multiplier3(Params(kOpRegApply), in, out, 1.0f, nullptr, 0.0f);
    return 0;
}

void Regularization::ApplyFused(VecField* out, const VecField* v, float alpha, const VecField* add, float beta) {
    if (m_Opt->beta[0] == 0.0f) {
        if (add) {
            launch_copy(add->m_X1, out->m_X1, 3 * m_Opt->domain.nl, m_Opt->stream);
            if (beta != 1.0f) launch_scale(out->m_X1, beta, 3 * m_Opt->domain.nl, m_Opt->stream);
        } else {
            out->SetValue(0.0f);
        }
        return;
    }
    const float* in[3] = {v->m_X1, v->m_X2, v->m_X3};
    float* o[3] = {out->m_X1, out->m_X2, out->m_X3};
    const float* a[3] = {add ? add->m_X1 : nullptr, add ? add->m_X2 : nullptr, add ? add->m_X3 : nullptr};
    m_Spec->Multiplier3(Params(kOpRegApply), in, o, alpha, add ? a : nullptr, beta);
}

int Regularization::ApplyInverse(VecField* Ainvx, const VecField* x, bool applysqrt) {
    if (m_Opt->beta[0] == 0.0f) return Ainvx->Copy(x);  // src/RegularizationH1.cpp:336-340
    const float* in[3] = {x->m_X1, x->m_X2, x->m_X3};
    float* out[3] = {Ainvx->m_X1, Ainvx->m_X2, Ainvx->m_X3};
    m_Spec->Multiplier3(Params(applysqrt ? kOpRegInvSqrt : kOpRegInverse), in, out, 1.0f, nullptr, 0.0f);
    return 0;
}

int Regularization::EvaluateFunctional(double* R, const VecField* v) {
    *R = 0.0;
    if (m_Opt->beta[0] == 0.0f) return 0;
    const int64_t nl = m_Opt->domain.nl;
    const double hd = m_Opt->GetLebesgueMeasure();
    if (!m_Work) m_Work.reset(new VecField(m_Opt));
    float* w[3] = {m_Work->m_X1, m_Work->m_X2, m_Work->m_X3};
    const int norm = Norm();
    double acc = 0.0;
    double dots[Reducer::kSlots];
    if (norm == kRegH1 || norm == kRegH1SN) {
        // Σ_i ‖∇v_i‖² (src/RegularizationH1SN.cpp:84-141); the nine products
        // (and ‖v‖² for H1) are read back together and summed in order
        for (int c = 0; c < 3; ++c) {
            const float* vc = c == 0 ? v->m_X1 : (c == 1 ? v->m_X2 : v->m_X3);
            m_Spec->Grad(w, vc);
            for (int k = 0; k < 3; ++k) m_Red->DotTo(w[k], w[k], nl, 3 * c + k);
        }
        if (norm == kRegH1) m_Red->DotTo(v->m_X1, v->m_X1, 3 * nl, 9);
        m_Red->Fetch(dots, norm == kRegH1 ? 10 : 9);
        for (int i = 0; i < 9; ++i) acc += dots[i];
        if (norm == kRegH1) {
            // ½·hd·β0·(H1v + β1·L2v), src/RegularizationH1.cpp:140-147
            *R = 0.5 * hd * m_Opt->beta[0] * (acc + m_Opt->beta[1] * dots[9]);
        } else {
            *R = 0.5 * hd * m_Opt->beta[0] * acc;
        }
        return 0;
    }
    const float* in[3] = {v->m_X1, v->m_X2, v->m_X3};
    SpecParams sp = Params(norm == kRegH2 ? kOpH2Functional : kOpScaledLaplace);
    if (norm == kRegH2) {
        // src/RegularizationH2.cpp:80-81,128: multiplier scale·√β0·(lapik + √β1)
        sp.beta0 = std::sqrt(m_Opt->beta[0]);
        sp.beta1 = std::sqrt(m_Opt->beta[1]);
    }
    m_Spec->Multiplier3(sp, in, w, 1.0f, nullptr, 0.0f);
    for (int k = 0; k < 3; ++k) m_Red->DotTo(w[k], w[k], nl, k);
    m_Red->Fetch(dots, 3);
    for (int k = 0; k < 3; ++k) acc += dots[k];
    if (norm == kRegH2)
        *R = 0.5 * hd * acc;                         // src/RegularizationH2.cpp:164-166
    else
        *R = 0.5 * m_Opt->beta[0] * hd * acc;        // src/RegularizationH2SN.cpp:163
    return 0;
}

int Regularization::GetExtremeEigValsInvOp(ScalarType& emin, ScalarType& emax) {
    float w[3];
    for (int i = 0; i < 3; ++i) w[i] = static_cast<float>(m_Opt->domain.n[i]) / 2.0f;
    float regop = -(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    const float b1 = m_Opt->beta[0], b2 = m_Opt->beta[1];
    switch (Norm()) {
        case kRegH1:  // src/RegularizationH1.cpp:440-448
            regop = b1 * (-regop + b2);
            emin = 1.0f / regop;
            emax = 1.0f / b2;
            break;
        case kRegH2:  // src/RegularizationH2.cpp:439-446
            regop = b1 * (regop * regop) + b2;
            emin = 1.0f / regop;
            emax = 1.0f / b2;
            break;
        case kRegH1SN:  // src/RegularizationH1SN.cpp:418-424
            regop = -b1 * regop + b2;
            emin = 1.0f / regop;
            emax = 1.0f / b2;
            break;
        default:  // src/RegularizationH2SN.cpp:447-455 ("by definition; it's 1/0")
            regop = b1 * (regop * regop);
            emin = 1.0f / regop;
            emax = 1.0f;
            break;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// SemiLagrangian (src/SemiLagrangian.cpp)
// ---------------------------------------------------------------------------
SemiLagrangian::SemiLagrangian(const RegOpt* opt) : m_Opt(opt), m_Scratch(1) {}

int SemiLagrangian::Which(const std::string& flag) const {
    if (flag == "state") return 0;
    if (flag == "adjoint") return 1;
    throw Error(kErrInvalidArgument, "flag wrong");  // src/SemiLagrangian.cpp:221-223
}

int SemiLagrangian::ComputeTrajectory(const VecField* v, const std::string& flag) {
    const int w = Which(flag);
    m_X[w].resize(3 * static_cast<size_t>(m_Opt->domain.nl));
    const float* vc[3] = {v->m_X1, v->m_X2, v->m_X3};
    const float sigma = w == 0 ? 1.0f : -1.0f;
    if (m_Opt->rkorder == 4) {
        const int64_t nl = m_Opt->domain.nl;
        const GatherDims g = gather_dims(m_Opt->domain);
        const float ht = m_Opt->GetTimeStepSize();
        m_F.resize(3 * static_cast<size_t>(nl));
        m_P[w].resize(3 * static_cast<size_t>(nl));
        const int dmax = lds_gather_max_displacement();
        int db = -1;
        {
            KernelScope k(m_Opt->prof, "gather_traj_rk4_plan", 24.0 * nl);
            db = launch_rk2_predictor_plan(m_Opt->domain, vc, sigma, 0.5f * ht, m_P[w].get(), m_Scratch.get(),
                                           m_Opt->stream);
        }
        if (vc[1] - vc[0] == nl && vc[2] - vc[1] == nl) {
            m_W.resize(3 * static_cast<size_t>(nl));
            m_F2.resize(3 * static_cast<size_t>(nl));
            // stage sums: stage 1 writes f into m_F, stage 2 reads m_F and
            // writes m_F2, stage 3 reads m_F2
            float* const fbuf[4] = {nullptr, m_F.get(), m_F2.get(), nullptr};
            for (int stage = 1; stage <= 3 && db >= 0 && db <= dmax; ++stage) {
                {
                    // 3 × (plan 12 + gathered v 4 + aux 4 + next point 4 (+ f 4)) per point
                    KernelScope k(m_Opt->prof, "gather_traj_rk4", (stage < 3 ? 84.0 : 72.0) * nl);
                    launch_lds_traj_rk4_stage(stage, m_Opt->domain, m_P[w].get(), db, vc, sigma, ht,
                                              fbuf[stage - 1], fbuf[stage], m_W.get(), m_Opt->stream);
                }
                // bytes: point components read (12) + plan (12) (+ X interleaved (12) at the end)
                KernelScope k(m_Opt->prof, "gather_plan", (stage < 3 ? 24.0 : 36.0) * nl);
                db = launch_make_plan_soa(g, m_W.get(), stage == 3 ? m_X[w].get() : nullptr, m_P[w].get(),
                                          m_Scratch.get(), m_Opt->stream);
                if (stage == 3) {
                    m_D[w] = db;
                    m_Valid[w] = true;
                    return 0;
                }
            }
        }
        // bytes: v read (12), v gathered 3× (36), f and X written, read and
        // rewritten (12 + 12 + 24 + 24 + 12 + 12) per point
        KernelScope k(m_Opt->prof, "gather_traj_rk4", 144.0 * m_Opt->domain.nl);
        launch_trajectory_rk4(m_Opt->domain, vc, sigma, m_Opt->GetTimeStepSize(), m_X[w].get(), m_F.get(),
                              m_Opt->stream);
    } else {
        CLR_REQUIRE(m_Opt->rkorder == 2, kErrInvalidArgument, "rk order must be 2 or 4");
        const int64_t nl = m_Opt->domain.nl;
        const GatherDims g = gather_dims(m_Opt->domain);
        m_P[w].resize(3 * static_cast<size_t>(nl));
        int dpred = -1;
        {
            // the Euler predictor's plan: v read (12) + plan written (12) per point
            KernelScope k(m_Opt->prof, "gather_traj_rk2_plan", 24.0 * nl);
            dpred = launch_rk2_predictor_plan(m_Opt->domain, vc, sigma, m_Opt->GetTimeStepSize(), m_P[w].get(),
                                              m_Scratch.get(), m_Opt->stream);
        }
        if (dpred >= 0 && dpred <= lds_gather_max_displacement() && vc[1] - vc[0] == nl && vc[2] - vc[1] == nl) {
            m_W.resize(3 * static_cast<size_t>(nl));
            {
                // 3 × (plan 12 + gathered v 4 + v at the point 4 + X component 4) per point
                KernelScope k(m_Opt->prof, "gather_traj_rk2", 72.0 * nl);
                launch_lds_traj_rk2_corrector(m_Opt->domain, m_P[w].get(), dpred, vc, sigma,
                                              m_Opt->GetTimeStepSize(), m_W.get(), m_Opt->stream);
            }
            // bytes: X components read (12) + X interleaved (12) + plan (12) per point
            KernelScope k(m_Opt->prof, "gather_plan", 36.0 * nl);
            m_D[w] = launch_make_plan_soa(g, m_W.get(), m_X[w].get(), m_P[w].get(), m_Scratch.get(), m_Opt->stream);
            m_Valid[w] = true;
            return 0;
        }
        // bytes: v read (12) + v gathered (12) + X written (12) per point
        KernelScope k(m_Opt->prof, "gather_traj_rk2", 36.0 * nl);
        launch_trajectory_rk2(m_Opt->domain, vc, sigma, m_Opt->GetTimeStepSize(), m_X[w].get(), m_Opt->stream);
    }
    m_P[w].resize(3 * static_cast<size_t>(m_Opt->domain.nl));
    {
        // bytes: X read (12) + packed plan written (12) per point
        KernelScope k(m_Opt->prof, "gather_plan", 24.0 * m_Opt->domain.nl);
        m_D[w] = launch_make_plan(gather_dims(m_Opt->domain), m_X[w].get(), m_P[w].get(), m_Scratch.get(),
                                  m_Opt->stream);
    }
    m_Valid[w] = true;
    return 0;
}

TrajectoryPlan SemiLagrangian::Plan(const std::string& flag) const {
    const int w = Which(flag);
    CLR_REQUIRE(m_Valid[w], kErrState, "trajectory '" + flag + "' has not been computed");
    return TrajectoryPlan{m_P[w].get(), m_X[w].get(), m_D[w]};
}

bool SemiLagrangian::HasTrajectory(const std::string& flag) const { return m_Valid[Which(flag)]; }

const float* SemiLagrangian::Trajectory(const std::string& flag) const {
    const int w = Which(flag);
    CLR_REQUIRE(m_Valid[w], kErrState, "trajectory '" + flag + "' has not been computed");
    return m_X[w].get();
}

void SemiLagrangian::Gather(float* xo, const float* xi, const std::string& flag) {
    const TrajectoryPlan tp = Plan(flag);
    const GatherDims g = gather_dims(m_Opt->domain);
    const float* in[1] = {xi};
    float* out[1] = {xo};
    // bytes: plan 12 + gathered field 4 + output 4 per point
    KernelScope k(m_Opt->prof, "gather_scalar", 20.0 * g.nl);
    launch_plan_gather(tp, g, in, 1, out, m_Opt->stream);
}

void SemiLagrangian::Gather3(const float* const in[3], float* const out[3], const std::string& flag) {
    const TrajectoryPlan tp = Plan(flag);
    const GatherDims g = gather_dims(m_Opt->domain);
    const bool spaced = (in[1] - in[0] == in[2] - in[1]) && (out[1] - out[0] == out[2] - out[1]);
    if (!spaced) {
        for (int d = 0; d < 3; ++d) Gather(out[d], in[d], flag);
        return;
    }
    // bytes: 3 × (plan 12 + gathered field 4 + output 4) per point
    KernelScope k(m_Opt->prof, "gather_vector", 60.0 * g.nl);
    launch_plan_gather(tp, g, in, 3, out, m_Opt->stream);
}

int SemiLagrangian::Interpolate(float* xo, const float* xi, const std::string& flag) {
    const GatherDims g = gather_dims(m_Opt->domain);
    if (xo == xi) {
        // in place: gather into the kept work buffer, then copy back
        // (stream-ordered, so the buffer is free for the next call)
        if (m_Tmp.size() < static_cast<size_t>(g.nl)) m_Tmp.resize(static_cast<size_t>(g.nl));
        Gather(m_Tmp.get(), xi, flag);
        launch_copy(m_Tmp.get(), xo, g.nl, m_Opt->stream);
        return 0;
    }
    Gather(xo, xi, flag);
    return 0;
}

int SemiLagrangian::Interpolate(VecField* vo, const VecField* vi, const std::string& flag) {
    const float* in[3] = {vi->m_X1, vi->m_X2, vi->m_X3};
    if (vo->m_X1 == vi->m_X1) {
        const int64_t nl = m_Opt->domain.nl;
        if (m_Tmp.size() < static_cast<size_t>(3 * nl)) m_Tmp.resize(static_cast<size_t>(3 * nl));
        float* out[3] = {m_Tmp.get(), m_Tmp.get() + nl, m_Tmp.get() + 2 * nl};
        Gather3(in, out, flag);
        vo->SetComponents(m_Tmp.get());
        return 0;
    }
    float* out[3] = {vo->m_X1, vo->m_X2, vo->m_X3};
    Gather3(in, out, flag);
    return 0;
}

// ---------------------------------------------------------------------------
// CLAIRE (src/CLAIRE.cpp, GN + semi-Lagrangian + nc = 1)
// ---------------------------------------------------------------------------
CLAIRE::CLAIRE(const RegOpt& opt) : m_Opt(opt) {
    CLR_REQUIRE(m_Opt.nt >= 1, kErrInvalidArgument, "nt must be >= 1");
    CLR_REQUIRE(m_Opt.nc == 1, kErrInvalidArgument, "only nc = 1 is supported");
    const size_t nl = static_cast<size_t>(m_Opt.domain.nl);
    const size_t nt = static_cast<size_t>(m_Opt.nt);
    m_Red.reset(new Reducer(&m_Opt));
    m_Spectral.reset(new SpectralOps(&m_Opt));
    switch (m_Opt.regnorm) {  // CLAIREBase::SetupRegularization, src/CLAIREBase.cpp:701-790
        case kRegH1: m_Regularization.reset(new RegularizationH1(&m_Opt, m_Spectral.get(), m_Red.get())); break;
        case kRegH2: m_Regularization.reset(new RegularizationH2(&m_Opt, m_Spectral.get(), m_Red.get())); break;
        case kRegH1SN: m_Regularization.reset(new RegularizationH1SN(&m_Opt, m_Spectral.get(), m_Red.get())); break;
        case kRegH2SN: m_Regularization.reset(new RegularizationH2SN(&m_Opt, m_Spectral.get(), m_Red.get())); break;
        default: throw Error(kErrInvalidArgument, "regularization norm not implemented");
    }
    m_SL.reset(new SemiLagrangian(&m_Opt));
    m_VelocityField.reset(new VecField(&m_Opt));
    m_IncVelocityField.reset(new VecField(&m_Opt));
    m_BodyForce.reset(new VecField(&m_Opt));
    m_IncBodyForce.reset(new VecField(&m_Opt));
    m_WorkVecField1.reset(new VecField(&m_Opt));
    m_TemplateImage.resize(nl);
    m_ReferenceImage.resize(nl);
    m_State.resize((nt + 1) * nl);
    m_Adjoint.resize(nl);
    m_IncState[0].resize(nl);
    m_IncState[1].resize(nl);
    m_IncAdjointT0.resize(nl);
    m_GradM.resize((nt + 1) * 3 * nl);
    m_GradMX.resize(nt * 3 * nl);
    m_S.resize(nt * nl);
    m_Lam.resize((nt + 1) * nl);
    m_Scalar.resize(nl);
    m_AdjCoeff.resize(nl);
    m_VtX.resize(3 * nl);
}

CLAIRE::~CLAIRE() {
    DropGraphs();
    if (m_CaptureStream) (void)hipStreamDestroy(m_CaptureStream);
}

void CLAIRE::DropGraphs() {
    for (auto& g : m_Graphs)
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
    m_Graphs.clear();
}

int CLAIRE::SetMatVecGraphs(bool on) {
    m_Opt.graphs = on ? 1 : 0;
    if (!on) DropGraphs();
    return 0;
}

size_t CLAIRE::DeviceBytes() const {
    const size_t nl = static_cast<size_t>(m_Opt.domain.nl), nt = static_cast<size_t>(m_Opt.nt);
    const size_t floats = 5 * 3 * nl + 2 * nl + (nt + 1) * nl + nl + 3 * nl + 4 * nl + (nt + 1) * 3 * nl +
                          nt * 3 * nl + nt * nl + nt * nl + nl + 2 * 3 * nl;
    const size_t halves = m_Opt.precision == 1 ? (2 * nt + 1) * 3 * nl : 0;
    const size_t scales = m_Opt.precision == 1 ? (2 * nt + 1) * 3 * static_cast<size_t>(bfp16_blocks(nl)) : 0;
    // the corrector / stage-point output, plus the two RK4 stage sums
    const size_t rk4 = 3 * nl + (m_Opt.rkorder == 4 ? 6 * nl : 0);
    const size_t cont = m_Opt.pdetype == 0 ? nl : 0;      // continuity-equation coefficient
    const size_t rk2 = m_Opt.pdesolver == 0 ? 6 * nl : 0;  // RK2 integrator work
    return (floats + rk4 + cont + rk2 + scales) * 4 + halves * 2 + 3 * static_cast<size_t>(m_Opt.domain.nlc) * 8;
}

void CLAIRE::RequireVelocity() const { CLR_REQUIRE(m_HaveVelocity, kErrState, "velocity field not set"); }

int CLAIRE::SetControlVariable(const float* v) {
    m_VelocityField->SetComponents(v);
    m_HaveVelocity = true;
    m_TrajValid = m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
    ++m_GraphEpoch;
    // CLAIREBase::IsVelocityZero (src/CLAIREBase.cpp:578-605): ∞-norm == 0
    m_VelocityIsZero = m_Red->AbsMax(m_VelocityField->m_X1, 3 * m_Opt.domain.nl) == 0.0;
    return 0;
}

int CLAIRE::SetIncControlVariable(const float* vtilde) {
    m_IncVelocityField->SetComponents(vtilde);
    m_HaveIncVelocity = true;
    return 0;
}

int CLAIRE::SetTemplateImage(const float* mT) {
    CLR_REQUIRE(mT != nullptr, kErrInvalidArgument, "null template image");
    CLR_HIP_CHECK(hipMemcpyAsync(m_TemplateImage.get(), mT, sizeof(float) * m_Opt.domain.nl, hipMemcpyDeviceToDevice,
                                 m_Opt.stream));
    m_HaveTemplate = true;
    m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
    return 0;
}

int CLAIRE::SetReferenceImage(const float* mR) {
    CLR_REQUIRE(mR != nullptr, kErrInvalidArgument, "null reference image");
    CLR_HIP_CHECK(hipMemcpyAsync(m_ReferenceImage.get(), mR, sizeof(float) * m_Opt.domain.nl, hipMemcpyDeviceToDevice,
                                 m_Opt.stream));
    m_HaveReference = true;
    return 0;
}

void CLAIRE::EnsureTrajectories() {
    RequireVelocity();
    if (m_TrajValid) return;
    m_SL->ComputeTrajectory(m_VelocityField.get(), "state");
    m_SL->ComputeTrajectory(m_VelocityField.get(), "adjoint");
    // div v and the adjoint update coefficient at every point
    // (src/CLAIRE.cpp:2614-2635 / 3907-3927 compute ∇·v and (∇·v)(X) per solve)
    const float* vc[3] = {m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3};
    m_Spectral->Div(m_Scalar.get(), vc);
    if (m_Opt.regmodel == kModelStokes) {
        // incompressible adjoints are pure transport (src/CLAIREStokes.cpp:169-406): c = 1
        launch_fill(m_AdjCoeff.get(), 1.0f, m_Opt.domain.nl, m_Opt.stream);
    } else {
        // bytes: plan 12 + div v gathered 4 + div v 4 + C 4 per point
        KernelScope k(m_Opt.prof, "gather_adjoint_coeff", 24.0 * m_Opt.domain.nl);
        launch_plan_adjoint_coeff(m_SL->Plan("adjoint"), gather_dims(m_Opt.domain), m_Scalar.get(),
                                  m_Opt.GetTimeStepSize(), 1.0f, m_AdjCoeff.get(), m_Opt.stream);
    }
    if (m_Opt.pdetype == 0) {
        // continuity-equation state update m^{j+1} = c·m^j(X_s) with
        // c = 1 − ht/2·(dX + (1 − ht·dX)·d), d = ∇·v, dX = d(X_s): the α = −1
        // form of the coefficient on the state trajectory
        // (SolveContinuityEquationSL, src/CLAIRE.cpp:2813-2835)
        m_StateCoeff.resize(static_cast<size_t>(m_Opt.domain.nl));
        KernelScope k(m_Opt.prof, "gather_state_coeff", 24.0 * m_Opt.domain.nl);
        launch_plan_adjoint_coeff(m_SL->Plan("state"), gather_dims(m_Opt.domain), m_Scalar.get(),
                                  m_Opt.GetTimeStepSize(), -1.0f, m_StateCoeff.get(), m_Opt.stream);
    }
    m_TrajValid = true;
    ++m_GraphEpoch;
}

int CLAIRE::SolveStateEquation() {
    RequireVelocity();
    CLR_REQUIRE(m_HaveTemplate, kErrState, "template image not set");
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    // the SL solves and the continuity equation need the trajectories
    if (m_Opt.pdesolver == 2 || m_Opt.pdetype == 0) EnsureTrajectories();
    float* m = m_State.get();
    launch_copy(m_TemplateImage.get(), m, nl, m_Opt.stream);  // m^0 = mT (src/CLAIRE.cpp:1896)
    if (m_VelocityIsZero) {
        for (int j = 1; j <= nt; ++j) launch_copy(m, m + j * nl, nl, m_Opt.stream);  // src/CLAIRE.cpp:1909-1920
    } else if (m_Opt.pdetype == 0) {
        // SolveContinuityEquationSL (src/CLAIRE.cpp:2735-2854)
        const TrajectoryPlan tp = m_SL->Plan("state");
        const GatherDims g = gather_dims(m_Opt.domain);
        for (int j = 0; j < nt; ++j) {
            KernelScope k(m_Opt.prof, "continuity_step", 24.0 * nl);
            launch_plan_adjoint_step(tp, g, m_StateCoeff.get(), m + j * nl, m + (j + 1) * nl, m_Opt.stream);
        }
    } else if (m_Opt.pdesolver == 0) {
        Rk2StateSteps();  // SolveStateEquationRK2 (src/CLAIRE.cpp:1995-2105)
    } else {
        for (int j = 0; j < nt; ++j)  // src/CLAIRE.cpp:2153-2163
            m_SL->Gather(m + (j + 1) * nl, m + j * nl, "state");
    }
    m_StateValid = true;
    m_GradMValid = m_GradMXValid = m_HalfValid = false;
    return 0;
}

void CLAIRE::EnsureStateGradients() {
    if (!m_StateValid) SolveStateEquation();
    if (m_GradMValid) return;
    const int64_t nl = m_Opt.domain.nl;
    for (int j = 0; j <= m_Opt.nt; ++j) {
        float* gj = m_GradM.get() + 3 * j * nl;
        float* g[3] = {gj, gj + nl, gj + 2 * nl};
        m_Spectral->Grad(g, m_State.get() + j * nl);
    }
    m_GradMValid = true;
    m_GradMXValid = m_HalfValid = false;
}

void CLAIRE::EnsureGradientsAtStateTrajectory() {
    EnsureStateGradients();
    EnsureTrajectories();
    if (m_GradMXValid) return;
    const int64_t nl = m_Opt.domain.nl;
    for (int j = 0; j < m_Opt.nt; ++j) {  // (∇m^j)(X_s), src/CLAIRE.cpp:3243-3247
        const float* gj = m_GradM.get() + 3 * j * nl;
        float* gx = m_GradMX.get() + 3 * j * nl;
        const float* in[3] = {gj, gj + nl, gj + 2 * nl};
        float* out[3] = {gx, gx + nl, gx + 2 * nl};
        m_SL->Gather3(in, out, "state");
    }
    m_GradMXValid = true;
}

void CLAIRE::EnsureHalfGradients() {
    EnsureGradientsAtStateTrajectory();
    if (m_HalfValid) return;
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const int64_t nblk = bfp16_blocks(nl);
    m_GradM16.resize(static_cast<size_t>((nt + 1) * 3 * nl));
    m_GradMX16.resize(static_cast<size_t>(nt * 3 * nl));
    m_GradMScale.resize(static_cast<size_t>((nt + 1) * 3 * nblk));
    m_GradMXScale.resize(static_cast<size_t>(nt * 3 * nblk));
    launch_to_bfp16(m_GradM.get(), nl, (nt + 1) * 3, GradMBfp(), m_Opt.stream);
    launch_to_bfp16(m_GradMX.get(), nl, nt * 3, GradMXBfp(), m_Opt.stream);
    m_HalfValid = true;
}

int CLAIRE::ComputeDetDefGrad(bool inverse, float* jac, double stats[3]) {
    // DeformationFields::ComputeDetDefGrad / ComputeDetDefGradSL
    // (src/DeformationFields.cpp:315-388, 636-740): J⁰ = 1, then nt SL steps
    // along the state trajectory J^{j+1} = C·J^j(X_s) with
    // C = 1 + ht/2·(α·dX + α·(1 + ht·α·dX)·d), d = ∇·v, dX = d(X_s), α = ±1.
    RequireVelocity();
    EnsureTrajectories();
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const GatherDims g = gather_dims(m_Opt.domain);
    // work space of the problem, free outside a solve: ∇·v in m_Scalar, the
    // coefficient in the first slot of the inc-state decrements m_S
    float* divv = m_Scalar.get();
    float* coeff = m_S.get();
    const float* vc[3] = {m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3};
    m_Spectral->Div(divv, vc);
    const TrajectoryPlan tp = m_SL->Plan("state");
    {
        KernelScope k(m_Opt.prof, "gather_det_coeff", 24.0 * nl);
        launch_plan_adjoint_coeff(tp, g, divv, m_Opt.GetTimeStepSize(), inverse ? -1.0f : 1.0f, coeff, m_Opt.stream);
    }
    float* buf[2] = {m_IncState[0].get(), m_IncState[1].get()};
    launch_fill(buf[0], 1.0f, nl, m_Opt.stream);
    for (int j = 0; j < nt; ++j) {
        KernelScope k(m_Opt.prof, "det_step", 24.0 * nl);
        launch_plan_adjoint_step(tp, g, coeff, buf[j % 2], buf[(j + 1) % 2], m_Opt.stream);
    }
    const float* J = buf[nt % 2];
    stats[0] = m_Red->Min(J, nl);
    stats[1] = m_Red->Max(J, nl);
    stats[2] = m_Red->Sum(J, nl) / static_cast<double>(nl);
    if (jac) launch_copy(J, jac, nl, m_Opt.stream);
    CLR_HIP_CHECK(hipStreamSynchronize(m_Opt.stream));
    return 0;
}

int CLAIRE::SetPrecision(int precision) {
    CLR_REQUIRE(precision == 0 || precision == 1, kErrInvalidArgument, "precision must be 0 (FP32) or 1 (mixed)");
    if (precision != m_Opt.precision) ++m_GraphEpoch;
    m_Opt.precision = precision;
    return 0;
}

int CLAIRE::SetRKOrder(int order) {
    // src/RegOpt.cpp:1690-1693
    CLR_REQUIRE(order == 2 || order == 4, kErrInvalidArgument, "options for rk order are 2 and 4");
    if (order != m_Opt.rkorder) {
        m_Opt.rkorder = order;
        m_TrajValid = m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
    }
    return 0;
}

int CLAIRE::SetPDEType(int type) {
    CLR_REQUIRE(type == 0 || type == 1, kErrInvalidArgument, "pde type must be 0 (continuity) or 1 (transport)");
    if (type != m_Opt.pdetype) {
        m_Opt.pdetype = type;
        m_TrajValid = m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
        ++m_GraphEpoch;
    }
    return 0;
}

int CLAIRE::SetPDESolver(int type) {
    // RK2A ("stabilized RK2") has no transport solvers in the reference:
    // SolveStateEquation and the adjoint/incremental solves throw "PDE solver
    // not implemented" for it (src/CLAIRE.cpp:1927-1944, 2315-2330)
    CLR_REQUIRE(type == 0 || type == 2, kErrInvalidArgument,
                "pde solver must be 0 (rk2) or 2 (sl); rk2a has no transport solvers");
    if (type != m_Opt.pdesolver) {
        m_Opt.pdesolver = type;
        if (type == 0) m_Rk2Work.resize(static_cast<size_t>(6 * m_Opt.domain.nl));
        m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
        ++m_GraphEpoch;
    }
    return 0;
}

void CLAIRE::EnsureMatVecCaches() {
    if (m_Opt.pdesolver == 0) {
        EnsureStateGradients();
        return;
    }
    EnsureGradientsAtStateTrajectory();
    if (m_Opt.precision == 1) EnsureHalfGradients();
}

int CLAIRE::PrepareHessian() {
    if (m_Opt.pdesolver == 2) EnsureTrajectories();
    EnsureMatVecCaches();
    CLR_HIP_CHECK(hipStreamSynchronize(m_Opt.stream));
    return 0;
}

void CLAIRE::Rk2StateSteps() {
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const float ht = m_Opt.GetTimeStepSize();
    float* w = m_Rk2Work.get();
    float* gw[3] = {w, w + nl, w + 2 * nl};
    const Vec3Ptr G{{gw[0], gw[1], gw[2]}};
    const Vec3Ptr V{{m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3}};
    float* rhs0 = w + 3 * nl;
    float* mbar = w + 4 * nl;
    float* m = m_State.get();
    for (int j = 0; j < nt; ++j) {
        m_Spectral->Grad(gw, m + j * nl);  // ∇m^j
        {
            KernelScope k(m_Opt.prof, "rk2_state_stage", 36.0 * nl);
            launch_rk2_state_stage1(nl, ht, G, V, m + j * nl, rhs0, mbar, m_Opt.stream);
        }
        m_Spectral->Grad(gw, mbar);  // ∇m̄
        KernelScope k(m_Opt.prof, "rk2_state_stage", 36.0 * nl);
        launch_rk2_state_stage2(nl, 0.5f * ht, G, V, m + j * nl, rhs0, m + (j + 1) * nl, m_Opt.stream);
    }
}

void CLAIRE::Rk2IncState(const VecField* vt, float* lt0_out) {
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const float ht = m_Opt.GetTimeStepSize();
    const Vec3Ptr VT{{vt->m_X1, vt->m_X2, vt->m_X3}};
    auto grad_m = [&](int j) {
        const float* gj = m_GradM.get() + 3 * static_cast<int64_t>(j) * nl;
        return Vec3Ptr{{gj, gj + nl, gj + 2 * nl}};
    };
    if (m_VelocityIsZero) {
        // src/CLAIRE.cpp:3022-3052
        KernelScope k(m_Opt.prof, "rk2_inc_stage", 28.0 * nl);
        launch_rk2_inc_static(nl, nt, ht, grad_m(0), VT, lt0_out, true, m_Opt.stream);
        return;
    }
    float* w = m_Rk2Work.get();
    float* gw[3] = {w, w + nl, w + 2 * nl};
    const Vec3Ptr G{{gw[0], gw[1], gw[2]}};
    const Vec3Ptr V{{m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3}};
    float* rhs0 = w + 3 * nl;
    float* mtbar = w + 4 * nl;
    float* mt = m_IncState[0].get();  // m̃^j (GN: one slot, src/CLAIRE.cpp:3074-3078)
    for (int j = 0; j < nt; ++j) {
        const bool first = (j == 0), last = (j == nt - 1);
        if (!first) m_Spectral->Grad(gw, mt);  // ∇m̃^j (zero for j = 0)
        {
            KernelScope k(m_Opt.prof, "rk2_inc_stage", 48.0 * nl);
            launch_rk2_inc_stage1(nl, ht, !first, G, V, grad_m(j), VT, mt, rhs0, mtbar, m_Opt.stream);
        }
        m_Spectral->Grad(gw, mtbar);  // ∇m̃̄
        KernelScope k(m_Opt.prof, "rk2_inc_stage", 64.0 * nl);
        launch_rk2_inc_stage2(nl, 0.5f * ht, !first, G, V, grad_m(j + 1), VT, mt, rhs0, last ? lt0_out : mt, last,
                              m_Opt.stream);
    }
}

void CLAIRE::Rk2AdjointSteps() {
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const float ht = m_Opt.GetTimeStepSize();
    float* w = m_Rk2Work.get();
    float* ww[3] = {w, w + nl, w + 2 * nl};
    const float* wc[3] = {ww[0], ww[1], ww[2]};
    const Vec3Out W{{ww[0], ww[1], ww[2]}};
    const Vec3Ptr V{{m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3}};
    float* rhs0 = w + 3 * nl;
    float* rhs1 = w + 4 * nl;
    float* lam = m_Lam.get();
    {
        KernelScope k(m_Opt.prof, "rk2_adj_stage", 28.0 * nl);
        launch_rk2_adj_flux(nl, lam, V, W, m_Opt.stream);
    }
    for (int j = 0; j < nt; ++j) {
        const float* lj = lam + j * nl;
        m_Spectral->Div(rhs0, wc);  // ∇·(λ^j v)
        {
            KernelScope k(m_Opt.prof, "rk2_adj_stage", 32.0 * nl);
            launch_rk2_adj_mid(nl, ht, lj, rhs0, V, W, m_Opt.stream);
        }
        m_Spectral->Div(rhs1, wc);  // ∇·(λ̄ v)
        const bool next = j + 1 < nt;
        KernelScope k(m_Opt.prof, "rk2_adj_stage", (next ? 40.0 : 16.0) * nl);
        launch_rk2_adj_end(nl, 0.5f * ht, lj, rhs0, rhs1, lam + (j + 1) * nl, next, V, W, m_Opt.stream);
    }
}

// λ^0 given in m_Lam slot 0; writes b (×hd) and, when lam_out is set, λ at t = 0.
void CLAIRE::IncAdjointCore(const float* lt0, VecField* bt, bool half) {
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    const GatherDims g = gather_dims(m_Opt.domain);
    float* b[3] = {bt->m_X1, bt->m_X2, bt->m_X3};
    (void)lt0;
    if (m_VelocityIsZero) {
        KernelScope k(m_Opt.prof, "body_force_static", 20.0 * nl);
        launch_body_force_static(nl, m_Opt.GetLebesgueMeasure(), m_Lam.get(), m_GradM.get(), b, m_Opt.stream);
        return;
    }
    if (m_Opt.pdesolver == 0) {
        // SolveAdjointEquationRK2 / SolveIncAdjointEquationGNRK2
        // (src/CLAIRE.cpp:2365-2548, 3504-3683): the same λ^j·∇m^{nt−j}
        // trapezoid body force, λ^{j+1} by RK2 steps of the flux form
        Rk2AdjointSteps();
        half = false;
    } else {
        const TrajectoryPlan tp = m_SL->Plan("adjoint");
        // λ^{j+1} = c·λ^j(X_a) (src/CLAIRE.cpp:2668-2673, 3955-3962), j = 0..nt−1
        for (int j = 0; j < nt; ++j) {
            // bytes: plan 12 + c 4 + gathered λ 4 + output 4 per point
            KernelScope k(m_Opt.prof, "adjoint_step", 24.0 * nl);
            launch_plan_adjoint_step(tp, g, m_AdjCoeff.get(), m_Lam.get() + j * nl, m_Lam.get() + (j + 1) * nl,
                                     m_Opt.stream);
        }
    }
    if (half) {
        // bytes: λ^0..λ^nt 4(nt+1) + ∇m (16-bit) 6(nt+1) + b 12
        KernelScope k(m_Opt.prof, "body_force", (10.0 * (nt + 1) + 12.0) * nl);
        launch_body_force_bfp(nl, nt, m_Opt.GetTimeStepSize(), m_Opt.GetLebesgueMeasure(), m_Lam.get(), GradMBfp(),
                              b, m_Opt.stream);
        return;
    }
    // bytes: λ^0..λ^nt 4(nt+1) + ∇m 12(nt+1) + b 12
    KernelScope k(m_Opt.prof, "body_force", (16.0 * (nt + 1) + 12.0) * nl);
    launch_body_force(nl, nt, m_Opt.GetTimeStepSize(), m_Opt.GetLebesgueMeasure(), m_Lam.get(), m_GradM.get(), b,
                      m_Opt.stream);
}

// Incremental state: writes λ̃^0 = −m̃^nt into out (or m̃^nt when !negate_last).
void CLAIRE::IncStateCore(const VecField* vt, float* out, bool half) {
    const int64_t nl = m_Opt.domain.nl;
    const int nt = m_Opt.nt;
    if (m_Opt.pdesolver == 0) {
        // SolveIncStateEquationRK2 (src/CLAIRE.cpp:2962-3135)
        EnsureStateGradients();
        Rk2IncState(vt, out);
        return;
    }
    EnsureGradientsAtStateTrajectory();
    const GatherDims g = gather_dims(m_Opt.domain);
    const TrajectoryPlan ts = m_SL->Plan("state");
    const float* v[3] = {vt->m_X1, vt->m_X2, vt->m_X3};
    float* first = (nt == 1) ? out : m_IncState[0].get();
    float* vx[3] = {m_VtX.get(), m_VtX.get() + nl, m_VtX.get() + 2 * nl};
    // ṽ(X_s) (src/CLAIRE.cpp:3221)
    m_SL->Gather3(v, vx, "state");
    if (half) {
        // bytes: ṽX 12 + ṽ 12 + (∇m^j)(X_s) 6nt + ∇m^{j+1} 6nt (16-bit) + s_j 4nt
        KernelScope k(m_Opt.prof, "inc_state_rhs", (24.0 + 16.0 * nt) * nl);
        launch_inc_state_rhs_bfp(nl, nt, m_Opt.GetTimeStepSize(), vx, v, GradMXBfp(), GradMBfp(), m_S.get(), first,
                                 m_Opt.stream);
    } else {
        // bytes: ṽX 12 + ṽ 12 + (∇m^j)(X_s) 12nt + ∇m^{j+1} 12nt + s_j 4nt
        KernelScope k(m_Opt.prof, "inc_state_rhs", (24.0 + 28.0 * nt) * nl);
        launch_inc_state_rhs(nl, nt, m_Opt.GetTimeStepSize(), vx, v, m_GradMX.get(), m_GradM.get(), m_S.get(), first,
                             m_Opt.stream);
    }
    for (int j = 1; j < nt; ++j) {
        const float* src = m_IncState[(j - 1) % 2].get();
        const bool last = (j == nt - 1);
        float* dst = last ? out : m_IncState[j % 2].get();
        // bytes: plan 12 + gathered m̃ 4 + s_j 4 + out 4
        KernelScope k(m_Opt.prof, "inc_state_step", 24.0 * nl);
        launch_plan_inc_state_step(ts, g, src, m_S.get() + j * nl, dst, last ? 1 : 0, m_Opt.stream);
    }
}

int CLAIRE::SolveIncStateEquation() {
    CLR_REQUIRE(m_HaveIncVelocity, kErrState, "incremental velocity not set");
    if (!m_StateValid) SolveStateEquation();
    // λ̃⁰ = −m̃^nt lands in m_Lam; m̃^nt = −λ̃⁰ is kept for the API.
    IncStateCore(m_IncVelocityField.get(), m_Lam.get(), false);
    const int64_t nl = m_Opt.domain.nl;
    float* mt = m_IncState[0].get();
    launch_copy(m_Lam.get(), mt, nl, m_Opt.stream);
    launch_scale(mt, -1.0f, nl, m_Opt.stream);
    m_IncStateSlot = 0;
    return 0;
}

int CLAIRE::SolveIncAdjointFrom(const float* mtilde) {
    RequireVelocity();
    if (!m_StateValid) SolveStateEquation();
    EnsureMatVecCaches();
    const int64_t nl = m_Opt.domain.nl;
    // λ̃⁰ = −m̃ (src/DistanceMeasureSL2.cpp:266-272)
    launch_copy(mtilde, m_Lam.get(), nl, m_Opt.stream);
    launch_scale(m_Lam.get(), -1.0f, nl, m_Opt.stream);
    IncAdjointCore(m_Lam.get(), m_IncBodyForce.get(), false);
    ProjectField(m_IncBodyForce.get());  // src/CLAIRE.cpp:3467-3468
    // λ̃ at t = 0: the last slot (constant in time when v = 0)
    launch_copy(m_VelocityIsZero ? m_Lam.get() : m_Lam.get() + m_Opt.nt * nl, m_IncAdjointT0.get(), nl, m_Opt.stream);
    return 0;
}

int CLAIRE::SolveIncAdjointEquation() { return SolveIncAdjointFrom(m_IncState[m_IncStateSlot].get()); }

int CLAIRE::HessianMatVec(float* Hvtilde, const float* vtilde, bool scale) {
    RequireVelocity();
    if (!m_StateValid) SolveStateEquation();
    const VecField* vt = m_IncVelocityField.get();
    std::unique_ptr<VecField> view;
    if (vtilde) {
        view.reset(new VecField(&m_Opt, const_cast<float*>(vtilde)));
        vt = view.get();
    } else {
        CLR_REQUIRE(m_HaveIncVelocity, kErrState, "incremental velocity not set");
    }
    CLR_REQUIRE(Hvtilde != nullptr, kErrInvalidArgument, "null output vector");
    CLR_REQUIRE(Hvtilde != vt->m_X1, kErrInvalidArgument, "HessianMatVec cannot run in place");
    // Newton-invariant caches, computed eagerly (never inside a capture)
    EnsureMatVecCaches();
    if (!m_Opt.graphs || (m_Opt.prof && m_Opt.prof->enabled())) {
        HessianMatVecBody(Hvtilde, vt, scale);
        return 0;
    }
    MatVecGraph* entry = nullptr;
    m_Graphs.erase(std::remove_if(m_Graphs.begin(), m_Graphs.end(),
                                  [&](MatVecGraph& g) {
                                      if (g.epoch == m_GraphEpoch) return false;
                                      if (g.exec) (void)hipGraphExecDestroy(g.exec);
                                      return true;
                                  }),
                   m_Graphs.end());
    for (auto& g : m_Graphs)
        if (g.in == vt->m_X1 && g.out == Hvtilde && g.scale == scale) entry = &g;
    if (!entry) {
        if (m_Graphs.size() >= 4) {
            if (m_Graphs.front().exec) (void)hipGraphExecDestroy(m_Graphs.front().exec);
            m_Graphs.erase(m_Graphs.begin());
        }
        m_Graphs.push_back(MatVecGraph{vt->m_X1, Hvtilde, scale, m_GraphEpoch, 0, nullptr});
        entry = &m_Graphs.back();
    }
    if (entry->uses++ < 2) {
        HessianMatVecBody(Hvtilde, vt, scale);
        return 0;
    }
    if (!entry->exec) CaptureMatVec(Hvtilde, vt, scale, &entry->exec);
    CLR_HIP_CHECK(hipGraphLaunch(entry->exec, m_Opt.stream));
    return 0;
}

void CLAIRE::CaptureMatVec(float* Hvtilde, const VecField* vt, bool scale, hipGraphExec_t* exec) {
    CaptureGraph([&] { HessianMatVecBody(Hvtilde, vt, scale); }, exec);
}

void CLAIRE::HessianMatVecInline(float* Hvtilde, const float* vtilde, bool scale) {
    CLR_REQUIRE(vtilde != nullptr && Hvtilde != nullptr && Hvtilde != vtilde, kErrInvalidArgument,
                "HessianMatVec needs distinct input and output vectors");
    RequireVelocity();
    if (!m_StateValid) SolveStateEquation();
    EnsureMatVecCaches();
    VecField vt(&m_Opt, const_cast<float*>(vtilde));
    HessianMatVecBody(Hvtilde, &vt, scale);
}

void CLAIRE::CaptureGraph(const std::function<void()>& body, hipGraphExec_t* exec) {
    if (!m_CaptureStream) CLR_HIP_CHECK(hipStreamCreateWithFlags(&m_CaptureStream, hipStreamNonBlocking));
    const hipStream_t saved = m_Opt.stream;
    Profiler* const saved_prof = m_Opt.prof;
    m_Opt.stream = m_CaptureStream;
    m_Opt.prof = nullptr;
    hipGraph_t graph = nullptr;
    CLR_HIP_CHECK(hipStreamBeginCapture(m_CaptureStream, hipStreamCaptureModeThreadLocal));
    try {
        body();
    } catch (...) {
        m_Opt.stream = saved;
        m_Opt.prof = saved_prof;
        (void)hipStreamEndCapture(m_CaptureStream, &graph);
        if (graph) (void)hipGraphDestroy(graph);
        throw;
    }
    m_Opt.stream = saved;
    m_Opt.prof = saved_prof;
    CLR_HIP_CHECK(hipStreamEndCapture(m_CaptureStream, &graph));
    const hipError_t e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) *exec = nullptr;
    CLR_HIP_CHECK(e);
}

int CLAIRE::SetMatVecType(int type) {
    CLR_REQUIRE(type >= 0 && type <= 2, kErrInvalidArgument,
                "matvec type must be 0 (default), 1 (preconditioned) or 2 (symmetric preconditioned)");
    if (type != m_Opt.matvectype) {
        m_Opt.matvectype = type;
        ++m_GraphEpoch;
    }
    return 0;
}

void CLAIRE::HessianMatVecBody(float* Hvtilde, const VecField* vt, bool scale) {
    VecField hv(&m_Opt, Hvtilde);
    const bool half = MixedKernels();
    const float hd = m_Opt.GetLebesgueMeasure();
    const int64_t n3 = 3 * m_Opt.domain.nl;
    if (m_Opt.matvectype != 0) {
        // PrecondHessMatVec (src/CLAIRE.cpp:1392-1451): hd·ṽ + (βA)⁻¹b̃[ṽ];
        // PrecondHessMatVecSym (src/CLAIRE.cpp:1463-1540):
        // hd·ṽ + (βA)^{-1/2}b̃[(βA)^{-1/2}ṽ]
        const bool sym = m_Opt.matvectype == 2;
        const VecField* in = vt;
        if (sym) {
            m_Regularization->ApplyInverse(m_WorkVecField1.get(), vt, true);
            in = m_WorkVecField1.get();
        }
        IncStateCore(in, m_Lam.get(), half);
        IncAdjointCore(m_Lam.get(), m_IncBodyForce.get(), half);
        ProjectField(m_IncBodyForce.get());
        m_Regularization->ApplyInverse(&hv, m_IncBodyForce.get(), sym);
        launch_waxpy(hd, vt->m_X1, Hvtilde, Hvtilde, n3, m_Opt.stream);
        if (!scale) launch_scale(Hvtilde, 1.0f / hd, n3, m_Opt.stream);  // src/CLAIRE.cpp:1283-1286
        return;
    }
    // \tilde{m} (src/CLAIRE.cpp:1345) with λ̃⁰ = −m̃^nt fused into the last step
    IncStateCore(vt, m_Lam.get(), half);
    // \tilde{λ} and \tilde{b}·hd (src/CLAIRE.cpp:1348, 3307-3491)
    IncAdjointCore(m_Lam.get(), m_IncBodyForce.get(), half);
    ProjectField(m_IncBodyForce.get());  // K[b̃], src/CLAIRE.cpp:3467-3468
    // Hṽ = β·A[ṽ] + b̃ (src/CLAIRE.cpp:1356-1361); 1/hd when !scale (src/CLAIRE.cpp:1283-1286)
    const float a = scale ? 1.0f : 1.0f / m_Opt.GetLebesgueMeasure();
    m_Regularization->ApplyFused(&hv, vt, a, m_IncBodyForce.get(), a);
}

int CLAIRE::SolveAdjointEquation() {
    RequireVelocity();
    CLR_REQUIRE(m_HaveReference, kErrState, "reference image not set");
    if (!m_StateValid) SolveStateEquation();
    EnsureStateGradients();
    const int64_t nl = m_Opt.domain.nl;
    // λ¹ = mR − m^nt (src/DistanceMeasureSL2.cpp:198-205)
    launch_sub(m_ReferenceImage.get(), m_State.get() + m_Opt.nt * nl, m_Lam.get(), nl, m_Opt.stream);
    IncAdjointCore(m_Lam.get(), m_BodyForce.get(), false);
    ProjectField(m_BodyForce.get());  // src/CLAIRE.cpp:2337-2338
    launch_copy(m_VelocityIsZero ? m_Lam.get() : m_Lam.get() + m_Opt.nt * nl, m_Adjoint.get(), nl, m_Opt.stream);
    return 0;
}

void CLAIRE::ProjectField(VecField* b) {
    if (m_Opt.regmodel == kModelCompressible) return;
    SpecParams sp{};
    sp.op = m_Opt.regmodel == kModelStokes ? kOpProjStokes : kOpProjRelaxed;
    sp.beta0 = m_Opt.beta[0];
    sp.beta1 = m_Opt.beta[1];
    sp.beta2 = m_Opt.beta[2];
    sp.hd = m_Opt.GetLebesgueMeasure();
    sp.scale = m_Opt.ComputeFFTScale();
    float* c[3] = {b->m_X1, b->m_X2, b->m_X3};
    m_Spectral->Project(sp, c);
}

int CLAIRE::ApplyProjection(float* v) {
    CLR_REQUIRE(v != nullptr, kErrInvalidArgument, "null vector field");
    VecField view(&m_Opt, v);
    ProjectField(&view);
    return 0;
}

int CLAIRE::SetRegModel(int model) {
    CLR_REQUIRE(model == kModelCompressible || model == kModelStokes || model == kModelRelaxedStokes,
                kErrInvalidArgument, "registration model must be 0 (compressible), 1 (Stokes) or 2 (relaxed Stokes)");
    if (model != m_Opt.regmodel) {
        m_Opt.regmodel = model;
        m_TrajValid = m_StateValid = m_GradMValid = m_GradMXValid = m_HalfValid = false;
        ++m_GraphEpoch;
    }
    return 0;
}

int CLAIRE::EvaluateObjective(double* J, double* D, double* R, const float* v) {
    if (v) SetControlVariable(v);
    CLR_REQUIRE(m_HaveReference, kErrState, "reference image not set");
    SolveStateEquation();
    const int64_t nl = m_Opt.domain.nl;
    const double hd = m_Opt.GetLebesgueMeasure();
    // ½·hd·Σ(mR − m^nt)²/nc (src/DistanceMeasureSL2.cpp:124-147)
    *D = 0.5 * hd * m_Red->SqDist(m_ReferenceImage.get(), m_State.get() + m_Opt.nt * nl, nl);
    *R = 0.0;
    if (!m_VelocityIsZero) {
        m_Regularization->EvaluateFunctional(R, m_VelocityField.get());
        if (m_Opt.regmodel == kModelRelaxedStokes) *R += DivergencePenalty();
    }
    *J = *D + *R;
    return 0;
}

int CLAIRE::EvaluateGradient(float* g, const float* v) {
    if (v) SetControlVariable(v);
    if (!m_StateValid) SolveStateEquation();
    SolveAdjointEquation();
    VecField gv(&m_Opt, g);
    if (m_VelocityIsZero) return m_BodyForce->GetComponents(g);  // src/CLAIRE.cpp:960-965
    // g = β·A[v] + b (src/CLAIRE.cpp:1021-1026)
    m_Regularization->ApplyFused(&gv, m_VelocityField.get(), 1.0f, m_BodyForce.get(), 1.0f);
    return 0;
}

// CLAIREDivReg::EvaluteRegularizationDIV (src/CLAIREDivReg.cpp:190-260):
// ½·hd·β2·(‖∂3 d‖² + ‖∂2 d‖² + ‖∂1 d‖² + ‖d‖²), d = ∇·v (the reference
// stores the gradient with its components reversed and sums in that order).
double CLAIRE::DivergencePenalty() {
    const int64_t nl = m_Opt.domain.nl;
    const float* vc[3] = {m_VelocityField->m_X1, m_VelocityField->m_X2, m_VelocityField->m_X3};
    float* d = m_Scalar.get();
    m_Spectral->Div(d, vc);
    float* gd[3] = {m_WorkVecField1->m_X3, m_WorkVecField1->m_X2, m_WorkVecField1->m_X1};
    m_Spectral->Grad(gd, d);
    m_Red->DotTo(m_WorkVecField1->m_X1, m_WorkVecField1->m_X1, nl, 0);
    m_Red->DotTo(m_WorkVecField1->m_X2, m_WorkVecField1->m_X2, nl, 1);
    m_Red->DotTo(m_WorkVecField1->m_X3, m_WorkVecField1->m_X3, nl, 2);
    m_Red->DotTo(d, d, nl, 3);
    double r[4];
    m_Red->Fetch(r, 4);
    const double sum = ((r[0] + r[1]) + r[2]) + r[3];
    return 0.5 * static_cast<double>(m_Opt.GetLebesgueMeasure()) * static_cast<double>(m_Opt.beta[2]) * sum;
}

int CLAIRE::ApplyInvRegularizationOperator(float* Ainvx, const float* x, bool applysqrt) {
    // CLAIREBase::ApplyInvRegularizationOperator (src/CLAIREBase.cpp:987-1021)
    VecField out(&m_Opt, Ainvx), in(&m_Opt, const_cast<float*>(x));
    return m_Regularization->ApplyInverse(&out, &in, applysqrt);
}

}  // namespace reg
