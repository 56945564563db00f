// solver.cpp -- Gauss–Newton–Krylov driver on the device (see solver.hpp).
#include "solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <string>

namespace reg {

using namespace claire_amd;

Optimizer::Optimizer(CLAIRE* problem)
    : m_Prob(problem), m_N(3 * problem->Opt().domain.nl) {
    const size_t n = static_cast<size_t>(m_N);
    m_G.resize(n);
    m_D.resize(n);
    m_R.resize(n);
    m_Z.resize(n);
    m_P.resize(n);
    m_W.resize(n);
    m_Vt.resize(n);
    m_State.resize(1);
}

Optimizer::~Optimizer() { DropIterationGraph(); }

void Optimizer::DropIterationGraph() {
    if (m_IterGraph) (void)hipGraphExecDestroy(m_IterGraph);
    m_IterGraph = nullptr;
}

double Optimizer::Dot(const float* a, const float* b) { return m_Prob->Red()->Dot(a, b, m_N); }

namespace {

// The iteration table of the reference: the header printed by
// CLAIREInterface::Run (src/CLAIREInterface.cpp:722-726, separators of
// m_LineLength = 101 dashes, src/RegOpt.cpp:1276) and one line per
// iteration from OptimizationMonitor (src/TaoInterface.cpp:584-588):
// iteration, J/J0, D/D0, ‖g‖/‖g0‖, ‖g‖ and the step length, each ratio's
// denominator replaced by 1 when it is not positive.
void monitor_header() {
    const std::string line(101, '-');
    std::printf("%s\n %s  %-20s %-20s %-20s %-20s %-20s\n%s\n", line.c_str(), "iter", "objective (rel)",
                "mismatch (rel)", "||gradient||_2,rel", "||gradient||_2", "step", line.c_str());
    std::fflush(stdout);
}

void monitor_line(int iter, double J, double J0, double D, double D0, double g, double g0, double step) {
    J0 = J0 > 0.0 ? J0 : 1.0;
    D0 = D0 > 0.0 ? D0 : 1.0;
    g0 = g0 > 0.0 ? g0 : 1.0;
    char msg[256];
    std::snprintf(msg, sizeof(msg), "  %03d  %-20.12E %-20.12E %-20.12E %-20.12E %.6f", iter, J / J0, D / D0, g / g0,
                  g, step);
    std::printf("%-80s\n", msg);
    std::fflush(stdout);
}

}  // namespace

// KSPCG on H d = −g, zero initial guess, unpreconditioned residual
// norm: converged when ‖r‖ <= max(rtol·‖b‖, 1e-16), diverged when
// ‖r‖ >= 1e6·‖b‖, at most krylov_maxit − 1 iterations
// (KSPSetTolerances, src/Optimizer.cpp:298-307); stops on a non-positive
// curvature pᵀHp (indefinite operator).  Returns the iteration count; the
// solution is in m_D.
void Optimizer::Precond(const float* in, float* out, const GNOptions& opt) {
    // the symmetrically preconditioned operator is solved without a further
    // preconditioner (the reference pairs it with its two-level scheme)
    if (opt.precond && m_Prob->Opt().matvectype != 2)
        m_Prob->ApplyInvRegularizationOperator(out, in, false);
    else
        launch_copy(in, out, m_N, Stream());
}

// The body of one KSPCG iteration.  Step lengths live in m_State and are
// formed on the device from FP64 reduction slots with the host's arithmetic,
// so a run of iterations equals the host-driven loop bit for bit; once the
// solve has stopped, the updates are skipped (the matvec and preconditioner
// of a queued iteration still run, on vectors that no longer change the
// result).  z = M⁻¹r is applied before the convergence test so that ‖r‖²
// and rᵀz are reduced together.
void Optimizer::PcgIteration(bool inline_matvec, const GNOptions& opt) {
    float* d = m_D.get();
    float* r = m_R.get();
    float* z = m_Z.get();
    float* p = m_P.get();
    float* w = m_W.get();
    PcgState* s = m_State.get();
    Reducer* red = m_Prob->Red();
    if (inline_matvec)
        m_Prob->HessianMatVecInline(w, p, true);
    else
        m_Prob->HessianMatVec(w, p, true);
    // pᵀw and α in one reduction; both updates in one pass; rᵀr and rᵀz
    // in one pass with the stopping test and β (the same sums and scalars
    // as separate launches, bit for bit)
    red->DotAlpha(p, w, m_N, s);
    launch_axpy2_pcg(s, p, d, w, r, m_N, Stream());
    Precond(r, z, opt);
    red->DotsBeta(r, z, m_N, s);
    launch_xpay_pcg(s, p, z, m_N, Stream());
}

// KSPCG on H d = −g, zero initial guess, unpreconditioned residual
// norm: converged when ‖r‖ <= max(rtol·‖b‖, 1e-16), diverged when
// ‖r‖ >= 1e6·‖b‖, at most krylov_maxit − 1 iterations
// (KSPSetTolerances, src/Optimizer.cpp:298-307); stops on a non-positive
// curvature pᵀHp (indefinite operator).  Returns the iteration count; the
// solution is in m_D.
//
// Iterations are queued in runs between host checks of the device state.
// A run's length is the remainder predicted from the residual's decay over
// the previous run (at most krylov_block), so few queued iterations outlive
// the stopping test.  With matvec graphs on, one iteration is captured per
// graph epoch and the runs replay it.
int Optimizer::Pcg(const float* g, double rtol, const GNOptions& opt, GNReport& rep) {
    float* d = m_D.get();
    float* r = m_R.get();
    float* z = m_Z.get();
    float* p = m_P.get();
    const int matvec = m_Prob->Opt().matvectype;
    CLR_REQUIRE(matvec != 1, kErrInvalidArgument,
                "the preconditioned Hessian (PRECONDMATVEC) is not symmetric; this driver's Krylov method is CG");
    CLR_REQUIRE(opt.krylov_block >= 1, kErrInvalidArgument, "krylov_block must be >= 1");
    const bool sym = matvec == 2;
    launch_fill(d, 0.0f, m_N, Stream());
    launch_copy(g, r, m_N, Stream());
    launch_scale(r, -1.0f, m_N, Stream());
    // PreKrylovSolve (src/CLAIREBase.cpp:941-953): the right-hand side of the
    // symmetrically preconditioned system is (βA)^{-1/2}(−g)
    if (sym) m_Prob->ApplyInvRegularizationOperator(r, r, true);
    const double bnorm = std::sqrt(Dot(r, r));
    const double tol = std::max(rtol * bnorm, 1e-16);
    const int kmax = std::max(0, opt.krylov_maxit - 1);
    // PostKrylovSolve (src/CLAIREBase.cpp:959-969): d ← (βA)^{-1/2}d
    auto post = [&]() {
        if (sym) m_Prob->ApplyInvRegularizationOperator(d, d, true);
    };
    if (!(bnorm > tol) || kmax == 0) {
        post();
        return 0;
    }
    Precond(r, z, opt);
    m_Prob->Red()->DotTo(r, z, m_N, 2);
    launch_pcg_init(m_State.get(), m_Prob->Red()->Slots(), tol, bnorm, kmax, Stream());
    launch_copy(z, p, m_N, Stream());

    const bool graphs = m_Prob->MatVecGraphs() && !(m_Prob->Opt().prof && m_Prob->Opt().prof->enabled());
    const int precond_key = opt.precond ? 1 : 0;
    if (m_IterGraph && (m_IterEpoch != m_Prob->GraphEpoch() || m_IterPrecond != precond_key)) DropIterationGraph();
    int its = 0, run = 1, runs = 0;
    double rprev = bnorm;
    PcgState hs{};
    for (;;) {
        run = std::min(run, kmax - its);
        if (graphs && runs > 0 && run >= 3 && !m_IterGraph) {
            // the first run was eager (lazy allocations, gather tuning); the
            // caches are current, so the matvec is captured inline.  A solve
            // predicted to end within two more iterations pays no capture.
            m_Prob->PrepareHessian();
            m_Prob->CaptureGraph([&] { PcgIteration(true, opt); }, &m_IterGraph);
            m_IterEpoch = m_Prob->GraphEpoch();
            m_IterPrecond = precond_key;
        }
        for (int q = 0; q < run; ++q) {
            if (graphs && m_IterGraph)
                CLR_HIP_CHECK(hipGraphLaunch(m_IterGraph, Stream()));
            else
                PcgIteration(false, opt);
        }
        ++runs;
        CLR_HIP_CHECK(hipMemcpyAsync(&hs, m_State.get(), sizeof(PcgState), hipMemcpyDeviceToHost, Stream()));
        CLR_HIP_CHECK(hipStreamSynchronize(Stream()));
        its = hs.k;
        if (hs.done || its >= kmax) break;
        // iterations still needed if the residual keeps its decay rate
        int next = opt.krylov_block;
        if (hs.rnorm > 0.0 && hs.rnorm < rprev && run > 0) {
            const double rate = std::pow(hs.rnorm / rprev, 1.0 / run);
            const double need = std::log(tol / hs.rnorm) / std::log(rate);
            if (std::isfinite(need)) next = static_cast<int>(std::ceil(std::max(1.0, need)));
        }
        run = std::max(1, std::min(next, opt.krylov_block));
        rprev = hs.rnorm;
    }
    rep.matvecs += its;
    post();
    return its;
}

int Optimizer::Run(float* v, const GNOptions& opt, GNReport& rep, double* hist, int hist_len) {
    CLR_REQUIRE(v != nullptr, kErrInvalidArgument, "null velocity");
    CLR_REQUIRE(opt.maxit >= 1 && opt.krylov_maxit >= 1, kErrInvalidArgument, "iteration limits must be >= 1");
    rep = GNReport{};
    float* g = m_G.get();
    float* d = m_D.get();
    float* vt = m_Vt.get();

    // ‖g₀‖ at v = 0 (InitializeOptimization, src/CLAIRE.cpp:320-335)
    m_Zero.resize(static_cast<size_t>(m_N));
    launch_fill(m_Zero.get(), 0.0f, m_N, Stream());
    m_Prob->EvaluateGradient(g, m_Zero.get());
    rep.gradient_evals++;
    rep.gnorm0 = std::sqrt(Dot(g, g));
    const double g0 = rep.gnorm0 > 0.0 ? rep.gnorm0 : 1.0;

    double J, D, R;
    m_Prob->EvaluateObjective(&J, &D, &R, v);
    rep.objective_evals++;
    m_Prob->EvaluateGradient(g, nullptr);  // same velocity: reuses the state
    rep.gradient_evals++;
    double gnorm = std::sqrt(Dot(g, g));
    rep.J0 = J;
    rep.D0 = D;
    double kg0 = -1.0;  // ‖g‖ at the first Krylov solve (KrylovMethod.g0norm)
    double step = 1.0;
    int iter = 0;
    if (opt.monitor) monitor_header();
    for (;;) {
        double* h = (hist && iter < hist_len) ? hist + 8 * iter : nullptr;
        if (h) {
            h[0] = iter; h[1] = J; h[2] = D; h[3] = R; h[4] = gnorm; h[5] = gnorm / g0; h[6] = 0; h[7] = 0;
        }
        // TaoMonitor runs before the convergence test, the step length being
        // the last line search's (1 before the first)
        if (opt.monitor) monitor_line(iter, J, rep.J0, D, rep.D0, gnorm, rep.gnorm0, step);
        // CheckConvergenceGrad (src/TaoInterface.cpp:421-480): a non-finite J
        // or ‖g‖ first, then the step length; of the three tolerance tests
        // every one that holds sets the reason in turn, so the last wins
        // (GATOL over GTTOL, MAXITS over both).  TAO's iteration limit is
        // maxit − 1 (src/Optimizer.cpp:431).
        rep.reason = convergence_reason(J, gnorm, step, g0, iter, opt.gttol, opt.gatol, opt.maxit);
        if (rep.reason != kGNContinue) break;

        // Eisenstat–Walker forcing (src/KrylovInterface.cpp:120-160).  The
        // ratio is ‖g‖/‖g_first‖ for every operator type: with PRECONDMATVEC
        // and PRECONDMATVECSYM the reference takes ‖b‖ before PreKrylovSolve
        // transforms b (src/KrylovInterface.cpp:104-108, :114), and ‖b‖ = ‖g‖.
        if (kg0 < 0.0) kg0 = gnorm;
        double rtol = opt.krylov_rtol;
        if (opt.forcing != 0 && kg0 > 0.0) {
            const double q = gnorm / kg0;
            rtol = std::max(1e-16, std::min(0.5, opt.forcing == 1 ? q : std::sqrt(q)));
        }
        // the current velocity is v (set by the last objective/gradient call)
        const int its = Pcg(g, rtol, opt, rep);
        rep.pcg_iterations += its;
        if (h) h[6] = its;

        // descent check, Armijo line search, steepest-descent fallback
        double gd = Dot(g, d);
        bool newton = true;
        if (!(gd < 0.0)) {
            launch_copy(g, d, m_N, Stream());
            launch_scale(d, -1.0f, m_N, Stream());
            gd = -gnorm * gnorm;
            newton = false;
        }
        double alpha = 1.0, Jt = 0.0, Dt = 0.0, Rt = 0.0;
        bool ok = false;
        for (;;) {
            alpha = opt.linesearch ? opt.ls_stepinit : 1.0;
            const int trials = opt.linesearch ? 30 : 1;
            for (int k = 0; k < trials; ++k) {
                launch_waxpy(static_cast<float>(alpha), d, v, vt, m_N, Stream());
                m_Prob->EvaluateObjective(&Jt, &Dt, &Rt, vt);
                rep.objective_evals++;
                if (!opt.linesearch || (std::isfinite(Jt) && Jt <= J + opt.ls_sigma * alpha * gd)) {
                    ok = true;
                    break;
                }
                alpha *= 0.5;
            }
            if (ok || !newton) break;
            launch_copy(g, d, m_N, Stream());
            launch_scale(d, -1.0f, m_N, Stream());
            gd = -gnorm * gnorm;
            newton = false;
        }
        if (!ok) {
            rep.reason = kGNLineSearch;
            m_Prob->SetControlVariable(v);
            break;
        }
        launch_copy(vt, v, m_N, Stream());
        J = Jt; D = Dt; R = Rt;
        m_Prob->EvaluateGradient(g, nullptr);  // at the accepted trial velocity
        rep.gradient_evals++;
        gnorm = std::sqrt(Dot(g, g));
        step = alpha;
        if (h) h[7] = alpha;
        ++iter;
    }
    rep.iterations = iter;
    rep.J = J; rep.D = D; rep.R = R; rep.gnorm = gnorm;
    CLR_HIP_CHECK(hipStreamSynchronize(Stream()));
    return 0;
}

}  // namespace reg
