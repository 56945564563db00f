// solver.hpp -- Gauss–Newton–Krylov driver on the device.
//
// Restates the reference's default optimizer, PETSc TAO "nls" with a
// matrix-free Hessian and KSPCG (src/Optimizer.cpp:235-527), the
// Eisenstat–Walker forcing of PreKrylovSolve (src/KrylovInterface.cpp:79-175),
// the inverse-regularization preconditioner (src/Preconditioner.cpp:521-573)
// and the gradient stopping test CheckConvergenceGrad
// (src/TaoInterface.cpp:390-512).  All vectors stay on the device; inner
// products are FP64 reductions; the host only sees scalars.
#pragma once

#include <cmath>

#include "claire.hpp"

namespace reg {

struct GNOptions {
    int maxit = 50;              // OptPara.maxiter
    double gttol = 5e-2;         // OptPara.tol[2]: ‖g‖ <= gttol·‖g₀‖
    double gatol = 1e-6;         // OptPara.tol[0]: ‖g‖ <= gatol
    int krylov_maxit = 50;       // KrylovMethod.maxiter (KSP runs maxit − 1)
    double krylov_rtol = 1e-12;  // KrylovMethod.tol[0], used without forcing
    int forcing = 2;             // 0 none, 1 quadratic, 2 superlinear (SLFS)
    int precond = 1;             // 0 none, 1 inverse regularization (INVREG)
    int linesearch = 1;          // 0 unit step, 1 Armijo
    int monitor = 0;             // 1: print the iteration table (OptimizationMonitor)
    double ls_sigma = 1e-4;      // Armijo sufficient decrease (PETSc armijo sigma)
    double ls_stepinit = 1.0;    // first trial step (PETSc -tao_ls_stepinit)
    // CG iterations queued between two host checks of the device-resident
    // Krylov scalars (at most; the run is shortened to the predicted
    // remainder of the solve).  1 checks after every iteration.
    int krylov_block = 8;
};

struct GNReport {
    int iterations = 0, reason = 0, matvecs = 0, pcg_iterations = 0, objective_evals = 0, gradient_evals = 0;
    double J0 = 0, D0 = 0, gnorm0 = 0, J = 0, D = 0, R = 0, gnorm = 0;
};

// Stop reasons (TAO_CONVERGED_* > 0, TAO_DIVERGED_* < 0).
enum GNReason { kGNContinue = 0, kGNGttol = 1, kGNGatol = 2, kGNSteptol = 3, kGNMaxits = -1, kGNNan = -2,
                kGNLineSearch = -3 };

// The gradient stopping test CheckConvergenceGrad (src/TaoInterface.cpp:
// 390-512) with the default miniter = 0 (src/RegOpt.cpp:1158): a non-finite
// J or ‖g‖ stops first (TAO_DIVERGED_NAN), then a line-search step below
// 2^-10 (TAO_CONVERGED_STEPTOL); then ‖g‖ < gttol·‖g₀‖ (GTTOL), ‖g‖ < gatol
// (GATOL) and iter > maxit − 1 (MAXITS, TAO's limit being maxit − 1,
// src/Optimizer.cpp:431) each set the reason in that order, so the last that
// holds wins.  g0 is ‖g₀‖, or 1 when ‖g₀‖ = 0.
inline int convergence_reason(double J, double gnorm, double step, double g0, int iter, double gttol, double gatol,
                              int maxit) {
    if (!std::isfinite(J) || !std::isfinite(gnorm)) return kGNNan;
    if (step < 1.0 / 1024.0) return kGNSteptol;
    int reason = kGNContinue;
    if (gnorm < gttol * g0) reason = kGNGttol;
    if (gnorm < gatol) reason = kGNGatol;
    if (iter > maxit - 1) reason = kGNMaxits;
    return reason;
}

class Optimizer {
  public:
    explicit Optimizer(CLAIRE* problem);
    // v: initial guess on entry (device, flat [v1|v2|v3]), solution on exit.
    // history (optional): 8 doubles per iterate: iter, J, D, R, ‖g‖,
    // ‖g‖/‖g₀‖, PCG iterations and step length of the step taken from it.
    int Run(float* v, const GNOptions& opt, GNReport& rep, double* history, int history_len);

    ~Optimizer();

  private:
    double Dot(const float* a, const float* b);
    int Pcg(const float* g, double rtol, const GNOptions& opt, GNReport& rep);
    // One CG iteration (matvec, step length, updates, preconditioner, the
    // convergence test and the next direction) on device-resident scalars.
    void PcgIteration(bool inline_matvec, const GNOptions& opt);
    void Precond(const float* in, float* out, const GNOptions& opt);
    void DropIterationGraph();
    CLAIRE* m_Prob;
    IntType m_N;
    DeviceArray<claire_amd::PcgState> m_State;
    // A captured CG iteration, replayed while the problem's graph epoch and
    // the operator options stay the same.
    hipGraphExec_t m_IterGraph = nullptr;
    uint64_t m_IterEpoch = 0;
    int m_IterPrecond = -1;
    // the problem's stream at the time of the call (a batch driver may move
    // a problem to another stream between solves)
    hipStream_t Stream() const { return m_Prob->Opt().stream; }
    DeviceArray<float> m_G, m_D, m_R, m_Z, m_P, m_W, m_Vt, m_Zero;
};

}  // namespace reg
