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
};

struct GNReport {
    int iterations = 0, reason = 0, matvecs = 0, pcg_iterations = 0, objective_evals = 0, gradient_evals = 0;
    double J0 = 0, D0 = 0, gnorm0 = 0, J = 0, D = 0, R = 0, gnorm = 0;
};

// Stop reasons (TAO_CONVERGED_* > 0, TAO_DIVERGED_* < 0).
enum GNReason { kGNContinue = 0, kGNGttol = 1, kGNGatol = 2, kGNSteptol = 3, kGNMaxits = -1, kGNNan = -2,
                kGNLineSearch = -3 };

class Optimizer {
  public:
    explicit Optimizer(CLAIRE* problem);
    // v: initial guess on entry (device, flat [v1|v2|v3]), solution on exit.
    // history (optional): 8 doubles per iterate: iter, J, D, R, ‖g‖,
    // ‖g‖/‖g₀‖, PCG iterations and step length of the step taken from it.
    int Run(float* v, const GNOptions& opt, GNReport& rep, double* history, int history_len);

  private:
    double Dot(const float* a, const float* b);
    int Pcg(const float* g, double rtol, const GNOptions& opt, GNReport& rep);
    CLAIRE* m_Prob;
    IntType m_N;
    // the problem's stream at the time of the call (a batch driver may move
    // a problem to another stream between solves)
    hipStream_t Stream() const { return m_Prob->Opt().stream; }
    DeviceArray<float> m_G, m_D, m_R, m_Z, m_P, m_W, m_Vt, m_Zero;
};

}  // namespace reg
