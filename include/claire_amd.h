/*
 * claire_amd.h -- C ABI of the MI355X-native CLAIRE registration hot path.
 *
 * Drop-in boundary for the reference's Gauss–Newton–Krylov path
 * (GN, semi-Lagrangian transport, nc = 1, FP32 data).  All field pointers
 * are DEVICE pointers to FP32 data; vector fields are flat [v1|v2|v3] with
 * each component stored x3-fastest (the layout of the PETSc vectors TAO/KSP
 * hand to CLAIRE, VecField::SetComponents/GetComponents,
 * src/VecField.cpp:396-476).  Field pointers must be 16-byte aligned
 * (hipMalloc returns 256-byte aligned memory, and a component of a vector
 * field starts 4·N1·N2·N3 bytes after the previous one): the streaming
 * kernels and the FFT's real passes move 16 bytes per lane.  Work is ordered on the stream given to
 * clr_create (0 = the null stream).  Every function returns 0 on success or
 * a nonzero status (CLR_ERR_*) with a message from clr_last_error(); this
 * maps one-to-one onto PetscErrorCode at the binding.  A context is not
 * re-entrant (like the reference's problem object): calls on one context
 * must not overlap.
 */
#ifndef CLAIRE_AMD_H_
#define CLAIRE_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CLR_OK 0
#define CLR_ERR_INVALID_ARGUMENT 1
#define CLR_ERR_UNSUPPORTED_SIZE 2
#define CLR_ERR_HIP 3
#define CLR_ERR_STATE 4
#define CLR_ERR_OUT_OF_MEMORY 5

/* Regularization norms: RegularizationH1 / H2 / H1SN / H2SN
 * (src/RegularizationH1.cpp, src/RegularizationH2.cpp,
 *  src/RegularizationH1SN.cpp, src/RegularizationH2SN.cpp). */
#define CLR_REG_H1 0
#define CLR_REG_H2 1
#define CLR_REG_H1SN 2
#define CLR_REG_H2SN 3

/* Trajectory selector: "state" (σ = +1) / "adjoint" (σ = −1),
 * SemiLagrangian::ComputeTrajectory flag (src/SemiLagrangian.cpp:217-223). */
#define CLR_STATE 0
#define CLR_ADJOINT 1

typedef struct clr_ctx clr_ctx;

/* Last error message of the calling thread. */
const char* clr_last_error(void);

/* --- device memory helpers (plumbing for bindings without a GPU array type;
 * no reference counterpart: the reference's Vecs live in host memory) */
int clr_device_malloc(void** ptr, size_t bytes);
int clr_device_free(void* ptr);
int clr_memcpy_htod(void* dst, const void* src, size_t bytes);
int clr_memcpy_dtoh(void* dst, const void* src, size_t bytes);
int clr_memcpy_dtod(void* dst, const void* src, size_t bytes);
int clr_device_synchronize(void);

/* Create a problem object for an nx[0]×nx[1]×nx[2] periodic grid on
 * [0,2π)^3 with nt time steps, the given regularization norm and weights
 * beta[0..2] (RegOpt m_RegNorm.beta), on `stream` (a hipStream_t or NULL).
 * Replaces `new CLAIRE(opt)` + RegOpt::DoSetup/InitializeFFT
 * (src/RegOpt.cpp:906-1030; hx = 2π/nx src/RegOpt.cpp:935).
 * Supported grid sizes: powers of two from 8 to 512 per axis, and 48, 96,
 * 150, 192. */
int clr_create(clr_ctx** ctx, const int nx[3], int nt, int regnorm, const double beta[3], void* stream);
int clr_destroy(clr_ctx* ctx);
int clr_synchronize(clr_ctx* ctx);
/* Determinant of the deformation gradient of the current velocity by SL
 * transport: DeformationFields::ComputeDetDefGrad + ComputeDetDefGradSL
 * (src/DeformationFields.cpp:315-388, 636-740); inverse != 0 gives
 * det(∇y⁻¹) (RegFlags.invdefgrad).  jac (device, nl) may be NULL; stats
 * (host, 3 doubles) receive min, max and mean (Monitor.detdgradmin/max/mean). */
int clr_det_def_grad(clr_ctx* ctx, int inverse, float* jac, double* stats);
/* CLAIREBase::CheckBounds (src/CLAIREBase.cpp:1386-1446): sets the velocity
 * v (device, 3·nl), computes det(∇y^{±1}) and reports whether
 * min(J) <= bound or 1/max(J) <= bound (default bound 0.25, src/RegOpt.cpp:1254). */
int clr_check_bounds(clr_ctx* ctx, const float* v, int inverse, double bound, int* reached, double* stats);

/* NIfTI-1 single-file images (.nii, .nii.gz), replacing ReadWriteReg::ReadNII
 * and WriteNII (src/ReadWriteReg.cpp:794-1475).  Host memory.  The image's
 * (nx, ny, nz) is the grid's (n3, n2, n1) (src/ReadWriteReg.cpp:817-819), so
 * voxels come in the grid's flat order (x3 fastest), converted to float
 * without intensity scaling (src/ReadWriteReg.cpp:1083-1110).
 * clr_nifti_read: nx receives (n1, n2, n3); data (capacity floats) may be
 * NULL to query the size; header (348 bytes) may be NULL.
 * clr_nifti_write: FLOAT32 volume; header (348 bytes of a read image, as
 * nifti_copy_nim_info does, src/ReadWriteReg.cpp:1157-1170) and spacing of
 * (x1, x2, x3) may be NULL. */
int clr_nifti_read(const char* path, int nx[3], float* data, int64_t capacity, unsigned char* header);
int clr_nifti_write(const char* path, const int nx[3], const float* data, const float* spacing,
                    const unsigned char* header);
/* PETSc binary vectors (.bin), replacing ReadWriteReg::ReadBIN and WriteBIN
 * (src/ReadWriteReg.cpp:1476-1537, VecLoad/VecView on a binary viewer).
 * Host memory.  Big-endian: class id 1211214 and length as PetscInt, then the
 * float values in the grid's flat order.  Written with 8-byte header words
 * (the reference's --with-64-bit-indices PETSc, deps/build_libs.sh:176);
 * read with 8- or 4-byte header words.  clr_bin_read: data (capacity floats)
 * may be NULL to query the length n. */
int clr_bin_read(const char* path, int64_t* n, float* data, int64_t capacity);
int clr_bin_write(const char* path, int64_t n, const float* data);

/* Arithmetic form of the Hessian matvec.  CLR_PRECISION_FP32 (default) is
 * the reference's single-precision path.  CLR_PRECISION_MIXED stores the
 * Newton-invariant gradient fields ∇m^j and (∇m^j)(X_s) that the streaming
 * kernels read as 16-bit block floating point (int16 mantissas, one
 * power-of-two scale per 256 values; FP32 arithmetic, FP64 reductions
 * unchanged);
 * it has no counterpart in the reference, whose precision is fixed by the
 * PETSc build (ScalarType = PetscReal, include/CLAIREUtils.hpp:58-67). */
#define CLR_PRECISION_FP32 0
#define CLR_PRECISION_MIXED 1
int clr_set_precision(clr_ctx* ctx, int precision);
/* Order of the characteristic scheme of every later trajectory: 2 (default)
 * or 4.  Replaces RegOpt m_PDESolver.rkorder, set by "-rkorder <int>"
 * (src/RegOpt.cpp:674-676, checked at src/RegOpt.cpp:1690-1693) and read by
 * SemiLagrangian::ComputeTrajectory (src/SemiLagrangian.cpp:172-178).
 * Changing it drops the cached trajectories and everything computed on them. */
int clr_set_rk_order(clr_ctx* ctx, int order);
/* Constraint of every later state solve, replacing RegOpt
 * m_PDESolver.pdetype (include/RegOpt.hpp:42-46, default TRANSPORTEQ at
 * src/RegOpt.cpp:1089; clairetools sets CONTINUITYEQ for its Ravens maps,
 * apps/clairetools.cpp:469-470), read only by CLAIRE::SolveStateEquation
 * (src/CLAIRE.cpp:1923-1926):
 *   CLR_PDE_CONTINUITY  ∂t m + ∇·(m v) = 0 (SolveContinuityEquationSL,
 *                       src/CLAIRE.cpp:2735-2854);
 *   CLR_PDE_TRANSPORT   ∂t m + v·∇m = 0 (SolveStateEquationSL).
 * The values are the reference enum's.  As in the reference, the adjoint,
 * incremental and Hessian solves stay the transport forms. */
#define CLR_PDE_CONTINUITY 0
#define CLR_PDE_TRANSPORT 1
int clr_set_pde_type(clr_ctx* ctx, int type);
/* Time integrator of every later transport solve, replacing RegOpt
 * m_PDESolver.type (include/RegOpt.hpp:31-35), set by "-pdesolver rk2|sl"
 * (src/RegOpt.cpp:653-658, default SL at src/RegOpt.cpp:1083):
 *   CLR_PDE_SOLVER_RK2  explicit RK2 with spectral derivatives
 *                       (SolveStateEquationRK2, SolveAdjointEquationRK2,
 *                       SolveIncStateEquationRK2, SolveIncAdjointEquationGNRK2,
 *                       src/CLAIRE.cpp:1995-2105, 2365-2548, 2962-3135,
 *                       3504-3683); conditionally stable;
 *   CLR_PDE_SOLVER_SL   semi-Lagrangian (default).
 * The values are the reference enum's; its RK2A (1) has no transport solvers
 * and is rejected.  With RK2 the mixed precision form has no effect. */
#define CLR_PDE_SOLVER_RK2 0
#define CLR_PDE_SOLVER_SL 2
int clr_set_pde_solver(clr_ctx* ctx, int type);
/* Operator applied by clr_hessian_matvec, replacing m_KrylovMethod.matvectype
 * (include/RegOpt.hpp:94-98, dispatched in CLAIRE::HessianMatVec,
 * src/CLAIRE.cpp:1251-1277):
 *   CLR_MATVEC_DEFAULT      the GN Hessian R''ṽ + b̃ (HessMatVec);
 *   CLR_MATVEC_PRECOND      hd·ṽ + (βA)⁻¹b̃[ṽ] (PrecondHessMatVec,
 *                           src/CLAIRE.cpp:1392-1451; not symmetric);
 *   CLR_MATVEC_PRECOND_SYM  hd·ṽ + (βA)^{-1/2}b̃[(βA)^{-1/2}ṽ]
 *                           (PrecondHessMatVecSym, src/CLAIRE.cpp:1463-1540).
 * clr_solve_gn accepts DEFAULT and PRECOND_SYM; with PRECOND_SYM it applies
 * (βA)^{-1/2} to the Krylov right-hand side and solution
 * (PreKrylovSolve/PostKrylovSolve, src/CLAIREBase.cpp:941-969). */
#define CLR_MATVEC_DEFAULT 0
#define CLR_MATVEC_PRECOND 1
#define CLR_MATVEC_PRECOND_SYM 2
int clr_set_matvec_type(clr_ctx* ctx, int type);
/* Registration model, replacing the choice of problem class in
 * CLAIREInterface (src/CLAIREInterface.cpp:685-690, RegOpt m_RegModel):
 *   CLR_MODEL_COMPRESSIBLE   CLAIRE (default);
 *   CLR_MODEL_STOKES         CLAIREStokes: incompressible, the adjoint and
 *                            incremental adjoint are pure transport
 *                            (src/CLAIREStokes.cpp:169-406) and both body
 *                            forces are Leray-projected (ApplyProjection,
 *                            src/CLAIREStokes.cpp:407-545);
 *   CLR_MODEL_RELAXED_STOKES CLAIREDivReg: compressible adjoints, body forces
 *                            projected with the β-weighted multiplier
 *                            (src/CLAIREDivReg.cpp:313-460) and the divergence
 *                            penalty ½·hd·β2·(‖∇(∇·v)‖² + ‖∇·v‖²) added to the
 *                            objective (src/CLAIREDivReg.cpp:109-260); β2 is
 *                            beta[2] of clr_create.
 * Changing the model drops what was computed for the previous one. */
#define CLR_MODEL_COMPRESSIBLE 0
#define CLR_MODEL_STOKES 1
#define CLR_MODEL_RELAXED_STOKES 2
int clr_set_reg_model(clr_ctx* ctx, int model);
/* The model's ApplyProjection on a vector field v (device, 3·nl), in place:
 * v ← v + F⁻¹[g·scale·m·(g·v̂)] (CLAIREStokes::ApplyProjection,
 * CLAIREDivReg::ApplyProjection; a no-op for the compressible model,
 * src/CLAIRE.cpp:4044-4048). */
int clr_apply_projection(clr_ctx* ctx, float* v);
/* Launch mode of clr_hessian_matvec, with no counterpart in the reference
 * (whose matvec is a sequence of host calls behind MatShell MATOP_MULT,
 * src/Optimizer.cpp:434-437).  on = 1: each (input, output, scale) triple
 * runs eagerly on its first two uses after the velocity, trajectory or
 * precision changed, is captured into a HIP graph on its third use and is
 * replayed from then on (at most 4 triples are kept); on = 0 (default):
 * always eager.  Results are the same either way.  clr_matvec_graph_count
 * reports the graphs currently held. */
int clr_set_matvec_graphs(clr_ctx* ctx, int on);
int clr_matvec_graph_count(clr_ctx* ctx, int* count);
/* Device memory held by the context (bytes); no reference counterpart. */
int clr_device_bytes(clr_ctx* ctx, size_t* bytes);

/* CLAIREBase::SetControlVariable (src/CLAIREBase.cpp:393) --
 * sets v (3·nl) and invalidates trajectories and Newton-invariant caches. */
int clr_set_velocity(clr_ctx* ctx, const float* v);
/* CLAIREBase::SetIncControlVariable (src/CLAIREBase.cpp:488) */
int clr_set_inc_velocity(clr_ctx* ctx, const float* vtilde);
/* CLAIREBase::SetTemplateImage / SetReferenceImage (src/CLAIREBase.cpp:304, :277) */
int clr_set_template_image(clr_ctx* ctx, const float* mT);
int clr_set_reference_image(clr_ctx* ctx, const float* mR);

/* CLAIRE::SolveStateEquation (src/CLAIRE.cpp:1861-1984, SL 2116-2170).
 * m_out: (nt+1)·nl time series, or NULL to keep it on the context only. */
int clr_solve_state(clr_ctx* ctx, float* m_out);
/* CLAIRE::SolveAdjointEquation (src/CLAIRE.cpp:2181-2356, SL 2550-2725):
 * λ at t = 0 (nl, may be NULL) and the body force b scaled by hd (3·nl, may
 * be NULL).  Runs the state solve first when it is stale. */
int clr_solve_adjoint(clr_ctx* ctx, float* lambda_out, float* b_out);
/* CLAIRE::SolveIncStateEquation (src/CLAIRE.cpp:2866-2950, SL 3147-3294):
 * m̃ at t = 1 for the incremental velocity vtilde (3·nl; NULL = the one set
 * with clr_set_inc_velocity). */
int clr_solve_inc_state(clr_ctx* ctx, const float* vtilde, float* mtilde_out);
/* CLAIRE::SolveIncAdjointEquation (src/CLAIRE.cpp:3307-3491, GN SL
 * 3849-4015) with final condition λ̃ = −m̃ (src/DistanceMeasureSL2.cpp:
 * 223-276): λ̃ at t = 0 (may be NULL) and b̃ scaled by hd (3·nl, may be NULL). */
int clr_solve_inc_adjoint(clr_ctx* ctx, const float* mtilde, float* ltilde_out, float* btilde_out);

/* CLAIRE::HessianMatVec(Vec Hvtilde, Vec vtilde, bool scale)
 * (src/CLAIRE.cpp:1218-1303 → HessMatVec 1311-1379), the GN Hessian behind
 * the TAO/KSP shell matrix callback reg::HessianMatVec(Mat,Vec,Vec)
 * (src/TaoInterface.cpp:137-155).  Hv must not alias vtilde (PETSc MatMult
 * rule).  scale = 0 divides the result by hd. */
int clr_hessian_matvec(clr_ctx* ctx, float* Hv, const float* vtilde, int scale);
/* Precompute the Newton-invariant data the Hessian uses (trajectories,
 * state, ∇m^j, (∇m^j)(X_s), adjoint update coefficient).  Optional: the
 * first clr_hessian_matvec after clr_set_velocity does it otherwise. */
int clr_prepare_hessian(clr_ctx* ctx);

/* CLAIRE::EvaluateObjective (src/CLAIRE.cpp:793-865) behind
 * reg::EvaluateObjective(Tao,Vec,PetscReal*,void*) (src/TaoInterface.cpp:40-61).
 * v may be NULL (use the current velocity). */
int clr_evaluate_objective(clr_ctx* ctx, const float* v, double* J, double* D, double* R);
/* CLAIRE::EvaluateGradient (src/CLAIRE.cpp:871-1008, L2 gradient 1015-1042)
 * behind reg::EvaluateGradient(Tao,Vec,Vec,void*) (src/TaoInterface.cpp:67-87). */
int clr_evaluate_gradient(clr_ctx* ctx, const float* v, float* g);
/* CLAIREBase::ApplyInvRegularizationOperator (src/CLAIREBase.cpp:987-1021),
 * the INVREG preconditioner behind reg::PrecondMatVec(PC,Vec,Vec)
 * (src/TaoInterface.cpp:162-180, src/Preconditioner.cpp:549-575).
 * out may alias in. */
int clr_apply_inv_reg(clr_ctx* ctx, const float* in, float* out, int applysqrt);

/* Gauss–Newton–Krylov registration on the device: replaces Optimizer::Run
 * + TaoSolve (src/Optimizer.cpp:448-527) with the reference's default
 * setup: TAO "nls" with the matrix-free Hessian (MatShell, src/Optimizer.cpp:
 * 433-437), KSPCG on the unpreconditioned residual with at most
 * krylov_maxit − 1 iterations (src/Optimizer.cpp:298-311), the INVREG
 * preconditioner (src/Preconditioner.cpp:521-573), Eisenstat–Walker forcing
 * (PreKrylovSolve, src/KrylovInterface.cpp:79-175), Armijo line search and
 * the gradient stopping test CheckConvergenceGrad (src/TaoInterface.cpp:
 * 390-512) with ‖g₀‖ taken at v = 0 (src/CLAIRE.cpp:320-335).
 * Template and reference images must be set.  v (device, 3·nl) holds the
 * initial guess on entry and the solution on exit.  history (host, may be
 * NULL) receives 8 doubles per iterate, up to history_len iterates:
 * iter, J, D, R, ‖g‖, ‖g‖/‖g₀‖, PCG iterations and step length of the step
 * taken from that iterate.  Inner products are FP64. */
typedef struct clr_gn_options {
    int maxit;           /* OptPara.maxiter (50) */
    double gttol;        /* OptPara.tol[2] (5e-2): stop when ‖g‖ < gttol·‖g₀‖ */
    double gatol;        /* OptPara.tol[0] (1e-6): stop when ‖g‖ < gatol */
    int krylov_maxit;    /* KrylovMethod.maxiter (50) */
    double krylov_rtol;  /* KrylovMethod.tol[0] (1e-12), used when forcing = 0 */
    int forcing;         /* 0 none, 1 quadratic, 2 superlinear (default) */
    int precond;         /* 0 none, 1 inverse regularization (default) */
    int linesearch;      /* 0 unit step, 1 Armijo (default) */
    int monitor;         /* 0 silent (default); 1 print the reference's iteration
                            table to stdout: the header of CLAIREInterface::Run
                            (src/CLAIREInterface.cpp:722-726) and one line per
                            iteration as OptimizationMonitor
                            (src/TaoInterface.cpp:525-592); batch solves stay
                            silent */
    double ls_sigma;     /* sufficient-decrease constant of the Armijo line
                            search, J(v + αd) <= J + σ·α·gᵀd (PETSc's
                            TAOLINESEARCHARMIJO, selected at src/Optimizer.cpp:390,
                            option -tao_ls_armijo_sigma, default 1e-4); in (0, 1) */
    double ls_stepinit;  /* first trial step of every line search (PETSc's
                            TaoLineSearch initial step, option
                            -tao_ls_stepinit, default 1); halved on each
                            rejected trial; > 0 */
    int krylov_block;    /* most CG iterations queued between two host checks
                            of the Krylov scalars, which stay on the device
                            (8); the queue is cut to the remainder predicted
                            from the residual's decay, and iterations queued
                            past the stopping test change nothing.  1 checks
                            after every iteration.  No counterpart in the
                            reference (PETSc's KSPCG checks on the host). */
} clr_gn_options;

#define CLR_GN_GTTOL 1      /* ‖g‖ < gttol·‖g₀‖ (TAO_CONVERGED_GTTOL) */
#define CLR_GN_GATOL 2      /* ‖g‖ < gatol (TAO_CONVERGED_GATOL) */
#define CLR_GN_STEPTOL 3    /* line-search step < 2^-10 (TAO_CONVERGED_STEPTOL) */
#define CLR_GN_MAXITS (-1)  /* iteration limit (TAO_DIVERGED_MAXITS) */
#define CLR_GN_NAN (-2)     /* objective or ‖g‖ not finite (TAO_DIVERGED_NAN) */
#define CLR_GN_LINESEARCH (-3)  /* no step satisfied the Armijo condition */

typedef struct clr_gn_report {
    int iterations, reason, matvecs, pcg_iterations, objective_evals, gradient_evals;
    double J0, D0, gnorm0, J, D, R, gnorm;
} clr_gn_report;

/* Fills the reference's defaults (RegOpt::Initialize, src/RegOpt.cpp:1068-1180). */
int clr_gn_options_default(clr_gn_options* opt);
int clr_solve_gn(clr_ctx* ctx, float* v, const clr_gn_options* opt, clr_gn_report* report, double* history,
                 int history_len);
/* Many independent registrations at once (no counterpart in the reference,
 * which runs one registration per MPI job): the Gauss–Newton–Krylov solve of
 * clr_solve_gn on each of `count` distinct contexts, spread over `workers`
 * host threads, each with its own non-blocking HIP stream, so that the
 * solves of small grids overlap on the device.  v[i] (device, 3·nl of
 * context i) is the initial guess on entry and the solution on exit;
 * reports[i] receives context i's report and status[i] (may be NULL) its
 * error code.  Every solve is attempted; the return value is the first
 * failing entry's code.  Each context runs on its worker's stream during the
 * call and is back on its own stream afterwards; work already queued on its
 * own stream is waited for first.  Results equal those of clr_solve_gn. */
int clr_solve_gn_batch(clr_ctx* const* ctxs, float* const* v, int count, const clr_gn_options* opt,
                       clr_gn_report* reports, int* status, int workers);

/* Regularization::EvaluateGradient == HessianMatVec (3·nl → 3·nl):
 * src/RegularizationH1.cpp:164,277, RegularizationH1SN.cpp:158,264,
 * RegularizationH2.cpp:178,285, RegularizationH2SN.cpp:182,293. */
int clr_reg_apply(clr_ctx* ctx, const float* v, float* out);
/* Regularization::EvaluateFunctional: src/RegularizationH1.cpp:64,
 * RegularizationH1SN.cpp:64, RegularizationH2.cpp:64, RegularizationH2SN.cpp:66. */
int clr_reg_functional(clr_ctx* ctx, const float* v, double* R);
/* Regularization::GetExtremeEigValsInvOp: src/RegularizationH1.cpp:427,
 * RegularizationH1SN.cpp:404, RegularizationH2.cpp:426, RegularizationH2SN.cpp:435. */
int clr_reg_eigvals_inv(clr_ctx* ctx, double* emin, double* emax);

/* SemiLagrangian::ComputeTrajectory (RK2, src/SemiLagrangian.cpp:192-288)
 * for the current velocity; X_out: nl×3 interleaved, normalized by 2π. */
int clr_trajectory(clr_ctx* ctx, int flag, float* X_out);
/* SemiLagrangian::Interpolate scalar (src/SemiLagrangian.cpp:536-594) and
 * vector (:632-728) along the current trajectory.  out may alias in. */
int clr_interp_scalar(clr_ctx* ctx, int flag, const float* in, float* out);
int clr_interp_vector(clr_ctx* ctx, int flag, const float* in, float* out);
/* Tricubic interpolation of f at arbitrary query points X (nl×3 normalized),
 * Interp3_Plan::interpolate + optimized_interp3_ghost_xyz_p
 * (src/Interp3_Plan.cpp:699-845, src/interp3.cpp:2492-2719). */
int clr_interp_points(clr_ctx* ctx, const float* f, const float* X, float* out);

/* accfft_grad_t / accfft_divergence_t (AccFFT, not vendored; e.g.
 * src/CLAIRE.cpp:2622,2457) and accfft_execute_r2c_t / c2r_t.  The complex
 * arrays are nx0·nx1·(nx2/2+1) interleaved (re,im) FP32; c2r is
 * unnormalized and leaves its input intact. */
int clr_grad(clr_ctx* ctx, const float* f, float* g);
int clr_div(clr_ctx* ctx, const float* v, float* d);
/* clr_grad without a context and without device allocation: g = ∇f on an
 * nx grid (g: [g1|g2|g3], 3·nl floats), launched on `stream` (null: the
 * default stream) of the current device, with results bit for bit equal to
 * clr_grad's on the same grid.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The three half spectra
 * live in a 32 MiB device array per device that every call shares (grids
 * with 3·nx0·nx1·(nx2/2+1)·8 bytes of spectra up to 32 MiB − 12 KiB, e.g.
 * 128³ or 128×150×128; larger grids return CLR_ERR_UNSUPPORTED_SIZE), so
 * calls that overlap on one device race.  The first call on a device looks
 * the array up and writes its twiddle tables synchronously. */
int clr_grad_noalloc(const int nx[3], const float* f, float* g, void* stream);
int clr_fft_r2c(clr_ctx* ctx, const float* in, float* out_complex);
int clr_fft_c2r(clr_ctx* ctx, const float* in_complex, float* out);

/* FP64 inner product / 2-norm of n FP32 values: PETSc's VecTDot / VecNorm
 * as the reference calls them (e.g. src/CLAIRE.cpp:286,334). */
int clr_dot(clr_ctx* ctx, const float* a, const float* b, int64_t n, double* out);
int clr_norm2(clr_ctx* ctx, const float* a, int64_t n, double* out);

/* Kernel timing with HIP events on the context stream, in place of the
 * reference's host timers (RegOpt::StartTimer/StopTimer/ResetTimers,
 * src/RegOpt.cpp:2585,2610,2519).  While enabled every
 * kernel launch is bracketed by events; clr_profile_get returns, for entry
 * i (0-based, sorted by name), the kernel class name, accumulated time in
 * ms, launch count and the algorithmic bytes of those launches.  Returns
 * CLR_ERR_INVALID_ARGUMENT past the last entry. */
int clr_profile_enable(clr_ctx* ctx, int on);
int clr_profile_reset(clr_ctx* ctx);
int clr_profile_get(clr_ctx* ctx, int i, char* name, int name_len, double* ms, long* count, double* bytes);

#ifdef __cplusplus
}
#endif

#endif /* CLAIRE_AMD_H_ */
