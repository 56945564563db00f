// capi.cpp -- the extern "C" boundary (include/claire_amd.h).
#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstring>
#include <memory>
#include <new>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/claire_amd.h"
#include "claire.hpp"
#include "nifti.hpp"
#include "petscbin.hpp"
#include "solver.hpp"

using claire_amd::Error;

struct clr_ctx {
    std::unique_ptr<reg::Profiler> prof;
    std::unique_ptr<reg::CLAIRE> claire;
    std::unique_ptr<reg::Optimizer> optimizer;  // work vectors kept across solves
    std::unique_ptr<claire_amd::FFTPlan3D> plan;  // for the raw FFT entry points
    reg::DeviceArray<claire_amd::c32> spec;
};

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return CLR_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "host allocation failed";
        return CLR_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_err = e.what();
        return CLR_ERR_INVALID_ARGUMENT;
    }
}

void need(const void* p, const char* what) {
    if (!p) throw Error(claire_amd::kErrInvalidArgument, std::string("null pointer: ") + what);
}

reg::CLAIRE& C(clr_ctx* ctx) {
    need(ctx, "context");
    return *ctx->claire;
}

const reg::RegOpt& O(clr_ctx* ctx) { return C(ctx).Opt(); }

void check_hip_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) throw Error(claire_amd::kErrHip, std::string("kernel launch failed: ") + hipGetErrorString(e));
}

reg::GNOptions gn_options(const clr_gn_options& opt) {
    reg::GNOptions o;
    o.maxit = opt.maxit;
    o.gttol = opt.gttol;
    o.gatol = opt.gatol;
    o.krylov_maxit = opt.krylov_maxit;
    o.krylov_rtol = opt.krylov_rtol;
    o.forcing = opt.forcing;
    o.precond = opt.precond;
    o.monitor = opt.monitor;
    o.linesearch = opt.linesearch;
    CLR_REQUIRE(opt.ls_sigma > 0.0 && opt.ls_sigma < 1.0, claire_amd::kErrInvalidArgument,
                "line-search sigma must be in (0, 1)");
    o.ls_sigma = opt.ls_sigma;
    CLR_REQUIRE(opt.ls_stepinit > 0.0 && std::isfinite(opt.ls_stepinit), claire_amd::kErrInvalidArgument,
                "initial line-search step must be positive");
    o.ls_stepinit = opt.ls_stepinit;
    CLR_REQUIRE(opt.krylov_block >= 1, claire_amd::kErrInvalidArgument, "krylov_block must be >= 1");
    o.krylov_block = opt.krylov_block;
    return o;
}

void run_gn(clr_ctx* ctx, float* v, const reg::GNOptions& o, clr_gn_report* report, double* history,
            int history_len) {
    if (!ctx->optimizer) ctx->optimizer.reset(new reg::Optimizer(&C(ctx)));
    reg::GNReport r;
    ctx->optimizer->Run(v, o, r, history, history_len);
    check_hip_launch();
    *report = clr_gn_report{r.iterations, r.reason, r.matvecs, r.pcg_iterations, r.objective_evals,
                            r.gradient_evals, r.J0, r.D0, r.gnorm0, r.J, r.D, r.R, r.gnorm};
}
}  // namespace

extern "C" {

const char* clr_last_error(void) { return g_err.c_str(); }

int clr_device_malloc(void** ptr, size_t bytes) {
    return guard([&] {
        need(ptr, "ptr");
        *ptr = nullptr;
        CLR_HIP_CHECK(hipMalloc(ptr, bytes));
    });
}
int clr_device_free(void* ptr) {
    return guard([&] {
        if (ptr) CLR_HIP_CHECK(hipFree(ptr));
    });
}
int clr_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guard([&] { CLR_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
int clr_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guard([&] { CLR_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}
int clr_memcpy_dtod(void* dst, const void* src, size_t bytes) {
    return guard([&] { CLR_HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice)); });
}
int clr_device_synchronize(void) {
    return guard([&] { CLR_HIP_CHECK(hipDeviceSynchronize()); });
}

int clr_create(clr_ctx** out, const int nx[3], int nt, int regnorm, const double beta[3], void* stream) {
    return guard([&] {
        need(out, "ctx");
        need(nx, "nx");
        need(beta, "beta");
        *out = nullptr;
        CLR_REQUIRE(regnorm >= 0 && regnorm <= 3, claire_amd::kErrInvalidArgument, "unknown regularization norm");
        CLR_REQUIRE(nt >= 1, claire_amd::kErrInvalidArgument, "nt must be >= 1");
        for (int a = 0; a < 3; ++a)
            CLR_REQUIRE(claire_amd::fft_size_supported(nx[a]), claire_amd::kErrUnsupportedSize,
                        "grid size " + std::to_string(nx[a]) + " not supported (powers of two 8..512)");
        std::unique_ptr<clr_ctx> c(new clr_ctx());
        hipStream_t st = static_cast<hipStream_t>(stream);
        c->prof.reset(new reg::Profiler(st));
        reg::RegOpt opt;
        opt.domain = claire_amd::make_dims(nx);
        opt.nt = nt;
        opt.nc = 1;
        opt.regnorm = regnorm;
        for (int i = 0; i < 3; ++i) opt.beta[i] = static_cast<float>(beta[i]);
        opt.stream = st;
        opt.prof = c->prof.get();
        c->claire.reset(new reg::CLAIRE(opt));
        check_hip_launch();
        *out = c.release();
    });
}

int clr_destroy(clr_ctx* ctx) {
    return guard([&] {
        if (!ctx) return;
        (void)hipStreamSynchronize(ctx->claire ? ctx->claire->Opt().stream : nullptr);
        delete ctx;
    });
}

int clr_synchronize(clr_ctx* ctx) {
    return guard([&] { CLR_HIP_CHECK(hipStreamSynchronize(O(ctx).stream)); });
}

int clr_device_bytes(clr_ctx* ctx, size_t* bytes) {
    return guard([&] {
        need(bytes, "bytes");
        *bytes = C(ctx).DeviceBytes();
    });
}

int clr_set_velocity(clr_ctx* ctx, const float* v) {
    return guard([&] {
        need(v, "v");
        C(ctx).SetControlVariable(v);
        check_hip_launch();
    });
}
int clr_set_inc_velocity(clr_ctx* ctx, const float* vt) {
    return guard([&] {
        need(vt, "vtilde");
        C(ctx).SetIncControlVariable(vt);
        check_hip_launch();
    });
}
int clr_set_template_image(clr_ctx* ctx, const float* mT) {
    return guard([&] { C(ctx).SetTemplateImage(mT); });
}
int clr_set_reference_image(clr_ctx* ctx, const float* mR) {
    return guard([&] { C(ctx).SetReferenceImage(mR); });
}

int clr_solve_state(clr_ctx* ctx, float* m_out) {
    return guard([&] {
        reg::CLAIRE& c = C(ctx);
        c.SolveStateEquation();
        check_hip_launch();
        if (m_out) {
            const size_t n = static_cast<size_t>(c.Opt().nt + 1) * c.Opt().domain.nl;
            CLR_HIP_CHECK(hipMemcpyAsync(m_out, c.StateVariable(), n * sizeof(float), hipMemcpyDeviceToDevice,
                                         c.Opt().stream));
        }
    });
}

int clr_solve_adjoint(clr_ctx* ctx, float* lambda_out, float* b_out) {
    return guard([&] {
        reg::CLAIRE& c = C(ctx);
        c.SolveAdjointEquation();
        check_hip_launch();
        const size_t nl = static_cast<size_t>(c.Opt().domain.nl);
        if (lambda_out)
            CLR_HIP_CHECK(hipMemcpyAsync(lambda_out, c.AdjointVariable(), nl * sizeof(float), hipMemcpyDeviceToDevice,
                                         c.Opt().stream));
        if (b_out) c.BodyForce()->GetComponents(b_out);
    });
}

int clr_solve_inc_state(clr_ctx* ctx, const float* vtilde, float* mtilde_out) {
    return guard([&] {
        reg::CLAIRE& c = C(ctx);
        if (vtilde) c.SetIncControlVariable(vtilde);
        c.SolveIncStateEquation();
        check_hip_launch();
        if (mtilde_out)
            CLR_HIP_CHECK(hipMemcpyAsync(mtilde_out, c.IncStateVariable(), c.Opt().domain.nl * sizeof(float),
                                         hipMemcpyDeviceToDevice, c.Opt().stream));
    });
}

int clr_solve_inc_adjoint(clr_ctx* ctx, const float* mtilde, float* ltilde_out, float* btilde_out) {
    return guard([&] {
        reg::CLAIRE& c = C(ctx);
        if (mtilde)
            c.SolveIncAdjointFrom(mtilde);
        else
            c.SolveIncAdjointEquation();
        check_hip_launch();
        if (ltilde_out)
            CLR_HIP_CHECK(hipMemcpyAsync(ltilde_out, c.IncAdjointVariable(), c.Opt().domain.nl * sizeof(float),
                                         hipMemcpyDeviceToDevice, c.Opt().stream));
        if (btilde_out) c.IncBodyForce()->GetComponents(btilde_out);
    });
}

int clr_hessian_matvec(clr_ctx* ctx, float* Hv, const float* vtilde, int scale) {
    return guard([&] {
        C(ctx).HessianMatVec(Hv, vtilde, scale != 0);
        check_hip_launch();
    });
}

int clr_prepare_hessian(clr_ctx* ctx) {
    return guard([&] {
        C(ctx).PrepareHessian();
        check_hip_launch();
    });
}

int clr_evaluate_objective(clr_ctx* ctx, const float* v, double* J, double* D, double* R) {
    return guard([&] {
        need(J, "J");
        double d = 0, r = 0, j = 0;
        C(ctx).EvaluateObjective(&j, &d, &r, v);
        *J = j;
        if (D) *D = d;
        if (R) *R = r;
    });
}

int clr_evaluate_gradient(clr_ctx* ctx, const float* v, float* g) {
    return guard([&] {
        need(g, "g");
        C(ctx).EvaluateGradient(g, v);
        check_hip_launch();
    });
}

int clr_nifti_read(const char* path, int nx[3], float* data, int64_t capacity, unsigned char* header) {
    return guard([&] {
        need(path, "path");
        need(nx, "nx");
        claire_amd::nifti::read(path, nx, data, capacity, header);
    });
}

int clr_nifti_write(const char* path, const int nx[3], const float* data, const float* spacing,
                    const unsigned char* header) {
    return guard([&] {
        need(path, "path");
        need(nx, "nx");
        need(data, "data");
        claire_amd::nifti::write(path, nx, data, spacing, header);
    });
}

int clr_bin_read(const char* path, int64_t* n, float* data, int64_t capacity) {
    return guard([&] {
        need(path, "path");
        need(n, "n");
        claire_amd::petscbin::read(path, n, data, capacity);
    });
}

int clr_bin_write(const char* path, int64_t n, const float* data) {
    return guard([&] {
        need(path, "path");
        if (n > 0) need(data, "data");
        claire_amd::petscbin::write(path, n, data);
    });
}

int clr_det_def_grad(clr_ctx* ctx, int inverse, float* jac, double* stats) {
    return guard([&] {
        need(stats, "stats");
        C(ctx).ComputeDetDefGrad(inverse != 0, jac, stats);
        check_hip_launch();
    });
}

int clr_check_bounds(clr_ctx* ctx, const float* v, int inverse, double bound, int* reached, double* stats) {
    return guard([&] {
        need(v, "v");
        need(reached, "reached");
        double st[3];
        C(ctx).SetControlVariable(v);
        C(ctx).ComputeDetDefGrad(inverse != 0, nullptr, st);
        *reached = (st[0] <= bound || 1.0 / st[1] <= bound) ? 1 : 0;
        if (stats) std::copy(st, st + 3, stats);
        check_hip_launch();
    });
}

int clr_set_precision(clr_ctx* ctx, int precision) {
    return guard([&] { C(ctx).SetPrecision(precision); });
}

int clr_set_matvec_type(clr_ctx* ctx, int type) {
    return guard([&] { C(ctx).SetMatVecType(type); });
}

int clr_set_reg_model(clr_ctx* ctx, int model) {
    return guard([&] { C(ctx).SetRegModel(model); });
}

int clr_apply_projection(clr_ctx* ctx, float* v) {
    return guard([&] {
        need(v, "v");
        C(ctx).ApplyProjection(v);
        check_hip_launch();
    });
}

int clr_set_matvec_graphs(clr_ctx* ctx, int on) {
    return guard([&] { C(ctx).SetMatVecGraphs(on != 0); });
}

int clr_matvec_graph_count(clr_ctx* ctx, int* count) {
    return guard([&] {
        need(count, "count");
        *count = C(ctx).MatVecGraphCount();
    });
}

int clr_set_rk_order(clr_ctx* ctx, int order) {
    return guard([&] { C(ctx).SetRKOrder(order); });
}

int clr_set_pde_type(clr_ctx* ctx, int type) {
    return guard([&] { C(ctx).SetPDEType(type); });
}

int clr_set_pde_solver(clr_ctx* ctx, int type) {
    return guard([&] { C(ctx).SetPDESolver(type); });
}

int clr_gn_options_default(clr_gn_options* opt) {
    return guard([&] {
        need(opt, "opt");
        const reg::GNOptions d;
        *opt = clr_gn_options{d.maxit,   d.gttol,      d.gatol,   d.krylov_maxit, d.krylov_rtol,
                              d.forcing, d.precond,    d.linesearch, d.monitor,   d.ls_sigma,
                              d.ls_stepinit, d.krylov_block};
    });
}

int clr_solve_gn(clr_ctx* ctx, float* v, const clr_gn_options* opt, clr_gn_report* report, double* history,
                 int history_len) {
    return guard([&] {
        need(v, "v");
        need(opt, "opt");
        need(report, "report");
        CLR_REQUIRE(history_len >= 0, claire_amd::kErrInvalidArgument, "negative history length");
        run_gn(ctx, v, gn_options(*opt), report, history, history ? history_len : 0);
    });
}

int clr_solve_gn_batch(clr_ctx* const* ctxs, float* const* v, int count, const clr_gn_options* opt,
                       clr_gn_report* reports, int* status, int workers) {
    return guard([&] {
        need(opt, "opt");
        CLR_REQUIRE(count >= 0, claire_amd::kErrInvalidArgument, "negative batch size");
        CLR_REQUIRE(workers >= 1, claire_amd::kErrInvalidArgument, "workers must be >= 1");
        if (count == 0) return;
        need(ctxs, "ctxs");
        need(v, "v");
        need(reports, "reports");
        std::set<const clr_ctx*> seen;
        for (int i = 0; i < count; ++i) {
            need(ctxs[i], "context");
            need(ctxs[i]->claire.get(), "context");
            need(v[i], "v");
            CLR_REQUIRE(seen.insert(ctxs[i]).second, claire_amd::kErrInvalidArgument,
                        "a context appears twice in the batch");
        }
        reg::GNOptions o = gn_options(*opt);
        o.monitor = 0;  // concurrent solves would interleave their tables
        int dev = 0;
        CLR_HIP_CHECK(hipGetDevice(&dev));
        std::atomic<int> next{0};
        std::vector<int> rc(count, CLR_OK);
        std::vector<std::string> err(count);
        auto worker = [&]() {
            hipStream_t st = nullptr;
            const bool ready = hipSetDevice(dev) == hipSuccess &&
                               hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
            for (int i = next.fetch_add(1); i < count; i = next.fetch_add(1)) {
                if (!ready) {
                    rc[i] = claire_amd::kErrHip;
                    err[i] = "could not create a stream for the batch worker";
                    continue;
                }
                reg::CLAIRE& c = *ctxs[i]->claire;
                const hipStream_t own = c.Opt().stream;
                try {
                    CLR_HIP_CHECK(hipStreamSynchronize(own));  // work queued on the context's stream first
                    c.SetStream(st);
                    run_gn(ctxs[i], v[i], o, &reports[i], nullptr, 0);
                    CLR_HIP_CHECK(hipStreamSynchronize(st));
                } catch (const Error& e) {
                    rc[i] = e.code;
                    err[i] = e.what();
                } catch (const std::exception& e) {
                    rc[i] = claire_amd::kErrInvalidArgument;
                    err[i] = e.what();
                }
                (void)hipStreamSynchronize(st);
                c.SetStream(own);
            }
            if (st) (void)hipStreamDestroy(st);
        };
        std::vector<std::thread> pool;
        const int nthreads = std::min(workers, count);
        for (int w = 0; w < nthreads; ++w) pool.emplace_back(worker);
        for (auto& t : pool) t.join();
        int first = -1;
        for (int i = 0; i < count; ++i) {
            if (status) status[i] = rc[i];
            if (rc[i] != CLR_OK && first < 0) first = i;
        }
        if (first >= 0) throw Error(rc[first], "batch entry " + std::to_string(first) + ": " + err[first]);
    });
}

int clr_apply_inv_reg(clr_ctx* ctx, const float* in, float* out, int applysqrt) {
    return guard([&] {
        need(in, "in");
        need(out, "out");
        C(ctx).ApplyInvRegularizationOperator(out, in, applysqrt != 0);
        check_hip_launch();
    });
}

int clr_reg_apply(clr_ctx* ctx, const float* v, float* out) {
    return guard([&] {
        need(v, "v");
        need(out, "out");
        reg::CLAIRE& c = C(ctx);
        reg::VecField o(&c.Opt(), out), in(&c.Opt(), const_cast<float*>(v));
        c.Reg()->HessianMatVec(&o, &in);
        check_hip_launch();
    });
}

int clr_reg_functional(clr_ctx* ctx, const float* v, double* R) {
    return guard([&] {
        need(v, "v");
        need(R, "R");
        reg::CLAIRE& c = C(ctx);
        reg::VecField in(&c.Opt(), const_cast<float*>(v));
        c.Reg()->EvaluateFunctional(R, &in);
    });
}

int clr_reg_eigvals_inv(clr_ctx* ctx, double* emin, double* emax) {
    return guard([&] {
        need(emin, "emin");
        need(emax, "emax");
        float lo = 0, hi = 0;
        C(ctx).Reg()->GetExtremeEigValsInvOp(lo, hi);
        *emin = lo;
        *emax = hi;
    });
}

int clr_trajectory(clr_ctx* ctx, int flag, float* X_out) {
    return guard([&] {
        need(X_out, "X");
        reg::CLAIRE& c = C(ctx);
        const std::string f = flag == CLR_STATE ? "state" : (flag == CLR_ADJOINT ? "adjoint" : "?");
        if (f == "?") throw Error(claire_amd::kErrInvalidArgument, "flag wrong");
        c.EnsureTrajectories();
        CLR_HIP_CHECK(hipMemcpyAsync(X_out, c.SL()->Trajectory(f), 3 * c.Opt().domain.nl * sizeof(float),
                                     hipMemcpyDeviceToDevice, c.Opt().stream));
    });
}

namespace {
void ensure_traj(reg::CLAIRE& c, int flag) {
    const std::string f = flag == CLR_STATE ? "state" : (flag == CLR_ADJOINT ? "adjoint" : "?");
    if (f == "?") throw Error(claire_amd::kErrInvalidArgument, "flag wrong");
    c.EnsureTrajectories();
}
}  // namespace

int clr_interp_scalar(clr_ctx* ctx, int flag, const float* in, float* out) {
    return guard([&] {
        need(in, "in");
        need(out, "out");
        reg::CLAIRE& c = C(ctx);
        ensure_traj(c, flag);
        c.SL()->Interpolate(out, in, flag == CLR_STATE ? "state" : "adjoint");
        check_hip_launch();
    });
}

int clr_interp_vector(clr_ctx* ctx, int flag, const float* in, float* out) {
    return guard([&] {
        need(in, "in");
        need(out, "out");
        reg::CLAIRE& c = C(ctx);
        ensure_traj(c, flag);
        reg::VecField o(&c.Opt(), out), i(&c.Opt(), const_cast<float*>(in));
        c.SL()->Interpolate(&o, &i, flag == CLR_STATE ? "state" : "adjoint");
        check_hip_launch();
    });
}

int clr_interp_points(clr_ctx* ctx, const float* f, const float* X, float* out) {
    return guard([&] {
        need(f, "f");
        need(X, "X");
        need(out, "out");
        CLR_REQUIRE(out != f, claire_amd::kErrInvalidArgument, "interpolation output must not alias the field");
        reg::CLAIRE& c = C(ctx);
        reg::KernelScope k(c.Opt().prof, "gather_scalar", 20.0 * c.Opt().domain.nl);
        claire_amd::launch_gather1(claire_amd::gather_dims(c.Opt().domain), f, X, out, c.Opt().stream);
        check_hip_launch();
    });
}

int clr_grad(clr_ctx* ctx, const float* f, float* g) {
    return guard([&] {
        need(f, "f");
        need(g, "g");
        reg::CLAIRE& c = C(ctx);
        const int64_t nl = c.Opt().domain.nl;
        float* o[3] = {g, g + nl, g + 2 * nl};
        c.Spectral()->Grad(o, f);
        check_hip_launch();
    });
}

int clr_div(clr_ctx* ctx, const float* v, float* d) {
    return guard([&] {
        need(v, "v");
        need(d, "d");
        reg::CLAIRE& c = C(ctx);
        const int64_t nl = c.Opt().domain.nl;
        const float* in[3] = {v, v + nl, v + 2 * nl};
        c.Spectral()->Div(d, in);
        check_hip_launch();
    });
}

int clr_grad_noalloc(const int nx[3], const float* f, float* g, void* stream) {
    return guard([&] {
        need(nx, "nx");
        need(f, "f");
        need(g, "g");
        claire_amd::grad_noalloc(nx, f, g, static_cast<hipStream_t>(stream));
        check_hip_launch();
    });
}

int clr_fft_r2c(clr_ctx* ctx, const float* in, float* out_complex) {
    return guard([&] {
        need(in, "in");
        need(out_complex, "out");
        reg::CLAIRE& c = C(ctx);
        c.Spectral()->R2C(in, reinterpret_cast<claire_amd::c32*>(out_complex));
        check_hip_launch();
    });
}

int clr_fft_c2r(clr_ctx* ctx, const float* in_complex, float* out) {
    return guard([&] {
        need(in_complex, "in");
        need(out, "out");
        reg::CLAIRE& c = C(ctx);
        const int64_t nlc = c.Opt().domain.nlc;
        ctx->spec.resize(static_cast<size_t>(nlc));
        CLR_HIP_CHECK(hipMemcpyAsync(ctx->spec.get(), in_complex, nlc * sizeof(claire_amd::c32), hipMemcpyDeviceToDevice,
                                     c.Opt().stream));
        c.Spectral()->C2R(ctx->spec.get(), out);
        check_hip_launch();
    });
}

int clr_dot(clr_ctx* ctx, const float* a, const float* b, int64_t n, double* out) {
    return guard([&] {
        need(a, "a");
        need(b, "b");
        need(out, "out");
        CLR_REQUIRE(n >= 0, claire_amd::kErrInvalidArgument, "negative length");
        *out = C(ctx).Red()->Dot(a, b, n);
    });
}

int clr_norm2(clr_ctx* ctx, const float* a, int64_t n, double* out) {
    return guard([&] {
        need(a, "a");
        need(out, "out");
        CLR_REQUIRE(n >= 0, claire_amd::kErrInvalidArgument, "negative length");
        // a sum of squares: non-negative, or NaN when an entry is not finite
        *out = std::sqrt(C(ctx).Red()->Dot(a, a, n));
    });
}

int clr_profile_enable(clr_ctx* ctx, int on) {
    return guard([&] {
        need(ctx, "context");
        if (!on) ctx->prof->collect();
        ctx->prof->enable(on != 0);
    });
}

int clr_profile_reset(clr_ctx* ctx) {
    return guard([&] {
        need(ctx, "context");
        ctx->prof->reset();
    });
}

int clr_profile_get(clr_ctx* ctx, int i, char* name, int name_len, double* ms, long* count, double* bytes) {
    return guard([&] {
        need(ctx, "context");
        ctx->prof->collect();
        const auto& e = ctx->prof->entries();
        CLR_REQUIRE(i >= 0 && i < static_cast<int>(e.size()), claire_amd::kErrInvalidArgument, "no such profile entry");
        auto it = e.begin();
        std::advance(it, i);
        if (name && name_len > 0) {
            std::strncpy(name, it->first.c_str(), static_cast<size_t>(name_len - 1));
            name[name_len - 1] = '\0';
        }
        if (ms) *ms = it->second.ms;
        if (count) *count = it->second.count;
        if (bytes) *bytes = it->second.bytes;
    });
}

}  // extern "C"
