// oracle_capi.cpp -- extern "C" entry points of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY (see claire_oracle.hpp).  Each function exists in
// a double ("_f64", the accuracy oracle) and a float ("_f32", the
// single-precision reference CPU path) flavour.  All arrays are host memory;
// vector fields are passed as three component pointers, time series as one
// contiguous (nt+1)·nl block.  Return value 0 = OK, 1 = error (message via
// oracle_last_error).
#include "claire_oracle.hpp"

#include <chrono>
#include <string>

namespace {
thread_local std::string g_err;

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}
}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// Order of the characteristic scheme (2 or 4) used by every solve that
// follows (RegOpt m_PDESolver.rkorder).
int oracle_set_rk_order(int order) {
    if (order != 2 && order != 4) {
        g_err = "rk order must be 2 or 4";
        return 1;
    }
    oracle::rk_order() = order;
    return 0;
}

// Constraint of every later state solve: 0 continuity equation, 1 transport
// equation (RegOpt m_PDESolver.pdetype).
int oracle_set_pde_type(int t) {
    if (t != 0 && t != 1) {
        g_err = "pde type must be 0 (continuity) or 1 (transport)";
        return 1;
    }
    oracle::pde_type() = t;
    return 0;
}

// Time integrator of every later transport solve: 0 RK2, 2 SL
// (RegOpt m_PDESolver.type).
int oracle_set_pde_solver(int t) {
    if (t != 0 && t != 2) {
        g_err = "pde solver must be 0 (rk2) or 2 (sl)";
        return 1;
    }
    oracle::pde_solver() = t;
    return 0;
}

// Hessian operator type of every later matvec (0 default, 1 preconditioned,
// 2 symmetric preconditioned).
int oracle_set_matvec_type(int t) {
    if (t < 0 || t > 2) {
        g_err = "matvec type must be 0, 1 or 2";
        return 1;
    }
    oracle::matvec_type() = t;
    return 0;
}

// Registration model (0 compressible, 1 Stokes, 2 relaxed Stokes) and the
// β of its projection, used by every solve that follows.
int oracle_set_reg_model(int model, const double* beta) {
    if (model < 0 || model > 2) {
        g_err = "model must be 0, 1 or 2";
        return 1;
    }
    oracle::reg_model() = model;
    for (int i = 0; i < 3; ++i) oracle::reg_model_beta()[i] = beta ? beta[i] : 0.0;
    return 0;
}

void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

#define ORACLE_DEFINE(T, SFX)                                                                                      \
    /* Times `reps` Hessian matvecs (trajectories precomputed outside the timed region). */                       \
    int oracle_hessian_bench_##SFX(const int* nx, int nt, int norm, const double* beta, const T* v1, const T* v2,  \
                                   const T* v3, const T* m, const T* w1, const T* w2, const T* w3, T* h1, T* h2,   \
                                   T* h3, int reps, double* seconds) {                                             \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* v[3] = {v1, v2, v3};                                                                           \
            const T* w[3] = {w1, w2, w3};                                                                           \
            T* h[3] = {h1, h2, h3};                                                                                 \
            const T ht = T(1) / static_cast<T>(nt);                                                                 \
            std::vector<T> Xs(3 * g.nl), Xa(3 * g.nl);                                                              \
            oracle::trajectory(g, v, ht, +1, Xs.data());                                                            \
            oracle::trajectory(g, v, ht, -1, Xa.data());                                                            \
            for (int r = 0; r < reps; ++r) {                                                                        \
                const auto t0 = std::chrono::steady_clock::now();                                                   \
                oracle::hessian_matvec_x(g, nt, norm, b, v, Xs.data(), Xa.data(), m, w, h, true);                   \
                seconds[r] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();         \
            }                                                                                                       \
        });                                                                                                        \
    }                                                                                                              \
                                                                                      \
    int oracle_fft1d_##SFX(int n, const T* in_ri, T* out_ri, int sign) {                                           \
        return guard([&] {                                                                                         \
            oracle::FFT1D<T> f(n);                                                                                  \
            f.exec(reinterpret_cast<const std::complex<T>*>(in_ri), 1, reinterpret_cast<std::complex<T>*>(out_ri),  \
                   sign);                                                                                           \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_r2c_##SFX(const int* nx, const T* in, T* out_ri) {                                                  \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            oracle::r2c3d(g, in, reinterpret_cast<std::complex<T>*>(out_ri));                                       \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_c2r_##SFX(const int* nx, const T* in_ri, T* out) {                                                  \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            oracle::c2r3d(g, reinterpret_cast<const std::complex<T>*>(in_ri), out);                                 \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_grad_##SFX(const int* nx, const T* f, T* g1, T* g2, T* g3) {                                        \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            oracle::grad(g, f, g1, g2, g3);                                                                         \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_div_##SFX(const int* nx, const T* v1, const T* v2, const T* v3, T* d) {                             \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            oracle::div(g, v1, v2, v3, d);                                                                          \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_reg_apply_##SFX(const int* nx, int norm, const double* beta, const T* v1, const T* v2, const T* v3, \
                               T* o1, T* o2, T* o3) {                                                              \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* in[3] = {v1, v2, v3};                                                                          \
            T* out[3] = {o1, o2, o3};                                                                               \
            oracle::reg_apply(g, norm, b, in, out);                                                                 \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_reg_inverse_##SFX(const int* nx, int norm, const double* beta, const T* v1, const T* v2,           \
                                 const T* v3, T* o1, T* o2, T* o3, int applysqrt) {                                \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* in[3] = {v1, v2, v3};                                                                          \
            T* out[3] = {o1, o2, o3};                                                                               \
            oracle::reg_inverse(g, norm, b, in, out, applysqrt != 0);                                              \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_reg_functional_##SFX(const int* nx, int norm, const double* beta, const T* v1, const T* v2,        \
                                    const T* v3, double* R) {                                                      \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* in[3] = {v1, v2, v3};                                                                          \
            *R = oracle::reg_functional(g, norm, b, in);                                                            \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_reg_eigvals_inv_##SFX(const int* nx, int norm, const double* beta, double* emin, double* emax) {    \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            T lo, hi;                                                                                               \
            oracle::reg_eigvals_inv(g, norm, b, &lo, &hi);                                                          \
            *emin = lo;                                                                                             \
            *emax = hi;                                                                                             \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_interp_##SFX(const int* nx, const T* f, const T* X, T* out) {                                      \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            oracle::interp(g, f, X, out);                                                                           \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_interp_points_##SFX(const int* nx, const T* f, int64_t npts, const T* X, T* out) {                 \
        return guard([&] {                                                                                         \
            for (int64_t i = 0; i < npts; ++i) out[i] = oracle::interp_point(f, nx, X + 3 * i);                    \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_project_##SFX(const int* nx, int model, const double* beta, T* b1, T* b2, T* b3) {                  \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T bt[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            T* b[3] = {b1, b2, b3};                                                                                 \
            oracle::project(g, model, bt, b);                                                                       \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_trajectory_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, int sigma, T* X) {  \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            oracle::trajectory(g, v, T(1) / static_cast<T>(nt), sigma, X);                                         \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_solve_state_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, const T* mT,       \
                                 T* m) {                                                                           \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            oracle::solve_state(g, nt, v, mT, m);                                                                   \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_solve_adjoint_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, const T* m,      \
                                   const T* mR, T* lambda, T* b1, T* b2, T* b3) {                                  \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            T* b[3] = {b1, b2, b3};                                                                                 \
            oracle::solve_adjoint(g, nt, v, m, mR, lambda, b);                                                      \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_solve_inc_state_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, const T* m,    \
                                     const T* w1, const T* w2, const T* w3, T* mtilde) {                           \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            const T* w[3] = {w1, w2, w3};                                                                           \
            oracle::solve_inc_state(g, nt, v, m, w, mtilde);                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_solve_inc_adjoint_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, const T* m,  \
                                       const T* mtilde, T* ltilde, T* b1, T* b2, T* b3) {                          \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            T* b[3] = {b1, b2, b3};                                                                                 \
            oracle::solve_inc_adjoint(g, nt, v, m, mtilde, ltilde, b);                                              \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_hessian_matvec_##SFX(const int* nx, int nt, int norm, const double* beta, const T* v1,             \
                                    const T* v2, const T* v3, const T* m, const T* w1, const T* w2, const T* w3,   \
                                    T* h1, T* h2, T* h3, int scale) {                                              \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* v[3] = {v1, v2, v3};                                                                           \
            const T* w[3] = {w1, w2, w3};                                                                           \
            T* h[3] = {h1, h2, h3};                                                                                 \
            oracle::hessian_matvec(g, nt, norm, b, v, m, w, h, scale != 0);                                        \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_gradient_##SFX(const int* nx, int nt, int norm, const double* beta, const T* v1, const T* v2,      \
                              const T* v3, const T* mT, const T* mR, T* g1, T* g2, T* g3) {                        \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* v[3] = {v1, v2, v3};                                                                           \
            T* go[3] = {g1, g2, g3};                                                                                \
            oracle::gradient(g, nt, norm, b, v, mT, mR, go);                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_objective_##SFX(const int* nx, int nt, int norm, const double* beta, const T* v1, const T* v2,     \
                               const T* v3, const T* mT, const T* mR, double* J, double* D, double* R) {           \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            const T* v[3] = {v1, v2, v3};                                                                           \
            oracle::objective(g, nt, norm, b, v, mT, mR, J, D, R);                                                  \
        });                                                                                                        \
    }                                                                                                              \
    int oracle_det_def_grad_##SFX(const int* nx, int nt, const T* v1, const T* v2, const T* v3, int inverse,      \
                                  T* jac, double* stats) {                                                         \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T* v[3] = {v1, v2, v3};                                                                           \
            oracle::det_def_grad(g, nt, v, inverse != 0, jac, stats, stats + 1, stats + 2);                         \
        });                                                                                                        \
    }                                                                                                              \
    /* iopts: maxit, krylov_maxit, forcing, precond, linesearch; dopts: gttol, gatol, krylov_rtol;              */ \
    /* irep: iterations, reason, matvecs, pcg_iterations, objective_evals, gradient_evals;                      */ \
    /* drep: J0, D0, gnorm0, J, D, R, gnorm; v is the initial guess on entry and the solution on exit.          */ \
    int oracle_gn_solve_##SFX(const int* nx, int nt, int norm, const double* beta, T* v1, T* v2, T* v3,          \
                              const T* mT, const T* mR, const int* iopts, const double* dopts, int* irep,          \
                              double* drep, double* hist, int hist_len) {                                          \
        return guard([&] {                                                                                         \
            oracle::Grid<T> g(nx);                                                                                  \
            const T b[3] = {static_cast<T>(beta[0]), static_cast<T>(beta[1]), static_cast<T>(beta[2])};           \
            T* v[3] = {v1, v2, v3};                                                                                 \
            oracle::GNOptions o;                                                                                    \
            o.maxit = iopts[0];                                                                                     \
            o.krylov_maxit = iopts[1];                                                                              \
            o.forcing = iopts[2];                                                                                   \
            o.precond = iopts[3];                                                                                   \
            o.linesearch = iopts[4];                                                                                \
            o.gttol = dopts[0];                                                                                     \
            o.gatol = dopts[1];                                                                                     \
            o.krylov_rtol = dopts[2];                                                                               \
            oracle::GNReport r;                                                                                     \
            oracle::gn_solve(g, nt, norm, b, v, mT, mR, o, r, hist, hist_len);                                      \
            const int ir[6] = {r.iterations, r.reason, r.matvecs, r.pcg_iterations, r.objective_evals,             \
                               r.gradient_evals};                                                                   \
            const double dr[7] = {r.J0, r.D0, r.gnorm0, r.J, r.D, r.R, r.gnorm};                                    \
            std::copy(ir, ir + 6, irep);                                                                            \
            std::copy(dr, dr + 7, drep);                                                                            \
        });                                                                                                        \
    }

ORACLE_DEFINE(double, f64)
ORACLE_DEFINE(float, f32)

}  // extern "C"
