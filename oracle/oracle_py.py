"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
benchmark's CPU-baseline leg, always as the checker.  Product code never
imports this module.

Every function takes numpy arrays in the reference's layouts:
  scalar field       (n1, n2, n3), x3 fastest
  vector field       (3, n1, n2, n3)  (VecField SoA, include/VecField.hpp:93-95)
  time series        (nt+1, n1, n2, n3)
  half spectrum      (n1, n2, n3//2+1) complex
  query points       (n1, n2, n3, 3) normalized by 2π, interleaved
and a precision flag: 'f64' (accuracy oracle) or 'f32' (reference CPU path).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

REG = {"h1": 0, "h2": 1, "h1sn": 2, "h2sn": 3}


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.oracle_last_error.restype = ctypes.c_char_p
    return _lib


def set_num_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def set_rk_order(order: int) -> None:
    """Characteristic scheme of every later solve: 2 (default) or 4."""
    if lib().oracle_set_rk_order(int(order)) != 0:
        raise ValueError(lib().oracle_last_error().decode())


class rk_order:
    """Context manager that runs a block with the given characteristic order."""

    def __init__(self, order: int):
        self.order = int(order)

    def __enter__(self):
        set_rk_order(self.order)
        return self

    def __exit__(self, *exc):
        set_rk_order(2)
        return False


PDE_TYPES = {"continuity": 0, "transport": 1}


def set_pde_type(name: str) -> None:
    """Constraint of every later state solve: "transport" (default) or
    "continuity"."""
    if lib().oracle_set_pde_type(PDE_TYPES[name]) != 0:
        raise ValueError(lib().oracle_last_error().decode())


class pde_type:
    """Context manager that runs a block with the given state constraint."""

    def __init__(self, name: str):
        self.name = name

    def __enter__(self):
        set_pde_type(self.name)
        return self

    def __exit__(self, *exc):
        set_pde_type("transport")
        return False


PDE_SOLVERS = {"rk2": 0, "sl": 2}


def set_pde_solver(name: str) -> None:
    """Time integrator of every later transport solve: "sl" (default) or
    "rk2"."""
    if lib().oracle_set_pde_solver(PDE_SOLVERS[name]) != 0:
        raise ValueError(lib().oracle_last_error().decode())


class pde_solver:
    """Context manager that runs a block with the given time integrator."""

    def __init__(self, name: str):
        self.name = name

    def __enter__(self):
        set_pde_solver(self.name)
        return self

    def __exit__(self, *exc):
        set_pde_solver("sl")
        return False


MODELS = {"compressible": 0, "stokes": 1, "relaxed_stokes": 2}


def set_reg_model(model, beta=(0.0, 0.0, 0.0)) -> None:
    """Registration model of every later solve ("compressible", "stokes",
    "relaxed_stokes") and the β of its projection."""
    if lib().oracle_set_reg_model(int(MODELS[model]), _beta(beta)) != 0:
        raise ValueError(lib().oracle_last_error().decode())


class reg_model:
    """Context manager that runs a block with the given registration model."""

    def __init__(self, model, beta=(0.0, 0.0, 0.0)):
        self.model, self.beta = model, beta

    def __enter__(self):
        set_reg_model(self.model, self.beta)
        return self

    def __exit__(self, *exc):
        set_reg_model("compressible")
        return False


MATVEC_TYPES = {"default": 0, "precond": 1, "precond_sym": 2}


class matvec_type:
    """Context manager: Hessian operator type of the matvecs in a block."""

    def __init__(self, kind):
        self.kind = kind

    def __enter__(self):
        if lib().oracle_set_matvec_type(MATVEC_TYPES[self.kind]) != 0:
            raise ValueError(lib().oracle_last_error().decode())
        return self

    def __exit__(self, *exc):
        lib().oracle_set_matvec_type(0)
        return False


def project(b, model, beta, prec="f64"):
    """The model's ApplyProjection on a vector field (3, n1, n2, n3)."""
    out = np.array(_c(b, prec), copy=True)
    _call("oracle_project", prec, _nx(out.shape[1:]), ctypes.c_int(MODELS[model]), _beta(beta), _p(out[0]),
          _p(out[1]), _p(out[2]))
    return out


def _dt(prec):
    return np.float64 if prec == "f64" else np.float32


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _nx(shape):
    return (ctypes.c_int * 3)(*[int(s) for s in shape])


def _beta(beta):
    b = list(beta) + [0.0] * (3 - len(beta))
    return (ctypes.c_double * 3)(*b[:3])


def _call(name, prec, *args):
    fn = getattr(lib(), f"{name}_{prec}")
    rc = fn(*args)
    if rc != 0:
        raise RuntimeError(f"{name}_{prec}: {lib().oracle_last_error().decode()}")


def _c(a, prec):
    return np.ascontiguousarray(a, dtype=_dt(prec))


def fft1d(x, sign=-1, prec="f64"):
    dt = _dt(prec)
    x = np.ascontiguousarray(x, dtype=np.complex128 if prec == "f64" else np.complex64)
    out = np.empty_like(x)
    _call("oracle_fft1d", prec, ctypes.c_int(x.size), _p(x.view(dt)), _p(out.view(dt)), ctypes.c_int(sign))
    return out


def r2c(f, prec="f64"):
    f = _c(f, prec)
    n = f.shape
    ct = np.complex128 if prec == "f64" else np.complex64
    out = np.empty((n[0], n[1], n[2] // 2 + 1), dtype=ct)
    _call("oracle_r2c", prec, _nx(n), _p(f), _p(out))
    return out


def c2r(fh, n, prec="f64"):
    ct = np.complex128 if prec == "f64" else np.complex64
    fh = np.ascontiguousarray(fh, dtype=ct)
    out = np.empty(tuple(n), dtype=_dt(prec))
    _call("oracle_c2r", prec, _nx(n), _p(fh), _p(out))
    return out


def grad(f, prec="f64"):
    f = _c(f, prec)
    out = np.empty((3,) + f.shape, dtype=_dt(prec))
    _call("oracle_grad", prec, _nx(f.shape), _p(f), _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def div(v, prec="f64"):
    v = _c(v, prec)
    out = np.empty(v.shape[1:], dtype=_dt(prec))
    _call("oracle_div", prec, _nx(v.shape[1:]), _p(v[0]), _p(v[1]), _p(v[2]), _p(out))
    return out


def reg_apply(v, regnorm, beta, prec="f64"):
    v = _c(v, prec)
    out = np.empty_like(v)
    _call("oracle_reg_apply", prec, _nx(v.shape[1:]), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(out[0]), _p(out[1]), _p(out[2]))
    return out


def reg_inverse(v, regnorm, beta, applysqrt=False, prec="f64"):
    v = _c(v, prec)
    out = np.empty_like(v)
    _call("oracle_reg_inverse", prec, _nx(v.shape[1:]), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(out[0]), _p(out[1]), _p(out[2]), ctypes.c_int(int(applysqrt)))
    return out


def reg_functional(v, regnorm, beta, prec="f64"):
    v = _c(v, prec)
    r = ctypes.c_double()
    _call("oracle_reg_functional", prec, _nx(v.shape[1:]), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), ctypes.byref(r))
    return r.value


def reg_eigvals_inv(n, regnorm, beta, prec="f64"):
    lo, hi = ctypes.c_double(), ctypes.c_double()
    _call("oracle_reg_eigvals_inv", prec, _nx(n), ctypes.c_int(REG[regnorm]), _beta(beta),
          ctypes.byref(lo), ctypes.byref(hi))
    return lo.value, hi.value


def interp(f, X, prec="f64"):
    """Interpolate f at query points X of shape (..., 3) (normalized by 2π)."""
    f = _c(f, prec)
    X = _c(X, prec)
    npts = X.size // 3
    out = np.empty(X.shape[:-1], dtype=_dt(prec))
    _call("oracle_interp_points", prec, _nx(f.shape), _p(f), ctypes.c_int64(npts), _p(X), _p(out))
    return out


def trajectory(v, nt, sigma, prec="f64"):
    v = _c(v, prec)
    n = v.shape[1:]
    X = np.empty(tuple(n) + (3,), dtype=_dt(prec))
    _call("oracle_trajectory", prec, _nx(n), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          ctypes.c_int(sigma), _p(X))
    return X


def solve_state(v, mT, nt, prec="f64"):
    v = _c(v, prec)
    mT = _c(mT, prec)
    m = np.empty((nt + 1,) + mT.shape, dtype=_dt(prec))
    _call("oracle_solve_state", prec, _nx(mT.shape), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          _p(mT), _p(m))
    return m


def solve_adjoint(v, m, mR, nt, prec="f64"):
    v, m, mR = _c(v, prec), _c(m, prec), _c(mR, prec)
    lam = np.empty(mR.shape, dtype=_dt(prec))
    b = np.empty((3,) + mR.shape, dtype=_dt(prec))
    _call("oracle_solve_adjoint", prec, _nx(mR.shape), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          _p(m), _p(mR), _p(lam), _p(b[0]), _p(b[1]), _p(b[2]))
    return lam, b


def solve_inc_state(v, m, vt, nt, prec="f64"):
    v, m, vt = _c(v, prec), _c(m, prec), _c(vt, prec)
    mt = np.empty(v.shape[1:], dtype=_dt(prec))
    _call("oracle_solve_inc_state", prec, _nx(v.shape[1:]), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          _p(m), _p(vt[0]), _p(vt[1]), _p(vt[2]), _p(mt))
    return mt


def solve_inc_adjoint(v, m, mt, nt, prec="f64"):
    v, m, mt = _c(v, prec), _c(m, prec), _c(mt, prec)
    lt = np.empty(mt.shape, dtype=_dt(prec))
    b = np.empty((3,) + mt.shape, dtype=_dt(prec))
    _call("oracle_solve_inc_adjoint", prec, _nx(mt.shape), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          _p(m), _p(mt), _p(lt), _p(b[0]), _p(b[1]), _p(b[2]))
    return lt, b


def hessian_matvec(v, m, vt, nt, regnorm, beta, scale=True, prec="f64"):
    v, m, vt = _c(v, prec), _c(m, prec), _c(vt, prec)
    h = np.empty_like(vt)
    _call("oracle_hessian_matvec", prec, _nx(v.shape[1:]), ctypes.c_int(nt), ctypes.c_int(REG[regnorm]),
          _beta(beta), _p(v[0]), _p(v[1]), _p(v[2]), _p(m), _p(vt[0]), _p(vt[1]), _p(vt[2]),
          _p(h[0]), _p(h[1]), _p(h[2]), ctypes.c_int(int(scale)))
    return h


def gradient(v, mT, mR, nt, regnorm, beta, prec="f64"):
    v, mT, mR = _c(v, prec), _c(mT, prec), _c(mR, prec)
    g = np.empty_like(v)
    _call("oracle_gradient", prec, _nx(mT.shape), ctypes.c_int(nt), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(mT), _p(mR), _p(g[0]), _p(g[1]), _p(g[2]))
    return g


def objective(v, mT, mR, nt, regnorm, beta, prec="f64"):
    v, mT, mR = _c(v, prec), _c(mT, prec), _c(mR, prec)
    J, D, R = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    _call("oracle_objective", prec, _nx(mT.shape), ctypes.c_int(nt), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(mT), _p(mR), ctypes.byref(J), ctypes.byref(D), ctypes.byref(R))
    return J.value, D.value, R.value


def hessian_bench(v, m, vt, nt, regnorm, beta, reps=1, prec="f32"):
    """Time `reps` Hessian matvecs of the CPU reference path (trajectories
    computed before the timed region); returns (H, list of per-matvec
    seconds)."""
    v, m, vt = _c(v, prec), _c(m, prec), _c(vt, prec)
    h = np.empty_like(vt)
    sec = (ctypes.c_double * max(1, int(reps)))()
    _call("oracle_hessian_bench", prec, _nx(v.shape[1:]), ctypes.c_int(nt), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(m), _p(vt[0]), _p(vt[1]), _p(vt[2]), _p(h[0]), _p(h[1]), _p(h[2]),
          ctypes.c_int(reps), sec)
    return h, [float(s) for s in sec[: max(0, int(reps))]]


GN_REASONS = {0: "continue", 1: "gttol", 2: "gatol", 3: "steptol", -1: "maxits", -2: "nan", -3: "linesearch"}


def gn_solve(v0, mT, mR, nt, regnorm, beta, maxit=50, gttol=5e-2, gatol=1e-6, krylov_maxit=50,
             krylov_rtol=1e-12, forcing=2, precond=1, linesearch=1, prec="f64", hist_len=128):
    """Gauss–Newton–Krylov registration (the reference's default TAO NLS +
    KSPCG + INVREG + Eisenstat–Walker + Armijo).  Returns (v, report, history)
    with history rows (iter, J, D, R, ‖g‖, ‖g‖/‖g₀‖, pcg_its, step)."""
    v = np.ascontiguousarray(np.array(v0, dtype=_dt(prec), copy=True))
    mT, mR = _c(mT, prec), _c(mR, prec)
    iopts = (ctypes.c_int * 5)(maxit, krylov_maxit, forcing, precond, linesearch)
    dopts = (ctypes.c_double * 3)(gttol, gatol, krylov_rtol)
    irep = (ctypes.c_int * 6)()
    drep = (ctypes.c_double * 7)()
    hist = np.zeros((hist_len, 8), dtype=np.float64)
    _call("oracle_gn_solve", prec, _nx(mT.shape), ctypes.c_int(nt), ctypes.c_int(REG[regnorm]), _beta(beta),
          _p(v[0]), _p(v[1]), _p(v[2]), _p(mT), _p(mR), iopts, dopts, irep, drep, _p(hist), ctypes.c_int(hist_len))
    rep = dict(zip(["iterations", "reason", "matvecs", "pcg_iterations", "objective_evals", "gradient_evals"],
                   list(irep)))
    rep.update(zip(["J0", "D0", "gnorm0", "J", "D", "R", "gnorm"], list(drep)))
    rep["reason"] = GN_REASONS.get(rep["reason"], rep["reason"])
    return v, rep, hist[: min(rep["iterations"] + 1, hist_len)]


def det_def_grad(v, nt, inverse=False, prec="f64"):
    """det(∇y) (or det(∇y⁻¹)) by SL transport; returns (J, (min, max, mean))."""
    v = _c(v, prec)
    n = v.shape[1:]
    jac = np.empty(n, dtype=_dt(prec))
    stats = (ctypes.c_double * 3)()
    _call("oracle_det_def_grad", prec, _nx(n), ctypes.c_int(nt), _p(v[0]), _p(v[1]), _p(v[2]),
          ctypes.c_int(int(inverse)), _p(jac), stats)
    return jac, tuple(stats)
