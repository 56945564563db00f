"""claire_amd -- MI355X-native CLAIRE registration hot path (Python binding).

A thin ctypes layer over the C ABI in include/claire_amd.h, implemented by
the in-tree HIP library claire_amd/libclaire_amd.so (hand-written gfx950
kernels + C++ host classes mirroring the reference's CLAIRE,
SemiLagrangian, Regularization and VecField).  There is no CPU fallback:
if the library is missing or no GPU is usable the calls raise.

Array conventions (host numpy arrays, FP32):
  scalar field   (n1, n2, n3), x3 fastest
  vector field   (3, n1, n2, n3), the flat [v1|v2|v3] solver vector
  time series    (nt+1, n1, n2, n3)
  query points   (n1, n2, n3, 3), normalized by 2π
  half spectrum  (n1, n2, n3//2+1) complex64
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclaire_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "claire_amd.h")

REG = {"h1": 0, "h2": 1, "h1sn": 2, "h2sn": 3}
STATE, ADJOINT = 0, 1

_lib = None


class ClaireError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libclaire_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ClaireError(-1, f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.clr_last_error.restype = ctypes.c_char_p
    return _lib


def declared_symbols():
    """Names of the functions declared in include/claire_amd.h."""
    with open(HEADER_PATH) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(clr_\w+)\s*\(", text, flags=re.M)))


def _check(rc):
    if rc != 0:
        raise ClaireError(rc, lib().clr_last_error().decode())


def _vp(p):
    return ctypes.c_void_p(p)


def nifti_read(path):
    """Reads a NIfTI-1 volume (.nii / .nii.gz): returns (array of shape
    (n1, n2, n3) = (nz, ny, nx), float32 in the grid's flat order, and the
    348 header bytes)."""
    nx = (ctypes.c_int * 3)()
    hdr = (ctypes.c_ubyte * 348)()
    _check(lib().clr_nifti_read(str(path).encode(), nx, None, ctypes.c_int64(0), hdr))
    out = np.empty(tuple(nx), dtype=np.float32)
    _check(lib().clr_nifti_read(str(path).encode(), nx, out.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_int64(out.size), hdr))
    return out, bytes(hdr)


def nifti_write(path, a, spacing=None, header=None):
    """Writes a FLOAT32 NIfTI-1 volume from an array of shape (n1, n2, n3);
    `header` (348 bytes of a read image) and `spacing` of (x1, x2, x3) are
    optional."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.ndim != 3:
        raise ValueError("expected a 3D array")
    nx = (ctypes.c_int * 3)(*a.shape)
    sp = (ctypes.c_float * 3)(*spacing) if spacing is not None else None
    hd = (ctypes.c_ubyte * 348)(*header) if header is not None else None
    _check(lib().clr_nifti_write(str(path).encode(), nx, a.ctypes.data_as(ctypes.c_void_p), sp, hd))


def bin_read(path, shape=None):
    """Reads a PETSc binary vector (.bin): float32 values in file order,
    reshaped to `shape` (e.g. the grid's (n1, n2, n3)) when given."""
    n = ctypes.c_int64()
    _check(lib().clr_bin_read(str(path).encode(), ctypes.byref(n), None, ctypes.c_int64(0)))
    out = np.empty(n.value, dtype=np.float32)
    _check(lib().clr_bin_read(str(path).encode(), ctypes.byref(n), out.ctypes.data_as(ctypes.c_void_p),
                              ctypes.c_int64(out.size)))
    return out.reshape(shape) if shape is not None else out


def bin_write(path, a):
    """Writes an array as a PETSc binary vector (.bin) of float32 values in
    the array's flat order."""
    a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
    _check(lib().clr_bin_write(str(path).encode(), ctypes.c_int64(a.size), a.ctypes.data_as(ctypes.c_void_p)))


class GNOptions(ctypes.Structure):
    """clr_gn_options (include/claire_amd.h)."""
    _fields_ = [("maxit", ctypes.c_int), ("gttol", ctypes.c_double), ("gatol", ctypes.c_double),
                ("krylov_maxit", ctypes.c_int), ("krylov_rtol", ctypes.c_double), ("forcing", ctypes.c_int),
                ("precond", ctypes.c_int), ("linesearch", ctypes.c_int), ("monitor", ctypes.c_int),
                ("ls_sigma", ctypes.c_double), ("ls_stepinit", ctypes.c_double), ("krylov_block", ctypes.c_int)]


class GNReport(ctypes.Structure):
    """clr_gn_report (include/claire_amd.h)."""
    _fields_ = [(n, ctypes.c_int) for n in ("iterations", "reason", "matvecs", "pcg_iterations",
                                            "objective_evals", "gradient_evals")] + \
               [(n, ctypes.c_double) for n in ("J0", "D0", "gnorm0", "J", "D", "R", "gnorm")]


MODELS = {"compressible": 0, "stokes": 1, "relaxed_stokes": 2}
PDE_TYPES = {"continuity": 0, "transport": 1}
PDE_SOLVERS = {"rk2": 0, "sl": 2}
GN_REASONS = {0: "continue", 1: "gttol", 2: "gatol", 3: "steptol", -1: "maxits", -2: "nan", -3: "linesearch"}


def solve_gn_batch(problems, v0s=None, workers=4, **options):
    """Gauss–Newton–Krylov registrations of many contexts at once
    (clr_solve_gn_batch): `workers` host threads, one HIP stream each.
    Returns one (v, report dict) per problem; raises on the first failure
    after every solve was attempted."""
    n = len(problems)
    opt = GNOptions()
    _check(lib().clr_gn_options_default(ctypes.byref(opt)))
    for k, val in options.items():
        if not hasattr(opt, k):
            raise TypeError(f"unknown option {k}")
        setattr(opt, k, val)
    vs = []
    for i, c in enumerate(problems):
        d = c.buf("gn_v", (3,) + c.nx)
        d.upload(np.zeros((3,) + c.nx, np.float32) if v0s is None else v0s[i])
        vs.append(d)
    ctxs = (ctypes.c_void_p * n)(*[c.ctx.value for c in problems])
    vp = (ctypes.c_void_p * n)(*[d.ptr for d in vs])
    reps = (GNReport * n)()
    status = (ctypes.c_int * n)()
    _check(lib().clr_solve_gn_batch(ctxs, vp, ctypes.c_int(n), ctypes.byref(opt), reps, status,
                                    ctypes.c_int(int(workers))))
    out = []
    for i in range(n):
        r = {name: getattr(reps[i], name) for name, _ in GNReport._fields_}
        r["reason"] = GN_REASONS.get(r["reason"], r["reason"])
        out.append((vs[i].download(), r))
    return out


def grad_noalloc(f, stream=None):
    """Spectral gradient of f (shape (n1, n2, n3)) through clr_grad_noalloc,
    the context-free entry point that allocates no device memory; returns
    the (3, n1, n2, n3) FP32 gradient.  The device buffers of the argument
    and result are allocated here, by the caller, as the entry point
    requires."""
    f = np.ascontiguousarray(f, dtype=np.float32)
    nx = (ctypes.c_int * 3)(*f.shape)
    d = DeviceArray.from_host(f)
    out = DeviceArray((3,) + f.shape)
    try:
        _check(lib().clr_grad_noalloc(nx, _vp(d.ptr), _vp(out.ptr), _vp(stream or 0)))
        return out.download()
    finally:
        d.free()
        out.free()


class DeviceArray:
    """Device allocation owned by Python (clr_device_malloc/free)."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().clr_device_malloc(ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 4))))
        self.ptr = p.value

    @classmethod
    def from_host(cls, a, dtype=np.float32):
        a = np.ascontiguousarray(a, dtype=dtype)
        d = cls(a.shape, dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes, (a.shape, self.shape)
        _check(lib().clr_memcpy_htod(_vp(self.ptr), a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes)))

    def download(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().clr_memcpy_dtoh(out.ctypes.data_as(ctypes.c_void_p), _vp(self.ptr), ctypes.c_size_t(self.nbytes)))
        return out

    def free(self):
        if self.ptr:
            lib().clr_device_free(_vp(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Claire:
    """One registration problem (GN, semi-Lagrangian, nc = 1) on the GPU."""

    def __init__(self, nx, nt=4, regnorm="h1sn", beta=(1e-2, 1e-4, 0.0), stream=None, precision="fp32",
                 rk_order=2, graphs=False, model="compressible", pde="transport", pde_solver="sl"):
        self.nx = tuple(int(n) for n in nx)
        self.nt = int(nt)
        self.nl = int(np.prod(self.nx))
        self.nh = self.nx[2] // 2 + 1
        b = list(beta) + [0.0] * (3 - len(beta))
        ctx = ctypes.c_void_p()
        _check(lib().clr_create(ctypes.byref(ctx), (ctypes.c_int * 3)(*self.nx), ctypes.c_int(self.nt),
                                ctypes.c_int(REG[regnorm]), (ctypes.c_double * 3)(*b[:3]), _vp(stream or 0)))
        self.ctx = ctx
        self._bufs = {}
        if precision != "fp32":
            self.set_precision(precision)
        if rk_order != 2:
            self.set_rk_order(rk_order)
        if graphs:
            self.set_matvec_graphs(True)
        if model != "compressible":
            self.set_reg_model(model)
        if pde != "transport":
            self.set_pde_type(pde)
        if pde_solver != "sl":
            self.set_pde_solver(pde_solver)

    def close(self):
        if getattr(self, "ctx", None):
            lib().clr_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scratch device buffers keyed by role -----------------------------
    def buf(self, key, shape, dtype=np.float32):
        d = self._bufs.get(key)
        if d is None or d.shape != tuple(shape) or d.dtype != np.dtype(dtype):
            d = DeviceArray(shape, dtype)
            self._bufs[key] = d
        return d

    def _in(self, key, a, shape):
        if isinstance(a, DeviceArray):
            return a
        d = self.buf(key, shape)
        d.upload(np.asarray(a, dtype=np.float32).reshape(shape))
        return d

    def sync(self):
        _check(lib().clr_synchronize(self.ctx))

    def device_bytes(self):
        n = ctypes.c_size_t()
        _check(lib().clr_device_bytes(self.ctx, ctypes.byref(n)))
        return n.value

    # -- problem data -------------------------------------------------------
    def set_velocity(self, v):
        d = self._in("v", v, (3,) + self.nx)
        _check(lib().clr_set_velocity(self.ctx, _vp(d.ptr)))

    def set_velocity_device(self, d):
        """set_velocity from a DeviceArray already in HBM."""
        _check(lib().clr_set_velocity(self.ctx, _vp(d.ptr)))

    def set_inc_velocity(self, vt):
        d = self._in("vt", vt, (3,) + self.nx)
        _check(lib().clr_set_inc_velocity(self.ctx, _vp(d.ptr)))

    def set_template_image(self, mT):
        d = self._in("mT", mT, self.nx)
        _check(lib().clr_set_template_image(self.ctx, _vp(d.ptr)))

    def set_reference_image(self, mR):
        d = self._in("mR", mR, self.nx)
        _check(lib().clr_set_reference_image(self.ctx, _vp(d.ptr)))

    # -- solves -----------------------------------------------------------------
    def solve_state(self):
        out = self.buf("m", (self.nt + 1,) + self.nx)
        _check(lib().clr_solve_state(self.ctx, _vp(out.ptr)))
        return out.download()

    def solve_adjoint(self):
        lam = self.buf("lam", self.nx)
        b = self.buf("b", (3,) + self.nx)
        _check(lib().clr_solve_adjoint(self.ctx, _vp(lam.ptr), _vp(b.ptr)))
        return lam.download(), b.download()

    def solve_inc_state(self, vt):
        d = self._in("vt", vt, (3,) + self.nx)
        out = self.buf("mt", self.nx)
        _check(lib().clr_solve_inc_state(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def solve_inc_adjoint(self, mt):
        d = self._in("mt_in", mt, self.nx)
        lt = self.buf("lt", self.nx)
        bt = self.buf("bt", (3,) + self.nx)
        _check(lib().clr_solve_inc_adjoint(self.ctx, _vp(d.ptr), _vp(lt.ptr), _vp(bt.ptr)))
        return lt.download(), bt.download()

    def prepare_hessian(self):
        _check(lib().clr_prepare_hessian(self.ctx))

    def hessian_matvec_device(self, out, vt, scale=True):
        _check(lib().clr_hessian_matvec(self.ctx, _vp(out.ptr), _vp(vt.ptr), ctypes.c_int(int(scale))))

    def hessian_matvec(self, vt, scale=True):
        d = self._in("hv_in", vt, (3,) + self.nx)
        out = self.buf("hv", (3,) + self.nx)
        self.hessian_matvec_device(out, d, scale)
        return out.download()

    def evaluate_objective(self, v=None):
        vp = _vp(self._in("v", v, (3,) + self.nx).ptr) if v is not None else _vp(0)
        J, D, R = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().clr_evaluate_objective(self.ctx, vp, ctypes.byref(J), ctypes.byref(D), ctypes.byref(R)))
        return J.value, D.value, R.value

    def evaluate_gradient(self, v=None):
        vp = _vp(self._in("v", v, (3,) + self.nx).ptr) if v is not None else _vp(0)
        g = self.buf("g", (3,) + self.nx)
        _check(lib().clr_evaluate_gradient(self.ctx, vp, _vp(g.ptr)))
        return g.download()

    def det_def_grad(self, inverse=False):
        """det(∇y) (or det(∇y⁻¹)) of the current velocity; returns
        (J, (min, max, mean))."""
        J = self.buf("jac", self.nx)
        st = (ctypes.c_double * 3)()
        _check(lib().clr_det_def_grad(self.ctx, ctypes.c_int(int(inverse)), _vp(J.ptr), st))
        return J.download(), tuple(st)

    def check_bounds(self, v, inverse=False, bound=0.25):
        """CLAIREBase::CheckBounds: (bound reached, (min, max, mean))."""
        d = self._in("v", v, (3,) + self.nx)
        reached = ctypes.c_int()
        st = (ctypes.c_double * 3)()
        _check(lib().clr_check_bounds(self.ctx, _vp(d.ptr), ctypes.c_int(int(inverse)), ctypes.c_double(bound),
                                      ctypes.byref(reached), st))
        return bool(reached.value), tuple(st)

    def set_precision(self, precision):
        """"fp32" (the reference's path) or "mixed" (16-bit block-floating-point
        storage of the Newton-invariant gradient fields in the Hessian matvec)."""
        _check(lib().clr_set_precision(self.ctx, ctypes.c_int({"fp32": 0, "mixed": 1}[precision])))

    def set_matvec_type(self, kind):
        """"default" (HessMatVec), "precond" (PrecondHessMatVec) or
        "precond_sym" (PrecondHessMatVecSym)."""
        _check(lib().clr_set_matvec_type(self.ctx, ctypes.c_int({"default": 0, "precond": 1, "precond_sym": 2}[kind])))

    def set_reg_model(self, model):
        """"compressible" (CLAIRE), "stokes" (CLAIREStokes) or
        "relaxed_stokes" (CLAIREDivReg)."""
        _check(lib().clr_set_reg_model(self.ctx, ctypes.c_int(MODELS[model])))

    def apply_projection(self, v):
        """The model's ApplyProjection of a vector field (3, n1, n2, n3)."""
        d = self._in("proj", v, (3,) + self.nx)
        _check(lib().clr_apply_projection(self.ctx, _vp(d.ptr)))
        return d.download()

    def set_matvec_graphs(self, on):
        """Replay Hessian matvecs as captured HIP graphs (same results)."""
        _check(lib().clr_set_matvec_graphs(self.ctx, ctypes.c_int(int(bool(on)))))

    def matvec_graph_count(self):
        n = ctypes.c_int()
        _check(lib().clr_matvec_graph_count(self.ctx, ctypes.byref(n)))
        return n.value

    def set_rk_order(self, order):
        """Characteristic scheme of every later trajectory: 2 or 4 (the
        reference's -rkorder option)."""
        _check(lib().clr_set_rk_order(self.ctx, ctypes.c_int(int(order))))

    def set_pde_type(self, pde):
        """Constraint of every later state solve: "transport" (default) or
        "continuity" (the reference's m_PDESolver.pdetype)."""
        if pde not in PDE_TYPES:
            raise ValueError(f"pde type must be one of {sorted(PDE_TYPES)}")
        _check(lib().clr_set_pde_type(self.ctx, ctypes.c_int(PDE_TYPES[pde])))

    def set_pde_solver(self, solver):
        """Time integrator of every later transport solve: "sl" (default) or
        "rk2" (the reference's -pdesolver option)."""
        if solver not in PDE_SOLVERS:
            raise ValueError(f"pde solver must be one of {sorted(PDE_SOLVERS)}")
        _check(lib().clr_set_pde_solver(self.ctx, ctypes.c_int(PDE_SOLVERS[solver])))

    def solve_gn(self, v0=None, hist_len=128, **options):
        """Gauss–Newton–Krylov registration (clr_solve_gn) from v0 (zero by
        default).  Options override the reference defaults (maxit, gttol,
        gatol, krylov_maxit, krylov_rtol, forcing, precond, linesearch,
        monitor, ls_sigma, ls_stepinit, krylov_block).
        Returns (v, report dict, history rows (iter, J, D, R, ‖g‖, ‖g‖/‖g₀‖,
        pcg_its, step))."""
        opt = GNOptions()
        _check(lib().clr_gn_options_default(ctypes.byref(opt)))
        for k, val in options.items():
            if not hasattr(opt, k):
                raise TypeError(f"unknown option {k}")
            setattr(opt, k, val)
        v = self.buf("gn_v", (3,) + self.nx)
        v.upload(np.zeros((3,) + self.nx, np.float32) if v0 is None else v0)
        rep = GNReport()
        hist = np.zeros((hist_len, 8), dtype=np.float64)
        _check(lib().clr_solve_gn(self.ctx, _vp(v.ptr), ctypes.byref(opt), ctypes.byref(rep),
                                  hist.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(hist_len)))
        r = {name: getattr(rep, name) for name, _ in GNReport._fields_}
        r["reason"] = GN_REASONS.get(r["reason"], r["reason"])
        return v.download(), r, hist[: min(r["iterations"] + 1, hist_len)]

    # -- regularization -----------------------------------------------------
    def apply_inv_reg(self, x, applysqrt=False):
        d = self._in("x", x, (3,) + self.nx)
        out = self.buf("ax", (3,) + self.nx)
        _check(lib().clr_apply_inv_reg(self.ctx, _vp(d.ptr), _vp(out.ptr), ctypes.c_int(int(applysqrt))))
        return out.download()

    def reg_apply(self, v):
        d = self._in("x", v, (3,) + self.nx)
        out = self.buf("ax", (3,) + self.nx)
        _check(lib().clr_reg_apply(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def reg_functional(self, v):
        d = self._in("x", v, (3,) + self.nx)
        r = ctypes.c_double()
        _check(lib().clr_reg_functional(self.ctx, _vp(d.ptr), ctypes.byref(r)))
        return r.value

    def reg_eigvals_inv(self):
        lo, hi = ctypes.c_double(), ctypes.c_double()
        _check(lib().clr_reg_eigvals_inv(self.ctx, ctypes.byref(lo), ctypes.byref(hi)))
        return lo.value, hi.value

    # -- semi-Lagrangian building blocks --------------------------------------
    def trajectory(self, flag=STATE):
        out = self.buf("X", self.nx + (3,))
        _check(lib().clr_trajectory(self.ctx, ctypes.c_int(flag), _vp(out.ptr)))
        return out.download()

    def interp_scalar(self, f, flag=STATE):
        d = self._in("f", f, self.nx)
        out = self.buf("fo", self.nx)
        _check(lib().clr_interp_scalar(self.ctx, ctypes.c_int(flag), _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def interp_vector(self, v, flag=STATE):
        d = self._in("f3", v, (3,) + self.nx)
        out = self.buf("fo3", (3,) + self.nx)
        _check(lib().clr_interp_vector(self.ctx, ctypes.c_int(flag), _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def interp_points(self, f, X):
        d = self._in("f", f, self.nx)
        x = self._in("Xq", X, self.nx + (3,))
        out = self.buf("fo", self.nx)
        _check(lib().clr_interp_points(self.ctx, _vp(d.ptr), _vp(x.ptr), _vp(out.ptr)))
        return out.download()

    # -- spectral ---------------------------------------------------------------
    def grad(self, f):
        d = self._in("f", f, self.nx)
        out = self.buf("g3", (3,) + self.nx)
        _check(lib().clr_grad(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def div(self, v):
        d = self._in("f3", v, (3,) + self.nx)
        out = self.buf("fo", self.nx)
        _check(lib().clr_div(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def fft_r2c(self, f):
        d = self._in("f", f, self.nx)
        out = self.buf("spec", (self.nx[0], self.nx[1], self.nh), np.complex64)
        _check(lib().clr_fft_r2c(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def fft_c2r(self, fh):
        shape = (self.nx[0], self.nx[1], self.nh)
        d = self.buf("spec_in", shape, np.complex64)
        d.upload(np.asarray(fh, dtype=np.complex64).reshape(shape))
        out = self.buf("fo", self.nx)
        _check(lib().clr_fft_c2r(self.ctx, _vp(d.ptr), _vp(out.ptr)))
        return out.download()

    def dot(self, a, b):
        da = DeviceArray.from_host(a)
        db = DeviceArray.from_host(b)
        r = ctypes.c_double()
        _check(lib().clr_dot(self.ctx, _vp(da.ptr), _vp(db.ptr), ctypes.c_int64(da.nbytes // 4), ctypes.byref(r)))
        return r.value

    def norm2(self, a):
        da = DeviceArray.from_host(a)
        r = ctypes.c_double()
        _check(lib().clr_norm2(self.ctx, _vp(da.ptr), ctypes.c_int64(da.nbytes // 4), ctypes.byref(r)))
        return r.value

    # -- kernel timing ---------------------------------------------------------
    def profile_enable(self, on=True):
        _check(lib().clr_profile_enable(self.ctx, ctypes.c_int(int(on))))

    def profile_reset(self):
        _check(lib().clr_profile_reset(self.ctx))

    def profile(self):
        """{kernel class: (ms, count, algorithmic bytes)} accumulated so far."""
        out = {}
        i = 0
        name = ctypes.create_string_buffer(128)
        while True:
            ms, cnt, by = ctypes.c_double(), ctypes.c_long(), ctypes.c_double()
            rc = lib().clr_profile_get(self.ctx, ctypes.c_int(i), name, ctypes.c_int(128), ctypes.byref(ms),
                                       ctypes.byref(cnt), ctypes.byref(by))
            if rc != 0:
                break
            out[name.value.decode()] = (ms.value, cnt.value, by.value)
            i += 1
        return out
