"""Pins for the oracle's interpolation, trajectory and transport solves (CPU).

Independent checks: the textbook Lagrange basis Π_{l≠k}(t−l)/(k−l)
evaluated in pure Python, polynomial exactness, integer shifts, O(h⁴)
convergence, constant-velocity closed forms, the v = 0 branches of the
reference (src/CLAIRE.cpp:1909-1920, 2249-2312, 3367-3421), linearity and
the gradient Taylor test (src/OptimizationProblem.cpp:157-239).
"""
import functools

import numpy as np
import pytest

from claire_amd import synth
from conftest import relerr


def _grid(shape):
    h = [2 * np.pi / n for n in shape]
    return np.meshgrid(*[np.arange(n) * h[a] for a, n in enumerate(shape)], indexing="ij")


def _lagrange_textbook(t):
    nodes = [0.0, 1.0, 2.0, 3.0]
    out = []
    for k in range(4):
        v = 1.0
        for l in range(4):
            if l != k:
                v *= (t - nodes[l]) / (nodes[k] - nodes[l])
        out.append(v)
    return out


def _interp_textbook(f, q):
    """Tricubic interpolation at normalized point q (periodic), from the
    textbook Lagrange basis on the nodes floor(qN)−1 … floor(qN)+2."""
    n = f.shape
    idx, w = [], []
    for a in range(3):
        u = (q[a] % 1.0) * n[a]
        i0 = int(np.floor(u)) - 1
        t = u - i0
        idx.append([(i0 + k) % n[a] for k in range(4)])
        w.append(_lagrange_textbook(t))
    val = 0.0
    for a in range(4):
        for b in range(4):
            for c in range(4):
                val += w[0][a] * w[1][b] * w[2][c] * f[idx[0][a], idx[1][b], idx[2][c]]
    return val


def test_interp_matches_textbook_lagrange(oracle):
    rng = np.random.default_rng(0)
    shape = (8, 12, 16)
    f = rng.standard_normal(shape)
    q = rng.uniform(-0.3, 1.3, (40, 3))
    got = oracle.interp(f, q)
    for i in range(q.shape[0]):
        assert abs(got[i] - _interp_textbook(f, q[i])) < 1e-12


def test_interp_partition_of_unity_and_nodes(oracle):
    rng = np.random.default_rng(1)
    shape = (16, 16, 16)
    q = rng.uniform(-0.05, 1.05, (200, 3))
    assert np.abs(oracle.interp(np.ones(shape), q) - 1.0).max() < 1e-13
    f = rng.standard_normal(shape)
    nodes = np.stack(np.meshgrid(*[np.arange(n) / n for n in shape], indexing="ij"), -1)
    assert np.abs(oracle.interp(f, nodes) - f).max() < 1e-12


def test_interp_integer_shift(oracle):
    rng = np.random.default_rng(2)
    shape = (16, 8, 32)
    f = rng.standard_normal(shape)
    g = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1).astype(np.float64)
    shift = np.array([3, -2, 5])
    q = (g + shift) / np.array(shape)
    expect = np.roll(f, shift=tuple(-shift), axis=(0, 1, 2))
    assert np.abs(oracle.interp(f, q) - expect).max() < 1e-12


def test_interp_exact_for_cubics(oracle):
    shape = (32, 32, 32)
    rng = np.random.default_rng(3)
    i = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), 0).astype(np.float64)
    c = rng.standard_normal((3, 4))
    f = sum(np.polyval(c[a], i[a] / 10.0) for a in range(3)) + (i[0] / 10) ** 3 * (i[1] / 10) * (i[2] / 10) ** 2
    # query points whose stencils stay inside the grid (no periodic wrap)
    u = rng.uniform(3.0, 27.0, (300, 3))
    exact = sum(np.polyval(c[a], u[:, a] / 10.0) for a in range(3)) + (u[:, 0] / 10) ** 3 * (u[:, 1] / 10) * (u[:, 2] / 10) ** 2
    got = oracle.interp(f, u / 32.0)
    assert np.abs(got - exact).max() < 1e-10


def test_interp_fourth_order_convergence(oracle):
    rng = np.random.default_rng(4)
    q = rng.uniform(0, 1, (500, 3))
    errs = []
    for n in (16, 32, 64):
        x1, x2, x3 = _grid((n, n, n))
        f = np.sin(x1) * np.cos(2 * x2) * np.sin(x3 + 0.3)
        X = 2 * np.pi * q
        exact = np.sin(X[:, 0]) * np.cos(2 * X[:, 1]) * np.sin(X[:, 2] + 0.3)
        errs.append(np.abs(oracle.interp(f, q) - exact).max())
    assert errs[0] / errs[1] > 12 and errs[1] / errs[2] > 12


@pytest.mark.parametrize("sigma", [1, -1])
def test_trajectory_constant_velocity(oracle, sigma):
    shape = (8, 16, 8)
    nt = 4
    c = np.array([0.3, -0.7, 1.1])
    v = np.stack([np.full(shape, c[a]) for a in range(3)])
    X = oracle.trajectory(v, nt, sigma)
    x = np.stack(_grid(shape), -1)
    expect = (x - sigma * c / nt) / (2 * np.pi)
    assert np.abs(X - expect).max() < 1e-14


def _coupled_velocity(y1, y2):
    """A smooth, periodic, nonlinear velocity independent of x3."""
    return np.stack([1.5 * np.sin(y2) + 0.3 * np.cos(2 * y1), -1.2 * np.sin(y1) * np.cos(y2),
                     0.5 * np.cos(y1 + y2)])


@functools.lru_cache(maxsize=None)
def _exact_foot_points(shape, ht, sigma, substeps=100):
    """Foot point of dy/dt = −σ·v(y), y(0) = x, after time ht, by a fine RK4
    on the analytic velocity; returned normalized by 2π like the trajectory."""
    y = np.stack(_grid(shape)).astype(np.float64)
    h = ht / substeps
    f = lambda z: -sigma * _coupled_velocity(z[0], z[1])
    for _ in range(substeps):
        k1 = f(y)
        k2 = f(y + 0.5 * h * k1)
        k3 = f(y + 0.5 * h * k2)
        k4 = f(y + h * k3)
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return np.moveaxis(y, 0, -1) / (2 * np.pi)


@pytest.mark.parametrize("order,rate", [(2, 8.0), (4, 32.0)])
def test_trajectory_local_order(oracle, order, rate):
    """One characteristic step has local error O(ht^3) for RK2 and O(ht^5)
    for RK4 (src/SemiLagrangian.cpp:220-295 and :297-459): halving ht divides
    the error by 8 and by 32.  A wrong stage weight or stage point in the RK4
    restatement drops it to a lower order."""
    shape = (128, 128, 8)
    x1, x2, _ = _grid(shape)
    v = _coupled_velocity(x1, x2)
    for sigma in (1, -1):
        errs = []
        for nt in (1, 2, 4):
            with oracle.rk_order(order):
                X = oracle.trajectory(v, nt, sigma)
            errs.append(np.abs(X - _exact_foot_points(shape, 1.0 / nt, sigma)).max())
        ratios = [errs[i] / errs[i + 1] for i in range(len(errs) - 1)]
        assert all(0.8 * rate < r < 1.3 * rate for r in ratios), (order, sigma, errs, ratios)


@pytest.mark.parametrize("sigma", [1, -1])
def test_rk4_trajectory_constant_velocity_and_state(oracle, sigma):
    """RK4 is exact for a constant velocity, and whole-cell shifts transport
    the template exactly, as with RK2."""
    shape = (8, 16, 8)
    nt = 4
    c = np.array([0.3, -0.7, 1.1])
    v = np.stack([np.full(shape, c[a]) for a in range(3)])
    with oracle.rk_order(4):
        X = oracle.trajectory(v, nt, sigma)
    x = np.stack(_grid(shape), -1)
    assert np.abs(X - (x - sigma * c / nt) / (2 * np.pi)).max() < 1e-14
    shape = (16, 16, 16)
    h = 2 * np.pi / 16
    shift = np.array([1, 2, -1])
    v = np.stack([np.full(shape, shift[a] * h * nt) for a in range(3)])
    mT = synth.gaussian_blobs(shape, count=5, seed=3).astype(np.float64)
    with oracle.rk_order(4):
        m = oracle.solve_state(v, mT, nt)
    for j in range(nt + 1):
        assert np.abs(m[j] - np.roll(mT, shift=tuple(j * shift), axis=(0, 1, 2))).max() < 1e-12


def test_rk_order_rejects_other_orders(oracle):
    with pytest.raises(ValueError):
        oracle.set_rk_order(3)


def test_state_translation_by_whole_cells(oracle):
    shape = (16, 16, 16)
    nt = 4
    h = 2 * np.pi / 16
    shift = np.array([1, 2, -1])            # cells per time step
    v = np.stack([np.full(shape, shift[a] * h * nt) for a in range(3)])
    mT = synth.gaussian_blobs(shape, count=5, seed=3).astype(np.float64)
    m = oracle.solve_state(v, mT, nt)
    for j in range(nt + 1):
        expect = np.roll(mT, shift=tuple(j * shift), axis=(0, 1, 2))
        assert np.abs(m[j] - expect).max() < 1e-12


def test_zero_velocity_closed_forms(oracle):
    shape = (16, 16, 16)
    nt = 4
    hd = (2 * np.pi / 16) ** 3
    v = np.zeros((3,) + shape)
    mT = synth.gaussian_blobs(shape, count=5, seed=4).astype(np.float64)
    mR = synth.gaussian_blobs(shape, count=5, seed=5).astype(np.float64)
    vt = synth.bandlimited_velocity(shape, kmax=3, cfl=1.0, seed=6).astype(np.float64)
    m = oracle.solve_state(v, mT, nt)
    assert all(np.array_equal(m[j], mT) for j in range(nt + 1))
    gm = oracle.grad(mT)
    lam, b = oracle.solve_adjoint(v, m, mR, nt)
    assert np.abs(lam - (mR - mT)).max() < 1e-14
    assert relerr(b, hd * (mR - mT) * gm) < 1e-13
    mt = oracle.solve_inc_state(v, m, vt, nt)
    gdotv = np.sum(gm * vt, axis=0)
    assert relerr(mt, -gdotv) < 1e-12
    lt, bt = oracle.solve_inc_adjoint(v, m, mt, nt)
    assert relerr(bt, hd * gdotv * gm) < 1e-12
    for norm in ("h1sn", "h2"):
        H = oracle.hessian_matvec(v, m, vt, nt, norm, (1e-2, 1e-4, 0.0))
        expect = oracle.reg_apply(vt, norm, (1e-2, 1e-4, 0.0)) + hd * gdotv * gm
        assert relerr(H, expect) < 1e-12


@pytest.fixture(scope="module")
def problem():
    shape = (16, 16, 16)
    nt = 4
    v = synth.bandlimited_velocity(shape, kmax=3, cfl=1.5, nt=nt, seed=7).astype(np.float64)
    mT = synth.gaussian_blobs(shape, count=6, seed=8).astype(np.float64)
    mR = synth.gaussian_blobs(shape, count=6, seed=9).astype(np.float64)
    return shape, nt, v, mT, mR


def test_hessian_linearity(oracle, problem):
    shape, nt, v, mT, mR = problem
    m = oracle.solve_state(v, mT, nt)
    a = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, seed=10).astype(np.float64)
    b = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, seed=11).astype(np.float64)
    H = lambda x: oracle.hessian_matvec(v, m, x, nt, "h1sn", (1e-2, 1e-4, 0.0))
    assert relerr(H(2.0 * a - 3.0 * b), 2.0 * H(a) - 3.0 * H(b)) < 1e-12


def test_hessian_symmetry_converges(oracle):
    """The GN Hessian of the SL scheme (optimize-then-discretize) is symmetric
    up to discretization error (HessianSymmetryCheck, src/OptimizationProblem.cpp:
    439-491): the normalized asymmetry must fall at least 8× per grid doubling
    and the Hessian must be positive on a test direction.  A wrong sign, time
    index or adjoint term would leave an O(1) asymmetry."""
    nt = 4
    beta = (1e-2, 1e-4, 0.0)
    asym = []
    for n in (16, 32):
        shape = (n, n, n)
        v = synth.bandlimited_velocity(shape, kmax=3, cfl=1.5, nt=nt, seed=7).astype(np.float64)
        mT = synth.benchmark_template(shape, dtype=np.float64)
        m = oracle.solve_state(v, mT, nt)
        a = synth.bandlimited_velocity(shape, kmax=3, cfl=1.0, seed=12).astype(np.float64)
        b = synth.bandlimited_velocity(shape, kmax=3, cfl=1.0, seed=13).astype(np.float64)
        Ha = oracle.hessian_matvec(v, m, a, nt, "h1sn", beta)
        Hb = oracle.hessian_matvec(v, m, b, nt, "h1sn", beta)
        scale = 0.5 * (np.linalg.norm(Ha) * np.linalg.norm(b) + np.linalg.norm(a) * np.linalg.norm(Hb))
        asym.append(abs(float(np.sum(Ha * b)) - float(np.sum(a * Hb))) / scale)
        assert float(np.sum(Ha * a)) > 0
    assert asym[0] < 1e-3 and asym[1] < asym[0] / 8


def test_gradient_taylor(oracle, problem):
    """J(v + εw) − J(v) − ε⟨g, w⟩ must shrink like ε² until the SL
    discretization error dominates."""
    shape, nt, v, mT, mR = problem
    beta = (1e-2, 1e-4, 0.0)
    w = synth.bandlimited_velocity(shape, kmax=3, cfl=1.0, seed=14).astype(np.float64)
    J0 = oracle.objective(v, mT, mR, nt, "h1sn", beta)[0]
    g = oracle.gradient(v, mT, mR, nt, "h1sn", beta)
    gw = float(np.sum(g * w))
    for eps in (1e-1, 3e-2):
        dJ = oracle.objective(v + eps * w, mT, mR, nt, "h1sn", beta)[0] - J0
        assert abs(dJ - eps * gw) < 0.05 * abs(dJ)


def test_fp32_twin_tracks_fp64(oracle, problem):
    shape, nt, v, mT, mR = problem
    beta = (1e-2, 1e-4, 0.0)
    m64 = oracle.solve_state(v, mT, nt)
    m32 = oracle.solve_state(v, mT, nt, prec="f32")
    assert relerr(m32, m64) < 1e-5
    vt = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, seed=15).astype(np.float64)
    h64 = oracle.hessian_matvec(v, m64, vt, nt, "h1sn", beta)
    h32 = oracle.hessian_matvec(v, m32, vt, nt, "h1sn", beta, prec="f32")
    assert relerr(h32, h64) < 1e-5
    g64 = oracle.gradient(v, mT, mR, nt, "h1sn", beta)
    g32 = oracle.gradient(v, mT, mR, nt, "h1sn", beta, prec="f32")
    assert relerr(g32, g64) < 1e-4


def test_hessian_is_derivative_of_gradient_at_zero_residual(oracle):
    """With mR = m^nt(v) the adjoint vanishes and the GN Hessian is the exact
    Hessian, so Hṽ must equal (g(v+εṽ) − g(v−εṽ))/2ε up to the SL
    discretization error, which shrinks ~8× per grid refinement (measured
    2.5e-2 at 16³, 3.0e-3 at 32³, 4.4e-4 at 64³).  A dropped or mis-signed
    term of the incremental solves gives an O(1) mismatch (the
    regularization part alone is 0.8 away)."""
    beta = (1e-2, 1e-4, 0.0)
    nt = 4
    errs = []
    for n in (16, 32):
        shape = (n, n, n)
        x1, x2, x3 = _grid(shape)
        mT = (np.sin(x1) ** 2 + np.sin(x2) ** 2 + np.sin(x3) ** 2) / 3
        v = synth.bandlimited_velocity(shape, kmax=2, cfl=0.5, nt=nt, seed=3).astype(np.float64)
        m = oracle.solve_state(v, mT, nt)
        mR = m[nt].copy()
        vt = synth.bandlimited_velocity(shape, kmax=2, cfl=0.5, nt=nt, seed=4).astype(np.float64)
        H = oracle.hessian_matvec(v, m, vt, nt, "h1sn", beta)
        eps = 1e-3
        fd = (oracle.gradient(v + eps * vt, mT, mR, nt, "h1sn", beta)
              - oracle.gradient(v - eps * vt, mT, mR, nt, "h1sn", beta)) / (2 * eps)
        errs.append(relerr(H, fd))
    assert errs[0] < 5e-2 and errs[1] < 5e-3 and errs[1] < errs[0] / 4


def test_gauss_newton_krylov_invariants(oracle):
    """GN–Krylov on the reference's synthetic problem 4 at 16³: the objective
    decreases monotonically, ‖g‖ falls below gttol·‖g₀‖ (the stop reason),
    and the FP32 twin takes the same Newton and PCG iterations as FP64."""
    shape, nt, beta = (16, 16, 16), 4, (1e-2, 1e-4, 0.0)
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    mR = synth.rescale(oracle.solve_state(vstar, mT, nt)[nt])
    v, rep, hist = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, nt, "h1sn", beta, maxit=10)
    assert rep["reason"] == "gttol" and rep["gnorm"] < 5e-2 * rep["gnorm0"]
    assert np.all(np.diff(hist[:, 1]) < 0)  # J_k decreasing
    v32, rep32, hist32 = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, nt, "h1sn", beta, maxit=10, prec="f32")
    assert rep32["iterations"] == rep["iterations"] and rep32["pcg_iterations"] == rep["pcg_iterations"]
    assert abs(rep32["J"] - rep["J"]) < 1e-5 * rep["J"]
    assert relerr(v32, v) < 1e-4


def _flow_determinant(shape, A, sign, steps=400):
    """exp(sign·∫₀¹ (∇·v)(y(s)) ds) along dy/ds = −v(y), y(0) = x, for the
    analytic v = A·(sin x1 cos x2, sin x2 cos x3, sin x3 cos x1), by RK4."""
    return _flow(shape, A, sign, steps)[0]


def _flow(shape, A, sign, steps=400):
    """(exp(sign·∫₀¹ (∇·v)(y(s)) ds), y(1)) along dy/ds = −v(y), y(0) = x,
    for v = A·(sin x1 cos x2, sin x2 cos x3, sin x3 cos x1), by RK4."""
    x = np.stack(np.meshgrid(*[np.arange(n) * 2 * np.pi / n for n in shape], indexing="ij")).reshape(3, -1)

    def vel(y):
        return A * np.stack([np.sin(y[0]) * np.cos(y[1]), np.sin(y[1]) * np.cos(y[2]), np.sin(y[2]) * np.cos(y[0])])

    def dv(y):
        return A * (np.cos(y[0]) * np.cos(y[1]) + np.cos(y[1]) * np.cos(y[2]) + np.cos(y[2]) * np.cos(y[0]))

    h, y, integral = 1.0 / steps, x.copy(), 0.0
    for _ in range(steps):
        k1 = -vel(y)
        k2 = -vel(y + h / 2 * k1)
        k3 = -vel(y + h / 2 * k2)
        k4 = -vel(y + h * k3)
        integral = integral + h / 6 * (dv(y) + 2 * dv(y + h / 2 * k1) + 2 * dv(y + h / 2 * k2) + dv(y + h * k3))
        y = y + h / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
    return np.exp(sign * integral).reshape(shape), y.reshape((3,) + tuple(shape))


@pytest.mark.parametrize("inverse", [False, True])
def test_det_def_grad_converges_to_flow_determinant(oracle, inverse):
    """det(∇y) by SL (src/DeformationFields.cpp:636-740) converges at second
    order in ht to the exact determinant of the characteristic flow,
    exp(±∫∇·v) along the back-traced path; v = 0 gives J = 1 up to the
    round-off of the foot points (x/2π is not exactly i/N)."""
    shape, A = (32, 32, 32), 0.5
    x1, x2, x3 = _grid(shape)
    v = A * np.stack([np.sin(x1) * np.cos(x2), np.sin(x2) * np.cos(x3), np.sin(x3) * np.cos(x1)])
    exact = _flow_determinant(shape, A, -1.0 if inverse else 1.0)
    errs = [np.abs(oracle.det_def_grad(v, nt, inverse)[0] - exact).max() for nt in (4, 8, 16)]
    assert errs[2] < 2e-2 and errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3
    J0, st = oracle.det_def_grad(np.zeros((3,) + shape), 4, inverse)
    assert np.abs(J0 - 1.0).max() < 1e-12 and np.allclose(st, 1.0, rtol=0, atol=1e-12)


def test_registration_golden_reproduces(oracle):
    """The stored FP64 registration results (tests/golden) come from the
    current oracle and inputs: the quick 64³ case is recomputed here."""
    import importlib.util
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("mk", os.path.join(here, "make_registration_golden.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    with open(os.path.join(here, "registration_phantom.json")) as f:
        golden = json.load(f)
    case, ref = golden["phantom64_h2"]["case"], golden["phantom64_h2"]["oracle_fp64"]
    mT, mR = mk.inputs(case)
    v, rep, _ = oracle.gn_solve(np.zeros((3,) + mT.shape), mT, mR, case["nt"], case["regnorm"],
                                (case["beta"], 0.0, 0.0), maxit=case["maxit"], gttol=case["gttol"])
    assert rep["reason"] == ref["reason"] and rep["iterations"] == ref["iterations"]
    for k in ("J", "D", "R", "gnorm", "gnorm0"):
        assert abs(rep[k] - ref[k]) <= 1e-9 * abs(ref[k]), k


@pytest.mark.parametrize("norm", ["h1", "h1sn", "h2sn"])
def test_preconditioned_matvecs_are_compositions(oracle, norm):
    """PrecondHessMatVec = (βA)⁻¹H and PrecondHessMatVecSym =
    (βA)^{-1/2}H(βA)^{-1/2} (src/CLAIRE.cpp:1380-1540), because (βA)⁻¹R'' is
    hd·I on the modes the operators act on: checked against the default
    matvec and the inverse regularization, on zero-mean fields without
    Nyquist content."""
    shape, nt, beta = (16, 16, 16), 4, (1e-2, 1e-4, 0.0)
    v = synth.bandlimited_velocity(shape, kmax=3, cfl=1.0, nt=nt, seed=21).astype(np.float64)
    mT = synth.gaussian_blobs(shape, count=5, seed=22).astype(np.float64)
    # FP64 band-limited field: no content on the Nyquist indices, where the
    # seminorms' multiplier vanishes but their inverse does not
    vt = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, nt=nt, seed=23, dtype=np.float64)
    vt -= vt.mean(axis=(1, 2, 3), keepdims=True)
    m = oracle.solve_state(v, mT, nt)
    H = oracle.hessian_matvec(v, m, vt, nt, norm, beta)
    with oracle.matvec_type("precond"):
        P = oracle.hessian_matvec(v, m, vt, nt, norm, beta)
    assert relerr(P, oracle.reg_inverse(H, norm, beta)) < 1e-12
    half = oracle.reg_inverse(vt, norm, beta, applysqrt=True)
    Hh = oracle.hessian_matvec(v, m, half, nt, norm, beta)
    with oracle.matvec_type("precond_sym"):
        S = oracle.hessian_matvec(v, m, vt, nt, norm, beta)
    assert relerr(S, oracle.reg_inverse(Hh, norm, beta, applysqrt=True)) < 1e-12


def test_gauss_newton_with_the_symmetric_preconditioned_operator(oracle):
    """PRECONDMATVECSYM in the GN driver: (βA)^{-1/2} on the right-hand side
    and the solution; the Newton steps decrease J and reach gttol."""
    shape, nt, beta = (16, 16, 16), 4, (1e-2, 1e-4, 0.0)
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    mR = synth.rescale(oracle.solve_state(vstar, mT, nt)[nt])
    with oracle.matvec_type("precond_sym"):
        v, rep, hist = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, nt, "h1sn", beta, maxit=20, gttol=5e-2)
    assert rep["reason"] == "gttol" and rep["J"] < rep["J0"]
    assert np.all(np.diff(hist[:, 1]) <= 0)
    with oracle.matvec_type("precond"):
        with pytest.raises(RuntimeError, match="nonsymmetric"):
            oracle.gn_solve(np.zeros((3,) + shape), mT, mR, nt, "h1sn", beta, maxit=2)


def _continuity_initial(x):
    return 1.0 + 0.4 * np.sin(x[0] + x[1]) * np.cos(x[2])


def test_continuity_state_converges_to_characteristic_solution(oracle):
    """Continuity-equation state (src/CLAIRE.cpp:2735-2854) against the
    exact solution of ∂t m + ∇·(m v) = 0 by characteristics,
    m(x, 1) = m0(y(1))·exp(−∫₀¹ (∇·v)(y(s)) ds), y the back-traced path:
    second order in ht until the tricubic interpolation error of the 32³
    grid (≈1e-3, accumulated over the nt steps) takes over; a dropped or
    sign-flipped source term is O(1)."""
    shape, A = (32, 32, 32), 0.5
    x = np.stack(_grid(shape))
    v = A * np.stack([np.sin(x[0]) * np.cos(x[1]), np.sin(x[1]) * np.cos(x[2]), np.sin(x[2]) * np.cos(x[0])])
    decay, foot = _flow(shape, A, -1.0)
    exact = _continuity_initial(foot) * decay
    m0 = _continuity_initial(x)
    with oracle.pde_type("continuity"):
        errs = [np.abs(oracle.solve_state(v, m0, nt)[nt] - exact).max() for nt in (4, 8, 16)]
    assert errs[1] < errs[0] / 3 and errs[2] < errs[1] and errs[2] < 2e-3
    transport = np.abs(oracle.solve_state(v, m0, 16)[16] - exact).max()
    assert transport > 10 * errs[2]  # the transport solve misses the source term


def test_continuity_state_conserves_mass_and_reduces_to_transport(oracle):
    """∫m(·, t) is invariant under the continuity equation (to the
    discretization error, which falls with nt); for ∇·v = 0 the source terms
    vanish and the solve equals the transport solve (src/CLAIRE.cpp:2153-2163)
    to round-off; m0 ≡ 1 gives det(∇y⁻¹), whose SL update has the same form
    (src/DeformationFields.cpp:714-721); v = 0 copies m0."""
    shape = (32, 32, 32)
    x = np.stack(_grid(shape))
    m0 = _continuity_initial(x)
    v = 0.5 * np.stack([np.sin(x[0]) * np.cos(x[1]), np.sin(x[1]) * np.cos(x[2]), np.sin(x[2]) * np.cos(x[0])])
    with oracle.pde_type("continuity"):
        drift = [abs(oracle.solve_state(v, m0, nt)[nt].mean() / m0.mean() - 1.0) for nt in (4, 16)]
        vfree = 0.7 * np.stack([np.sin(x[1]), np.cos(x[2]), np.sin(x[0] + 0.3)])
        cont = oracle.solve_state(vfree, m0, 4)
        ones = oracle.solve_state(v, np.ones(shape), 4)[4]
        still = oracle.solve_state(np.zeros((3,) + shape), m0, 4)
    assert drift[1] < 1e-3 and drift[1] < drift[0]
    assert np.abs(cont - oracle.solve_state(vfree, m0, 4)).max() < 1e-12
    assert np.abs(ones - oracle.det_def_grad(v, 4, True)[0]).max() < 1e-12
    assert np.array_equal(still, np.broadcast_to(m0, still.shape))
    with pytest.raises(KeyError):
        oracle.set_pde_type("burgers")
