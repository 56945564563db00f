"""GPU parity of the explicit RK2 transport solvers (-pdesolver rk2,
clr_set_pde_solver) against the oracle run with the same integrator.

Bars as in test_gpu_parity.py: plain FP32 within 1e-5 relative L2 of the FP64
oracle, or no less accurate than the oracle's FP32 twin (the reference's
single-precision path) where FP32 conditioning puts 1e-5 out of reach.
"""
import numpy as np
import pytest

from claire_amd import Claire, synth
from conftest import relerr
from test_gpu_parity import SHAPES, assert_fp32_parity

pytestmark = pytest.mark.gpu

BETA = (1e-2, 1e-4, 0.0)
NT = 4


def _data(shape, seed):
    # RK2 is conditionally stable: CFL 0.5 on a band-limited velocity
    v = synth.bandlimited_velocity(shape, kmax=min(4, min(shape) // 4), cfl=0.5, nt=NT, seed=seed)
    mT = synth.gaussian_blobs(shape, count=8, seed=seed + 1)
    mR = synth.gaussian_blobs(shape, count=8, seed=seed + 2)
    vt = synth.bandlimited_velocity(shape, kmax=4, cfl=0.5, nt=NT, seed=seed + 3)
    return v, mT, mR, vt


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_rk2_solves_match_oracle(shape, oracle):
    """State, adjoint (λ at t = 0 and b), incremental state and adjoint, the
    Hessian matvec, objective and gradient, all with the RK2 integrator."""
    print(f"\n[gpu case] rk2 solves {shape}", flush=True)
    v, mT, mR, vt = _data(shape, seed=100)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde_solver="rk2")
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        c.set_reference_image(mR)
        m = c.solve_state()
        lam, b = c.solve_adjoint()
        mt = c.solve_inc_state(vt)
        lt, bt = c.solve_inc_adjoint(mt)
        H = c.hessian_matvec(vt)
        J = c.evaluate_objective()
        g = c.evaluate_gradient()
    finally:
        c.close()
    with oracle.pde_solver("rk2"):
        m32 = oracle.solve_state(v, mT, NT, prec="f32")
        m64 = oracle.solve_state(v, mT, NT)
        lam32, b32 = oracle.solve_adjoint(v, m32, mR, NT, prec="f32")
        lam64, b64 = oracle.solve_adjoint(v, m64, mR, NT)
        mt32 = oracle.solve_inc_state(v, m32, vt, NT, prec="f32")
        mt64 = oracle.solve_inc_state(v, m64, vt, NT)
        lt32, bt32 = oracle.solve_inc_adjoint(v, m32, mt, NT, prec="f32")
        lt64, bt64 = oracle.solve_inc_adjoint(v, m64, mt, NT)
        H32 = oracle.hessian_matvec(v, m32, vt, NT, "h1sn", BETA, prec="f32")
        H64 = oracle.hessian_matvec(v, m64, vt, NT, "h1sn", BETA)
        J64 = oracle.objective(v, mT, mR, NT, "h1sn", BETA)
        g32 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA, prec="f32")
        g64 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA)
    for name, gpu, r32, r64 in (("m", m, m32, m64), ("lam", lam, lam32, lam64), ("b", b, b32, b64),
                                ("mt", mt, mt32, mt64), ("lt", lt, lt32, lt64), ("bt", bt, bt32, bt64),
                                ("H", H, H32, H64), ("g", g, g32, g64)):
        print(f"{name}: gpu {relerr(gpu, r64):.2e} fp32 oracle {relerr(r32, r64):.2e}")
        assert_fp32_parity(gpu, r32, r64)
    assert abs(J[0] - J64[0]) <= 1e-5 * abs(J64[0])


def test_rk2_zero_velocity_and_switching(oracle):
    """v = 0 under RK2 (m^j = mT; m̃ = −nt·ht·∇mT·ṽ, src/CLAIRE.cpp:3022-3052)
    and switching the integrator back to SL recomputes the state."""
    shape = (16, 16, 16)
    print(f"\n[gpu case] rk2 zero velocity {shape}", flush=True)
    _, mT, mR, vt = _data(shape, seed=110)
    v = synth.bandlimited_velocity(shape, kmax=3, cfl=0.5, nt=NT, seed=111)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde_solver="rk2")
    try:
        c.set_velocity(np.zeros((3,) + shape, np.float32))
        c.set_template_image(mT)
        c.set_reference_image(mR)
        m = c.solve_state()
        mt = c.solve_inc_state(vt)
        c.set_velocity(v)
        m_rk2 = c.solve_state()
        c.set_pde_solver("sl")
        m_sl = c.solve_state()
        with pytest.raises(Exception):
            c.set_pde_solver("rk2a")
    finally:
        c.close()
    assert np.array_equal(m, np.broadcast_to(mT, m.shape))
    gm = oracle.grad(mT.astype(np.float64))
    assert relerr(mt, -np.sum(gm * vt, axis=0)) < 1e-5
    with oracle.pde_solver("rk2"):
        assert relerr(m_rk2, oracle.solve_state(v, mT, NT)) < 1e-5
    assert relerr(m_sl, oracle.solve_state(v, mT, NT)) < 1e-5


def test_rk2_matvec_graph_replay():
    """The RK2 Hessian matvec (spectral derivatives inside every step)
    captured as a HIP graph and replayed equals the eager matvec bit for bit,
    on every replay, with eager work between replays."""
    from claire_amd import DeviceArray

    shape = (32, 16, 64)
    print(f"\n[gpu case] rk2 matvec graphs {shape}", flush=True)
    v, mT, mR, vt = _data(shape, seed=120)
    eager = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde_solver="rk2")
    graph = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde_solver="rk2", graphs=True)
    din = DeviceArray.from_host(vt)
    dout = DeviceArray((3,) + shape)
    try:
        for c in (eager, graph):
            c.set_velocity(v)
            c.set_template_image(mT)
            c.set_reference_image(mR)
        ref = eager.hessian_matvec(vt)
        for k in range(5):
            graph.hessian_matvec_device(dout, din)
            graph.sync()
            assert np.array_equal(dout.download(), ref), k
            eager.solve_state()
        assert graph.matvec_graph_count() == 1
    finally:
        din.free()
        dout.free()
        eager.close()
        graph.close()


def test_rk2_gauss_newton_matches_oracle(oracle):
    """GN–Krylov with the RK2 integrator on problem 4 at 32³ against the FP64
    oracle run with the same integrator, its FP32 twin as the yardstick:
    stop reason, final mismatch, ‖g‖/‖g₀‖ and J at 3 digits."""
    from test_gpu_registration import _close, _mismatch

    shape = (32, 32, 32)
    print(f"\n[gpu case] rk2 GN registration {shape}", flush=True)
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    with oracle.pde_solver("rk2"):
        mR = synth.rescale(oracle.solve_state(vstar, mT, NT)[NT])
        v64, rep64, _ = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, NT, "h1sn", BETA, maxit=10, gttol=5e-2)
        v32, rep32, _ = oracle.gn_solve(np.zeros((3,) + shape, np.float32), mT, mR, NT, "h1sn", BETA, maxit=10,
                                        gttol=5e-2, prec="f32")
        m1_64 = oracle.solve_state(v64, mT, NT)[NT]
        m1_32 = oracle.solve_state(v32.astype(np.float64), mT, NT)[NT]
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde_solver="rk2")
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, _ = c.solve_gn(maxit=10, gttol=5e-2)
        c.set_velocity(v)
        m1 = c.solve_state()[NT]
    finally:
        c.close()
    mis, mis64, mis32 = _mismatch(m1, mT, mR), _mismatch(m1_64, mT, mR), _mismatch(m1_32, mT, mR)
    red, red64, red32 = (r["gnorm"] / r["gnorm0"] for r in (rep, rep64, rep32))
    for tag, r, mm, gg in (("gpu", rep, mis, red), ("fp32 oracle", rep32, mis32, red32),
                           ("fp64 oracle", rep64, mis64, red64)):
        print(f"{tag}: its {r['iterations']} pcg {r['pcg_iterations']} mismatch {mm:.6e} |g|/|g0| {gg:.6e} "
              f"J {r['J']:.6e} ({r['reason']})")
    assert rep["reason"] == rep64["reason"]
    assert _close(mis, mis64, mis32)
    assert _close(red, red64, red32, floor=5e-5)
    assert _close(rep["J"], rep64["J"], rep32["J"])


@pytest.mark.parametrize("nt", [1, 3])
def test_rk2_time_step_counts(nt, oracle):
    """RK2 state, gradient and Hessian matvec with nt other than 4 (the first
    incremental step without ∇m̃, the trapezoid weights)."""
    shape = (16, 32, 16)
    print(f"\n[gpu case] rk2 nt={nt} {shape}", flush=True)
    v = synth.bandlimited_velocity(shape, kmax=3, cfl=0.5, nt=nt, seed=130)
    mT = synth.gaussian_blobs(shape, count=6, seed=131)
    mR = synth.gaussian_blobs(shape, count=6, seed=132)
    vt = synth.bandlimited_velocity(shape, kmax=3, cfl=0.5, nt=nt, seed=133)
    c = Claire(shape, nt=nt, regnorm="h1sn", beta=BETA, pde_solver="rk2")
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        c.set_reference_image(mR)
        m = c.solve_state()
        g = c.evaluate_gradient()
        H = c.hessian_matvec(vt)
    finally:
        c.close()
    with oracle.pde_solver("rk2"):
        m32, m64 = oracle.solve_state(v, mT, nt, prec="f32"), oracle.solve_state(v, mT, nt)
        assert_fp32_parity(m, m32, m64)
        assert_fp32_parity(g, oracle.gradient(v, mT, mR, nt, "h1sn", BETA, prec="f32"),
                           oracle.gradient(v, mT, mR, nt, "h1sn", BETA))
        assert_fp32_parity(H, oracle.hessian_matvec(v, m32, vt, nt, "h1sn", BETA, prec="f32"),
                           oracle.hessian_matvec(v, m64, vt, nt, "h1sn", BETA))
