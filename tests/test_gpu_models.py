"""GPU parity of the registration models (clr_set_reg_model,
clr_apply_projection) against the oracle run with the same model:
CLAIREStokes (incompressible: pure-transport adjoints, Leray-projected body
forces) and CLAIREDivReg (relaxed: β-weighted projection, divergence penalty
in the objective).  Bars as in test_gpu_parity: 1e-5 relative L2 against the
FP64 oracle, or no worse than twice the reference FP32 path where FP32
conditioning rules that out."""
import numpy as np
import pytest

from claire_amd import Claire, synth
from conftest import relerr

pytestmark = pytest.mark.gpu

NT = 4
BETA = (1e-2, 1e-4, 5e-2)
SHAPES = [(16, 16, 16), (32, 16, 64), (48, 96, 150)]
MODELS = ["stokes", "relaxed_stokes"]


def assert_fp32_parity(gpu, ref32, ref64, tol=1e-5):
    e_gpu, e_ref = relerr(gpu, ref64), relerr(ref32, ref64)
    assert e_gpu <= max(tol, 2.0 * e_ref), (e_gpu, e_ref)


def _data(shape, seed):
    v = synth.bandlimited_velocity(shape, kmax=min(6, min(shape) // 4), cfl=1.5, nt=NT, seed=seed)
    mT = synth.gaussian_blobs(shape, count=8, seed=seed + 1)
    mR = synth.gaussian_blobs(shape, count=8, seed=seed + 2)
    vt = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, nt=NT, seed=seed + 3)
    return v, mT, mR, vt


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_projection(shape, model, oracle):
    print(f"\n[gpu case] projection {model} {shape}", flush=True)
    b = synth.bandlimited_velocity(shape, kmax=min(8, min(shape) // 2 - 1), cfl=1.0, nt=NT, seed=90)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, model=model)
    try:
        got = c.apply_projection(b)
    finally:
        c.close()
    assert_fp32_parity(got, oracle.project(b, model, BETA, prec="f32"), oracle.project(b, model, BETA))


@pytest.mark.parametrize("model", MODELS)
@pytest.mark.parametrize("shape", SHAPES[:2], ids=lambda s: "x".join(map(str, s)))
def test_model_solves_hessian_objective_gradient(shape, model, oracle):
    """Adjoint (λ and the projected body force), Hessian matvec, objective and
    gradient of the model against the oracle with the same model."""
    print(f"\n[gpu case] model {model} {shape}", flush=True)
    v, mT, mR, vt = _data(shape, seed=100)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, model=model)
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        c.set_reference_image(mR)
        lam, b = c.solve_adjoint()
        H = c.hessian_matvec(vt)
        J, D, R = c.evaluate_objective(v)
        g = c.evaluate_gradient(v)
    finally:
        c.close()
    with oracle.reg_model(model, BETA):
        m32 = oracle.solve_state(v, mT, NT, prec="f32")
        m64 = oracle.solve_state(v, mT, NT)
        l32, b32 = oracle.solve_adjoint(v, m32, mR, NT, prec="f32")
        l64, b64 = oracle.solve_adjoint(v, m64, mR, NT)
        H32 = oracle.hessian_matvec(v, m32, vt, NT, "h1sn", BETA, prec="f32")
        H64 = oracle.hessian_matvec(v, m64, vt, NT, "h1sn", BETA)
        J32, D32, R32 = oracle.objective(v, mT, mR, NT, "h1sn", BETA, prec="f32")
        g32 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA, prec="f32")
        g64 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA)
    assert relerr(lam, l32) < 1e-5
    assert_fp32_parity(b, b32, b64)
    assert_fp32_parity(H, H32, H64)
    assert abs(D - D32) < 1e-5 * abs(D32) and abs(R - R32) < 1e-5 * abs(R32)
    assert_fp32_parity(g, g32, g64)


def test_stokes_adjoint_differs_from_compressible():
    """The models are really different paths: on a compressible velocity the
    incompressible adjoint (no divergence term) differs from CLAIRE's."""
    shape = (32, 16, 64)
    v, mT, mR, vt = _data(shape, seed=110)
    out = {}
    for model in ("compressible", "stokes"):
        c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, model=model)
        try:
            c.set_velocity(v)
            c.set_template_image(mT)
            c.set_reference_image(mR)
            out[model] = c.solve_adjoint()[0]
        finally:
            c.close()
    assert relerr(out["stokes"], out["compressible"]) > 1e-4


@pytest.mark.parametrize("model", MODELS)
def test_model_registration(model, oracle):
    """Gauss–Newton–Krylov with the model at 32³ (synthetic problem 4) against
    the FP64 oracle, with the FP32 reference path as yardstick."""
    shape = (32, 32, 32)
    print(f"\n[gpu case] GN registration {model} {shape}", flush=True)
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    mR = synth.rescale(oracle.solve_state(vstar, mT, NT)[NT])
    with oracle.reg_model(model, BETA):
        v64, rep64, _ = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, NT, "h1sn", BETA, maxit=10)
        v32, rep32, _ = oracle.gn_solve(np.zeros((3,) + shape, np.float32), mT, mR, NT, "h1sn", BETA, maxit=10,
                                        prec="f32")
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, model=model)
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, _ = c.solve_gn(maxit=10)
    finally:
        c.close()
    for tag, r in (("gpu", rep), ("fp32 oracle", rep32), ("fp64 oracle", rep64)):
        print(f"{tag}: its {r['iterations']} pcg {r['pcg_iterations']} J {r['J']:.6e} "
              f"|g|/|g0| {r['gnorm'] / r['gnorm0']:.6e} ({r['reason']})")
    assert rep["reason"] == rep64["reason"]
    assert abs(rep["J"] - rep64["J"]) <= max(1e-3 * rep64["J"], 2.0 * abs(rep32["J"] - rep64["J"]))
    red, red64, red32 = (r["gnorm"] / r["gnorm0"] for r in (rep, rep64, rep32))
    assert abs(red - red64) <= max(1e-3 * red64, 5e-5, 2.0 * abs(red32 - red64))


@pytest.mark.parametrize("kind", ["precond", "precond_sym"])
@pytest.mark.parametrize("scale", [True, False])
def test_preconditioned_matvec_types(kind, scale, oracle):
    """PrecondHessMatVec / PrecondHessMatVecSym (clr_set_matvec_type) against
    the oracle with the same operator type."""
    shape = (32, 16, 64)
    print(f"\n[gpu case] matvec {kind} scale={scale}", flush=True)
    v, mT, mR, vt = _data(shape, seed=120)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_matvec_type(kind)
        c.set_velocity(v)
        c.set_template_image(mT)
        H = c.hessian_matvec(vt, scale=scale)
    finally:
        c.close()
    with oracle.matvec_type(kind):
        H32 = oracle.hessian_matvec(v, oracle.solve_state(v, mT, NT, prec="f32"), vt, NT, "h1sn", BETA, scale=scale,
                                    prec="f32")
        H64 = oracle.hessian_matvec(v, oracle.solve_state(v, mT, NT), vt, NT, "h1sn", BETA, scale=scale)
    assert_fp32_parity(H, H32, H64)


def test_registration_with_the_symmetric_preconditioned_operator(oracle):
    shape = (32, 32, 32)
    print(f"\n[gpu case] GN registration precond_sym {shape}", flush=True)
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    mR = synth.rescale(oracle.solve_state(vstar, mT, NT)[NT])
    beta = (1e-2, 1e-4, 0.0)
    with oracle.matvec_type("precond_sym"):
        v64, rep64, _ = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, NT, "h1sn", beta, maxit=10)
        v32, rep32, _ = oracle.gn_solve(np.zeros((3,) + shape, np.float32), mT, mR, NT, "h1sn", beta, maxit=10,
                                        prec="f32")
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=beta)
    try:
        c.set_matvec_type("precond_sym")
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, _ = c.solve_gn(maxit=10)
        c.set_matvec_type("precond")
        from claire_amd import ClaireError

        with pytest.raises(ClaireError, match="not symmetric"):
            c.solve_gn(maxit=2)
    finally:
        c.close()
    for tag, r in (("gpu", rep), ("fp32 oracle", rep32), ("fp64 oracle", rep64)):
        print(f"{tag}: its {r['iterations']} pcg {r['pcg_iterations']} J {r['J']:.6e} "
              f"|g|/|g0| {r['gnorm'] / r['gnorm0']:.6e} ({r['reason']})")
    assert rep["reason"] == rep64["reason"]
    assert abs(rep["J"] - rep64["J"]) <= max(1e-3 * rep64["J"], 2.0 * abs(rep32["J"] - rep64["J"]))


@pytest.mark.parametrize("model", MODELS)
def test_models_with_rk2_integrator(model, oracle):
    """-pdesolver rk2 with the other models: CLAIREStokes overrides only the
    SL adjoint solves, so its RK2 adjoints keep the ∇·v terms; the projection
    and (relaxed model) the divergence penalty still apply."""
    shape = (32, 16, 32)
    print(f"\n[gpu case] model {model} rk2 {shape}", flush=True)
    v = synth.bandlimited_velocity(shape, kmax=4, cfl=0.5, nt=NT, seed=140)
    mT = synth.gaussian_blobs(shape, count=8, seed=141)
    mR = synth.gaussian_blobs(shape, count=8, seed=142)
    vt = synth.bandlimited_velocity(shape, kmax=4, cfl=0.5, nt=NT, seed=143)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, model=model, pde_solver="rk2")
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        c.set_reference_image(mR)
        lam, b = c.solve_adjoint()
        H = c.hessian_matvec(vt)
        J, D, R = c.evaluate_objective(v)
        g = c.evaluate_gradient(v)
    finally:
        c.close()
    with oracle.reg_model(model, BETA), oracle.pde_solver("rk2"):
        m32, m64 = oracle.solve_state(v, mT, NT, prec="f32"), oracle.solve_state(v, mT, NT)
        b32, b64 = oracle.solve_adjoint(v, m32, mR, NT, prec="f32")[1], oracle.solve_adjoint(v, m64, mR, NT)[1]
        H32 = oracle.hessian_matvec(v, m32, vt, NT, "h1sn", BETA, prec="f32")
        H64 = oracle.hessian_matvec(v, m64, vt, NT, "h1sn", BETA)
        J64, D64, R64 = oracle.objective(v, mT, mR, NT, "h1sn", BETA)
        g32 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA, prec="f32")
        g64 = oracle.gradient(v, mT, mR, NT, "h1sn", BETA)
    assert_fp32_parity(b, b32, b64)
    assert_fp32_parity(H, H32, H64)
    assert abs(J - J64) < 1e-5 * abs(J64) and abs(R - R64) < 1e-5 * abs(R64)
    assert_fp32_parity(g, g32, g64)


def test_continuity_with_rk4_and_rk2(oracle):
    """The continuity-equation state is an SL solve whatever the integrator
    (src/CLAIRE.cpp:1923-1926): with -rkorder 4 it follows the RK4
    characteristic, with -pdesolver rk2 it still uses the SL trajectory."""
    shape = (32, 32, 16)
    print(f"\n[gpu case] continuity rk4 / pdesolver rk2 {shape}", flush=True)
    v = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, nt=NT, seed=150)
    mT = synth.gaussian_blobs(shape, count=8, seed=151)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA, pde="continuity", rk_order=4)
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        m4 = c.solve_state()
        c.set_rk_order(2)
        c.set_pde_solver("rk2")
        m2 = c.solve_state()
    finally:
        c.close()
    with oracle.pde_type("continuity"):
        with oracle.rk_order(4):
            assert_fp32_parity(m4, oracle.solve_state(v, mT, NT, prec="f32"), oracle.solve_state(v, mT, NT))
        with oracle.pde_solver("rk2"):
            assert_fp32_parity(m2, oracle.solve_state(v, mT, NT, prec="f32"), oracle.solve_state(v, mT, NT))
