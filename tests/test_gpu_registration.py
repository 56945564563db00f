"""GPU Gauss–Newton–Krylov registration (clr_solve_gn) against the FP64
oracle's restatement of the same driver, on the reference's synthetic
problem 4 (template (sin²x1+sin²x2+sin²x3)/3, reference image = template
transported by v = (sin x3·cos x2·sin x2, …), H1SN β = 1e-2, nt = 4; the
benchmark configuration "plain CLAIRE, 64³, GNK with PCG").

The parity bar is the one the benchmark states for full registrations: final
mismatch and gradient-norm reduction equal at 3 digits."""
import numpy as np
import pytest

from claire_amd import Claire, synth

pytestmark = pytest.mark.gpu

NT = 4
BETA = (1e-2, 1e-4, 0.0)


def _problem(shape, oracle):
    mT, vstar = synth.smooth_template_problem(shape, np.float64)
    mR = synth.rescale(oracle.solve_state(vstar, mT, NT)[NT])
    return mT, mR


def _mismatch(m1, mT, mR):
    return float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))


def _close(a, a64, a32, rel=1e-3, floor=0.0):
    """3 digits against FP64, or no further from FP64 than twice the
    reference's FP32 path is (where FP32 round-off changes the Krylov
    iteration counts of the reference itself)."""
    return abs(a - a64) <= max(rel * abs(a64), floor, 2.0 * abs(a32 - a64))


@pytest.mark.parametrize("n,maxit,gttol,regnorm", [(32, 5, 1e-12, "h1sn"), (64, 50, 5e-2, "h1sn"), (32, 50, 5e-2, "h1"),
                                                     (32, 50, 5e-2, "h2sn")])
def test_gn_registration_matches_oracle(n, maxit, gttol, regnorm, oracle):
    """GPU solve against the FP64 oracle, with the oracle's FP32 twin (the
    reference's single-precision path) as the yardstick where FP32 round-off
    alone moves the result: with H1 and β1 = 1e-4 the inverse-regularization
    preconditioner amplifies the mean mode by 1/(β0·β1) = 1e6 and the FP32
    Krylov solves of the reference path stop later than the FP64 ones.

    The deep case (gttol 1e-12) stops at maxit = 5, before the noise floor:
    run further, the FP64 solve itself ends on either the step-length test or
    a failed line search, depending on the host's libm (1e-16 differences in
    J decide it)."""
    shape = (n, n, n)
    print(f"\n[gpu case] GN registration {shape} {regnorm} maxit={maxit} gttol={gttol}", flush=True)
    mT, mR = _problem(shape, oracle)
    v64, rep64, hist64 = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, NT, regnorm, BETA, maxit=maxit, gttol=gttol)
    v32, rep32, hist32 = oracle.gn_solve(np.zeros((3,) + shape, np.float32), mT, mR, NT, regnorm, BETA, maxit=maxit,
                                         gttol=gttol, prec="f32")
    c = Claire(shape, nt=NT, regnorm=regnorm, beta=BETA)
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, hist = c.solve_gn(maxit=maxit, gttol=gttol)
        c.set_velocity(v)
        m1 = c.solve_state()[NT]
    finally:
        c.close()
    m1_64 = oracle.solve_state(v64, mT, NT)[NT]
    m1_32 = oracle.solve_state(v32.astype(np.float64), mT, NT)[NT]
    mis, mis64, mis32 = _mismatch(m1, mT, mR), _mismatch(m1_64, mT, mR), _mismatch(m1_32, mT, mR)
    red, red64 = rep["gnorm"] / rep["gnorm0"], rep64["gnorm"] / rep64["gnorm0"]
    red32 = rep32["gnorm"] / rep32["gnorm0"]
    for tag, r, m, g in (("gpu", rep, mis, red), ("fp32 oracle", rep32, mis32, red32), ("fp64 oracle", rep64, mis64, red64)):
        print(f"{tag}: its {r['iterations']} pcg {r['pcg_iterations']} mismatch {m:.6e} |g|/|g0| {g:.6e} "
              f"J {r['J']:.6e} ({r['reason']})")
    assert rep["reason"] == rep64["reason"]
    assert _close(mis, mis64, mis32)
    assert _close(red, red64, red32, floor=1e-3 * gttol)
    assert _close(rep["J"], rep64["J"], rep32["J"])


def _golden():
    import importlib.util
    import json
    import os

    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    spec = importlib.util.spec_from_file_location("make_registration_golden",
                                                  os.path.join(here, "make_registration_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(os.path.join(here, "registration_phantom.json")) as f:
        return mod, json.load(f)


_GOLDEN_MOD, _GOLDEN = _golden()


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("name", sorted(_GOLDEN))
def test_phantom_registration_matches_fp64_golden(name, precision):
    """Ellipsoid-phantom registrations against the FP64 oracle's stored end
    results (tests/golden/make_registration_golden.py): 128³ with H2SN
    β = 1e-3, and 64³ and 128³ (the full-registration configuration) with
    RegularizationH2 as written, whose line search fails after one Newton
    step in the oracle as well.  Plain FP32 and the
    mixed form: stop reason, final mismatch, ‖g‖/‖g₀‖ and J at 3 digits.

    In the H2 case the second line search finds no descent in FP64, and in
    FP32 whether a trial step α < 2⁻¹⁰ passes the Armijo test is decided by
    round-off in J (the required decrease 1e-4·α·|gᵀd| is below FP32 noise in
    the state): an FP32 solve then either fails the line search or accepts
    such a step and stops on the step-length test one iteration later (or
    one earlier).  Both are the same stall: either reason within one
    iteration passes, and the end state is held to the same 3 digits."""
    case, ref = _GOLDEN[name]["case"], _GOLDEN[name]["oracle_fp64"]
    print(f"\n[gpu case] phantom registration {name} {precision}", flush=True)
    mT, mR = _GOLDEN_MOD.inputs(case)
    shape = mT.shape
    c = Claire(shape, nt=case["nt"], regnorm=case["regnorm"], beta=(case["beta"], 0.0, 0.0), precision=precision)
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, hist = c.solve_gn(maxit=case["maxit"], gttol=case["gttol"])
        c.set_velocity(v)
        m1 = c.solve_state()[case["nt"]]
    finally:
        c.close()
    mis = _mismatch(m1, mT, mR)
    red, red64 = rep["gnorm"] / rep["gnorm0"], ref["gnorm"] / ref["gnorm0"]
    print(f"gpu {precision}: its {rep['iterations']} pcg {rep['pcg_iterations']} mismatch {mis:.6e} "
          f"|g|/|g0| {red:.6e} J {rep['J']:.6e} ({rep['reason']})")
    print(f"fp64 oracle: its {ref['iterations']} pcg {ref['pcg_iterations']} mismatch {ref['mismatch']:.6e} "
          f"|g|/|g0| {red64:.6e} J {ref['J']:.6e} ({ref['reason']})")
    stall = ("linesearch", "steptol")
    if ref["reason"] in stall:
        assert rep["reason"] in stall and abs(rep["iterations"] - ref["iterations"]) <= 1
    else:
        assert rep["reason"] == ref["reason"]
    assert abs(mis - ref["mismatch"]) <= 1e-3 * ref["mismatch"]
    assert abs(red - red64) <= 1e-3 * red64 or abs(red - red64) <= 1e-3 * case["gttol"]
    assert abs(rep["J"] - ref["J"]) <= 1e-3 * ref["J"]


def test_gn_monitor_prints_the_reference_table(capfd, oracle):
    """monitor = 1 prints the reference's iteration table (header of
    src/CLAIREInterface.cpp:722-726, lines of src/TaoInterface.cpp:584-588):
    one line per iteration with J/J0, D/D0, ‖g‖/‖g0‖, ‖g‖ and the step,
    agreeing with the returned history; the default stays silent."""
    shape = (16, 16, 16)
    print(f"\n[gpu case] GN monitor {shape}", flush=True)
    mT, mR = _problem(shape, oracle)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        capfd.readouterr()
        c.solve_gn(maxit=5)
        assert capfd.readouterr().out == ""
        v, rep, hist = c.solve_gn(maxit=5, monitor=1)
        out = capfd.readouterr().out.splitlines()
    finally:
        c.close()
    assert out[0] == "-" * 101 and out[2] == "-" * 101
    assert out[1].split() == ["iter", "objective", "(rel)", "mismatch", "(rel)", "||gradient||_2,rel",
                              "||gradient||_2", "step"]
    rows = [line.split() for line in out[3:]]
    assert len(rows) == rep["iterations"] + 1
    for k, r in enumerate(rows):
        assert len(out[3 + k]) >= 80 and r[0] == f"{k:03d}"
        it, J, D, R, g, grel, its, step = hist[k]
        expect = (J / rep["J0"], D / rep["D0"], grel, g, 1.0 if k == 0 else hist[k - 1][7])
        for got, want in zip(map(float, r[1:]), expect):
            assert abs(got - want) <= 1e-10 * abs(want) + 5e-7, (k, got, want)


def _exit_problem(shape):
    mT = synth.gaussian_blobs(shape, count=6, seed=40)
    v = synth.bandlimited_velocity(shape, kmax=4, cfl=1.0, nt=NT, seed=41)
    return mT, v


def test_gn_exit_on_zero_initial_gradient():
    """mR = mT: the gradient at v = 0 vanishes, so ‖g₀‖ = 0 and the gradient
    test stops at iteration 0 with GATOL (‖g‖ < gatol overrides the GTTOL
    test that also holds; CheckConvergenceGrad, src/TaoInterface.cpp:452-480)."""
    shape = (16, 16, 16)
    print(f"\n[gpu case] GN exit gatol {shape}", flush=True)
    mT, _ = _exit_problem(shape)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_template_image(mT)
        c.set_reference_image(mT)
        v, rep, hist = c.solve_gn(maxit=10)
    finally:
        c.close()
    assert rep["reason"] == "gatol"
    assert rep["iterations"] == 0 and rep["matvecs"] == 0
    assert rep["gnorm0"] == 0.0 and rep["gnorm"] == 0.0
    assert not np.any(v)


def test_gn_exit_on_nan_objective():
    """A NaN voxel in the template makes J NaN: TAO_DIVERGED_NAN before any
    Newton step (src/TaoInterface.cpp:424-436)."""
    shape = (16, 16, 16)
    print(f"\n[gpu case] GN exit nan {shape}", flush=True)
    mT, v = _exit_problem(shape)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        mR = c.solve_state()[NT]
        bad = mT.copy()
        bad[3, 5, 7] = np.nan
        c.set_template_image(bad)
        c.set_reference_image(mR)
        _, rep, _ = c.solve_gn(maxit=10)
    finally:
        c.close()
    assert rep["reason"] == "nan"
    assert rep["iterations"] == 0 and rep["matvecs"] == 0
    assert not np.isfinite(rep["J0"])


def test_gn_exit_on_small_step():
    """A first trial step of 2⁻¹¹ (PETSc's -tao_ls_stepinit) is accepted by
    the Armijo test along the Newton direction, so the next gradient test
    stops on the step length (step < 2⁻¹⁰: TAO_CONVERGED_STEPTOL,
    src/TaoInterface.cpp:445-453) after one iteration."""
    shape = (16, 16, 16)
    print(f"\n[gpu case] GN exit steptol {shape}", flush=True)
    mT, v = _exit_problem(shape)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        mR = c.solve_state()[NT]
        c.set_reference_image(mR)
        _, rep, hist = c.solve_gn(maxit=10, ls_stepinit=2.0 ** -11)
    finally:
        c.close()
    assert rep["reason"] == "steptol"
    assert rep["iterations"] == 1
    assert hist[0][7] == 2.0 ** -11
    assert rep["J"] < rep["J0"]


@pytest.mark.parametrize("n", [1, 3, 4, 1023, 4097, 3 * 32 ** 3 + 2])
def test_dot_and_norm2_against_fp64_sums(n):
    """clr_dot / clr_norm2 (PETSc VecTDot / VecNorm in the reference) against
    FP64 sums of the same FP32 values, over lengths that leave 1–3 tail
    elements after the 16-byte loads, and through a pointer one float past a
    16-byte boundary (the scalar-load path)."""
    import ctypes

    import claire_amd

    rng = np.random.default_rng(n)
    a = rng.standard_normal(n + 1).astype(np.float32)
    b = rng.standard_normal(n + 1).astype(np.float32)
    c = Claire((8, 8, 8), nt=1, regnorm="h1sn", beta=BETA)
    try:
        d, dn = c.dot(a[:n], b[:n]), c.norm2(a[:n])
        exact_d = float(np.dot(a[:n].astype(np.float64), b[:n].astype(np.float64)))
        exact_n = float(np.linalg.norm(a[:n].astype(np.float64)))
        assert abs(d - exact_d) <= 1e-12 * float(np.dot(np.abs(a[:n]), np.abs(b[:n]))) + 1e-300
        assert abs(dn - exact_n) <= 1e-12 * exact_n
        da = claire_amd.DeviceArray.from_host(a)
        db = claire_amd.DeviceArray.from_host(b)
        r = ctypes.c_double()
        assert claire_amd.lib().clr_dot(c.ctx, ctypes.c_void_p(da.ptr + 4), ctypes.c_void_p(db.ptr + 4),
                                        ctypes.c_int64(n), ctypes.byref(r)) == 0
        exact_off = float(np.dot(a[1:].astype(np.float64), b[1:].astype(np.float64)))
        assert abs(r.value - exact_off) <= 1e-12 * float(np.dot(np.abs(a[1:]), np.abs(b[1:]))) + 1e-300
        assert claire_amd.lib().clr_dot(c.ctx, ctypes.c_void_p(da.ptr), ctypes.c_void_p(db.ptr),
                                        ctypes.c_int64(-1), ctypes.byref(r)) == 1
        bad = a[:n].copy()
        bad[n // 2] = np.nan
        assert np.isnan(c.norm2(bad))
    finally:
        c.close()


def test_queued_krylov_iterations_equal_host_checked_ones():
    """The CG scalars stay on the device and iterations are queued in runs
    (krylov_block) and, with matvec graphs on, replayed from one captured
    iteration: the velocity, the iteration counts and J equal those of the
    solve that checks on the host after every iteration, bit for bit (an
    H1 phantom pair whose Krylov solves take tens of iterations)."""
    shape = (32, 32, 32)
    print(f"\n[gpu case] queued Krylov runs {shape}", flush=True)
    mT = synth.ellipsoid_phantom(shape, count=12, seed=51, radii=(3.0, 9.0))
    v = synth.bandlimited_velocity(shape, kmax=6, cfl=2.0, nt=NT, seed=52)
    out = {}
    for name, graphs, block in (("host", False, 1), ("queued", False, 8), ("graphs", True, 8), ("graphs1", True, 1)):
        c = Claire(shape, nt=NT, regnorm="h1", beta=(1e-2, 1e-4, 0.0), graphs=graphs)
        try:
            c.set_velocity(v)
            c.set_template_image(mT)
            mR = c.solve_state()[NT]
            c.set_reference_image(mR)
            out[name] = c.solve_gn(krylov_block=block)
        finally:
            c.close()
    v0, rep0, _ = out["host"]
    print(f"host-checked: {rep0['iterations']} its, {rep0['pcg_iterations']} PCG, J {rep0['J']:.9e}")
    assert rep0["pcg_iterations"] > 3 * rep0["iterations"]
    for name in ("queued", "graphs", "graphs1"):
        v1, rep1, _ = out[name]
        assert np.array_equal(v1, v0), name
        for k in ("iterations", "pcg_iterations", "reason", "J", "gnorm", "matvecs"):
            assert rep1[k] == rep0[k], (name, k, rep1[k], rep0[k])
