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


@pytest.mark.parametrize("n,maxit,gttol", [(32, 10, 1e-12), (64, 50, 5e-2)])
def test_gn_registration_matches_oracle(n, maxit, gttol, oracle):
    shape = (n, n, n)
    print(f"\n[gpu case] GN registration {shape} maxit={maxit} gttol={gttol}", flush=True)
    mT, mR = _problem(shape, oracle)
    v64, rep64, hist64 = oracle.gn_solve(np.zeros((3,) + shape), mT, mR, NT, "h1sn", BETA, maxit=maxit, gttol=gttol)
    c = Claire(shape, nt=NT, regnorm="h1sn", beta=BETA)
    try:
        c.set_template_image(mT.astype(np.float32))
        c.set_reference_image(mR.astype(np.float32))
        v, rep, hist = c.solve_gn(maxit=maxit, gttol=gttol)
        c.set_velocity(v)
        m1 = c.solve_state()[NT]
    finally:
        c.close()
    m1_64 = oracle.solve_state(v64, mT, NT)[NT]
    mis, mis64 = _mismatch(m1, mT, mR), _mismatch(m1_64, mT, mR)
    red, red64 = rep["gnorm"] / rep["gnorm0"], rep64["gnorm"] / rep64["gnorm0"]
    print(f"gpu: its {rep['iterations']} pcg {rep['pcg_iterations']} mismatch {mis:.6e} |g|/|g0| {red:.6e} "
          f"({rep['reason']})")
    print(f"fp64 oracle: its {rep64['iterations']} pcg {rep64['pcg_iterations']} mismatch {mis64:.6e} "
          f"|g|/|g0| {red64:.6e} ({rep64['reason']})")
    assert rep["reason"] == rep64["reason"]
    assert abs(mis - mis64) <= 1e-3 * mis64
    assert abs(red - red64) <= 1e-3 * red64 or abs(red - red64) <= 1e-3 * 5e-2
    assert abs(rep["J"] - rep64["J"]) <= 1e-3 * rep64["J"]
