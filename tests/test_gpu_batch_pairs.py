"""Pairs of the 48-registration batch (H1, β = (1e-2, 1e-4), nt = 4; pair 0
at 64³, pair 32 at 96³, pair 40 at 128³) against the stored end states of
the oracle (tests/golden/make_batch_golden.py): the FP64 solve and three
FP32 solves of the reference's single-precision path, on the pair and on
two copies whose reference image differs by one FP32 ulp at 50 voxels.

These solves are ill-conditioned: every Krylov solve stops at its iteration
limit (49) on an operator whose preconditioned spectrum spans about four
decades, so the end state moves by about 1% for a change of the inputs or
of the round-off at the 1e-7 level -- the FP64 solve itself moves that much
when the FP32 reference image is computed by a different (FP32) state solve.
The FP32 solves therefore differ from FP64 by up to ~1.5%, and three-digit
agreement is out of reach of any FP32 implementation.  The GPU is held to
the same stop reason and Newton iteration count as FP64 (within one), and
to a final mismatch and gradient reduction no further from FP64 than twice
the furthest of the reference FP32 path's three solves (the GPU solve is a
fourth FP32 sample of the same spread, and a fourth sample exceeds the
largest of three with probability 1/4)."""
import importlib.util
import json
import os

import numpy as np
import pytest

from claire_amd import Claire

pytestmark = pytest.mark.gpu

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden():
    path = os.path.join(HERE, "batch_pairs.json")
    if not os.path.exists(path):
        return None, {}
    spec = importlib.util.spec_from_file_location("make_batch_golden", os.path.join(HERE, "make_batch_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    with open(path) as f:
        return mod, json.load(f)["pairs"]


_MOD, _PAIRS = _golden()


@pytest.mark.parametrize("pair", sorted(_PAIRS, key=int))
def test_batch_pair_against_fp64_and_fp32_spread(pair, oracle):
    g = _PAIRS[pair]
    i = int(pair)
    print(f"\n[gpu case] batch pair {i} ({g['n']}^3)", flush=True)
    mT, mR = _MOD.inputs(i)
    c = Claire(mT.shape, nt=_MOD.NT, regnorm="h1", beta=_MOD.BETA, graphs=True)
    try:
        c.set_template_image(mT)
        c.set_reference_image(mR)
        v, rep, _ = c.solve_gn()
    finally:
        c.close()
    m1 = oracle.solve_state(v, mT, _MOD.NT)[_MOD.NT]
    mis = float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))
    red = rep["gnorm"] / rep["gnorm0"]
    ref = g["oracle_fp64"]
    spread_mis = max(abs(s["mismatch"] - ref["mismatch"]) for s in g["fp32_spread"])
    spread_red = max(abs(s["grad_reduction"] - ref["grad_reduction"]) for s in g["fp32_spread"])
    print(f"gpu: its {rep['iterations']} pcg {rep['pcg_iterations']} mismatch {mis:.6e} |g|/|g0| {red:.6e} "
          f"({rep['reason']})")
    print(f"fp64: its {ref['iterations']} pcg {ref['pcg_iterations']} mismatch {ref['mismatch']:.6e} "
          f"|g|/|g0| {ref['grad_reduction']:.6e}; FP32 path within {spread_mis:.2e} / {spread_red:.2e}")
    assert rep["reason"] == ref["reason"]
    assert abs(rep["iterations"] - ref["iterations"]) <= 1
    assert abs(mis - ref["mismatch"]) <= max(1e-3 * ref["mismatch"], 2.0 * spread_mis)
    assert abs(red - ref["grad_reduction"]) <= max(1e-3 * ref["grad_reduction"], 2.0 * spread_red)
