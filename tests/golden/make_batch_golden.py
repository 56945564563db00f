#!/usr/bin/env python3
"""Writes batch_pairs.json: end results of the oracle's Gauss–Newton–Krylov
driver on three pairs of the 48-registration batch (pair 0 at 64³, pair 32,
the first 96³ one, and pair 40, the first 128³ one), which take minutes on
the CPU:

* "oracle_fp64": the FP64 oracle from v = 0;
* "fp32_spread": the FP32 oracle (the reference's single-precision path)
  on the same pair and on two copies of it whose reference image is moved
  by one FP32 ulp at 50 seeded voxels.  The FP64 solve does not notice such
  a perturbation (its mismatch moves in the 5th digit), but the FP32 solves
  end up to ~1.5% apart: with H1, β = (1e-2, 1e-4), the Krylov solves stop
  at their iteration limit on a badly conditioned system, and the FP32
  round-off of the matvec decides the truncated directions.  The spread of
  these FP32 end states is the yardstick for any FP32 implementation.

Inputs are regenerated from the pair index as the batch builds them
(claire_amd.synth): a 30-ellipsoid phantom (seed 1000 + i) warped by a
band-limited velocity (|w|∞ <= 8, CFL 2, nt = 4, seed 2000 + i); the
reference image is the FP64 oracle's state solve.

Usage: python tests/golden/make_batch_golden.py [--threads N] [--pairs I ...]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from claire_amd import synth  # noqa: E402
from oracle import oracle_py as oracle  # noqa: E402

NT = 4
BETA = (1e-2, 1e-4, 0.0)
SIZES = [64] * 32 + [96] * 8 + [128] * 8
PAIRS = [0, 32, 40]
PERTURBATIONS = [1, 2]


def inputs(i):
    n = SIZES[i]
    shape = (n, n, n)
    mT = synth.ellipsoid_phantom(shape, count=30, seed=1000 + i)
    v = synth.bandlimited_velocity(shape, kmax=8, cfl=2.0, nt=NT, seed=2000 + i)
    mR = oracle.solve_state(v.astype(np.float64), mT.astype(np.float64), NT)[NT].astype(np.float32)
    return mT, mR


def perturbed(mR, k):
    rng = np.random.default_rng(k)
    r = mR.copy()
    idx = rng.integers(0, r.size, 50)
    r.flat[idx] = np.nextafter(r.flat[idx], np.float32(2))
    return r


def solve(mT, mR, prec):
    v0 = np.zeros((3,) + mT.shape, np.float64 if prec == "f64" else np.float32)
    v, rep, _ = oracle.gn_solve(v0, mT, mR, NT, "h1", BETA, prec=prec)
    m1 = oracle.solve_state(v, mT, NT)[NT]
    mis = float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))
    return {"iterations": rep["iterations"], "pcg_iterations": rep["pcg_iterations"], "reason": rep["reason"],
            "mismatch": mis, "grad_reduction": rep["gnorm"] / rep["gnorm0"], "J": rep["J"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--pairs", type=int, nargs="*", default=PAIRS)
    args = ap.parse_args()
    if args.threads:
        oracle.set_num_threads(args.threads)
    path = os.path.join(HERE, "batch_pairs.json")
    out = {"nt": NT, "beta": list(BETA), "pairs": {}}
    if os.path.exists(path):
        with open(path) as f:
            out = json.load(f)
    for i in args.pairs:
        t0 = time.time()
        mT, mR = inputs(i)
        res = {"pair": i, "n": SIZES[i], "oracle_fp64": solve(mT, mR, "f64"), "fp32_spread": []}
        for k in [0] + PERTURBATIONS:
            r = solve(mT, perturbed(mR, k) if k else mR, "f32")
            r["perturbation"] = k
            res["fp32_spread"].append(r)
        out["pairs"][str(i)] = res
        print(f"pair {i} ({SIZES[i]}^3): fp64 mismatch {res['oracle_fp64']['mismatch']:.6e}, fp32 "
              f"{[round(r['mismatch'], 6) for r in res['fp32_spread']]}, {time.time() - t0:.0f}s", flush=True)
        with open(path, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
